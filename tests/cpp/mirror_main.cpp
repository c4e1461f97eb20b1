// Drives the C++ mirror (louvain_hip.hpp) the way the reference's main.cxx
// drives inc/louvain.hxx: Static on x, then ND, DS and DF on y from a given
// previous state.  Input and output are raw little-endian arrays:
//   in:  S, nnz_x, x.offsets[S+1] (u64), x.keys (u32), x.weights (f32), x.exists[S] (u8),
//        nnz_y, y.offsets, y.keys, y.weights, y.exists,
//        nd, del.u, del.v, del.w, ni, ins.u, ins.v, ins.w,
//        q[S] (u32), qvtot[S] (f64), qctot[S] (f64)
//   out: for Static, ND, DS, DF: iterations, passes (i32), membership[S] (u32),
//        vertexWeight[S], communityWeight[S] (f64)
// Exit status: 0, or 3 with the lv_status printed when the library reports one.
#include <cstdio>
#include <cstdint>
#include <tuple>
#include <vector>

#include "louvain_hip.hpp"

namespace {

// The part of DiGraph's interface (inc/Graph.hxx:59-169) the mirror uses.
struct CsrGraph {
  using key_type = uint32_t;
  std::vector<uint64_t> off;
  std::vector<uint32_t> keys;
  std::vector<float> w;
  std::vector<uint8_t> ex;
  size_t span() const { return ex.size(); }
  bool hasVertex(uint32_t u) const { return u < ex.size() && ex[u]; }
  size_t degree(uint32_t u) const { return off[u + 1] - off[u]; }
  template <class F>
  void forEachEdge(uint32_t u, F f) const {
    for (uint64_t i = off[u]; i < off[u + 1]; ++i) f(keys[i], w[i]);
  }
};

template <class T>
bool get(FILE* f, T* p, size_t n) {
  return fread(p, sizeof(T), n, f) == n;
}

bool read_graph(FILE* f, uint64_t S, CsrGraph& g) {
  uint64_t nnz = 0;
  if (!get(f, &nnz, 1)) return false;
  g.off.resize(S + 1);
  g.keys.resize(nnz);
  g.w.resize(nnz);
  g.ex.resize(S);
  return get(f, g.off.data(), S + 1) && get(f, g.keys.data(), nnz) && get(f, g.w.data(), nnz) &&
         get(f, g.ex.data(), S);
}

bool read_batch(FILE* f, std::vector<std::tuple<uint32_t, uint32_t, float>>& b) {
  uint64_t n = 0;
  if (!get(f, &n, 1)) return false;
  std::vector<uint32_t> u(n), v(n);
  std::vector<float> w(n);
  if (!(get(f, u.data(), n) && get(f, v.data(), n) && get(f, w.data(), n))) return false;
  b.resize(n);
  for (uint64_t i = 0; i < n; ++i) b[i] = std::make_tuple(u[i], v[i], w[i]);
  return true;
}

template <class R>
void put(FILE* f, const R& r) {
  int32_t it = r.iterations, ps = r.passes;
  fwrite(&it, 4, 1, f);
  fwrite(&ps, 4, 1, f);
  fwrite(r.membership.data(), sizeof(uint32_t), r.membership.size(), f);
  fwrite(r.vertexWeight.data(), sizeof(double), r.vertexWeight.size(), f);
  fwrite(r.communityWeight.data(), sizeof(double), r.communityWeight.size(), f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  uint64_t S = 0;
  CsrGraph x, y;
  std::vector<std::tuple<uint32_t, uint32_t, float>> dels, ins;
  bool ok = get(f, &S, 1) && read_graph(f, S, x) && read_graph(f, S, y) && read_batch(f, dels) && read_batch(f, ins);
  std::vector<uint32_t> q(S);
  std::vector<double> qvtot(S), qctot(S);
  ok = ok && get(f, q.data(), S) && get(f, qvtot.data(), S) && get(f, qctot.data(), S);
  fclose(f);
  if (!ok) {
    fprintf(stderr, "short input\n");
    return 2;
  }
  using namespace louvain_hip;
  try {
    LouvainOptions o;
    auto a = louvainStaticHip(x, o);
    auto b = louvainNaiveDynamicHip(y, dels, ins, q, qvtot, qctot, o);
    auto c = louvainDynamicDeltaScreeningHip(y, dels, ins, q, qvtot, qctot, o);
    auto d = louvainDynamicFrontierHip(y, dels, ins, q, qvtot, qctot, o);
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 2;
    put(g, a);
    put(g, b);
    put(g, c);
    put(g, d);
    fclose(g);
  } catch (const Error& e) {
    printf("lv_status %d: %s\n", e.status(), e.what());
    return 3;
  }
  return 0;
}
