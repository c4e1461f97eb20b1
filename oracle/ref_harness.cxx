// Test infrastructure only: a C-ABI wrapper around the reference's own
// header-only OpenMP Louvain (puzzlef/louvain-communities-openmp-dynamic,
// inc/*.hxx), compiled in place from /root/reference by oracle/Makefile into
// oracle/_ref/libref.so.  It is the checker used to write the golden fixtures
// under tests/golden/ and to pin the C restatement in oracle/louvain_oracle.c.
// Nothing in the product links or loads it.
//
// Graph input is a (possibly gapped) CSR with rows sorted by neighbour id,
// which is the order DiGraph keeps its LazyBitset rows in
// (inc/_bitset.hxx:245-262).  Batches are SoA (u, v, w) arrays, turned into
// the reference's vector<tuple<K,K,V>> here.
#include <cstdint>
#include <cstring>
#include <tuple>
#include <vector>
#include <omp.h>
#include "inc/main.hxx"

using namespace std;
using K = uint32_t;
using V = float;
using W = double;

extern "C" {
struct ref_options {
  int32_t repeat;
  double resolution, tolerance, aggregation_tolerance, tolerance_drop;
  int32_t max_iterations, max_passes;
};
struct ref_out {
  int32_t iterations, passes;
  float time, marking_time, initialization_time, first_pass_time, local_move_time, aggregation_time;
  uint64_t affected_vertices;
};
}

static DiGraph<K, None, V> build_digraph(uint64_t S, const uint64_t* off, const uint32_t* deg,
                                         const uint32_t* keys, const float* w, const uint8_t* ex) {
  DiGraph<K, None, V> x;
  x.respan(S);
  for (uint64_t u = 0; u < S; ++u)
    if (!ex || ex[u]) x.addVertex(K(u));
  for (uint64_t u = 0; u < S; ++u) {
    if (ex && !ex[u]) continue;
    uint64_t d = deg ? deg[u] : off[u + 1] - off[u];
    for (uint64_t i = off[u]; i < off[u] + d; ++i) x.addEdge(K(u), keys[i], w[i]);
  }
  x.update();
  return x;
}

static DiGraphCsr<K, None, W> build_csr64(uint64_t S, const uint64_t* off, const uint32_t* deg,
                                          const uint32_t* keys, const double* w) {
  uint64_t E = off[S];
  DiGraphCsr<K, None, W> y(S, E);
  for (uint64_t u = 0; u <= S; ++u) y.offsets[u] = off[u];
  for (uint64_t u = 0; u < S; ++u) y.degrees[u] = K(deg ? deg[u] : off[u + 1] - off[u]);
  for (uint64_t i = 0; i < E; ++i) { y.edgeKeys[i] = keys[i]; y.edgeValues[i] = w[i]; }
  return y;
}

static vector<tuple<K, K, V>> batch(const uint32_t* u, const uint32_t* v, const float* w, uint64_t n) {
  vector<tuple<K, K, V>> a;
  a.reserve(n);
  for (uint64_t i = 0; i < n; ++i) a.push_back(make_tuple(u[i], v[i], w[i]));
  return a;
}

static LouvainOptions opts(const ref_options* o) {
  return LouvainOptions(o->repeat, o->resolution, o->tolerance, o->aggregation_tolerance,
                        o->tolerance_drop, o->max_iterations, o->max_passes);
}

template <class R>
static void copy_result(const R& a, uint32_t* memb, double* vtot, double* ctot, ref_out* out) {
  memcpy(memb, a.membership.data(), a.membership.size() * sizeof(K));
  memcpy(vtot, a.vertexWeight.data(), a.vertexWeight.size() * sizeof(W));
  memcpy(ctot, a.communityWeight.data(), a.communityWeight.size() * sizeof(W));
  out->iterations = a.iterations;
  out->passes = a.passes;
  out->time = a.time;
  out->marking_time = a.markingTime;
  out->initialization_time = a.initializationTime;
  out->first_pass_time = a.firstPassTime;
  out->local_move_time = a.localMoveTime;
  out->aggregation_time = a.aggregationTime;
  out->affected_vertices = a.affectedVertices;
}

extern "C" {

void ref_set_threads(int t) { omp_set_num_threads(t); }

double ref_edge_weight(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                       const float* w, const uint8_t* ex) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  return edgeWeightOmp(x);
}

double ref_modularity(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                      const float* w, const uint8_t* ex, const uint32_t* memb, double R) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  double M = edgeWeightOmp(x) / 2;
  auto fc = [&](auto u) { return memb[u]; };
  return modularityBy(x, fc, M, R);
}

// approach: 0 static, 1 naive-dynamic, 2 delta-screening, 3 dynamic frontier.
int ref_louvain(int approach, uint64_t S, const uint64_t* off, const uint32_t* deg,
                const uint32_t* keys, const float* w, const uint8_t* ex,
                const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                const uint32_t* q, const double* qvtot, const double* qctot,
                const ref_options* o, uint32_t* memb, double* vtot, double* ctot, ref_out* out) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  auto lo = opts(o);
  if (approach == 0) {
    auto a = louvainStaticOmp(x, lo);
    copy_result(a, memb, vtot, ctot, out);
    return 0;
  }
  auto dels = batch(du, dv, dw, nd);
  auto inss = batch(iu, iv, iw, ni);
  vector<K> qq(q, q + S);
  vector<W> qv(qvtot, qvtot + S), qc(qctot, qctot + S);
  if (approach == 1) {
    auto a = louvainNaiveDynamicOmp(x, dels, inss, qq, qv, qc, lo);
    copy_result(a, memb, vtot, ctot, out);
  } else if (approach == 2) {
    auto a = louvainDynamicDeltaScreeningOmp(x, dels, inss, qq, qv, qc, lo);
    copy_result(a, memb, vtot, ctot, out);
  } else if (approach == 3) {
    auto a = louvainDynamicFrontierOmp(x, dels, inss, qq, qv, qc, lo);
    copy_result(a, memb, vtot, ctot, out);
  } else return -1;
  return 0;
}

// Phase: incremental K / Sigma_tot update (louvainUpdateWeightsFromOmpU).
void ref_update_weights(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                        const float* w, const uint8_t* ex,
                        const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                        const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                        const uint32_t* vcom, double* vtot, double* ctot) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  auto dels = batch(du, dv, dw, nd);
  auto inss = batch(iu, iv, iw, ni);
  vector<K> c(vcom, vcom + S);
  vector<W> vt(vtot, vtot + S), ct(ctot, ctot + S);
  louvainUpdateWeightsFromOmpU(vt, ct, x, dels, inss, c);
  memcpy(vtot, vt.data(), S * sizeof(W));
  memcpy(ctot, ct.data(), S * sizeof(W));
}

// Phase: DF frontier marking (louvainAffectedVerticesFrontierOmpW).
void ref_frontier(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                  const float* w, const uint8_t* ex,
                  const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                  const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                  const uint32_t* vcom, uint8_t* vaff) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  auto dels = batch(du, dv, dw, nd);
  auto inss = batch(iu, iv, iw, ni);
  vector<K> c(vcom, vcom + S);
  vector<char> a(S);
  louvainAffectedVerticesFrontierOmpW(a, x, dels, inss, c);
  for (uint64_t u = 0; u < S; ++u) vaff[u] = uint8_t(a[u]);
}

// Phase: delta-screening affected set (louvainAffectedVerticesDeltaScreeningOmpW).
void ref_delta_screening(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                         const float* w, const uint8_t* ex,
                         const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                         const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                         const uint32_t* vcom, const double* vtot, const double* ctot, double M, double R,
                         uint8_t* vertices, uint8_t* neighbors, uint8_t* communities) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  auto dels = batch(du, dv, dw, nd);
  auto inss = batch(iu, iv, iw, ni);
  vector<K> c(vcom, vcom + S);
  vector<W> vt(vtot, vtot + S), ct(ctot, ctot + S);
  vector<char> a(S), b(S), d(S);
  int T = omp_get_max_threads();
  vector<vector<K>*> vcs(T);
  vector<vector<W>*> vcout(T);
  louvainAllocateHashtablesW(vcs, vcout, S);
  louvainAffectedVerticesDeltaScreeningOmpW(a, b, d, vcs, vcout, x, dels, inss, c, vt, ct, M, R);
  louvainFreeHashtablesW(vcs, vcout);
  for (uint64_t u = 0; u < S; ++u) { vertices[u] = a[u]; neighbors[u] = b[u]; communities[u] = d[u]; }
}

// Phase: per-vertex best community for a fixed (vcom, vtot, ctot) snapshot,
// i.e. the clear/scan/choose steps of louvainMoveOmpW without the move.
// weights64 selects a super-graph (f64 weights, DiGraphCsr) input.
void ref_choose(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                const void* w, int weights64, const uint8_t* ex, const uint32_t* vcom,
                const double* vtot, const double* ctot, double M, double R,
                uint32_t* best_c, double* best_e) {
  vector<K> c(vcom, vcom + S);
  vector<W> vt(vtot, vtot + S), ct(ctot, ctot + S);
  vector<K> vcs;
  vector<W> vcout(S);
  auto run = [&](const auto& x) {
    for (uint64_t u = 0; u < S; ++u) {
      best_c[u] = 0; best_e[u] = 0;
      if (!x.hasVertex(K(u))) continue;
      louvainClearScanW(vcs, vcout);
      louvainScanCommunitiesW(vcs, vcout, x, K(u), c);
      auto [bc, be] = louvainChooseCommunity(x, K(u), c, vt, ct, vcs, vcout, M, R);
      best_c[u] = bc; best_e[u] = be;
    }
  };
  if (weights64) run(build_csr64(S, off, deg, keys, (const double*)w));
  else run(build_digraph(S, off, deg, keys, (const float*)w, ex));
}

// Phase: community existence + renumbering (louvainCommunityExistsOmpW,
// louvainRenumberCommunitiesOmpW).  vcom is renumbered in place; returns CN.
uint64_t ref_renumber(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                      const void* w, int weights64, const uint8_t* ex, uint32_t* vcom, uint32_t* cext) {
  vector<K> c(vcom, vcom + S), a(S), bufk(omp_get_max_threads());
  uint64_t CN = 0;
  auto run = [&](const auto& x) {
    CN = louvainCommunityExistsOmpW(a, x, c);
    louvainRenumberCommunitiesOmpW(c, a, bufk, x);
  };
  if (weights64) run(build_csr64(S, off, deg, keys, (const double*)w));
  else run(build_digraph(S, off, deg, keys, (const float*)w, ex));
  memcpy(vcom, c.data(), S * sizeof(K));
  memcpy(cext, a.data(), S * sizeof(K));
  return CN;
}

// Phase: aggregation given a renumbered membership with CN communities
// (louvainCommunityVerticesOmpW + louvainAggregateOmpW).  Output is the
// gapped super-graph: zoff[CN+1], zdeg[CN], zkeys/zw of capacity off[S].
void ref_aggregate(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                   const void* w, int weights64, const uint8_t* ex, const uint32_t* vcom, uint64_t CN,
                   uint64_t* zoff, uint32_t* zdeg, uint32_t* zkeys, double* zw,
                   uint32_t* coff, uint32_t* cedg) {
  int T = omp_get_max_threads();
  uint64_t E = off[S];
  vector<K> c(vcom, vcom + S), bufk(T);
  vector<size_t> bufs(T);
  vector<vector<K>*> vcs(T);
  vector<vector<W>*> vcout(T);
  louvainAllocateHashtablesW(vcs, vcout, S);
  DiGraphCsr<K, None, None, K> cv(S, S);
  DiGraphCsr<K, None, W> z(S, E);
  auto run = [&](const auto& x) {
    cv.respan(CN); z.respan(CN);
    louvainCommunityVerticesOmpW(cv.offsets, cv.degrees, cv.edgeKeys, bufk, x, c);
    louvainAggregateOmpW(z.offsets, z.degrees, z.edgeKeys, z.edgeValues, bufs, vcs, vcout, x, c, cv.offsets, cv.edgeKeys);
  };
  if (weights64) run(build_csr64(S, off, deg, keys, (const double*)w));
  else run(build_digraph(S, off, deg, keys, (const float*)w, ex));
  louvainFreeHashtablesW(vcs, vcout);
  for (uint64_t i = 0; i <= CN; ++i) { zoff[i] = z.offsets[i]; coff[i] = cv.offsets[i]; }
  for (uint64_t i = 0; i < CN; ++i) zdeg[i] = z.degrees[i];
  for (uint64_t i = 0; i < E; ++i) { zkeys[i] = z.edgeKeys[i]; zw[i] = z.edgeValues[i]; }
  for (uint64_t i = 0; i < S; ++i) cedg[i] = cv.edgeKeys[i];
}

// Batch tidy and update: tidyBatchUpdateU (inc/batch.hxx:202-209) of the raw
// lists against graph x, then applyBatchUpdateOmpU (inc/batch.hxx:241-248)
// on a copy of x.  The tidy lists go to tdu/tdv/tdw (room for nd) and
// tiu/tiv/tiw (room for ni), their sizes to *tnd / *tni; the updated graph to
// yoff[S+1] / ykeys / yw (room for ycap edges, rows sorted by key) and
// yex[S].  Ids must be below S.  Returns y's edge count, or -1 if ycap is
// too small.
int64_t ref_tidy_update(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                        const float* w, const uint8_t* ex,
                        const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                        const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                        uint32_t* tdu, uint32_t* tdv, float* tdw, uint64_t* tnd,
                        uint32_t* tiu, uint32_t* tiv, float* tiw, uint64_t* tni,
                        uint64_t* yoff, uint32_t* ykeys, float* yw, uint8_t* yex, uint64_t ycap) {
  auto x = build_digraph(S, off, deg, keys, w, ex);
  auto dels = batch(du, dv, dw, nd);
  auto inss = batch(iu, iv, iw, ni);
  tidyBatchUpdateU(dels, inss, x);
  *tnd = dels.size();
  *tni = inss.size();
  for (size_t i = 0; i < dels.size(); ++i) { tdu[i] = get<0>(dels[i]); tdv[i] = get<1>(dels[i]); tdw[i] = get<2>(dels[i]); }
  for (size_t i = 0; i < inss.size(); ++i) { tiu[i] = get<0>(inss[i]); tiv[i] = get<1>(inss[i]); tiw[i] = get<2>(inss[i]); }
  auto y = duplicate(x);
  applyBatchUpdateOmpU(y, dels, inss);
  uint64_t e = 0;
  yoff[0] = 0;
  for (uint64_t u = 0; u < S; ++u) {
    yex[u] = y.hasVertex(K(u)) ? 1 : 0;
    bool room = true;
    if (u < y.span())
      y.forEachEdge(K(u), [&](K v, V wt) {
        if (e >= ycap) { room = false; return; }
        ykeys[e] = v; yw[e] = wt; ++e;
      });
    if (!room) return -1;
    yoff[u + 1] = e;
  }
  return int64_t(e);
}

}  // extern "C"
