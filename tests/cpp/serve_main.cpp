// Drives the serving loop through the C ABI alone (lv_server_* in
// louvain_hip.h): resident graphs with given starting states, every request
// submitted at once, polled first without waiting, then to completion.
// Input and output are raw little-endian arrays:
//   in:  G; per graph: S, nnz, offsets[S+1] (u64), keys (u32), weights (f32), exists[S] (u8),
//        q[S] (u32), qvtot[S] (f64), qctot[S] (f64);
//        R; per request: graph (u32), nd, del.u, del.v, del.w, ni, ins.u, ins.v, ins.w
//   out: per request, in submission order: status, iterations, passes (i32), modularity,
//        started_ms, done_ms (f64), tidy deletions, insertions (u64), membership[S] (u32),
//        vertexWeight[S], communityWeight[S] (f64);
//        then the stats (7 x u64), pending polls seen (u64), and per graph its final
//        state from lv_server_graph_state
//   argv: in out slots memory_budget
// Exit status: 0; 3 with the lv_status printed when the library reports one;
// 4 when a poll contract check fails.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "louvain_hip.h"

namespace {

struct Graph {
  uint64_t S = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> keys;
  std::vector<float> w;
  std::vector<uint8_t> ex;
  std::vector<uint32_t> q;
  std::vector<double> vt, ct;
};

struct Batch {
  std::vector<uint32_t> u, v;
  std::vector<float> w;
};

struct Req {
  uint32_t graph = 0;
  Batch del, ins;
  uint64_t ticket = 0;
};

template <class T>
bool get(FILE* f, std::vector<T>& v, uint64_t n) {
  v.resize(n);
  return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

bool get1(FILE* f, uint64_t& x) { return fread(&x, sizeof x, 1, f) == 1; }

bool read_batch(FILE* f, Batch& b) {
  uint64_t n = 0;
  return get1(f, n) && get(f, b.u, n) && get(f, b.v, n) && get(f, b.w, n);
}

lv_batch view(const Batch& b) {
  lv_batch x;
  x.u = b.u.data();
  x.v = b.v.data();
  x.w = b.w.data();
  x.size = b.u.size();
  return x;
}

int fail(lv_server* s, int rc, const char* what) {
  printf("lv_status %d: %s: %s\n", rc, what, s ? lv_server_last_error(s) : lv_status_string(rc));
  if (s) lv_server_destroy(s);
  return 3;
}

template <class T>
void put(FILE* f, const T* p, size_t n) {
  fwrite(p, sizeof(T), n, f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s in out slots memory_budget\n", argv[0]);
    return 2;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 2;
  uint64_t G = 0, R = 0;
  if (!get1(in, G)) return 2;
  std::vector<Graph> gs(G);
  for (Graph& g : gs) {
    uint64_t nnz = 0;
    if (!(get1(in, g.S) && get1(in, nnz) && get(in, g.off, g.S + 1) && get(in, g.keys, nnz) && get(in, g.w, nnz) &&
          get(in, g.ex, g.S) && get(in, g.q, g.S) && get(in, g.vt, g.S) && get(in, g.ct, g.S)))
      return 2;
  }
  if (!get1(in, R)) return 2;
  std::vector<Req> rs(R);
  for (Req& r : rs) {
    if (fread(&r.graph, sizeof r.graph, 1, in) != 1 || !read_batch(in, r.del) || !read_batch(in, r.ins)) return 2;
  }
  fclose(in);

  lv_server_options o;
  lv_server_options_default(&o);
  o.slots = atoi(argv[3]);
  o.memory_budget = strtoull(argv[4], nullptr, 10);
  o.modularity = 1;
  o.keep_states = 1;
  lv_server* s = nullptr;
  int rc = lv_server_create(&o, &s);
  if (rc != LV_OK) return fail(nullptr, rc, "lv_server_create");
  std::vector<uint32_t> ids(G);
  for (uint64_t k = 0; k < G; ++k) {
    const Graph& g = gs[k];
    lv_graph x;
    x.span = g.S;
    x.offsets = g.off.data();
    x.degrees = nullptr;
    x.keys = g.keys.data();
    x.weights = g.w.data();
    x.weights_f64 = nullptr;
    x.exists = g.ex.data();
    rc = lv_server_add_graph(s, &x, g.q.data(), g.vt.data(), g.ct.data(), &ids[k]);
    if (rc != LV_OK) return fail(s, rc, "lv_server_add_graph");
  }
  for (Req& r : rs) {
    const lv_batch d = view(r.del), i = view(r.ins);
    rc = lv_server_submit(s, ids[r.graph], &d, &i, &r.ticket);
    if (rc != LV_OK) return fail(s, rc, "lv_server_submit");
  }
  FILE* out = fopen(argv[2], "wb");
  if (!out) return 2;
  uint64_t pending_seen = 0;
  for (Req& r : rs) {
    const uint64_t S = gs[r.graph].S;
    std::vector<uint32_t> memb(S);
    std::vector<double> vt(S), ct(S);
    lv_completion c;
    c.result.membership = memb.data();
    c.result.vertex_weight = vt.data();
    c.result.community_weight = ct.data();
    rc = lv_server_poll(s, r.ticket, 0, &c);
    if (rc == LV_ERR_PENDING) {
      ++pending_seen;
      rc = lv_server_poll(s, r.ticket, -1, &c);
    }
    if (rc != LV_OK) return fail(s, rc, "lv_server_poll");
    if (c.ticket != r.ticket || c.graph_id != ids[r.graph] || !(c.submitted_ms <= c.started_ms) ||
        !(c.started_ms <= c.done_ms)) {
      printf("completion of ticket %llu does not match its request\n", (unsigned long long)r.ticket);
      lv_server_destroy(s);
      return 4;
    }
    lv_completion again;
    again.result.membership = nullptr;
    again.result.vertex_weight = nullptr;
    again.result.community_weight = nullptr;
    if (lv_server_poll(s, r.ticket, 0, &again) != LV_ERR_INVALID) {
      printf("ticket %llu reported twice\n", (unsigned long long)r.ticket);
      lv_server_destroy(s);
      return 4;
    }
    const int32_t head[3] = {c.status, c.result.iterations, c.result.passes};
    const double t[3] = {c.modularity, c.started_ms, c.done_ms};
    const uint64_t sz[2] = {c.deletions, c.insertions};
    put(out, head, 3);
    put(out, t, 3);
    put(out, sz, 2);
    put(out, memb.data(), S);
    put(out, vt.data(), S);
    put(out, ct.data(), S);
  }
  lv_completion none;
  if (lv_server_poll(s, 0, 0, &none) != LV_ERR_INVALID || lv_server_poll(s, R + 1, 0, &none) != LV_ERR_INVALID) {
    printf("a ticket never issued was accepted\n");
    lv_server_destroy(s);
    return 4;
  }
  rc = lv_server_drain(s);
  if (rc != LV_OK) return fail(s, rc, "lv_server_drain");
  lv_server_stats st;
  rc = lv_server_get_stats(s, &st);
  if (rc != LV_OK) return fail(s, rc, "lv_server_get_stats");
  const uint64_t stats[7] = {st.submitted,       st.completed,      st.failed,    st.resident_bytes,
                             st.workspace_bytes, st.inflight_bytes, st.peak_bytes};
  put(out, stats, 7);
  put(out, &pending_seen, 1);
  for (uint64_t k = 0; k < G; ++k) {
    const uint64_t S = gs[k].S;
    std::vector<uint32_t> memb(S);
    std::vector<double> vt(S), ct(S);
    rc = lv_server_graph_state(s, ids[k], memb.data(), vt.data(), ct.data());
    if (rc != LV_OK) return fail(s, rc, "lv_server_graph_state");
    put(out, memb.data(), S);
    put(out, vt.data(), S);
    put(out, ct.data(), S);
  }
  fclose(out);
  lv_server_destroy(s);
  return 0;
}
