/*
 * louvain_hip.hpp -- C++ mirror of the reference's four Louvain entry points
 * over the C ABI of liblouvain_hip.so (louvain_hip.h).
 *
 *   louvainStaticHip(x, o)                                   louvainStaticOmp, inc/louvain.hxx:1285
 *   louvainNaiveDynamicHip(y, dels, ins, q, qvtot, qctot, o) louvainNaiveDynamicOmp, :1349
 *   louvainDynamicDeltaScreeningHip(...)                     louvainDynamicDeltaScreeningOmp, :1542
 *   louvainDynamicFrontierHip(...)                           louvainDynamicFrontierOmp, :1674
 *
 * Same parameter order and types as the reference: a graph G with the
 * interface of DiGraph (span, hasVertex, degree, forEachEdge;
 * inc/Graph.hxx:59-169), batches as vector<tuple<K, K, V>>, the previous
 * membership and weights as vectors, LouvainOptions (inc/louvain.hxx:30-63)
 * with the same fields and defaults, and a LouvainResult<K, double>
 * (inc/louvain.hxx:77-145) with the same fields.  A caller that includes this
 * header instead of inc/louvain.hxx and appends "Hip" to the four calls gets
 * the GPU path.
 *
 * Differences a caller can see: the reference has no error channel, this
 * mirror throws louvain_hip::Error (carrying the lv_status) when the
 * library reports one; each host thread uses its own lv_context (one HIP
 * stream, device buffers kept between calls) on device 0, or on the device
 * set with louvain_hip::setDevice before the thread's first call.
 */
#ifndef LOUVAIN_HIP_HPP
#define LOUVAIN_HIP_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "louvain_hip.h"

namespace louvain_hip {

// LouvainOptions, inc/louvain.hxx:30-63: fields, order and defaults.
struct LouvainOptions {
  int repeat;
  double resolution;
  double tolerance;
  double aggregationTolerance;
  double toleranceDrop;
  int maxIterations;
  int maxPasses;
  LouvainOptions(int repeat = 1, double resolution = 1, double tolerance = 1e-2, double aggregationTolerance = 1.0,
                 double toleranceDrop = 10, int maxIterations = 20, int maxPasses = 10)
      : repeat(repeat), resolution(resolution), tolerance(tolerance), aggregationTolerance(aggregationTolerance),
        toleranceDrop(toleranceDrop), maxIterations(maxIterations), maxPasses(maxPasses) {}
};

// LouvainResult<K, W>, inc/louvain.hxx:77-145: fields and meaning.
template <class K, class W>
struct LouvainResult {
  std::vector<K> membership;
  std::vector<W> vertexWeight;
  std::vector<W> communityWeight;
  int iterations = 0;
  int passes = 0;
  float time = 0;
  float markingTime = 0;
  float initializationTime = 0;
  float firstPassTime = 0;
  float localMoveTime = 0;
  float aggregationTime = 0;
  size_t affectedVertices = 0;
};

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const noexcept { return status_; }

 private:
  int status_;
};

namespace detail {

inline int& device_slot() {
  thread_local int device = 0;
  return device;
}

struct ContextDeleter {
  void operator()(lv_context* c) const noexcept { lv_context_destroy(c); }
};

inline lv_context* context() {
  thread_local std::unique_ptr<lv_context, ContextDeleter> ctx;
  if (!ctx) {
    lv_context* c = nullptr;
    int st = lv_context_create(device_slot(), &c);
    if (st != LV_OK) throw Error(st, std::string("lv_context_create: ") + lv_status_string(st));
    ctx.reset(c);
  }
  return ctx.get();
}

inline void check(int st, const char* what) {
  if (st == LV_OK) return;
  std::string msg = std::string(what) + ": " + lv_status_string(st);
  const char* last = lv_last_error(context());
  if (last && *last) msg += std::string(" (") + last + ")";
  throw Error(st, msg);
}

// The graph as the library's CSR: rows in forEachEdge order (sorted by
// neighbour id in DiGraph, inc/_bitset.hxx:245-262), absent ids empty.
struct Csr {
  std::vector<uint64_t> offsets;
  std::vector<uint32_t> keys;
  std::vector<float> weights;
  std::vector<uint8_t> exists;
  lv_graph view{};
};

template <class G>
void to_csr(const G& x, Csr& c) {
  const uint64_t S = x.span();
  c.offsets.assign(S + 1, 0);
  c.exists.assign(S, 0);
  for (uint64_t u = 0; u < S; ++u) {
    bool has = x.hasVertex(u);
    c.exists[u] = has ? 1 : 0;
    c.offsets[u + 1] = c.offsets[u] + (has ? (uint64_t)x.degree(u) : 0);
  }
  c.keys.resize(c.offsets[S]);
  c.weights.resize(c.offsets[S]);
  for (uint64_t u = 0; u < S; ++u) {
    if (!c.exists[u]) continue;
    uint64_t i = c.offsets[u];
    x.forEachEdge(u, [&](auto v, auto w) {
      c.keys[i] = (uint32_t)v;
      c.weights[i] = (float)w;
      ++i;
    });
  }
  c.view.span = S;
  c.view.offsets = c.offsets.data();
  c.view.degrees = nullptr;
  c.view.keys = c.keys.data();
  c.view.weights = c.weights.data();
  c.view.weights_f64 = nullptr;
  c.view.exists = c.exists.data();
}

// vector<tuple<K, K, V>> as three parallel arrays (the tuple's memory layout
// is the standard library's, not part of any ABI).
struct Edges {
  std::vector<uint32_t> u, v;
  std::vector<float> w;
  lv_batch view{};
};

template <class K, class V>
void to_edges(const std::vector<std::tuple<K, K, V>>& b, Edges& e) {
  e.u.resize(b.size());
  e.v.resize(b.size());
  e.w.resize(b.size());
  for (size_t i = 0; i < b.size(); ++i) {
    e.u[i] = (uint32_t)std::get<0>(b[i]);
    e.v[i] = (uint32_t)std::get<1>(b[i]);
    e.w[i] = (float)std::get<2>(b[i]);
  }
  e.view.u = e.u.data();
  e.view.v = e.v.data();
  e.view.w = e.w.data();
  e.view.size = b.size();
}

inline lv_options to_options(const LouvainOptions& o) {
  lv_options c;
  c.repeat = o.repeat;
  c.resolution = o.resolution;
  c.tolerance = o.tolerance;
  c.aggregation_tolerance = o.aggregationTolerance;
  c.tolerance_drop = o.toleranceDrop;
  c.max_iterations = o.maxIterations;
  c.max_passes = o.maxPasses;
  return c;
}

template <class K>
void finish(LouvainResult<K, double>& a, const std::vector<uint32_t>& memb, const lv_result& r) {
  a.membership.assign(memb.begin(), memb.end());
  a.iterations = r.iterations;
  a.passes = r.passes;
  a.time = r.time;
  a.markingTime = r.marking_time;
  a.initializationTime = r.initialization_time;
  a.firstPassTime = r.first_pass_time;
  a.localMoveTime = r.local_move_time;
  a.aggregationTime = r.aggregation_time;
  a.affectedVertices = (size_t)r.affected_vertices;
}

template <class G, class K, class V, class W>
LouvainResult<K, double> dynamic(int approach, const G& y, const std::vector<std::tuple<K, K, V>>& deletions,
                                 const std::vector<std::tuple<K, K, V>>& insertions, const std::vector<K>& q,
                                 const std::vector<W>& qvtot, const std::vector<W>& qctot, const LouvainOptions& o) {
  Csr g;
  to_csr(y, g);
  const uint64_t S = g.view.span;
  if (q.size() < S || qvtot.size() < S || qctot.size() < S)
    throw Error(LV_ERR_INVALID, "previous membership / weights shorter than the graph's span");
  Edges d, i;
  to_edges(deletions, d);
  to_edges(insertions, i);
  std::vector<uint32_t> qm(q.begin(), q.begin() + S), memb(S);
  std::vector<double> qv(qvtot.begin(), qvtot.begin() + S), qc(qctot.begin(), qctot.begin() + S);
  LouvainResult<K, double> a;
  a.vertexWeight.resize(S);
  a.communityWeight.resize(S);
  lv_result r{};
  r.membership = memb.data();
  r.vertex_weight = a.vertexWeight.data();
  r.community_weight = a.communityWeight.data();
  lv_options oc = to_options(o);
  int st = LV_ERR_INVALID;
  lv_context* c = context();
  switch (approach) {
    case LV_NAIVE_DYNAMIC:
      st = lv_louvain_naive_dynamic(c, &g.view, &d.view, &i.view, qm.data(), qv.data(), qc.data(), &oc, &r);
      break;
    case LV_DELTA_SCREENING:
      st = lv_louvain_delta_screening(c, &g.view, &d.view, &i.view, qm.data(), qv.data(), qc.data(), &oc, &r);
      break;
    case LV_DYNAMIC_FRONTIER:
      st = lv_louvain_dynamic_frontier(c, &g.view, &d.view, &i.view, qm.data(), qv.data(), qc.data(), &oc, &r);
      break;
  }
  check(st, "louvain dynamic");
  finish(a, memb, r);
  return a;
}

}  // namespace detail

// Device used by the calling thread's context (before its first call).
inline void setDevice(int device) { detail::device_slot() = device; }

// louvainStaticOmp, inc/louvain.hxx:1285-1296.
template <class G>
inline auto louvainStaticHip(const G& x, const LouvainOptions& o = {}) {
  using K = typename G::key_type;
  detail::Csr g;
  detail::to_csr(x, g);
  const uint64_t S = g.view.span;
  std::vector<uint32_t> memb(S);
  LouvainResult<K, double> a;
  a.vertexWeight.resize(S);
  a.communityWeight.resize(S);
  lv_result r{};
  r.membership = memb.data();
  r.vertex_weight = a.vertexWeight.data();
  r.community_weight = a.communityWeight.data();
  lv_options oc = detail::to_options(o);
  detail::check(lv_louvain_static(detail::context(), &g.view, &oc, &r), "lv_louvain_static");
  detail::finish(a, memb, r);
  return a;
}

// louvainNaiveDynamicOmp, inc/louvain.hxx:1349-1366.
template <class G, class K, class V, class W>
inline auto louvainNaiveDynamicHip(const G& y, const std::vector<std::tuple<K, K, V>>& deletions,
                                   const std::vector<std::tuple<K, K, V>>& insertions, const std::vector<K>& q,
                                   const std::vector<W>& qvtot, const std::vector<W>& qctot,
                                   const LouvainOptions& o = {}) {
  return detail::dynamic(LV_NAIVE_DYNAMIC, y, deletions, insertions, q, qvtot, qctot, o);
}

// louvainDynamicDeltaScreeningOmp, inc/louvain.hxx:1542-1565.
template <class G, class K, class V, class W>
inline auto louvainDynamicDeltaScreeningHip(const G& y, const std::vector<std::tuple<K, K, V>>& deletions,
                                            const std::vector<std::tuple<K, K, V>>& insertions,
                                            const std::vector<K>& q, const std::vector<W>& qvtot,
                                            const std::vector<W>& qctot, const LouvainOptions& o = {}) {
  return detail::dynamic(LV_DELTA_SCREENING, y, deletions, insertions, q, qvtot, qctot, o);
}

// louvainDynamicFrontierOmp, inc/louvain.hxx:1674-1690.
template <class G, class K, class V, class W>
inline auto louvainDynamicFrontierHip(const G& y, const std::vector<std::tuple<K, K, V>>& deletions,
                                      const std::vector<std::tuple<K, K, V>>& insertions, const std::vector<K>& q,
                                      const std::vector<W>& qvtot, const std::vector<W>& qctot,
                                      const LouvainOptions& o = {}) {
  return detail::dynamic(LV_DYNAMIC_FRONTIER, y, deletions, insertions, q, qvtot, qctot, o);
}

}  // namespace louvain_hip

#endif  // LOUVAIN_HIP_HPP
