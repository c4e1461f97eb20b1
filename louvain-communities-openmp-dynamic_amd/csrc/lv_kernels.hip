// HIP kernels of the MI355X Louvain library (gfx950, wave64).
//
// Hot path: the local-moving kernel (move_kernel) scores every neighbour
// community of a vertex with deltaModularity (inc/properties.hxx:253-256)
// after summing edge weights per community in an open-addressing hashtable.
// A vertex is handled by a group of G lanes; G and the table capacity come
// from the vertex's degree tier:
//   tier 0: deg <= 8     G = 8,    16 LDS slots   (32 vertices per block)
//   tier 1: deg <= 32    G = 16,   64 LDS slots   (16 per block)
//   tier 2: deg <= 128   G = 64,  256 LDS slots   (one wave per vertex)
//   tier 3: deg <= 1024  G = 256, 2048 LDS slots  (one workgroup per vertex)
//   tier 4: deg > 1024   G = 1024, HBM table of nextpow2(2 deg) slots
// A slot holds (key u32, first position u32, sum f64): the first position
// (edge index of the key's first appearance) reproduces the reference's
// first-seen tie-break (inc/louvain.hxx:409, :461) independently of the
// order in which lanes insert.  Aggregation reuses the same tables.
//
// Built with -ffp-contract=off: deltaModularity must evaluate without fused
// multiply-adds to match the reference's x86-64 arithmetic bit for bit.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "lv_internal.hpp"

namespace lv {

namespace {

constexpr int kBlock = 256;
// Grid cap of the reductions that end in one same-address atomic per block
// (those serialise, about 12 ns each): two blocks per CU still stream HBM.
constexpr uint64_t kReduceGrid = 512;

__host__ __device__ inline uint64_t grid_for(uint64_t n, uint64_t per_block, uint64_t cap = 1u << 20) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return g < cap ? g : cap;
}

__device__ __forceinline__ uint32_t slot_of(uint32_t c, uint64_t cap) {
  return (uint32_t)(((uint64_t)(c * 2654435761u) * cap) >> 32);
}

// deltaModularity, inc/properties.hxx:253-256 (same operation order).
__device__ __forceinline__ double delta_q(double vcout, double vdout, double vtot, double ctot, double dtot,
                                          double M, double R) {
  return (vcout - vdout) / M - R * vtot * (vtot + ctot - dtot) / (2 * M * M);
}

// Batched form: the M, community-0 id and skip flag of vertex u's graph.
__device__ __forceinline__ double vertex_M(const MoveArgs& a, uint32_t u) { return a.gid ? a.gM[a.gid[u]] : a.M; }
__device__ __forceinline__ uint32_t vertex_zero(const MoveArgs& a, uint32_t u) { return a.gid ? a.gzero[a.gid[u]] : 0u; }
__device__ __forceinline__ bool vertex_skipped(const MoveArgs& a, uint32_t u) { return a.gid && a.gmask[a.gid[u]]; }

__device__ __forceinline__ void tab_add(uint32_t* key, uint32_t* pos, double* val, uint64_t cap, uint32_t c,
                                        double w, uint32_t p) {
  uint32_t s = slot_of(c, cap);
  while (true) {
    uint32_t k = atomicCAS(&key[s], kEmpty, c);
    if (k == kEmpty || k == c) {
      unsafeAtomicAdd(&val[s], w);
      if (pos) atomicMin(&pos[s], p);
      return;
    }
    s = (uint32_t)((s + 1) & (cap - 1));
  }
}

__device__ __forceinline__ double tab_get(const uint32_t* key, const double* val, uint64_t cap, uint32_t c) {
  uint32_t s = slot_of(c, cap);
  while (true) {
    uint32_t k = key[s];
    if (k == c) return val[s];
    if (k == kEmpty) return 0;
    s = (uint32_t)((s + 1) & (cap - 1));
  }
}

// tab_add for the lanes of a wave with ok set (every lane of the wave calls
// it): the lanes whose key equals the first such lane's key c0 add their
// summed weight once, through that lane -- one atomic instead of one per lane
// when a community dominates -- and the others insert their own.  The lanes
// hold ascending positions p, so the first lane holding c0 has the smallest
// position of the lanes that share it.
__device__ __forceinline__ void tab_add_wave(uint32_t* key, uint32_t* pos, double* val, uint64_t cap, bool ok,
                                             uint32_t c, double w, uint32_t p) {
  const unsigned long long act = __ballot(ok);
  if (!act) return;
  const int lane = threadIdx.x % 64;
  const int lead = __ffsll((long long)act) - 1;
  const uint32_t c0 = __shfl(c, lead, 64);
  const bool same = ok && c == c0;
  double sum = same ? w : 0.0;
  for (int m = 32; m > 0; m >>= 1) sum += __shfl_xor(sum, m, 64);
  if (lane == lead) tab_add(key, pos, val, cap, c0, sum, p);
  else if (ok && !same) tab_add(key, pos, val, cap, c, w, p);
}

// Better candidate: larger gain, then earlier first appearance.
__device__ __forceinline__ bool better(double e, uint32_t p, double be, uint32_t bp) {
  return e > be || (e == be && p < bp);
}

template <int W>
__device__ __forceinline__ void group_best_shfl(double& be, uint32_t& bp, uint32_t& bc) {
#pragma unroll
  for (int m = W / 2; m > 0; m >>= 1) {
    double oe = __shfl_xor(be, m, W);
    uint32_t op = __shfl_xor(bp, m, W);
    uint32_t oc = __shfl_xor(bc, m, W);
    if (better(oe, op, be, bp)) { be = oe; bp = op; bc = oc; }
  }
}

// Block-wide best (every thread of the block must call it).
template <int BLOCK>
__device__ void block_best(double& be, uint32_t& bp, uint32_t& bc) {
  __shared__ double se[BLOCK / 64];
  __shared__ uint32_t sp[BLOCK / 64], sc[BLOCK / 64];
  group_best_shfl<64>(be, bp, bc);
  int w = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) { se[w] = be; sp[w] = bp; sc[w] = bc; }
  __syncthreads();
  be = se[0]; bp = sp[0]; bc = sc[0];
#pragma unroll
  for (int i = 1; i < BLOCK / 64; ++i)
    if (better(se[i], sp[i], be, bp)) { be = se[i]; bp = sp[i]; bc = sc[i]; }
  __syncthreads();
}

template <int BLOCK>
__device__ double block_sum(double x) {
  __shared__ double s[BLOCK / 64];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m, 64);
  if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = x;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / 64; ++i) t += s[i];
  __syncthreads();
  return t;
}

template <int BLOCK>
__device__ unsigned long long block_sum_u64(unsigned long long x) {
  __shared__ unsigned long long s[BLOCK / 64];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m, 64);
  if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = x;
  __syncthreads();
  unsigned long long t = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / 64; ++i) t += s[i];
  __syncthreads();
  return t;
}

// Slot of this block in a spread sum (kSpread slots, 128 B apart).
__device__ __forceinline__ uint64_t spread_slot() { return blockIdx.x % kSpread; }

// A block's counter totals into its spread slot (cnt: the tier's counters
// in slot 0).
__device__ __forceinline__ void add_counts(unsigned long long* cnt, unsigned long long v, unsigned long long e,
                                           unsigned long long c, unsigned long long m, unsigned long long me) {
  unsigned long long* s = cnt + spread_slot() * kCntStride;
  if (v) atomicAdd(&s[CNT_VERTICES], v);
  if (e) atomicAdd(&s[CNT_EDGES], e);
  if (c) atomicAdd(&s[CNT_CANDIDATES], c);
  if (m) atomicAdd(&s[CNT_MOVES], m);
  if (me) atomicAdd(&s[CNT_MOVED_EDGES], me);
}

// v joins the next sub-round: its bit, and its compaction segment's dirty
// mark (a segment without one is skipped by the compaction).
__device__ __forceinline__ void mark_same(const MoveArgs& a, uint32_t v) {
  __hip_atomic_fetch_or(&a.fsame[v >> 5], 1u << (v & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (a.wl_dirty) a.wl_dirty[v >> a.wl_shift] = 1;
}

// The worklist iteration serial (device-resident when the launch sequence
// is replayed from a captured graph).
__device__ __forceinline__ uint32_t cur_serial(const MoveArgs& a) { return a.serial_dev ? *a.serial_dev : a.serial; }

// vaff[v] = 1, as a moved vertex u flags its neighbour v (inc/louvain.hxx:602).
__device__ __forceinline__ void flag_vertex(uint8_t* vaff, uint32_t v, uint32_t u, const MoveArgs& a) {
  __hip_atomic_store(&vaff[v], (uint8_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // worklist mode: a vertex flagged by a lower id and not visited in this
  // iteration yet joins the next sub-round (its turn in the reference's
  // ascending sweep is still to come, inc/louvain.hxx:602 then :598)
  if (a.fsame && v > u && __hip_atomic_load(&a.stamp[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != cur_serial(a))
    mark_same(a, v);
}

// Hub tables: the slot of community c in the table at off (cap slots),
// inserted if absent and insert is set; -1 if absent or no room.
__device__ __forceinline__ int64_t ht_slot(uint32_t* key, uint64_t off, uint32_t cap, uint32_t c, bool insert) {
  uint32_t s = slot_of(c, cap);
  for (uint32_t probe = 0; probe < cap; ++probe) {
    uint32_t k = __hip_atomic_load(&key[off + s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == kEmpty) {
      if (!insert) return -1;
      k = atomicCAS(&key[off + s], kEmpty, c);
    }
    if (k == kEmpty || k == c) return (int64_t)(off + s);
    s = (s + 1) & (cap - 1);
  }
  return -1;
}

// x's table dropped (x is then evaluated from its edges); *ndrop counts
// each dropped table once.
__device__ __forceinline__ void ht_drop(uint32_t* cap, uint32_t x, unsigned long long* ndrop) {
  if (atomicExch(&cap[x], kHtDropped) > kHtDropped && ndrop) atomicAdd(ndrop, 1ull);
}

// count n and weight w of community c in x's table; a table with no room
// for c is dropped.
__device__ __forceinline__ void ht_add(uint32_t* cap, const uint64_t* off, uint32_t* key, int32_t* cnt, double* wt,
                                       unsigned long long* ndrop, uint32_t x, uint32_t c, int32_t n, double w) {
  const uint32_t cx = __hip_atomic_load(&cap[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (cx <= kHtDropped) return;
  const int64_t s = ht_slot(key, off[x], cx, c, n > 0);
  if (s < 0) {
    if (n > 0) ht_drop(cap, x, ndrop);
    return;
  }
  atomicAdd(&cnt[s], n);
  unsafeAtomicAdd(&wt[s], w);
}

// Neighbour v of x moved from community `from` to `to` over an edge of
// weight w: x's table follows.
__device__ __forceinline__ void ht_move(const MoveArgs& a, uint32_t x, double w, uint32_t from, uint32_t to) {
  if (!a.ht_cap) return;
  const uint32_t cx = __hip_atomic_load(&a.ht_cap[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (cx <= kHtDropped) return;
  const uint64_t off = a.ht_off[x];
  int64_t s = ht_slot(a.ht_key, off, cx, from, false);
  if (s >= 0) {
    atomicAdd(&a.ht_cnt[s], -1);
    unsafeAtomicAdd(&a.ht_w[s], -w);
  }
  s = ht_slot(a.ht_key, off, cx, to, true);
  if (s < 0) {
    ht_drop(a.ht_cap, x, a.ht_ndrop);
    return;
  }
  atomicAdd(&a.ht_cnt[s], 1);
  unsafeAtomicAdd(&a.ht_w[s], w);
}

// ht_move with x's table size (cx > kHtDropped) and first slot read: the
// slots of `from` and `to` are searched together, one key load of each per
// probe step (linear probing from each one's home slot, at most cx steps).
__device__ __forceinline__ void ht_move_cap(const MoveArgs& a, uint32_t x, uint32_t cx, uint64_t off, double w,
                                            uint32_t from, uint32_t to) {
  uint32_t sf = slot_of(from, cx), st = slot_of(to, cx);
  int64_t pf = -1, pt = -1;
  bool df = false, dt = false;
  for (uint32_t probe = 0; probe < cx && !(df && dt); ++probe) {
    const uint32_t kf = df ? 0u : __hip_atomic_load(&a.ht_key[off + sf], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t kt = dt ? 0u : __hip_atomic_load(&a.ht_key[off + st], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!df) {
      if (kf == from) {
        pf = (int64_t)(off + sf);
        df = true;
      } else if (kf == kEmpty) {
        df = true;
      } else {
        sf = (sf + 1) & (cx - 1);
      }
    }
    if (!dt) {
      if (kt == kEmpty) kt = atomicCAS(&a.ht_key[off + st], kEmpty, to);
      if (kt == kEmpty || kt == to) {
        pt = (int64_t)(off + st);
        dt = true;
      } else {
        st = (st + 1) & (cx - 1);
      }
    }
  }
  if (pf >= 0) {
    atomicAdd(&a.ht_cnt[pf], -1);
    unsafeAtomicAdd(&a.ht_w[pf], -w);
  }
  if (pt < 0) {
    ht_drop(a.ht_cap, x, a.ht_ndrop);
    return;
  }
  atomicAdd(&a.ht_cnt[pt], 1);
  unsafeAtomicAdd(&a.ht_w[pt], w);
}

// A vertex u that moved from community `from` to `to` (row at o, d edges)
// flags its neighbours (inc/louvain.hxx:602) and moves their hub tables
// with it.  Its G lanes take four rounds of G edges at a time: the
// round's key and weight loads, then its flags with the stamp and table-size
// loads, then the next-sub-round marks and table updates.
template <int G, class WT>
__device__ __forceinline__ void flag_neighbours(const MoveArgs& a, const WT* W, uint32_t u, uint64_t o, uint32_t d,
                                                int lane, uint32_t from, uint32_t to) {
  constexpr int R = 4;
  const uint32_t serial = a.fsame ? cur_serial(a) : 0;
  for (uint32_t b = 0; b < d; b += R * G) {
    uint32_t v[R], sv[R], cx[R];
    double w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t i = b + r * G + lane;
      v[r] = i < d ? a.g.keys[o + i] : u;
      w[r] = i < d && a.ht_cap ? (double)W[o + i] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      sv[r] = serial;
      cx[r] = 0;
      if (v[r] == u) continue;
      __hip_atomic_store(&a.vaff[v[r]], (uint8_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (a.fsame && v[r] > u) sv[r] = __hip_atomic_load(&a.stamp[v[r]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (a.ht_cap) cx[r] = __hip_atomic_load(&a.ht_cap[v[r]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // flagged by a lower id and not visited this iteration: next sub-round
      if (sv[r] != serial) mark_same(a, v[r]);
      if (cx[r] > kHtDropped) ht_move_cap(a, v[r], cx[r], a.ht_off[v[r]], w[r], from, to);
    }
  }
}

// ---------------------------------------------------------------- fills
template <class T>
__global__ void fill_kernel(T* a, T v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void iota_kernel(uint32_t* a, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = (uint32_t)i;
}

__global__ void degrees_kernel(const uint64_t* off, uint32_t* deg, uint64_t S) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < S; i += (uint64_t)gridDim.x * blockDim.x)
    deg[i] = (uint32_t)(off[i + 1] - off[i]);
}

// --------------------------------------------- per-vertex row sums (K, 2M)
// Vertex of this thread in the grid-stride step at `base` (steps of N =
// the grid's threads): thread t takes base + (t x P mod N), P a prime above
// any N (a bijection of [0, N)).  R-MAT-like graphs put their long rows at
// ids with many zero bits, so plain or power-of-two strided assignments give
// a few waves most of the long rows (each wave reads its long rows one after
// another); the scrambled one spreads them over the grid.
__device__ __forceinline__ uint64_t scrambled_vertex(uint64_t base) {
  const uint64_t N = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  return base + (t * 2654435761ull) % N;
}

// louvainVertexWeightsOmpW (inc/louvain.hxx:210-218).  Rows of up to 64
// edges are summed by one lane in row order (the reference's order); longer
// rows by the whole wave (lane i sums edges i, i + 64, ... in that order,
// then a butterfly over the lanes).  Eight loads are issued before their
// sums, which keep that order.
template <class WT>
__global__ __launch_bounds__(kBlock) void row_sum_kernel(GView g, double* vtot) {
  const WT* W = (const WT*)g.w;
  const int lane = __lane_id();
  for (uint64_t base = 0; base < g.S; base += (uint64_t)gridDim.x * kBlock) {  // grid-uniform
    const uint64_t u = scrambled_vertex(base);
    const bool live = u < g.S && (!g.ex || g.ex[u]);
    const uint64_t o = live ? g.off[u] : 0;
    const uint32_t d = live ? g.deg[u] : 0;
    const bool big = live && d > 64;
    if (live && !big) {
      double s = 0;
      for (uint32_t i0 = 0; i0 < d; i0 += 8) {
        double v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = i0 + r < d ? (double)W[o + i0 + r] : 0.0;
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (i0 + r < d) s += v[r];
      }
      vtot[u] += s;
    }
    unsigned long long mask = __ballot(big);
    while (mask) {
      const int j = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const uint64_t oj = __shfl(o, j);
      const uint32_t dj = __shfl(d, j);
      double s = 0;
      for (uint32_t i0 = 0; i0 < dj; i0 += 8 * 64) {
        double v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const uint32_t i = i0 + r * 64 + lane;
          v[r] = i < dj ? (double)W[oj + i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (i0 + r * 64 + lane < dj) s += v[r];
      }
#pragma unroll
      for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
      if (lane == j) vtot[u] += s;
    }
  }
}

// edgeWeightOmp (inc/properties.hxx:96-106): 2M, the sum of every live
// row's weights (in no fixed order, like the reference's reduction).  A
// wave reads its 64 rows one at a time, lane i taking edge i, eight rows in
// flight; edges past the 64th of a row eight 64-edge runs at a time.
template <class WT>
__global__ __launch_bounds__(kBlock) void edge_weight_kernel(GView g, double* total) {
  const WT* W = (const WT*)g.w;
  const int lane = __lane_id();
  double acc = 0;
  for (uint64_t base = 0; base < g.S; base += (uint64_t)gridDim.x * kBlock) {  // grid-uniform
    const uint64_t u = scrambled_vertex(base);
    const bool live = u < g.S && (!g.ex || g.ex[u]);
    const uint64_t o = live ? g.off[u] : 0;
    const uint32_t d = live ? g.deg[u] : 0;
#pragma unroll
    for (int j0 = 0; j0 < 64; j0 += 8) {
      double v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const uint64_t oj = __shfl(o, j0 + r);
        const uint32_t dj = __shfl(d, j0 + r);
        v[r] = (uint32_t)lane < dj ? (double)W[oj + lane] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) acc += v[r];
    }
    unsigned long long mask = __ballot(d > 64);
    while (mask) {
      const int j = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const uint64_t oj = __shfl(o, j);
      const uint32_t dj = __shfl(d, j);
      for (uint32_t i0 = 64; i0 < dj; i0 += 8 * 64) {
        double v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const uint32_t i = i0 + r * 64 + lane;
          v[r] = i < dj ? (double)W[oj + i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) acc += v[r];
      }
    }
  }
  const double t = block_sum<kBlock>(acc);
  if (threadIdx.x == 0) unsafeAtomicAdd(total, t);
}

// louvainInitializeOmpW (inc/louvain.hxx:285-293)
__global__ void initialize_kernel(GView g, uint32_t* vcom, double* ctot, const double* vtot) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < g.S; u += (uint64_t)gridDim.x * blockDim.x) {
    if (g.ex && !g.ex[u]) continue;
    vcom[u] = (uint32_t)u;
    ctot[u] = vtot[u];
  }
}

// louvainUpdateWeightsFromOmpU (inc/louvain.hxx:374-388): K[u] and
// Sigma_tot[vcom[u]] minus deletions, plus insertions.  f64 atomics; the
// reference's per-key batch order only matters for rounding.
// A wave takes 64 consecutive entries of one list (deletions in the first
// waves, insertions after); a list is sorted by source, so the entries of a
// source are consecutive lanes: the first lane of each run adds the run's
// sum (one atomic pair per source and wave instead of per entry -- an R-MAT
// batch has hundreds of entries on a hub).
__global__ __launch_bounds__(kBlock) void update_weights_kernel(const uint32_t* du, const float* dw, uint64_t nd,
                                                                const uint32_t* iu, const float* iw, uint64_t ni,
                                                                const uint32_t* vcom, double* vtot, double* ctot) {
  const uint64_t wd = (nd + 63) / 64, wn = wd + (ni + 63) / 64;  // waves of each list
  const int lane = threadIdx.x % 64;
  for (uint64_t wv = blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64; wv < wn;
       wv += (uint64_t)gridDim.x * (kBlock / 64)) {  // wave-uniform
    const bool del = wv < wd;
    const uint64_t i = (del ? wv : wv - wd) * 64 + lane, n = del ? nd : ni;
    const bool in = i < n;
    const uint32_t u = in ? (del ? du[i] : iu[i]) : 0xFFFFFFFFu;
    double s = in ? (del ? -(double)dw[i] : (double)iw[i]) : 0.0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {  // s: the sum over this lane's run from here on
      const double y = __shfl_down(s, d, 64);
      const uint32_t uy = __shfl_down(u, d, 64);
      if (lane + d < 64 && uy == u) s += y;
    }
    const uint32_t up = __shfl_up(u, 1, 64);
    if (in && (lane == 0 || up != u)) {
      unsafeAtomicAdd(&vtot[u], s);
      unsafeAtomicAdd(&ctot[vcom[u]], s);
    }
  }
}

// louvainAffectedVerticesFrontierOmpW (inc/louvain.hxx:1606-1624); vaff must
// be zero on entry.
__global__ void frontier_kernel(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* iu,
                                const uint32_t* iv, uint64_t ni, const uint32_t* vcom, uint8_t* vaff) {
  const uint64_t n = nd + ni;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < nd) {
      if (vcom[du[i]] == vcom[dv[i]]) vaff[du[i]] = 1;
    } else if (vcom[iu[i - nd]] != vcom[iv[i - nd]]) {
      vaff[iu[i - nd]] = 1;
    }
  }
}

// -------------------------------------------------- delta-screening marking
// louvainAffectedVerticesDeltaScreeningOmpW (inc/louvain.hxx:1441-1486).
__global__ void ds_deletions_kernel(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* vcom,
                                    uint8_t* vertices, uint8_t* neighbors, uint8_t* communities) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nd; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = du[i], v = dv[i];
    if (vcom[u] != vcom[v]) continue;
    vertices[u] = 1;
    neighbors[u] = 1;
    communities[vcom[v]] = 1;
  }
}

// One thread per insertion group (run of equal sources, in batch order),
// scanned sequentially exactly as the reference does: communities listed in
// first-seen order, sums in batch order, strict '>' from 0.  The group at
// batch index i uses table slots [2i, 2i + 2 len) and list slots [i, i + len).
// A union of graphs (gid set) takes each source's M from its own graph.
__global__ void ds_insertions_kernel(const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                                     const uint32_t* vcom, const double* vtot, const double* ctot, double M, double R,
                                     uint32_t* tkey, double* tval, uint32_t* tlist, uint8_t* vertices,
                                     uint8_t* neighbors, uint8_t* communities, const uint32_t* gid,
                                     const double* gM) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ni; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = iu[i];
    if (i > 0 && iu[i - 1] == u) continue;  // not a group head
    const double Mu = gid ? gM[gid[u]] : M;
    uint64_t len = 1;
    while (i + len < ni && iu[i + len] == u) ++len;
    uint64_t cap = 2 * len;
    uint32_t* key = tkey + 2 * i;
    double* val = tval + 2 * i;
    uint32_t* list = tlist + i;
    uint32_t d = vcom[u], cnt = 0;
    for (uint64_t j = i; j < i + len; ++j) {
      uint32_t c = vcom[iv[j]];
      if (c == d) continue;
      uint64_t s = ((uint64_t)(c * 2654435761u)) % cap;
      while (key[s] != kEmpty && key[s] != c) s = (s + 1) % cap;
      if (key[s] == kEmpty) { key[s] = c; val[s] = 0; }
      if (!val[s]) list[cnt++] = c;  // louvainScanCommunityW, inc/louvain.hxx:409
      val[s] += (double)iw[j];
    }
    uint32_t cmax = 0;
    double emax = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t c = list[k];
      if (c == d) continue;
      uint64_t s = ((uint64_t)(c * 2654435761u)) % cap;
      while (key[s] != c) s = (s + 1) % cap;
      double e = delta_q(val[s], 0.0, vtot[u], ctot[c], ctot[d], Mu, R);
      if (e > emax) { emax = e; cmax = c; }
    }
    if (emax <= 0) continue;
    vertices[u] = 1;
    neighbors[u] = 1;
    communities[cmax] = 1;
  }
}

// Sweep: neighbours of flagged vertices, members of flagged communities.
__global__ __launch_bounds__(kBlock) void ds_sweep_kernel(GView g, const uint32_t* vcom, const uint8_t* neighbors,
                                                          const uint8_t* communities, uint8_t* vertices) {
  int lane = __lane_id();
  for (uint64_t base = 0; base < g.S; base += (uint64_t)gridDim.x * kBlock) {  // grid-uniform
    uint64_t u = scrambled_vertex(base);
    bool live = u < g.S && (!g.ex || g.ex[u]);
    bool nb = live && neighbors[u];
    if (live && communities[vcom[u]]) vertices[u] = 1;
    unsigned long long mask = __ballot(nb);
    uint64_t o = live ? g.off[u] : 0;
    uint32_t d = live ? g.deg[u] : 0;
    while (mask) {
      int j = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      uint64_t oj = __shfl(o, j);
      uint32_t dj = __shfl(d, j);
      for (uint32_t i = lane; i < dj; i += 64) vertices[g.keys[oj + i]] = 1;
    }
  }
}

// Bytes of w equal to v (the zero bytes of w ^ vvvv).
__device__ __forceinline__ uint32_t bytes_equal(uint32_t w, uint8_t v) {
  const uint32_t x = w ^ (0x01010101u * v);
  return 4 - __popc((((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u);
}

// Entries of a[0 .. n) equal to v, added to *out: 16 bytes per load where a
// is 16-byte aligned, then the tail byte by byte.
__global__ __launch_bounds__(kBlock) void count_u8_kernel(const uint8_t* a, uint8_t v, uint64_t n,
                                                          unsigned long long* out) {
  unsigned long long c = 0;
  const uint64_t t = blockIdx.x * (uint64_t)kBlock + threadIdx.x, T = (uint64_t)gridDim.x * kBlock;
  const uint64_t n16 = ((uintptr_t)a & 15) ? 0 : n / 16;
  for (uint64_t i = t; i < n16; i += T) {
    const uint4 x = reinterpret_cast<const uint4*>(a)[i];
    c += bytes_equal(x.x, v) + bytes_equal(x.y, v) + bytes_equal(x.z, v) + bytes_equal(x.w, v);
  }
  for (uint64_t i = n16 * 16 + t; i < n; i += T) c += a[i] == v;
  c = block_sum_u64<kBlock>(c);
  if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

__global__ void check_sorted_kernel(const uint32_t* u, const uint32_t* v, uint64_t n, uint32_t S, int* bad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (u[i] >= S || v[i] >= S) { atomicOr(bad, 1); continue; }
    if (i > 0 && (u[i - 1] > u[i] || (u[i - 1] == u[i] && v[i - 1] >= v[i]))) atomicOr(bad, 2);
  }
}

// Graph rows (the live part of a gapped row): neighbour ids < S (bit 0) and
// strictly ascending (bit 1), the order DiGraph keeps (inc/_bitset.hxx:245-262).
// One wave per row.
__global__ __launch_bounds__(kBlock) void check_rows_kernel(const uint64_t* off, const uint32_t* deg,
                                                            const uint32_t* keys, uint64_t S, int* bad) {
  const int lane = threadIdx.x % 64;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t u = blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64; u < S; u += nw) {
    const uint64_t o = off[u];
    const uint32_t d = deg[u];
    int b = 0;
    for (uint32_t i = lane; i < d; i += 64) {
      uint32_t k = keys[o + i];
      if (k >= S) b |= 1;
      if (i + 1 < d && k >= keys[o + i + 1]) b |= 2;
    }
    if (b) atomicOr(bad, b);
  }
}

// Ungapped CSR: the same checks edge-parallel.  Bit i of `starts` is set when
// a row starts at edge i, so edge i and i + 1 share a row unless bit i + 1
// is set.
__global__ void row_starts_kernel(const uint64_t* off, uint64_t S, uint32_t* starts) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < S; u += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t o = off[u];
    if (off[u + 1] > o) atomicOr(&starts[o >> 5], 1u << (o & 31));
  }
}

__global__ void check_edges_kernel(const uint32_t* keys, uint64_t E, uint64_t S, const uint32_t* starts, int* bad) {
  int b = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t k = keys[i];
    if (k >= S) b |= 1;
    if (i + 1 < E && !((starts[(i + 1) >> 5] >> ((i + 1) & 31)) & 1u) && k >= keys[i + 1]) b |= 2;
  }
  if (__ballot(b != 0) && b) atomicOr(bad, b);
}

// ------------------------------------------------------- local-moving
// Shared mutable state (memberships, Sigma_tot, flags) is read and written
// at device scope so that a stream sees the moves other streams made.
template <class T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One vertex's neighbour-community table: phases of louvainClearScanW,
// louvainScanCommunitiesW and louvainChooseCommunity (inc/louvain.hxx:406-464)
// for a group of G lanes.  Called with LDS or HBM table pointers; after
// inlining each call site keeps native ds_* or global_* instructions.
template <int G>
__device__ __forceinline__ void table_clear(uint32_t* key, uint32_t* pos, double* val, uint64_t cap, int lane) {
  for (uint64_t s = lane; s < cap; s += G) { key[s] = kEmpty; pos[s] = kNoPos; val[s] = 0; }
}

constexpr int kScanRounds = 4;

// Order-free digest of the neighbour communities one lane read (its edges
// of the row, in the same lane-to-edge layout every time): a re-read whose
// digest differs saw a neighbour move.
__device__ __forceinline__ uint64_t comm_digest(uint32_t c) { return ((uint64_t)c + 1) * 0x9E3779B97F4A7C15ull; }

template <int G, class WT>
__device__ __forceinline__ void table_scan(const MoveArgs& a, const WT* W, uint32_t u, uint64_t o, uint32_t d,
                                           uint32_t* key, uint32_t* pos, double* val, uint64_t cap, int lane,
                                           uint64_t& dig) {
  // kScanRounds rounds of G edges at a time: their key and weight loads,
  // then their community gathers, then the inserts in round order (the
  // order of a plain per-round loop, so the sums are the same)
  constexpr int R = kScanRounds;
  for (uint32_t b = 0; b < d; b += R * G) {
    uint32_t v[R], c[R];
    double w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t i = b + r * G + lane;
      v[r] = i < d ? a.g.keys[o + i] : u;
      w[r] = i < d ? (double)W[o + i] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) c[r] = v[r] != u ? ld_dev(&a.vcom[v[r]]) : 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {  // self-loops are skipped (louvainScanCommunityW<SELF=false>)
      if (v[r] != u) tab_add(key, pos, val, cap, c[r], w[r], b + r * G + lane);
      dig += comm_digest(c[r]);
    }
  }
}

// The digest of the row's neighbour communities read again (table_scan's
// lane-to-edge layout, no table).
template <int G>
__device__ __forceinline__ uint64_t row_digest(const MoveArgs& a, uint32_t u, uint64_t o, uint32_t d, int lane) {
  constexpr int R = kScanRounds;
  uint64_t dig = 0;
  for (uint32_t b = 0; b < d; b += R * G) {
    uint32_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t i = b + r * G + lane;
      v[r] = i < d ? a.g.keys[o + i] : u;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) dig += comm_digest(v[r] != u ? ld_dev(&a.vcom[v[r]]) : 0);
  }
  return dig;
}

constexpr int kBestBatch = 2;

template <int G>
__device__ __forceinline__ void table_best(const MoveArgs& a, uint32_t cu, double Ku, double M, const uint32_t* key,
                                           const uint32_t* pos, const double* val, uint64_t cap, int lane, double& be,
                                           uint32_t& bp, uint32_t& bc, unsigned long long& c_cand) {
  const double Sd = ld_dev(&a.ctot[cu]);
  const double vd = tab_get(key, val, cap, cu);
  // kBestBatch of the lane's slots at a time: their Sigma_tot gathers are
  // issued together (the best is order-free: larger gain, then earlier
  // first position)
  constexpr int B = kBestBatch;
  for (uint64_t s0 = 0; s0 < cap; s0 += (uint64_t)B * G) {
    uint32_t c[B];
    double t[B];
#pragma unroll
    for (int r = 0; r < B; ++r) {
      const uint64_t s = s0 + (uint64_t)r * G + lane;
      c[r] = s < cap ? key[s] : kEmpty;
      c_cand += c[r] != kEmpty;
      if (c[r] == cu || (a.pick_less && c[r] > cu)) c[r] = kEmpty;
      t[r] = c[r] != kEmpty ? ld_dev(&a.ctot[c[r]]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
      if (c[r] == kEmpty) continue;
      const uint64_t s = s0 + (uint64_t)r * G + lane;
      const double e = delta_q(val[s], vd, Ku, t[r], Sd, M, a.R);
      const uint32_t p = pos[s];
      if (e > 0 && better(e, p, be, bp)) { be = e; bp = p; bc = c[r]; }
    }
  }
}

// MOVE_HYBRID_X: MOVE_HYBRID with validated moves and hubs visited in the
// sweep through their own tables (a separate instantiation: the extra state
// costs the plain kernel a wave per SIMD of occupancy).
enum { MOVE_LDS = 0, MOVE_HBM = 1, MOVE_HYBRID = 2, MOVE_HYBRID_X = 3 };

// louvainMoveOmpW body (inc/louvain.hxx:595-605): for each affected vertex,
// sum edge weights per neighbour community (self-loops skipped), choose
// the best positive deltaModularity (first-seen on ties), and unless it is
// community 0 move there (atomic Sigma_tot), flag the neighbours; clear the
// vertex's flag and add its gain to el.
//
// Scheduling follows the reference's OpenMP loop (schedule(dynamic, 2048),
// inc/louvain.hxx:594): the list is cut into chunks of a.chunk consecutive
// entries handed out in ascending order from a device counter (a.work); each
// group of G lanes walks its chunk SEQUENTIALLY, one vertex at a time, then
// takes the next free chunk.  The list is therefore swept in ascending order
// by a window of (resident groups x chunk) entries, and a vertex sees the
// moves of nearly every entry before it.
//
// MODE: MOVE_LDS    per-group LDS table of CAP slots (degree <= CAP / 2)
//       MOVE_HBM    per-block HBM table of nextpow2(2 deg) <= hcap slots (hubs)
//       MOVE_HYBRID one wave per stream: LDS table when 2 deg <= CAP, else the
//                   block's HBM region of gcap slots; vertices with degree above
//                   hub_min are left to the MOVE_HBM pre-pass.
//
// The body runs as block `bid` of `nblk` (a launch of its own, or the
// first blocks of the fused worklist launch).
// BT: the batched form's per-graph state (a.gid and what it indexes) is
// read; without it a.gid must be null (one graph: its registers are spared).
template <int BLOCK, int G, int CAP, class WT, int MODE_, bool BT = true>
__device__ __forceinline__ void move_body(const MoveArgs& a, uint32_t bid, uint32_t nblk) {
  constexpr bool EXTRAS = MODE_ == MOVE_HYBRID_X;
  // the batched form's graph of each entry read with it, its M and
  // community-0 id loaded ahead of their use (the validated kernel, short of
  // registers, looks them up where used)
  constexpr bool EARLY = BT && !EXTRAS;
  constexpr int MODE = EXTRAS ? MOVE_HYBRID : MODE_;
  constexpr int GPB = BLOCK / G;
  constexpr int LCAP = MODE == MOVE_HBM ? 1 : CAP;
  __shared__ uint32_t s_key[GPB * LCAP];
  __shared__ uint32_t s_pos[GPB * LCAP];
  __shared__ double s_val[GPB * LCAP];
  __shared__ uint64_t s_take;
  if (a.done && *a.done) return;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const uint64_t n = a.n_dev ? (uint64_t)*a.n_dev : a.n;
  // a.chunk > 0: chunks of a.chunk entries handed out in ascending order;
  // a.chunk == 0: ceil(n / streams) entries per chunk, chunk g to group g,
  // so the groups walk far-apart parts of the list at the same time.
  // a.interleave (chunk 0, no ctab): min(n, streams) groups, group g taking
  // entries g, g + N, g + 2N, ... -- an ascending list whose low entries
  // are the heavy ones (R-MAT ids) is then spread evenly over the groups.
  const bool dyn = a.chunk != 0;
  const bool ilv = a.interleave && !dyn && !a.ctab;
  const uint64_t streams = a.max_streams ? a.max_streams : 1;
  const uint64_t chunk = dyn ? a.chunk : ((n + streams - 1) / streams ? (n + streams - 1) / streams : 1);
  // a.ctab (id-order kernel, one group per block): chunk bounds per chunk
  const uint64_t nchunks = a.ctab ? (a.nct_dev ? (uint64_t)*a.nct_dev : a.nct)
                          : ilv ? (n < streams ? n : streams) : (n + chunk - 1) / chunk;
  uint64_t next = (uint64_t)bid * GPB;
  const WT* W = (const WT*)a.g.w;
  // LDS table of this group; HBM region of this block
  uint32_t* lkey = s_key + grp * LCAP;
  uint32_t* lpos = s_pos + grp * LCAP;
  double* lval = s_val + grp * LCAP;
  uint32_t *hkey = nullptr, *hpos = nullptr;
  double* hval = nullptr;
  if constexpr (MODE == MOVE_HBM) {
    uint64_t base = (uint64_t)bid * a.hcap;
    hkey = a.hkey + base;
    hpos = a.hpos + base;
    hval = a.hval + base;
  } else if constexpr (MODE == MOVE_HYBRID) {
    uint64_t base = ((uint64_t)bid * GPB + grp) * a.gcap;  // the group's region
    hkey = a.gkey + base;
    hpos = a.gpos + base;
    hval = a.gval + base;
  }
  // gains: the block's in el_acc; in the batched form (a.gel) el_acc holds
  // the group's gains in graph g_cur so far, added to the graph's gel slot
  // when the group leaves the graph or its chunk (one atomic per stretch of
  // visits, not per visit: all of a graph's visits add to one word)
  double el_acc = 0;
  uint32_t g_cur = 0;
  unsigned long long c_vert = 0, c_edge = 0, c_cand = 0, c_move = 0, c_medge = 0;
  bool worked = false;  // (block-uniform: the block took a chunk)
  for (;;) {
    // the block takes GPB consecutive chunks (one per group)
    uint64_t cb = next;
    if (dyn) {
      if (threadIdx.x == 0) s_take = atomicAdd(a.work, (unsigned long long)GPB);
      __syncthreads();
      cb = s_take;
      __syncthreads();
    } else {
      next += (uint64_t)nblk * GPB;
    }
    if (cb >= nchunks) break;
    worked = true;
    const uint64_t ch = cb + grp;
    // entries [c0, c0 + clen) of the list
    // (interleaved: entries ch + j * nchunks, j < clen)
    uint64_t c0 = ch * chunk, clen = chunk, step = 1;
    if (a.ctab) {
      c0 = ch < nchunks ? a.ctab[2 * ch] : 0;
      clen = ch < nchunks ? a.ctab[2 * ch + 1] - c0 : 0;
    } else if (ilv) {
      c0 = ch;
      step = nchunks;
      clen = ch < nchunks ? (n - ch + nchunks - 1) / nchunks : 0;
    }
    // One vertex (offset o, degree d, community cu, weight Ku) through the
    // group; every thread of the block calls it.  Its barriers are the
    // block's, or the wave's in the id-order kernel (one or more groups per
    // wave, each with its own LDS table and HBM region: the groups run
    // independently).
    auto group_sync = [&]() __attribute__((always_inline)) {
      if constexpr (MODE == MOVE_HYBRID && G <= 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      } else {
        __syncthreads();
      }
    };
    // (gu: u's graph in the batched form, read with the entry; its M and
    // community-0 id are loaded at once, ahead of their use)
    auto visit = [&](bool act, uint32_t u, uint64_t o, uint32_t d, uint32_t cu, double Ku,
                     uint32_t gu) __attribute__((always_inline)) {
      if (act && a.stamp && lane == 0) st_dev(&a.stamp[u], cur_serial(a));
      double Mu = a.M;
      uint32_t zu = 0;
      if constexpr (EARLY) {
        if (act && a.gid) {
          Mu = a.gM[gu];
          zu = a.gzero[gu];
        }
      }
      bool hbm = MODE == MOVE_HBM;
      uint64_t cap = CAP;
      uint32_t *tkey = hkey, *tpos = hpos;
      double* tval = hval;
      if (act) {
        if constexpr (MODE == MOVE_HYBRID) hbm = 2 * (uint64_t)d > CAP;
        // nextpow2(2 d) slots: of the block's HBM region, or (at least 16)
        // of the group's LDS table
        uint64_t want = hbm ? 2 : 16;
        while (want < 2 * (uint64_t)d) want <<= 1;
        cap = hbm ? want : (want < CAP ? want : CAP);
        if constexpr (EXTRAS) {
          if (a.xids && d > kHubMin) {  // a hub in the sweep: its own table
            uint64_t lo = 0, hi = a.nx;
            while (lo < hi) {
              const uint64_t mid = (lo + hi) / 2;
              if (a.xids[mid] < u) lo = mid + 1;
              else hi = mid;
            }
            if (lo < a.nx && a.xids[lo] == u) {
              const uint64_t base = a.xoff[lo];
              tkey = a.xkey + base;
              tpos = a.xpos + base;
              tval = a.xval + base;
            } else {
              act = false;  // (every live vertex above kHubMin is listed; never taken)
            }
          }
        }
      }
      double be = 0;
      uint32_t bp = kNoPos, bc = 0;
      // Validated moves (id-order kernel, a.validate): a vertex about to
      // move reads its neighbours' communities again; if one moved while the
      // vertex was being decided (a concurrent stream), the decision is
      // taken again from the new view (at most kRedo times).  One stream
      // never sees a change, so its visits are the plain ones.
      constexpr int kRedo = 2;
      for (int attempt = 0;; ++attempt) {
        if (act) {
          if (hbm) table_clear<G>(tkey, tpos, tval, cap, lane);
          else table_clear<G>(lkey, lpos, lval, cap, lane);
        }
        group_sync();
        uint64_t dig = 0;
        if (act) {
          if (hbm) table_scan<G, WT>(a, W, u, o, d, tkey, tpos, tval, cap, lane, dig);
          else table_scan<G, WT>(a, W, u, o, d, lkey, lpos, lval, cap, lane, dig);
        }
        group_sync();
        be = 0;
        bp = kNoPos;
        bc = 0;
        if (act) {
          if constexpr (EXTRAS) Mu = vertex_M(a, u);
          if (hbm) table_best<G>(a, cu, Ku, Mu, tkey, tpos, tval, cap, lane, be, bp, bc, c_cand);
          else table_best<G>(a, cu, Ku, Mu, lkey, lpos, lval, cap, lane, be, bp, bc, c_cand);
        }
        if constexpr (G <= 64) group_best_shfl<G>(be, bp, bc);
        else block_best<BLOCK>(be, bp, bc);
        if constexpr (EXTRAS && G == 64) {
          // (the group is the wave: the redo decision is uniform over it)
          const bool check = a.validate && act && !a.choose_only && attempt < kRedo && bp != kNoPos &&
                             bc != vertex_zero(a, u);
          if (check && __ballot(row_digest<G>(a, u, o, d, lane) != dig)) {
            group_sync();
            continue;
          }
        }
        break;
      }
      if (act) {
        if (a.choose_only) {
          if (lane == 0) { a.out_c[u] = bc; a.out_e[u] = be; }
        } else {
          // a candidate exists (bp set) and it is not community 0 of u's
          // graph (with no candidate bc is 0, as the reference's (0, 0))
          if (bp != kNoPos && bc != (EXTRAS ? vertex_zero(a, u) : zu)) {
            if (lane == 0) {
              unsafeAtomicAdd(&a.ctot[cu], -Ku);
              unsafeAtomicAdd(&a.ctot[bc], Ku);
              st_dev(&a.vcom[u], bc);
              ++c_move;
              c_medge += d;
            }
            flag_neighbours<G, WT>(a, W, u, o, d, lane, cu, bc);
          }
          if (lane == 0) {
            st_dev(&a.vaff[u], (uint8_t)0);
            if (BT && a.gel) {
              const uint32_t k = EARLY ? gu : a.gid[u];
              if (k != g_cur) {
                if (el_acc != 0) unsafeAtomicAdd(&a.gel[g_cur], el_acc);
                g_cur = k;
                el_acc = 0;
              }
            }
            el_acc += be;
            ++c_vert;
            c_edge += d;
          }
        }
      }
      group_sync();
    };
    if constexpr (MODE == MOVE_HYBRID) {
      if (a.sparse) {
        // G-entry windows (an entry per lane of the group): the flags of a
        // window are read at once and only the flagged entries are visited,
        // in order.  Sweeps (sparse 1) read the later flags again after
        // each visit, so a vertex flagged by an earlier one of the window is
        // still visited in this sweep (inc/louvain.hxx:602 then :598); a
        // worklist queue (sparse 2) holds flagged entries whose flags only
        // their own visits clear, and a flag set again meanwhile belongs to
        // the next sub-round or iteration, so its flags are read once.
        const int gsh = (int)(threadIdx.x % 64) / G * G;  // the group's first lane in its wave
        auto group_mask = [&](bool p) __attribute__((always_inline)) {
          const unsigned long long m = __ballot(p);
          if constexpr (G == 64) return m;
          else return (m >> gsh) & ((1ull << G) - 1);
        };
        for (uint64_t j0 = 0; j0 < clen; j0 += G) {
          // (a flagged entry's offset, degree, community and K are loaded
          // with the window; its community changes only when it is visited)
          const uint64_t idx = c0 + (j0 + lane) * step;
          const bool in = ch < nchunks && j0 + lane < clen && idx < n;
          uint32_t ul = 0, dl = 0, cl = 0, gl = 0;
          uint64_t ol = 0;
          double kl = 0;
          bool ok = false, pre = false;
          if (in) {
            ul = a.list ? a.list[idx] : (uint32_t)(a.base + idx);
            dl = a.g.deg[ul];
            if (EARLY && a.gid) gl = a.gid[ul];
            ok = (!a.g.ex || a.g.ex[ul]) && (a.choose_only || !a.fa || a.fa[ul]) && dl <= a.hub_min &&
                 !(EARLY ? a.gid && a.gmask[gl] : BT && vertex_skipped(a, ul));
          }
          bool live = ok && (a.choose_only || ld_dev(&a.vaff[ul]));
          const bool once = a.choose_only || a.sparse == 2;
          unsigned long long mask = group_mask(live);
          while (mask) {
            if (live && !pre) {
              ol = a.g.off[ul];
              cl = ld_dev(&a.vcom[ul]);
              kl = a.vtot[ul];
              pre = true;
            }
            const int b = __ffsll((long long)mask) - 1;
            visit(true, (uint32_t)__shfl(ul, b, G), (uint64_t)__shfl(ol, b, G), (uint32_t)__shfl(dl, b, G),
                  (uint32_t)__shfl(cl, b, G), __shfl(kl, b, G), (uint32_t)__shfl(gl, b, G));
            live = once ? live && lane > b : ok && lane > b && ld_dev(&a.vaff[ul]);
            mask = group_mask(live);
          }
        }
        if (BT && a.gel && el_acc != 0) {
          unsafeAtomicAdd(&a.gel[g_cur], el_acc);
          el_acc = 0;
        }
        continue;
      }
    }
    for (uint64_t j = 0; j < clen; ++j) {
      const uint64_t idx = c0 + j * step;
      bool act = false;
      uint32_t u = 0, gu = 0;
      if (ch < nchunks && idx < n) {
        u = a.list ? a.list[idx] : (uint32_t)(a.base + idx);
        act = (!a.g.ex || a.g.ex[u]) && (a.choose_only || (ld_dev(&a.vaff[u]) && (!a.fa || a.fa[u])));
        if constexpr (MODE == MOVE_HYBRID) act = act && a.g.deg[u] <= a.hub_min;
        if (EARLY && a.gid) gu = a.gid[u];
        act = act && !(EARLY ? a.gid && a.gmask[gu] : BT && vertex_skipped(a, u));
      }
      uint64_t o = 0;
      uint32_t d = 0, cu = 0;
      double Ku = 0;
      if (act) {
        o = a.g.off[u];
        d = a.g.deg[u];
        cu = ld_dev(&a.vcom[u]);
        Ku = a.vtot[u];
      }
      visit(act, u, o, d, cu, Ku, gu);
    }
    if (BT && a.gel && el_acc != 0) {
      unsafeAtomicAdd(&a.gel[g_cur], el_acc);
      el_acc = 0;
    }
  }
  if (!worked) return;  // nothing to add up (a block past a short list)
  if (!a.choose_only) {
    double t = block_sum<BLOCK>(el_acc);
    if (threadIdx.x == 0 && t != 0) unsafeAtomicAdd(a.el + spread_slot() * kElStride, t);
  }
  if (a.cnt) {
    unsigned long long v = block_sum_u64<BLOCK>(c_vert), e = block_sum_u64<BLOCK>(c_edge);
    unsigned long long c = block_sum_u64<BLOCK>(c_cand), m = block_sum_u64<BLOCK>(c_move);
    unsigned long long me = block_sum_u64<BLOCK>(c_medge);
    if (threadIdx.x == 0) add_counts(a.cnt, v, e, c, m, me);
  }
}

template <int BLOCK, int G, int CAP, class WT, int MODE_>
__global__ __launch_bounds__(BLOCK) void move_kernel(MoveArgs a) {
  move_body<BLOCK, G, CAP, WT, MODE_>(a, blockIdx.x, gridDim.x);
}

// The batched form's id-order kernel, held to three waves a SIMD (its
// per-graph lookups would otherwise take it past 168 registers, to two).
template <int BLOCK, int G, int CAP, class WT, int MODE_>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void move_bt_kernel(MoveArgs a) {
  move_body<BLOCK, G, CAP, WT, MODE_>(a, blockIdx.x, gridDim.x);
}

// The id-order kernel of one graph (a.gid null).
template <int BLOCK, int G, int CAP, class WT, int MODE_>
__global__ __launch_bounds__(BLOCK) void move_one_kernel(MoveArgs a) {
  move_body<BLOCK, G, CAP, WT, MODE_, false>(a, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------- hub batches
// Steps of louvainMoveOmpW's body (inc/louvain.hxx:599-605) for a batch of
// high-degree vertices (hubs) processed together: each with its own HBM table
// of nextpow2(2 deg) slots.  Its edges are cut into kHubChunk-edge chunks
// spread over the grid; each chunk is summed per community in an LDS table
// first, so a community that many edges share costs one HBM atomic per chunk.
// The best community is the largest positive gain, ties to the earliest
// first position (inc/louvain.hxx:461): one pass over the tables reduces
// every kTabChunk-slot chunk (and empties its slots, so the slab is clean
// for the next batch), the apply pass reduces a hub's chunk results; the
// chosen id is the community of the neighbour at that position.
constexpr uint64_t kHubChunk = kHubChunkEdges;  // edges per insert / flag chunk
static_assert(kHubChunk % kBlock == 0, "hub_insert_kernel: whole edges per thread");
constexpr uint32_t kHubLdsTab = 2 * kHubChunk;  // LDS slots for one chunk's communities
constexpr uint64_t kTabChunk = kHubTabChunk;  // table slots per best-pass chunk

// Block-wide exclusive scan of one value per thread; `total` gets the sum.
template <int BLOCK>
__device__ uint64_t block_scan_excl(uint64_t x, uint64_t& total) {
  __shared__ uint64_t s[BLOCK / 64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  uint64_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) s[w] = inc;
  __syncthreads();
  uint64_t before = 0;
  total = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / 64; ++i) {
    if (i < w) before += s[i];
    total += s[i];
  }
  __syncthreads();
  return before + inc - x;
}

// Insert into an HBM table whose keys, once set, never change: a key that
// is already there is found without a compare-and-swap.
__device__ __forceinline__ void tab_merge(uint32_t* key, uint32_t* pos, double* val, uint64_t cap, uint32_t c,
                                          double w, uint32_t p) {
  uint32_t s = slot_of(c, cap);
  while (true) {
    uint32_t k = __hip_atomic_load(&key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == kEmpty) k = atomicCAS(&key[s], kEmpty, c);
    if (k == kEmpty || k == c) {
      unsafeAtomicAdd(&val[s], w);
      if (pos) atomicMin(&pos[s], p);
      return;
    }
    s = (uint32_t)((s + 1) & (cap - 1));
  }
}

// tab_merge that probes at most cap slots: false if c found no slot (the
// table is full of other communities).
__device__ __forceinline__ bool tab_merge_bounded(uint32_t* key, uint32_t* pos, double* val, uint64_t cap,
                                                  uint32_t c, double w, uint32_t p) {
  uint32_t s = slot_of(c, cap);
  for (uint64_t probe = 0; probe < cap; ++probe) {
    uint32_t k = __hip_atomic_load(&key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == kEmpty) k = atomicCAS(&key[s], kEmpty, c);
    if (k == kEmpty || k == c) {
      unsafeAtomicAdd(&val[s], w);
      if (pos) atomicMin(&pos[s], p);
      return true;
    }
    s = (uint32_t)((s + 1) & (cap - 1));
  }
  return false;
}

__device__ __forceinline__ uint64_t list_len(const MoveArgs& a) { return a.n_dev ? (uint64_t)*a.n_dev : a.n; }

// One block: per listed vertex, active or not, its table slots
// (nextpow2(2 deg), at least kTabChunk) and edge chunks, scanned into
// hb_cap / hb_chunks (n + 1 entries, batch totals in hb_tot), and the owner
// of every edge chunk (hb_cown) and table chunk (hb_town); an active hub's
// row, community, that community's Sigma_tot and a zero own-community
// weight (summed by the insert pass); every hub starts as "did not move".
// The list is walked in tiles of kPrepBlock x kPrepItems entries, each thread
// taking kPrepItems consecutive ones; all of a tile's loads are issued
// before any store, so they overlap.
constexpr int kPrepBlock = 512, kPrepItems = 4;

__global__ __launch_bounds__(kPrepBlock) void hub_prep_kernel(MoveArgs a) {
  const uint64_t n = list_len(a);
  constexpr uint64_t tile = (uint64_t)kPrepBlock * kPrepItems;
  uint64_t run_cap = 0, run_chk = 0;
  for (uint64_t b = 0; b < n; b += tile) {  // block-uniform trip count
    const uint64_t k0 = b + (uint64_t)threadIdx.x * kPrepItems;
    uint32_t u[kPrepItems], d[kPrepItems], cu[kPrepItems];
    bool act[kPrepItems], full[kPrepItems];
    uint64_t o[kPrepItems];
    double sd[kPrepItems];
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) u[j] = k0 + j < n ? a.list[k0 + j] : 0;
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) {
      act[j] = full[j] = false;
      d[j] = 0;
      if (k0 + j >= n) continue;
      d[j] = a.g.deg[u[j]];
      act[j] = d[j] > a.hub_min && (!a.g.ex || a.g.ex[u[j]]) && !vertex_skipped(a, u[j]) &&
               (a.choose_only || (ld_dev(&a.vaff[u[j]]) && (!a.fa || a.fa[u[j]])));
      full[j] = !a.hb_over || a.hb_big[u[j]];
    }
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) {
      cu[j] = act[j] ? ld_dev(&a.vcom[u[j]]) : 0;
      o[j] = act[j] ? a.g.off[u[j]] : 0;
    }
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) sd[j] = act[j] ? ld_dev(&a.ctot[cu[j]]) : 0;
    uint64_t cap[kPrepItems], chk[kPrepItems], tc = 0, tk = 0;
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) {
      cap[j] = chk[j] = 0;
      if (!act[j]) continue;
      uint64_t c = kTabChunk;
      while (c < 2 * (uint64_t)d[j]) c <<= 1;
      // worklist batches: a hub's neighbours fall in few communities, so
      // its table starts at kHubTabCap slots; one that overflowed keeps
      // the full nextpow2(2 deg) from then on
      if (!full[j] && c > kHubTabCap) c = kHubTabCap;
      cap[j] = c;
      chk[j] = (d[j] + kHubChunk - 1) / kHubChunk;
      tc += cap[j];
      tk += chk[j];
    }
    uint64_t totc, totk;
    uint64_t pc = run_cap + block_scan_excl<kPrepBlock>(tc, totc);
    uint64_t pk = run_chk + block_scan_excl<kPrepBlock>(tk, totk);
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) {
      const uint64_t k = k0 + j;
      if (k >= n) continue;
      a.hb_bpos[k] = kNoPos;
      if (a.hb_over) a.hb_over[k] = 0;
      a.hb_cap[k] = pc;
      a.hb_chunks[k] = pk;
      if (act[j]) {
        if (a.stamp) st_dev(&a.stamp[u[j]], cur_serial(a));
        a.hb_cu[k] = cu[j];
        a.hb_sd[k] = sd[j];
        a.hb_vd[k] = 0;
        a.hb_off[k] = o[j];
        a.hb_deg[k] = d[j];
        for (uint64_t q = 0; q < chk[j]; ++q) a.hb_cown[pk + q] = (uint32_t)k;
        for (uint64_t t = 0; t < cap[j] / kTabChunk; ++t) a.hb_town[pc / kTabChunk + t] = (uint32_t)k;
      }
      pc += cap[j];
      pk += chk[j];
    }
    run_cap += totc;
    run_chk += totk;
  }
  if (threadIdx.x == 0) {
    a.hb_tot[0] = run_cap;
    a.hb_tot[1] = run_chk;
    a.hb_cap[n] = run_cap;
    a.hb_chunks[n] = run_chk;
  }
}

__device__ __forceinline__ uint64_t find_owner(const uint64_t* scan, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n - 1;  // largest k with scan[k] <= x
  while (lo < hi) {
    uint64_t mid = (lo + hi + 1) >> 1;
    if (scan[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class WT>
__global__ __launch_bounds__(kBlock) void hub_insert_kernel(MoveArgs a) {
  __shared__ uint32_t s_key[kHubLdsTab], s_pos[kHubLdsTab];
  __shared__ double s_val[kHubLdsTab];
  const uint64_t nchunks = a.hb_tot[1];
  const WT* W = (const WT*)a.g.w;
  for (uint64_t q = blockIdx.x; q < nchunks; q += gridDim.x) {
    for (uint32_t s = threadIdx.x; s < kHubLdsTab; s += kBlock) {
      s_key[s] = kEmpty;
      s_pos[s] = kNoPos;
      s_val[s] = 0;
    }
    __syncthreads();
    const uint32_t k = a.hb_cown[q];
    const uint32_t u = a.list[k];
    const uint64_t o = a.hb_off[k], d = a.hb_deg[k];
    uint64_t i0 = (q - a.hb_chunks[k]) * kHubChunk, i1 = min(d, i0 + kHubChunk);
    // the chunk's kHubChunk / kBlock edges per thread: every key and weight
    // load first, then every community gather, then the inserts (two
    // dependent memory round trips per chunk, not two per edge)
    constexpr int R = kHubChunk / kBlock;
    uint32_t v[R], c[R];
    double w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = i0 + (uint64_t)r * kBlock + threadIdx.x;
      v[r] = i < i1 ? a.g.keys[o + i] : u;
      w[r] = i < i1 ? (double)W[o + i] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) c[r] = v[r] != u ? ld_dev(&a.vcom[v[r]]) : 0;
#pragma unroll
    for (int r = 0; r < R; ++r)  // self-loops are skipped
      tab_add_wave(s_key, s_pos, s_val, kHubLdsTab, v[r] != u, c[r], w[r],
                   (uint32_t)(i0 + (uint64_t)r * kBlock + threadIdx.x));
    __syncthreads();
    uint64_t base = a.hb_cap[k], cap = a.hb_cap[k + 1] - base;
    const uint32_t cu = a.hb_cu[k];
    for (uint32_t s = threadIdx.x; s < kHubLdsTab; s += kBlock) {
      uint32_t c = s_key[s];
      if (c == kEmpty) continue;
      // a capped table that is full: the hub is redone with a full one (a
      // full table, nextpow2(2 deg) slots, always has room)
      if (!tab_merge_bounded(a.hb_key + base, a.hb_pos + base, a.hb_val + base, cap, c, s_val[s], s_pos[s]) &&
          a.hb_over)
        a.hb_over[k] = 1;
      if (c == cu) unsafeAtomicAdd(&a.hb_vd[k], s_val[s]);  // weight to the hub's own community
    }
    __syncthreads();
  }
}

// Best (gain, first position) of each kTabChunk-slot chunk of the tables;
// the slots are emptied as they are read.
__global__ __launch_bounds__(kBlock) void hub_best_kernel(MoveArgs a) {
  const uint64_t nt = a.hb_tot[0] / kTabChunk;
  for (uint64_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const uint64_t s0 = t * kTabChunk;
    const uint32_t k = a.hb_town[t];
    const uint32_t cu = a.hb_cu[k], u = a.list[k];
    const double vd = a.hb_vd[k], sd = a.hb_sd[k], Ku = a.vtot[u], Mu = vertex_M(a, u);
    double be = 0;
    uint32_t bp = kNoPos, bc = 0;
    for (uint64_t s = s0 + threadIdx.x; s < s0 + kTabChunk; s += kBlock) {
      uint32_t c = a.hb_key[s];
      if (c == kEmpty) continue;
      uint32_t p = a.hb_pos[s];
      double w = a.hb_val[s];
      a.hb_key[s] = kEmpty;
      a.hb_pos[s] = kNoPos;
      a.hb_val[s] = 0;
      if (c == cu || (a.pick_less && c > cu)) continue;
      double e = delta_q(w, vd, Ku, ld_dev(&a.ctot[c]), sd, Mu, a.R);
      if (e > 0 && better(e, p, be, bp)) {
        be = e;
        bp = p;
      }
    }
    block_best<kBlock>(be, bp, bc);
    if (threadIdx.x == 0) {
      a.hb_ce[t] = be;
      a.hb_cp[t] = bp;
    }
  }
}

// Per hub (one wave each): best over its chunks, then the move (or the
// choice, in choose-only mode) and its gain.
__global__ __launch_bounds__(kBlock) void hub_apply_kernel(MoveArgs a) {
  const uint64_t n = list_len(a);
  const int lane = threadIdx.x % 64;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  double el = 0;
  unsigned long long cv = 0, ce = 0, cm = 0, cme = 0;
  for (uint64_t k = blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64; k < n; k += nw) {
    const uint64_t c0 = a.hb_cap[k], c1 = a.hb_cap[k + 1];
    if (c0 == c1) continue;  // wave-uniform
    double be = 0;
    uint32_t bp = kNoPos, bcx = 0;
    for (uint64_t t = c0 / kTabChunk + lane; t < c1 / kTabChunk; t += 64) {
      double e = a.hb_ce[t];
      uint32_t p = a.hb_cp[t];
      if (p != kNoPos && better(e, p, be, bp)) {
        be = e;
        bp = p;
      }
    }
    group_best_shfl<64>(be, bp, bcx);
    if (lane != 0) continue;
    uint32_t u = a.list[k];
    if (a.hb_over && a.hb_over[k]) {
      // its capped table overflowed: no decision; it stays flagged, gets a
      // full table and joins the next sub-round
      a.hb_big[u] = 1;
      mark_same(a, u);
      continue;
    }
    double e = bp == kNoPos ? 0.0 : be;
    uint32_t bc = 0;
    if (bp != kNoPos) bc = ld_dev(&a.vcom[a.g.keys[a.g.off[u] + bp]]);
    uint32_t d = a.g.deg[u];
    if (a.choose_only) {
      a.out_c[u] = bc;
      a.out_e[u] = e;
      continue;
    }
    if (bp != kNoPos && bc != vertex_zero(a, u)) {
      uint32_t cu = a.hb_cu[k];
      double Ku = a.vtot[u];
      unsafeAtomicAdd(&a.ctot[cu], -Ku);
      unsafeAtomicAdd(&a.ctot[bc], Ku);
      st_dev(&a.vcom[u], bc);
      a.hb_bpos[k] = bp;  // tells the flag pass this hub moved
      ++cm;
      cme += d;
    }
    st_dev(&a.vaff[u], (uint8_t)0);
    if (a.gel) unsafeAtomicAdd(&a.gel[a.gid[u]], e);
    else el += e;
    ++cv;
    ce += d;
  }
  if (a.choose_only) return;
  double t = block_sum<kBlock>(el);
  if (threadIdx.x == 0 && t != 0) unsafeAtomicAdd(a.el + spread_slot() * kElStride, t);
  if (a.cnt) {
    unsigned long long v = block_sum_u64<kBlock>(cv), e2 = block_sum_u64<kBlock>(ce);
    unsigned long long m = block_sum_u64<kBlock>(cm), me = block_sum_u64<kBlock>(cme);
    if (threadIdx.x == 0) add_counts(a.cnt, v, e2, 0, m, me);
  }
}

// Neighbours of the hubs that moved are flagged (chunks spread over the grid).
__global__ __launch_bounds__(kBlock) void hub_flag_kernel(MoveArgs a) {
  const uint64_t nchunks = a.hb_tot[1];
  for (uint64_t q = blockIdx.x; q < nchunks; q += gridDim.x) {
    const uint32_t k = a.hb_cown[q];
    if (a.hb_bpos[k] == kNoPos) continue;  // uniform per block
    const uint32_t u = a.list[k];
    const uint64_t o = a.hb_off[k], d = a.hb_deg[k];
    // the hub moved from hb_cu[k] to the community the apply pass stored
    const uint32_t from = a.hb_cu[k], to = a.ht_cap ? ld_dev(&a.vcom[u]) : 0;
    uint64_t i0 = (q - a.hb_chunks[k]) * kHubChunk, i1 = min(d, i0 + kHubChunk);
    for (uint64_t b = i0; b < i1; b += kBlock) {
      uint64_t i = b + threadIdx.x;
      bool on = i < i1;
      uint32_t v = on ? a.g.keys[o + i] : u;
      if (on && v != u) {
        flag_vertex(a.vaff, v, u, a);
        if (a.ht_cap) ht_move(a, v, a.f64_weights ? ((const double*)a.g.w)[o + i] : (double)((const float*)a.g.w)[o + i],
                              from, to);
      }
    }
  }
}

// louvainMoveOmpW's body (inc/louvain.hxx:595-605) for worklist hubs of
// moderate degree (MIN < deg <= MAXD): one block of BLOCK threads per hub,
// its neighbour-community table in LDS.  Block b takes queue entries b,
// b + grid, b + 2 grid, ... of a.list (length *a.n_dev), reading the row,
// community and K of BLOCK of them at once ahead of their turn (they change
// only when the hub itself moves).  Per hub: sum edge weights per neighbour
// community (self-loops skipped, first position kept for the first-seen
// tie-break), pick the best positive gain, and unless it is community 0
// move there (atomic Sigma_tot) and flag the neighbours; clear the hub's
// flag and add its gain to el.
//
// The table has nextpow2(2 deg) slots up to SLOTS.  A hub of more than
// SLOTS / 2 edges may have more distinct neighbour communities than that:
// its distinct count is kept, and past SLOTS - BLOCK (the table then still
// has a free slot for every thread) the hub is left flagged, marked for a
// full table (hb_big) and sent to the next sub-round, where it runs in the
// multi-block batch.

// tab_add_wave with a count of the slots it creates.
__device__ __forceinline__ void tab_add_wave_counted(uint32_t* key, uint32_t* pos, double* val, uint32_t cap,
                                                     bool ok, uint32_t c, double w, uint32_t p, uint32_t* made) {
  const unsigned long long act = __ballot(ok);
  if (!act) return;
  const int lane = threadIdx.x % 64;
  const int lead = __ffsll((long long)act) - 1;
  const uint32_t c0 = __shfl(c, lead, 64);
  const bool same = ok && c == c0;
  double sum = same ? w : 0.0;
  for (int m = 32; m > 0; m >>= 1) sum += __shfl_xor(sum, m, 64);
  if (lane == lead) {
    c = c0;
    w = sum;
  } else if (!ok || same) {
    return;
  }
  uint32_t s = slot_of(c, cap);
  while (true) {
    const uint32_t k = atomicCAS(&key[s], kEmpty, c);
    if (k == kEmpty || k == c) {
      if (k == kEmpty) atomicAdd(made, 1u);
      unsafeAtomicAdd(&val[s], w);
      atomicMin(&pos[s], p);
      return;
    }
    s = (s + 1) & (cap - 1);
  }
}

template <int BLOCK, uint32_t SLOTS, uint32_t MAXD, class WT>
__global__ __launch_bounds__(BLOCK) void hub_mid_kernel(MoveArgs a) {
  constexpr uint32_t kDistinct = SLOTS - BLOCK;
  constexpr int NB = kScanBatch;  // edges per thread loaded at a time
  constexpr int AHEAD = 64;       // hubs read ahead at a time
  static_assert(SLOTS > BLOCK, "room for a slot per thread past the distinct limit");
  static_assert(AHEAD <= BLOCK, "one thread per hub read ahead");
  __shared__ uint32_t s_key[SLOTS], s_pos[SLOTS];
  __shared__ double s_val[SLOTS];
  __shared__ uint32_t p_u[AHEAD], p_d[AHEAD], p_cu[AHEAD];
  __shared__ uint64_t p_o[AHEAD];
  __shared__ double p_ku[AHEAD];
  __shared__ uint32_t s_made;
  const uint64_t n = list_len(a);
  const uint64_t G = gridDim.x;
  const WT* W = (const WT*)a.g.w;
  double el_acc = 0;
  unsigned long long c_vert = 0, c_edge = 0, c_cand = 0, c_move = 0, c_medge = 0;
  for (uint64_t base = blockIdx.x; base < n; base += G * AHEAD) {
    if (threadIdx.x < AHEAD) {  // thread j reads hub base + j G (degree 0: nothing to do)
      const uint64_t k = base + threadIdx.x * G;
      uint32_t u = 0, d = 0;
      if (k < n) {
        u = a.list[k];
        d = a.g.deg[u];
        const bool act = d > a.hub_min && d <= MAXD && (!a.g.ex || a.g.ex[u]) && !vertex_skipped(a, u) &&
                         (a.choose_only || (ld_dev(&a.vaff[u]) && (!a.fa || a.fa[u])));
        if (!act) d = 0;
      }
      p_u[threadIdx.x] = u;
      p_d[threadIdx.x] = d;
      if (d) {
        p_cu[threadIdx.x] = ld_dev(&a.vcom[u]);
        p_o[threadIdx.x] = a.g.off[u];
        p_ku[threadIdx.x] = a.vtot[u];
      }
    }
    __syncthreads();
    const uint64_t J = min((uint64_t)AHEAD, (n - base + G - 1) / G);
    for (uint64_t j = 0; j < J; ++j) {
      const uint32_t d = p_d[j];
      if (!d) continue;  // block-uniform
      const uint32_t u = p_u[j], cu = p_cu[j];
      const uint64_t o = p_o[j];
      const double Ku = p_ku[j];
      if (threadIdx.x == 0) {
        if (a.stamp) st_dev(&a.stamp[u], cur_serial(a));
        s_made = 0;
      }
      uint32_t cap = 16;
      while (cap < 2 * d && cap < SLOTS) cap <<= 1;
      const bool capped = cap < 2 * d;  // more edges than the table has room for as distinct keys
      for (uint32_t s = threadIdx.x; s < cap; s += BLOCK) {
        s_key[s] = kEmpty;
        s_pos[s] = kNoPos;
        s_val[s] = 0;
      }
      __syncthreads();
      // NB rounds of BLOCK edges at a time: every key and weight load, then
      // every community gather, then the inserts (a thread makes at most one
      // new slot per round after reading the distinct count, so a capped
      // table never fills).  The last batch's keys stay for the flags.
      bool over = false;
      uint32_t v[NB];
      double w[NB];
      for (uint32_t b0 = 0; b0 < d && !over; b0 += NB * BLOCK) {  // block-uniform
        uint32_t c[NB];
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          const uint32_t i = b0 + r * BLOCK + threadIdx.x;
          v[r] = i < d ? a.g.keys[o + i] : u;
          w[r] = i < d ? (double)W[o + i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) c[r] = v[r] != u ? ld_dev(&a.vcom[v[r]]) : 0;
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          if (b0 + r * BLOCK < d) {  // block-uniform
            if (capped && __hip_atomic_load(&s_made, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= kDistinct)
              over = true;
            // self-loops are skipped (louvainScanCommunityW<SELF=false>)
            tab_add_wave_counted(s_key, s_pos, s_val, cap, !over && v[r] != u, c[r], w[r],
                                 b0 + r * BLOCK + threadIdx.x, &s_made);
          }
        }
        over = __syncthreads_or(over);
      }
      if (over) {
        // too many distinct communities for the table: no decision here
        if (threadIdx.x == 0) {
          a.hb_big[u] = 1;
          mark_same(a, u);
        }
        __syncthreads();
        continue;
      }
      const double vd = tab_get(s_key, s_val, cap, cu), Sd = ld_dev(&a.ctot[cu]), Mu = vertex_M(a, u);
      double be = 0;
      uint32_t bp = kNoPos, bc = 0;
      for (uint32_t s = threadIdx.x; s < cap; s += BLOCK) {
        const uint32_t c = s_key[s];
        if (c == kEmpty) continue;
        ++c_cand;
        if (c == cu || (a.pick_less && c > cu)) continue;
        const double e = delta_q(s_val[s], vd, Ku, ld_dev(&a.ctot[c]), Sd, Mu, a.R);
        const uint32_t p = s_pos[s];
        if (e > 0 && better(e, p, be, bp)) {
          be = e;
          bp = p;
          bc = c;
        }
      }
      block_best<BLOCK>(be, bp, bc);
      if (a.choose_only) {
        if (threadIdx.x == 0) {
          a.out_c[u] = bc;
          a.out_e[u] = be;
        }
      } else {
        // a candidate exists and it is not community 0 of u's graph
        const bool moves = bp != kNoPos && bc != vertex_zero(a, u);
        if (moves) {
          if (threadIdx.x == 0) {
            unsafeAtomicAdd(&a.ctot[cu], -Ku);
            unsafeAtomicAdd(&a.ctot[bc], Ku);
            st_dev(&a.vcom[u], bc);
            ++c_move;
            c_medge += d;
          }
          if (d <= NB * BLOCK) {  // one batch: its keys and weights are still here
#pragma unroll
            for (int r = 0; r < NB; ++r)
              if (r * BLOCK + threadIdx.x < d && v[r] != u) {
                flag_vertex(a.vaff, v[r], u, a);
                ht_move(a, v[r], w[r], cu, bc);
              }
          } else {
            flag_neighbours<BLOCK, WT>(a, W, u, o, d, threadIdx.x, cu, bc);
          }
        }
        if (threadIdx.x == 0) {
          st_dev(&a.vaff[u], (uint8_t)0);
          if (a.gel) unsafeAtomicAdd(&a.gel[a.gid[u]], be);
          else el_acc += be;
          ++c_vert;
          c_edge += d;
        }
      }
      __syncthreads();  // the table is reused
    }
    __syncthreads();  // the read-ahead arrays are reused
  }
  if (a.choose_only) return;
  const double t = block_sum<BLOCK>(el_acc);
  if (threadIdx.x == 0 && t != 0) unsafeAtomicAdd(a.el + spread_slot() * kElStride, t);
  if (a.cnt) {
    unsigned long long v = block_sum_u64<BLOCK>(c_vert), e = block_sum_u64<BLOCK>(c_edge);
    unsigned long long c = block_sum_u64<BLOCK>(c_cand), m = block_sum_u64<BLOCK>(c_move);
    unsigned long long me = block_sum_u64<BLOCK>(c_medge);
    if (threadIdx.x == 0) add_counts(a.cnt, v, e, c, m, me);
  }
}

// louvainMoveOmpW's body (inc/louvain.hxx:595-605) for hubs with a
// neighbour-community table, one wave per hub from a.list (length
// *a.n_dev): the candidates and their weights come from the table (entries
// with a positive count), the gain is deltaModularity as for every vertex,
// and a tie for the best gain goes to the tied community that appears first
// in the hub's row (inc/louvain.hxx:461), found by walking the row.  A hub
// whose table was dropped since the queue was built joins the next
// sub-round (it is then queued by degree).
// (as block `bid` of `nblk`: a launch of its own, or the last blocks of the
// fused worklist launch)
template <class WT>
__device__ __forceinline__ void hub_tab_body(const MoveArgs& a, uint32_t bid, uint32_t nblk,
                                             uint32_t& mv_made) {
  constexpr int R = kHtMaxCap / 64;  // table slots per lane
  const int lane = threadIdx.x % 64;
  const uint64_t n = list_len(a);
  const uint64_t nw = (uint64_t)nblk * (kBlock / 64);
  const WT* W = (const WT*)a.g.w;
  if ((uint64_t)bid * (kBlock / 64) >= n) return;  // (block-uniform: no hub for this block)
  double el_acc = 0;
  unsigned long long c_vert = 0, c_edge = 0, c_cand = 0, c_move = 0, c_medge = 0;
  for (uint64_t k = (uint64_t)bid * (kBlock / 64) + threadIdx.x / 64; k < n; k += nw) {  // wave-uniform
    // the hub's scalars, read by lane 0 (its table may be dropped meanwhile)
    uint32_t u = 0, d = 0, cap = 0, act = 0;
    if (lane == 0) {
      u = a.list[k];
      d = a.g.deg[u];
      act = d > a.hub_min && (!a.g.ex || a.g.ex[u]) && !vertex_skipped(a, u) && ld_dev(&a.vaff[u]) &&
            (!a.fa || a.fa[u]);
      cap = act ? __hip_atomic_load(&a.ht_cap[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
      if (act && cap <= kHtDropped) mark_same(a, u);  // no table any more: evaluated from its edges next
    }
    u = __shfl(u, 0, 64);
    d = __shfl(d, 0, 64);
    cap = __shfl(cap, 0, 64);
    if (cap <= kHtDropped) continue;
    if (lane == 0 && a.stamp) st_dev(&a.stamp[u], cur_serial(a));
    const uint32_t cu = ld_dev(&a.vcom[u]);
    const uint64_t o = a.g.off[u], off = a.ht_off[u];
    const double Ku = a.vtot[u], Sd = ld_dev(&a.ctot[cu]), Mu = vertex_M(a, u);
    uint32_t ck[R];
    int32_t cn[R];
    double cw[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t sl = r * 64 + lane;
      ck[r] = sl < cap ? ld_dev(&a.ht_key[off + sl]) : kEmpty;
      cn[r] = sl < cap ? ld_dev(&a.ht_cnt[off + sl]) : 0;
      cw[r] = sl < cap ? ld_dev(&a.ht_w[off + sl]) : 0.0;
    }
    double vd = 0;  // weight to the hub's own community
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (ck[r] == cu && cn[r] > 0) vd = cw[r];
    for (int m = 32; m > 0; m >>= 1) vd += __shfl_xor(vd, m, 64);
    double e[R], be = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      e[r] = 0;
      if (ck[r] == kEmpty || cn[r] <= 0) continue;
      ++c_cand;
      if (ck[r] == cu || (a.pick_less && ck[r] > cu)) continue;
      e[r] = delta_q(cw[r], vd, Ku, ld_dev(&a.ctot[ck[r]]), Sd, Mu, a.R);
      if (e[r] > be) be = e[r];
    }
    for (int m = 32; m > 0; m >>= 1) be = fmax(be, __shfl_xor(be, m, 64));
    uint32_t bc = 0;
    if (be > 0) {
      // the communities with the best gain; one: it; several: the first in the row
      int ties = 0;
      uint32_t mine = kEmpty;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (e[r] == be) {
          ++ties;
          mine = ck[r];
        }
      for (int m = 32; m > 0; m >>= 1) ties += __shfl_xor(ties, m, 64);
      const unsigned long long hold = __ballot(mine != kEmpty);
      bc = __shfl(mine, __ffsll((long long)hold) - 1, 64);
      if (ties > 1) {
        bool found = false;
        for (uint32_t b0 = 0; b0 < d && !found; b0 += 64) {  // wave-uniform
          const uint32_t i = b0 + lane;
          const uint32_t v = i < d ? a.g.keys[o + i] : u;
          bool tied = false;
          uint32_t c = 0;
          if (v != u) {
            c = ld_dev(&a.vcom[v]);
            if (c != cu && !(a.pick_less && c > cu)) {
              const int64_t sl = ht_slot(a.ht_key, off, cap, c, false);
              tied = sl >= 0 && ld_dev(&a.ht_cnt[sl]) > 0 &&
                     delta_q(ld_dev(&a.ht_w[sl]), vd, Ku, ld_dev(&a.ctot[c]), Sd, Mu, a.R) == be;
            }
          }
          const unsigned long long tm = __ballot(tied);
          if (tm) {
            bc = __shfl(c, __ffsll((long long)tm) - 1, 64);
            found = true;
          }
        }
      }
    }
    // a candidate exists and it is not community 0 of u's graph
    if (be > 0 && bc != vertex_zero(a, u)) {
      uint32_t slot = kEmpty, pb = 0;
      if (lane == 0) {
        unsafeAtomicAdd(&a.ctot[cu], -Ku);
        unsafeAtomicAdd(&a.ctot[bc], Ku);
        st_dev(&a.vcom[u], bc);
        ++c_move;
        c_medge += d;
        // its neighbours are flagged by the grid-wide pass after this kernel
        if (a.mv_n) {
          slot = atomicAdd(a.mv_n, 1u);
          if (slot < a.mv_cap) {
            a.mv_u[slot] = u;
            st_dev(&a.mv_u[slot], u);
            st_dev(&a.mv_from[slot], cu);
            st_dev(&a.mv_to[slot], bc);
            st_dev(&a.mv_d[slot], d);
            st_dev(&a.mv_o[slot], o);
            if (a.mv_piece) {
              __builtin_amdgcn_s_waitcnt(0);  // (the entry is complete before its pieces exist)
              pb = atomicAdd(a.mv_n + 1, (d + kMvChunk - 1) / kMvChunk);
            }
          }
        }
      }
      slot = __shfl(slot, 0, 64);
      pb = __shfl(pb, 0, 64);
      // its row's pieces, for the blocks of this launch to take; with no
      // room for all of them, none (those below the cap are written empty)
      // and the wave flags the row itself
      const uint32_t np = (d + kMvChunk - 1) / kMvChunk;
      const bool listed = a.mv_n && slot < a.mv_cap;
      const bool own = !listed || (a.mv_piece && (uint64_t)pb + np > a.mv_pcap);
      if (listed && a.mv_piece) {
        for (uint32_t j = lane; j < np && pb + j < a.mv_pcap; j += 64)
          st_dev(&a.mv_piece[pb + j], own ? kMvSkip : (slot | j << 12));
        if (lane == 0 && !own) mv_made = 1;  // this block takes pieces before it ends
      }
      if (own) flag_neighbours<64, WT>(a, W, u, o, d, lane, cu, bc);
    }
    if (lane == 0) {
      st_dev(&a.vaff[u], (uint8_t)0);
      if (a.gel) unsafeAtomicAdd(&a.gel[a.gid[u]], be);
      else el_acc += be;
      ++c_vert;
      c_edge += cap;  // (its "edges": the table slots read, 16 B each)
    }
  }
  const double t = block_sum<kBlock>(el_acc);
  if (threadIdx.x == 0 && t != 0) unsafeAtomicAdd(a.el + spread_slot() * kElStride, t);
  if (a.cnt) {
    unsigned long long v = block_sum_u64<kBlock>(c_vert), ed = block_sum_u64<kBlock>(c_edge);
    unsigned long long c = block_sum_u64<kBlock>(c_cand), m = block_sum_u64<kBlock>(c_move);
    unsigned long long me = block_sum_u64<kBlock>(c_medge);
    if (threadIdx.x == 0) add_counts(a.cnt, v, ed, c, m, me);
  }
}

template <class WT>
__global__ __launch_bounds__(kBlock) void hub_tab_kernel(MoveArgs a) {
  __shared__ uint32_t s_made;
  hub_tab_body<WT>(a, blockIdx.x, gridDim.x, s_made);
}

// The pieces of the moved hubs' rows (h.mv_*), taken one at a time by the
// block until none is reserved: every block of the fused launch whose waves
// reserved pieces calls it when its own work is done, so every piece is
// taken before the launch ends (by its own block, or another that was
// draining then).  A taken piece's descriptor is cleared for the next launch.
// (Counters and descriptors are device-coherent atomics; a descriptor is
// written right after its reservation by a running wave, and the wait for
// it is bounded all the same.)
template <class WT>
__device__ __forceinline__ void drain_moved(const MoveArgs& h) {
  __shared__ uint32_t s_k, s_u, s_d, s_from, s_to;
  __shared__ uint64_t s_o;
  const WT* W = (const WT*)h.g.w;
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t k = kMvPending, t = atomicAdd(h.mv_n + 2, 0u);
      for (;;) {
        const uint32_t n = min(atomicAdd(h.mv_n + 1, 0u), h.mv_pcap);
        if (t >= n) break;
        const uint32_t old = atomicCAS(h.mv_n + 2, t, t + 1);
        if (old == t) {
          k = t;
          break;
        }
        t = old;
      }
      uint32_t dsc = kMvPending;
      if (k != kMvPending) {
        uint32_t spin = 0;
        while ((dsc = ld_dev(&h.mv_piece[k])) == kMvPending && ++spin < (1u << 20)) __builtin_amdgcn_s_sleep(1);
        st_dev(&h.mv_piece[k], kMvPending);
        if (dsc != kMvPending && dsc != kMvSkip) {
          const uint32_t m = dsc & 4095u;
          s_u = ld_dev(&h.mv_u[m]);
          s_d = ld_dev(&h.mv_d[m]);
          s_o = ld_dev(&h.mv_o[m]) + (uint64_t)(dsc >> 12) * kMvChunk;
          s_from = ld_dev(&h.mv_from[m]);
          s_to = ld_dev(&h.mv_to[m]);
          const uint32_t e = (dsc >> 12) * kMvChunk;
          s_d = e < s_d ? min(s_d - e, kMvChunk) : 0u;  // (edges of this piece)
        } else {
          s_d = 0;
        }
      }
      s_k = k;
    }
    __syncthreads();
    const uint32_t k = s_k, u = s_u, d = s_d, from = s_from, to = s_to;
    const uint64_t o = s_o;
    __syncthreads();
    if (k == kMvPending) break;
    if (d) flag_neighbours<kBlock, WT>(h, W, u, o, d, threadIdx.x, from, to);
  }
}

// One worklist sub-round in one launch when every hub has a table: blocks
// [0, nlo) run the id-order kernel over queue 0 (a, four one-wave streams a
// block), the others evaluate the tabled hubs of queue 4 (h), side by side
// as the reference's threads visit hubs and other vertices at once; a block
// whose hub waves moved hubs then flags the neighbours of moved hubs, a piece
// of a row at a time, until no piece is left.
template <class WT>
__global__ __launch_bounds__(kBlock) void wl_fused_kernel(MoveArgs a, MoveArgs h, uint32_t nlo, WlQueues qs,
                                                          const uint32_t* lens) {
  __shared__ uint32_t s_made;  // a wave of this block reserved moved-hub pieces
  if (threadIdx.x == 0) s_made = 0;
  __syncthreads();
  if (blockIdx.x < nlo) {
    move_body<kBlock, 64, kHybridLds, WT, MOVE_HYBRID, false>(a, blockIdx.x, nlo);
  } else {
    hub_tab_body<WT>(h, blockIdx.x - nlo, gridDim.x - nlo, s_made);
    // the hubs queued for the edge-scanning kernels (queues 1-3: tables
    // dropped), which are not launched with this one, go back to the next
    // sub-round, so none is lost before the host launches them again
    const uint64_t nt = (uint64_t)(gridDim.x - nlo) * kBlock;
    for (int q = 1; q <= 3; ++q)
      for (uint64_t i = (uint64_t)(blockIdx.x - nlo) * kBlock + threadIdx.x; i < lens[q]; i += nt)
        mark_same(h, qs.q[q][i]);
  }
  __syncthreads();
  if (h.mv_piece && s_made) drain_moved<WT>(h);
}

// Neighbours of the tabled hubs that moved in the kernel before (list
// a.mv_*, at most a.mv_cap of *a.mv_n): the rows, cut into chunks of one
// block round (4 kBlock edges), are spread over the grid together, so a
// list of many small hubs costs about what one hub does.  With carry, the
// hubs queued for the edge-scanning kernels (queues 1-3: tables dropped in
// this run) while those kernels are not launched go back to the next
// sub-round, so none is lost before the host launches them again.
template <class WT>
__global__ __launch_bounds__(kBlock) void moved_flag_kernel(MoveArgs a, WlQueues qs, const uint32_t* lens,
                                                            int carry) {
  constexpr uint32_t CH = 4 * kBlock;  // edges per chunk
  __shared__ uint64_t s_first[kBlock], s_o[kBlock];  // first chunk of each entry of the batch, its row
  __shared__ uint32_t s_u[kBlock], s_d[kBlock], s_from[kBlock], s_to[kBlock];
  const uint32_t n = min(*a.mv_n, a.mv_cap);
  const WT* W = (const WT*)a.g.w;
  for (uint32_t m0 = 0; m0 < n; m0 += kBlock) {  // block-uniform: kBlock entries at a time
    const uint32_t m = m0 + threadIdx.x;
    uint32_t u = 0, d = 0;
    if (m < n) {
      u = a.mv_u[m];
      d = a.mv_d[m];
      s_o[threadIdx.x] = a.mv_o[m];
      s_from[threadIdx.x] = a.mv_from[m];
      s_to[threadIdx.x] = a.mv_to[m];
    }
    uint64_t chunks;
    const uint64_t first = block_scan_excl<kBlock>((d + CH - 1) / CH, chunks);
    s_first[threadIdx.x] = first;
    s_u[threadIdx.x] = u;
    s_d[threadIdx.x] = d;
    __syncthreads();
    const uint32_t nb = min(n - m0, (uint32_t)kBlock);
    for (uint64_t k = blockIdx.x; k < chunks; k += gridDim.x) {  // block-uniform
      // the entry holding chunk k: the last with first <= k (entries
      // without edges hold no chunk)
      uint32_t lo = 0, hi = nb;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_first[mid] <= k) lo = mid;
        else hi = mid;
      }
      const uint32_t e = (uint32_t)(k - s_first[lo]) * CH, du = s_d[lo];
      flag_neighbours<kBlock, WT>(a, W, s_u[lo], s_o[lo] + e, min(du - e, CH), threadIdx.x, s_from[lo], s_to[lo]);
    }
    __syncthreads();
  }
  if (!carry) return;
  for (int q = 1; q <= 3; ++q) {
    const uint32_t nq = lens[q];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
      mark_same(a, qs.q[q][i]);
  }
}

// Sources of the batch's insertions that passed kHtMinDeg edges with it (a
// source's degree changes only with its own entries) and have no table: a
// table each from the spare slots (bump counter next[0], slots up to total),
// listed for a build; with no spare slot left, dropped (counted in
// next[1]).  A source listed many times is claimed once (0 -> kHtDropped).
__global__ void ht_new_hubs_kernel(const uint32_t* iu, uint64_t ni, const uint32_t* deg, uint32_t* cap,
                                   uint64_t* off, unsigned long long* next, uint64_t total, uint32_t* list,
                                   uint32_t* nlist) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ni; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = iu[t];
    if (deg[u] <= kHtMinDeg || cap[u] != 0) continue;
    if (atomicCAS(&cap[u], 0u, kHtDropped) != 0) continue;
    const unsigned long long s0 = atomicAdd(&next[0], (unsigned long long)kHtMaxCap);
    if (s0 + kHtMaxCap > total) {
      atomicAdd(&next[1], 1ull);
      continue;
    }
    off[u] = s0;
    __hip_atomic_store(&cap[u], kHtMaxCap, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    list[atomicAdd(nlist, 1u)] = u;
  }
}

// Hub tables: per vertex its slots (0 below kHtMinDeg edges) and a flag for
// the build list.
__global__ void ht_caps_kernel(const uint32_t* deg, uint64_t S, uint32_t* cap, uint64_t* cap64, uint8_t* flags) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < S; u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = deg[u];
    uint32_t c = 0;
    if (d > kHtMinDeg) {
      c = 64;
      while (c < 2 * d && c < kHtMaxCap) c <<= 1;
    }
    cap[u] = c;
    cap64[u] = c;
    flags[u] = c != 0;
  }
}

// One wave per listed hub: its neighbours (self-loops skipped) summed per
// community in an LDS table, then written to its slots; more than
// kHtMaxDistinct communities drop the table.
//
// CHECK: compare instead of write -- every listed vertex that has a table
// and fits the build gets counted in out[0], and in out[1] if its table
// differs from the fresh sums (a count, a weight beyond 1e-9 relative, or a
// community with a positive count that is not a neighbour's).
template <class WT, bool CHECK>
__global__ __launch_bounds__(kBlock) void ht_build_kernel(GView g, const uint32_t* vcom, const uint32_t* list,
                                                          const uint32_t* n_dev, const uint64_t* off, uint32_t* cap,
                                                          uint32_t* key, int32_t* cnt, double* wt,
                                                          unsigned long long* out, unsigned long long* ndrop) {
  constexpr int WPB = kBlock / 64;
  __shared__ uint32_t s_key[WPB][kHtMaxCap];
  __shared__ int32_t s_cnt[WPB][kHtMaxCap];
  __shared__ double s_w[WPB][kHtMaxCap];
  __shared__ uint32_t s_made[WPB];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const uint64_t n = *n_dev;
  const WT* W = (const WT*)g.w;
  for (uint64_t k = (uint64_t)blockIdx.x * WPB + wv; k < n; k += (uint64_t)gridDim.x * WPB) {  // wave-uniform
    const uint32_t u = list[k], c_u = cap[u];
    if (c_u <= kHtDropped) continue;
    for (uint32_t sl = lane; sl < c_u; sl += 64) {
      s_key[wv][sl] = kEmpty;
      s_cnt[wv][sl] = 0;
      s_w[wv][sl] = 0;
    }
    if (lane == 0) s_made[wv] = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const uint64_t o = g.off[u];
    const uint32_t d = g.deg[u];
    bool over = false;
    for (uint32_t b0 = 0; b0 < d && !over; b0 += 64 * kScanRounds) {  // wave-uniform
      uint32_t v[kScanRounds], c[kScanRounds];
      double w[kScanRounds];
#pragma unroll
      for (int r = 0; r < kScanRounds; ++r) {
        const uint32_t i = b0 + r * 64 + lane;
        v[r] = i < d ? g.keys[o + i] : u;
        w[r] = i < d ? (double)W[o + i] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < kScanRounds; ++r) c[r] = v[r] != u ? vcom[v[r]] : 0;
#pragma unroll
      for (int r = 0; r < kScanRounds; ++r) {
        // each lane makes at most one new slot per round after reading the
        // distinct count, so the table never fills
        if (__hip_atomic_load(&s_made[wv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) >= kHtMaxDistinct)
          over = true;
        const bool ok = !over && v[r] != u;
        uint32_t sl = slot_of(c[r], c_u);
        if (ok) {
          while (true) {
            const uint32_t kk = atomicCAS(&s_key[wv][sl], kEmpty, c[r]);
            if (kk == kEmpty || kk == c[r]) {
              if (kk == kEmpty) atomicAdd(&s_made[wv], 1u);
              atomicAdd(&s_cnt[wv][sl], 1);
              unsafeAtomicAdd(&s_w[wv][sl], w[r]);
              break;
            }
            sl = (sl + 1) & (c_u - 1);
          }
        }
      }
      over = __ballot(over) != 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint64_t base = off[u];
    if constexpr (CHECK) {
      if (over) continue;  // more communities than a build keeps: not compared
      bool bad = false;
      int fresh = 0, kept = 0;
      for (uint32_t sl = lane; sl < c_u; sl += 64) {
        const uint32_t fk = s_key[wv][sl];
        if (fk != kEmpty) {
          ++fresh;
          const int64_t t = ht_slot(key, base, c_u, fk, false);
          const double fw = s_w[wv][sl];
          if (t < 0 || cnt[t] != s_cnt[wv][sl] || fabs(wt[t] - fw) > 1e-9 * fmax(1.0, fabs(fw))) bad = true;
        }
        if (key[base + sl] != kEmpty && cnt[base + sl] > 0) ++kept;
        if (key[base + sl] != kEmpty && cnt[base + sl] < 0) bad = true;
      }
      for (int m = 32; m > 0; m >>= 1) {
        fresh += __shfl_xor(fresh, m, 64);
        kept += __shfl_xor(kept, m, 64);
      }
      bad = __ballot(bad) != 0 || fresh != kept;
      if (lane == 0) {
        atomicAdd(&out[0], 1ull);
        if (bad) atomicAdd(&out[1], 1ull);
      }
      continue;
    }
    if (over) {
      if (lane == 0) ht_drop(cap, u, ndrop);
      continue;
    }
    for (uint32_t sl = lane; sl < c_u; sl += 64) {
      key[base + sl] = s_key[wv][sl];
      cnt[base + sl] = s_cnt[wv][sl];
      wt[base + sl] = s_w[wv][sl];
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// The batch's edges into the tables of their first endpoints (both
// directions are listed): deletions out, insertions in, by the endpoint's
// community before any move.
__global__ void ht_batch_kernel(const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                                const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                                const uint32_t* vcom, const uint64_t* off, uint32_t* cap, uint32_t* key, int32_t* cnt,
                                double* wt, unsigned long long* ndrop) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nd + ni;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const bool del = t < nd;
    const uint64_t j = del ? t : t - nd;
    const uint32_t u = del ? du[j] : iu[j], v = del ? dv[j] : iv[j];
    if (u == v) continue;
    const double w = (double)(del ? dw[j] : iw[j]);
    ht_add(cap, off, key, cnt, wt, ndrop, u, vcom[v], del ? -1 : 1, del ? -w : w);
  }
}

// ------------------------------------------------------------ renumbering
// louvainCommunityExistsOmpW (inc/louvain.hxx:668-682): flags must be zero.
__global__ void exists_kernel(GView g, const uint32_t* vcom, uint32_t* flags) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < g.S; u += (uint64_t)gridDim.x * blockDim.x) {
    if (g.ex && !g.ex[u]) continue;
    flags[vcom[u]] = 1;
  }
}

// louvainLookupCommunitiesOmpU (inc/louvain.hxx:838-844)
__global__ void lookup_kernel(uint32_t* a, uint64_t n, const uint32_t* map) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = map[a[i]];
}

// ------------------------------------------------------------ aggregation
// louvainCountCommunityVerticesOmpW / louvainCommunityVerticesOmpW
// (inc/louvain.hxx:752-762, :800-812)
// Atomic increment of x[c] for each active lane, aggregated over the wave:
// the lanes sharing the first active lane's c add once (the others add
// alone); returns the value the lane's increment took, so lanes sharing c
// take consecutive values in lane order.
__device__ __forceinline__ uint32_t wave_inc(uint32_t* x, bool on, uint32_t c) {
  const unsigned long long act = __ballot(on);
  if (!act) return 0;
  const int lane = threadIdx.x % 64;
  const int lead = __ffsll((long long)act) - 1;
  const uint32_t c0 = __shfl(c, lead, 64);
  const bool same = on && c == c0;
  const unsigned long long grp = __ballot(same);
  uint32_t base = 0;
  if (lane == lead) base = atomicAdd(&x[c0], (uint32_t)__popcll(grp));
  base = __shfl(base, lead, 64);
  if (same) return base + (uint32_t)__popcll(grp & ((1ull << lane) - 1));
  return on ? atomicAdd(&x[c], 1u) : 0;
}

// Community offsets (the exclusive scan of the member counts that
// louvainCommunityVerticesOmpW fills, inc/louvain.hxx:749-767) read off the
// ids sorted by community: coff[c] is the first position whose key is at
// least c.  Position i writes the c in (key[i-1], key[i]]; position n stands
// for a key C past the last, so every c in 0..C is written once.
__global__ void member_bounds_kernel(const uint32_t* sk, uint64_t n, uint32_t C, uint32_t* coff) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t hi = i < n ? min(sk[i], C) : C;
    const int64_t lo = i ? (int64_t)min(sk[i - 1], C) : -1;
    for (int64_t c = lo + 1; c <= (int64_t)hi; ++c) coff[c] = (uint32_t)i;
  }
}

// Degrees of the members in member order (+ a trailing 0): its exclusive
// scan gives each community's flat edge range, which is also the
// super-graph row offsets of louvainCommunityTotalDegreeOmpW + scan
// (inc/louvain.hxx:712-722, :989-990).
__global__ void member_degrees_kernel(const uint32_t* cedg, uint64_t n, GView g, uint64_t* mdeg) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j <= n; j += (uint64_t)gridDim.x * blockDim.x)
    mdeg[j] = j < n ? g.deg[cedg[j]] : 0;
}

__global__ void gather_yoff_kernel(const uint32_t* coff, const uint64_t* mscan, uint64_t C, uint64_t* yoff) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c <= C; c += (uint64_t)gridDim.x * blockDim.x)
    yoff[c] = mscan[coff[c]];
}

// Member j of community range [m0, m1) holding flat edge f.
__device__ __forceinline__ uint64_t member_of(const uint64_t* mscan, uint64_t m0, uint64_t m1, uint64_t f) {
  uint64_t lo = m0, hi = m1 - 1;
  while (lo < hi) {
    uint64_t mid = (lo + hi + 1) >> 1;
    if (mscan[mid] <= f) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// louvainAggregateEdgesOmpW (inc/louvain.hxx:897-912): per community, sum
// edge weights by the neighbour's community over all members' edges
// (self-loops included) and write the (community, weight) pairs into row c
// of the gapped super-graph.  LDS tiers 0..3, one group per community.
//
// Row order is the reference's: communities in the order they are first
// seen over the members (ascending ids) and their edges (louvainScanCommunitiesW
// appends a community when first seen).  Each table slot keeps the first
// flat member-edge index r at which its community appeared; a bitmap over
// the community's member edges marks those r, and a slot's place in the row
// is the number of marked bits below its r.
template <int BLOCK, int G, int CAP, class WT>
__global__ __launch_bounds__(BLOCK) void aggregate_kernel(AggArgs a) {
  constexpr int GPB = BLOCK / G;
  constexpr int BW = CAP / 64 > 0 ? CAP / 64 : 1;  // bitmap words (member edges <= CAP / 2)
  __shared__ uint32_t s_key[GPB * CAP];
  __shared__ uint32_t s_pos[GPB * CAP];
  __shared__ double s_val[GPB * CAP];
  __shared__ uint32_t s_bm[GPB * BW];
  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const WT* W = (const WT*)a.g.w;
  uint32_t* tkey = s_key + grp * CAP;
  uint32_t* tpos = s_pos + grp * CAP;
  double* tval = s_val + grp * CAP;
  uint32_t* bm = s_bm + grp * BW;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < a.n; base += (uint64_t)gridDim.x * GPB) {
    const uint64_t idx = base + grp;
    const bool act = idx < a.n;
    uint32_t c = act ? a.list[idx] : 0;
    uint64_t F0 = 0, F1 = 0, m0 = 0, m1 = 0;
    if (act) {
      F0 = a.yoff[c];
      F1 = a.yoff[c + 1];
      m0 = a.coff[c];
      m1 = a.coff[c + 1];
      for (int s = lane; s < CAP; s += G) { tkey[s] = kEmpty; tpos[s] = kNoPos; tval[s] = 0; }
      for (int w = lane; w < BW; w += G) bm[w] = 0;
    }
    __syncthreads();
    if (act) {
      for (uint64_t f = F0 + lane; f < F1; f += G) {
        uint64_t j = member_of(a.mscan, m0, m1, f);
        uint32_t u = a.cedg[j];
        uint64_t i = a.g.off[u] + (f - a.mscan[j]);
        tab_add(tkey, tpos, tval, CAP, a.vcom[a.g.keys[i]], (double)W[i], (uint32_t)(f - F0));
      }
    }
    __syncthreads();
    if (act) {
      for (int s = lane; s < CAP; s += G)
        if (tkey[s] != kEmpty) atomicOr(&bm[tpos[s] / 32], 1u << (tpos[s] % 32));
    }
    __syncthreads();
    if (act) {
      double sum = 0;
      for (int s = lane; s < CAP; s += G) {
        uint32_t k = tkey[s];
        if (k == kEmpty) continue;
        const uint32_t r = tpos[s];
        uint32_t at = __popc(bm[r / 32] & ((1u << (r % 32)) - 1u));
        for (uint32_t w = 0; w < r / 32; ++w) at += __popc(bm[w]);
        a.zkeys[F0 + at] = k;
        a.zw[F0 + at] = tval[s];
        sum += tval[s];
      }
      if (sum != 0) unsafeAtomicAdd(&a.ztot[c], sum);
      if (lane == 0) {
        uint32_t n = 0;
        for (int w = 0; w < BW; ++w) n += __popc(bm[w]);
        a.zdeg[c] = n;
      }
    }
    __syncthreads();
  }
}

// Tier 4: communities with > 1024 member edges share one HBM table each,
// filled by many blocks (kAggHbmChunk flat member edges per block chunk).
// A chunk is summed per neighbour community in an LDS table first (the
// lanes of a wave that share the wave's first community add it once), then
// merged into the HBM table with one atomic per community of the chunk: a
// community whose edges mostly stay inside it costs one HBM atomic per
// chunk, not one per edge.  The chunk's members (their flat edge starts and
// row offsets) are staged in LDS, so an edge finds its member there.
constexpr uint32_t kAggLdsTab = 2 * kAggHbmChunk;
constexpr uint32_t kAggMembers = kAggHbmChunk;

template <class WT>
__global__ __launch_bounds__(kBlock) void aggregate_hbm_kernel(AggArgs a) {
  __shared__ uint32_t s_key[kAggLdsTab];
  __shared__ uint32_t s_pos[kAggLdsTab];
  __shared__ double s_val[kAggLdsTab];
  __shared__ uint64_t s_ms[kAggMembers];   // flat edge start of each member of the chunk
  __shared__ uint64_t s_off[kAggMembers];  // its row offset
  __shared__ uint64_t s_k, s_j0, s_nm;
  const WT* W = (const WT*)a.g.w;
  for (uint64_t q = blockIdx.x; q < a.nchunks; q += gridDim.x) {
    for (uint32_t s = threadIdx.x; s < kAggLdsTab; s += kBlock) {
      s_key[s] = kEmpty;
      s_pos[s] = kNoPos;
      s_val[s] = 0;
    }
    if (threadIdx.x == 0) {
      // tier-4 community owning chunk q, the chunk's first and last member
      const uint64_t k = find_owner(a.choff, a.n, q);
      const uint32_t c = a.list[k];
      const uint64_t F1 = a.yoff[c + 1], m0 = a.coff[c], m1 = a.coff[c + 1];
      const uint64_t f0 = a.yoff[c] + (q - a.choff[k]) * kAggHbmChunk;
      const uint64_t f1 = min(F1, f0 + kAggHbmChunk);
      s_k = k;
      s_j0 = member_of(a.mscan, m0, m1, f0);
      s_nm = member_of(a.mscan, m0, m1, f1 - 1) - s_j0 + 1;
    }
    __syncthreads();
    const uint64_t k = s_k, j0 = s_j0, nm = s_nm;
    const uint32_t c = a.list[k];
    const uint64_t F1 = a.yoff[c + 1], m0 = a.coff[c], m1 = a.coff[c + 1];
    const uint64_t f0 = a.yoff[c] + (q - a.choff[k]) * kAggHbmChunk;
    const uint64_t f1 = min(F1, f0 + kAggHbmChunk);
    // (a chunk spanning more members than LDS holds -- members without
    // edges in between -- searches the member scan in HBM)
    const bool staged = nm <= kAggMembers;
    if (staged) {
      for (uint64_t t = threadIdx.x; t < nm; t += kBlock) {
        s_ms[t] = a.mscan[j0 + t];
        s_off[t] = a.g.off[a.cedg[j0 + t]];
      }
    }
    __syncthreads();
    for (uint64_t fb = f0; fb < f1; fb += kBlock) {  // block-uniform trip count
      const uint64_t f = fb + threadIdx.x;
      const bool ok = f < f1;
      uint32_t key = 0;
      double w = 0;
      if (ok) {
        uint64_t i;
        if (staged) {
          uint64_t lo = 0, hi = nm - 1;  // last member starting at or before f
          while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) >> 1;
            if (s_ms[mid] <= f) lo = mid;
            else hi = mid - 1;
          }
          i = s_off[lo] + (f - s_ms[lo]);
        } else {
          const uint64_t j = member_of(a.mscan, m0, m1, f);
          i = a.g.off[a.cedg[j]] + (f - a.mscan[j]);
        }
        key = a.vcom[a.g.keys[i]];
        w = (double)W[i];
      }
      // (position relative to the community's first member edge)
      tab_add_wave(s_key, s_pos, s_val, kAggLdsTab, ok, key, w, (uint32_t)(f - a.yoff[c]));
    }
    __syncthreads();
    const uint64_t cap = a.hoff[k + 1] - a.hoff[k];
    uint32_t* hkey = a.hkey + a.hoff[k];
    uint32_t* hpos = a.hpos + a.hoff[k];
    double* hval = a.hval + a.hoff[k];
    for (uint32_t s = threadIdx.x; s < kAggLdsTab; s += kBlock) {
      const uint32_t ck = s_key[s];
      if (ck != kEmpty) tab_merge(hkey, hpos, hval, cap, ck, s_val[s], s_pos[s]);
    }
    __syncthreads();
  }
}

// Tier-4 rows in first-seen order: each community's member edges are covered
// by a bitmap of 32 words per chunk (bm, from word a.choff[k] * 32); a table
// slot marks the bit of its first position.  After an exclusive scan of the
// words' popcounts (bsc), a slot's place in the row is the number of marked
// bits of its community below its position.
__global__ void agg_mark_hbm_kernel(AggArgs a, uint64_t total, uint32_t* bm) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < total; s += (uint64_t)gridDim.x * blockDim.x) {
    if (a.hkey[s] == kEmpty) continue;
    const uint64_t k = find_owner(a.hoff, a.n, s);
    const uint32_t r = a.hpos[s];
    atomicOr(&bm[a.choff[k] * 32 + r / 32], 1u << (r % 32));
  }
}

__global__ void popc_words_kernel(const uint32_t* bm, uint64_t W, uint32_t* cnt) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w <= W; w += (uint64_t)gridDim.x * blockDim.x)
    cnt[w] = w < W ? (uint32_t)__popc(bm[w]) : 0u;
}

__global__ void agg_emit_hbm_kernel(AggArgs a, uint64_t total, const uint32_t* bm, const uint32_t* bsc) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < total; s += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t key = a.hkey[s];
    if (key == kEmpty) continue;
    const uint64_t k = find_owner(a.hoff, a.n, s);
    const uint32_t c = a.list[k];
    const uint32_t r = a.hpos[s];
    const uint64_t w0 = a.choff[k] * 32, w = w0 + r / 32;
    const uint32_t at = bsc[w] - bsc[w0] + (uint32_t)__popc(bm[w] & ((1u << (r % 32)) - 1u));
    const uint64_t F0 = a.yoff[c];
    a.zkeys[F0 + at] = key;
    a.zw[F0 + at] = a.hval[s];
    unsafeAtomicAdd(&a.ztot[c], a.hval[s]);
  }
}

__global__ void agg_deg_hbm_kernel(AggArgs a, const uint32_t* bsc) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < a.n; k += (uint64_t)gridDim.x * blockDim.x)
    a.zdeg[a.list[k]] = bsc[a.choff[k + 1] * 32] - bsc[a.choff[k] * 32];
}

// Sort keys of the members: community (C for a vertex that does not exist)
// and id; a stable sort by community leaves each community's members in
// ascending id order, as louvainCommunityVerticesOmpW lists them with one
// thread (inc/louvain.hxx:800-812).
__global__ void member_keys_kernel(GView g, const uint32_t* vcom, uint32_t C, uint32_t* key, uint32_t* id) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < g.S; u += (uint64_t)gridDim.x * blockDim.x) {
    key[u] = (!g.ex || g.ex[u]) ? vcom[u] : C;
    id[u] = (uint32_t)u;
  }
}

__global__ void t4_sizes_kernel(const uint64_t* yoff, const uint32_t* list, uint64_t n, uint64_t C, uint64_t chunk,
                                uint64_t* caps, uint64_t* chunks) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k <= n; k += (uint64_t)gridDim.x * blockDim.x) {
    if (k == n) { caps[k] = 0; chunks[k] = 0; continue; }
    uint32_t c = list[k];
    uint64_t td = yoff[c + 1] - yoff[c];
    uint64_t m = td < C ? td : C;
    uint64_t cap = 1;
    while (cap < 2 * m) cap <<= 1;
    caps[k] = cap;
    chunks[k] = (td + chunk - 1) / chunk;
  }
}

// ------------------------------------------------------------ tier lists
// Partition ids 0..n-1 (live ones) into degree tiers, ascending within each
// tier.  Pass 1 counts per block and tier (tier-major layout), an exclusive
// scan of the counts gives each block's base, pass 2 scatters in order.
// thr == 0: the five degree tiers; thr > 0: two classes, tier 4 = degree > thr.
__device__ __forceinline__ int dev_tier(uint64_t d, uint32_t thr) {
  if (thr) return d > thr ? 4 : 0;
  return d <= 8 ? 0 : d <= 32 ? 1 : d <= 128 ? 2 : d <= 1024 ? 3 : 4;
}

__global__ __launch_bounds__(kBlock) void tier_count_kernel(const uint64_t* off, const uint32_t* deg,
                                                            const uint8_t* ex, uint64_t n, uint32_t thr,
                                                            uint32_t* counts, uint32_t nblocks) {
  __shared__ uint32_t s[kTiers];
  if (threadIdx.x < kTiers) s[threadIdx.x] = 0;
  __syncthreads();
  uint64_t u = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (u < n && (!ex || ex[u])) {
    uint64_t d = deg ? deg[u] : off[u + 1] - off[u];
    atomicAdd(&s[dev_tier(d, thr)], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kTiers) counts[threadIdx.x * (uint64_t)nblocks + blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void tier_scatter_kernel(const uint64_t* off, const uint32_t* deg,
                                                              const uint8_t* ex, uint64_t n, uint32_t thr,
                                                              const uint32_t* base, uint32_t nblocks, uint32_t* ids) {
  __shared__ uint32_t s_wave[kBlock / 64][kTiers];
  uint64_t u = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  int t = -1;
  if (u < n && (!ex || ex[u])) {
    uint64_t d = deg ? deg[u] : off[u + 1] - off[u];
    t = dev_tier(d, thr);
  }
  int lane = __lane_id(), w = threadIdx.x / 64;
  uint32_t mine = 0;
  for (int k = 0; k < kTiers; ++k) {
    unsigned long long m = __ballot(t == k);
    if (lane == 0) s_wave[w][k] = (uint32_t)__popcll(m);
    if (t == k) mine = (uint32_t)__popcll(m & ((1ull << lane) - 1));
  }
  __syncthreads();
  if (t >= 0) {
    uint32_t before = 0;
    for (int j = 0; j < w; ++j) before += s_wave[j][t];
    ids[base[t * (uint64_t)nblocks + blockIdx.x] + before + mine] = (uint32_t)u;
  }
}

// ------------------------------------------------------------ worklists
// Split a queue into per-tier queues (tier t at tq + t * stride), by degree.
__global__ __launch_bounds__(kBlock) void bin_queue_kernel(const uint32_t* q, const uint32_t* qlen,
                                                           const uint32_t* deg, uint32_t thr, uint32_t* tq,
                                                           uint64_t stride, uint32_t* tlen) {
  const uint64_t n = *qlen;
  for (uint64_t b = blockIdx.x * (uint64_t)kBlock; b < n; b += (uint64_t)gridDim.x * kBlock) {
    uint64_t i = b + threadIdx.x;
    int t = -1;
    uint32_t u = 0;
    if (i < n) {
      u = q[i];
      t = dev_tier(deg[u], thr);
    }
    int lane = __lane_id();
    for (int k = 0; k < kTiers; ++k) {
      unsigned long long m = __ballot(t == k);
      if (!m) continue;
      int leader = __ffsll((long long)m) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(&tlen[k], (uint32_t)__popcll(m));
      base = __shfl(base, leader);
      if (t == k) tq[k * stride + base + __popcll(m & ((1ull << lane) - 1))] = u;
    }
  }
}

// Maximum over a kBlock-thread block (every thread gets it).
__device__ uint32_t block_max_u32(uint32_t m) {
  __shared__ uint32_t s[kBlock / 64];
#pragma unroll
  for (int k = 32; k > 0; k >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, k, 64));
  if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = m;
  __syncthreads();
  m = s[0];
#pragma unroll
  for (int i = 1; i < kBlock / 64; ++i) m = max(m, s[i]);
  __syncthreads();
  return m;
}

__global__ __launch_bounds__(kBlock) void max_deg_kernel(const uint32_t* ids, uint64_t n, const uint32_t* deg,
                                                         uint32_t* out) {
  uint32_t m = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    m = max(m, deg[ids[i]]);
  m = block_max_u32(m);
  if (threadIdx.x == 0 && m) atomicMax(out, m);
}

__global__ __launch_bounds__(kBlock) void max_u32_kernel(const uint32_t* a, uint64_t n, uint32_t* out) {
  uint32_t m = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    m = max(m, a[i]);
  m = block_max_u32(m);
  if (threadIdx.x == 0 && m) atomicMax(out, m);
}

__global__ void gather_u32_kernel(const uint32_t* a, const uint32_t* idx, uint64_t n, uint32_t* out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = a[idx[i]];
}

// Sum of nextpow2(2 deg) over the ids with degree above thr (hub table slots).
__global__ __launch_bounds__(kBlock) void caps_sum_kernel(const uint32_t* deg, uint64_t n, uint32_t thr,
                                                          unsigned long long* out) {
  unsigned long long acc = 0;
  for (uint64_t u = blockIdx.x * (uint64_t)kBlock + threadIdx.x; u < n; u += (uint64_t)gridDim.x * kBlock) {
    uint64_t d = deg[u];
    if (d <= thr) continue;
    uint64_t cap = 2;
    while (cap < 2 * d) cap <<= 1;
    acc += cap;
  }
  acc = block_sum_u64<kBlock>(acc);
  if (threadIdx.x == 0 && acc) atomicAdd(out, acc);
}

// Ordered worklist queues (inc/louvain.hxx:594-598 visits ids in ascending
// order): the ids whose flag byte is set, ascending, split into the
// kWlQueues queues.  With clear, the flags read are reset.  Counts per
// contiguous segment (one block each), then an ordered scatter.  A thread
// reads 16 consecutive flag bytes (one 16-byte load), so a block covers a
// 4096-id tile per step and needs one block-wide scan per tile.
// TB threads a block: a tile of 16 TB ids.  The launchers take TB =
// min(512, segment / 16), so a segment of up to 8192 ids is one tile (one
// round of loads, one scan); the tiles of a larger segment follow each other.
static_assert(kWlScanBlocks % kBlock == 0, "whole segments per block count pass");
template <int TB>
constexpr int wl_count_bits() { return TB <= 256 ? 13 : TB <= 512 ? 14 : 15; }  // 16 TB fits

// ids of block b: [b << shift, (b + 1) << shift), a multiple of the tile

// Flag words of ids id0 .. id0 + 15 (id0 a multiple of 16; 0 at or past hi;
// the flag array holds whole words up to S).
__device__ __forceinline__ void wl_words16(const uint8_t* f, uint64_t id0, uint64_t hi, uint32_t (&w)[4]) {
  if (id0 + 16 <= hi && !((uintptr_t)f & 15)) {
    const uint4 v = *reinterpret_cast<const uint4*>(f + id0);
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = id0 + 4 * k < hi ? reinterpret_cast<const uint32_t*>(f)[(id0 >> 2) + k] : 0u;
  }
}

// Bit j: id id0 + j is flagged (id0 a multiple of 16; 0 at or past hi).
// Byte flags (one per id, whole words up to S) or a bitmap (bit v of u32
// word v >> 5; its 16-bit halves read here).
template <bool BITS>
__device__ __forceinline__ uint32_t wl_mask16(const void* f, uint64_t id0, uint64_t hi) {
  if constexpr (BITS) {
    if (id0 >= hi) return 0u;
    uint32_t m = reinterpret_cast<const uint16_t*>(f)[id0 >> 4];
    if (id0 + 16 > hi) m &= (1u << (hi - id0)) - 1u;
    return m;
  } else {
    uint32_t w[4];
    wl_words16(static_cast<const uint8_t*>(f), id0, hi, w);
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if ((w[j / 4] >> (8 * (j % 4))) & 0xFFu) m |= 1u << j;
    return m;
  }
}

// Clear the flags of ids id0 .. id0 + 15 (below hi) whose mask was f16.
template <bool BITS>
__device__ __forceinline__ void wl_clear16(void* f, uint64_t id0, uint64_t hi, uint32_t f16) {
  if (!f16) return;
  if constexpr (BITS) {
    reinterpret_cast<uint16_t*>(f)[id0 >> 4] = 0;
  } else {
    uint8_t* b = static_cast<uint8_t*>(f);
    if (id0 + 16 <= hi && !((uintptr_t)b & 15)) {
      *reinterpret_cast<uint4*>(b + id0) = make_uint4(0u, 0u, 0u, 0u);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if ((f16 >> (4 * k)) & 0xFu) reinterpret_cast<uint32_t*>(b)[(id0 >> 2) + k] = 0u;
    }
  }
}

// Bit j of m[q]: id id0 + j (j < 16) is flagged (bit j of f16) and belongs
// to queue q (queue 0: degree <= t[0]; 1: <= t[1]; 2: <= t[2]; 3: above, or
// a 1/2 id marked in big; 4: an id of queues 1-3 with a table; t[0] == 0
// puts every id in queue 0).  The 16 degrees are read together when more
// than 4 of the ids are flagged, one at a time otherwise.
__device__ __forceinline__ void wl_bits16(uint32_t f16, const uint32_t* deg, uint64_t S, const WlSplit& sp,
                                          uint64_t id0, uint32_t (&m)[kWlQueues]) {
#pragma unroll
  for (int q = 0; q < kWlQueues; ++q) m[q] = 0;
  if (!f16) return;
  if (__popc(f16) <= 4) {
    for (uint32_t bits = f16; bits; bits &= bits - 1) {
      const int j = __ffs(bits) - 1;
      if (id0 + j >= S) continue;
      const uint32_t dj = sp.t[0] ? deg[id0 + j] : 0u;
      int q = dj <= sp.t[0] ? 0 : (dj <= sp.t[1] ? 1 : (dj <= sp.t[2] ? 2 : 3));
      if (q == 1 || q == 2) q = sp.big && sp.big[id0 + j] ? 3 : q;
      if (q && sp.tab && sp.tab[id0 + j] > kHtDropped) q = 4;
      m[q] |= 1u << j;
    }
    return;
  }
  uint32_t d[16];
  if (sp.t[0] && id0 + 16 <= S && !((uintptr_t)deg & 15)) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 v = reinterpret_cast<const uint4*>(deg + id0)[k];
      d[4 * k] = v.x;
      d[4 * k + 1] = v.y;
      d[4 * k + 2] = v.z;
      d[4 * k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) d[j] = sp.t[0] && id0 + j < S ? deg[id0 + j] : 0u;
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (id0 + j < S && ((f16 >> j) & 1u)) {
      int q = d[j] <= sp.t[0] ? 0 : (d[j] <= sp.t[1] ? 1 : (d[j] <= sp.t[2] ? 2 : 3));
      if (q == 1 || q == 2) q = sp.big && sp.big[id0 + j] ? 3 : q;
      if (q && sp.tab && sp.tab[id0 + j] > kHtDropped) q = 4;
      m[q] |= 1u << j;
    }
  }
}

// Sum over a wave (every lane gets it).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m, 64);
  return x;
}

// Per block (segment) and queue: the flagged ids it holds.
template <bool BITS, int TB>
__global__ __launch_bounds__(TB) void wl_count_kernel(const void* f, uint64_t S, const uint32_t* deg,
                                                      WlSplit sp, uint64_t* bsum) {
  __shared__ unsigned long long s_c[kWlQueues];
  const uint64_t lo = (uint64_t)blockIdx.x << sp.shift, hi = min(S, lo + (1ull << sp.shift));
  if (sp.dirty && !sp.dirty[blockIdx.x]) {  // no flag was set in this segment
    if (threadIdx.x < kWlQueues) bsum[threadIdx.x * kWlScanBlocks + blockIdx.x] = 0;
    return;
  }
  if (threadIdx.x < kWlQueues) s_c[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long c[kWlQueues] = {};
  for (uint64_t id0 = lo + 16 * (uint64_t)threadIdx.x; id0 < hi; id0 += 16 * (uint64_t)TB) {
    uint32_t m[kWlQueues];
    wl_bits16(wl_mask16<BITS>(f, id0, hi), deg, S, sp, id0, m);
#pragma unroll
    for (int q = 0; q < kWlQueues; ++q) c[q] += __popc(m[q]);
  }
#pragma unroll
  for (int q = 0; q < kWlQueues; ++q) {
    const unsigned long long t = wave_sum_u64(c[q]);
    if (threadIdx.x % 64 == 0 && t) atomicAdd(&s_c[q], t);
  }
  __syncthreads();
  if (threadIdx.x < kWlQueues) bsum[threadIdx.x * kWlScanBlocks + blockIdx.x] = s_c[threadIdx.x];
}

// Block b writes the ids of its segment to every queue from the sum of the
// counts of blocks 0 .. b-1 (each block adds them up itself; block 0 adds
// up all of them for the queue lengths).  One block-wide scan per 4096-id
// tile gives each thread its place in queues 0-3 at once (their tile counts,
// at most 4096 each, packed 13 bits apart); queue 4 takes a second scan in
// the tiles that hold any of its ids.  Writes stay below S, the queues'
// capacity.
template <bool BITS, int TB>
__global__ __launch_bounds__(TB) void wl_scatter_kernel(void* f, int clear, uint64_t S, const uint32_t* deg,
                                                        WlSplit sp, const uint64_t* bsum, WlQueues qs,
                                                        uint32_t* lens) {
  constexpr int CB = wl_count_bits<TB>();
  constexpr uint64_t CM = (1ull << CB) - 1;
  static_assert(16 * TB < (1 << CB) && 4 * CB <= 64 && kWlQueues == 5, "packed tile counts");
  __shared__ unsigned long long s_pre[kWlQueues];
  const uint64_t lo = (uint64_t)blockIdx.x << sp.shift, hi = min(S, lo + (1ull << sp.shift));
  const bool skip = sp.dirty && !sp.dirty[blockIdx.x];  // block-uniform: no flag in this segment
  if (skip && blockIdx.x != 0) return;
  if (threadIdx.x < kWlQueues) s_pre[threadIdx.x] = 0;
  __syncthreads();
  {
    const uint32_t nb = blockIdx.x == 0 ? kWlScanBlocks : blockIdx.x;
    unsigned long long loc[kWlQueues] = {};
    for (uint32_t j = threadIdx.x; j < nb; j += TB)
#pragma unroll
      for (int q = 0; q < kWlQueues; ++q) loc[q] += bsum[q * kWlScanBlocks + j];
#pragma unroll
    for (int q = 0; q < kWlQueues; ++q) {
      const unsigned long long t = wave_sum_u64(loc[q]);
      if (threadIdx.x % 64 == 0 && t) atomicAdd(&s_pre[q], t);
    }
  }
  __syncthreads();
  uint64_t r[kWlQueues];
#pragma unroll
  for (int q = 0; q < kWlQueues; ++q) r[q] = blockIdx.x == 0 ? 0 : s_pre[q];
  if (blockIdx.x == 0 && threadIdx.x < kWlQueues) lens[threadIdx.x] = (uint32_t)s_pre[threadIdx.x];
  if (blockIdx.x == 0 && threadIdx.x < sp.nreset && sp.reset) sp.reset[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && sp.ndrop) lens[kWlQueues] = (uint32_t)min(*sp.ndrop, 0xFFFFFFFFull);
  if (blockIdx.x == 0 && threadIdx.x == 0 && sp.el) {  // the iteration's gain so far, summed as the host does
    double t = 0;
    for (int sl = 0; sl < kSpread; ++sl) t += sp.el[(size_t)sl * kElStride];
    *reinterpret_cast<double*>(lens + 6) = t;
  }
  if (skip) return;
  if (sp.dirty && clear && threadIdx.x == 0) sp.dirty[blockIdx.x] = 0;
  for (uint64_t t0 = lo; t0 < hi; t0 += 16 * (uint64_t)TB) {  // block-uniform trip count
    const uint64_t id0 = t0 + 16 * (uint64_t)threadIdx.x;
    const uint32_t f16 = wl_mask16<BITS>(f, id0, hi);
    uint32_t m[kWlQueues];
    wl_bits16(f16, deg, S, sp, id0, m);
    uint64_t packed = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) packed |= (uint64_t)__popc(m[q]) << (CB * q);
    uint64_t tot;
    const uint64_t ex = block_scan_excl<TB>(packed, tot);
    uint64_t ex4 = 0, tot4 = 0;
    if (__syncthreads_or(m[4] != 0)) ex4 = block_scan_excl<TB>((uint64_t)__popc(m[4]), tot4);
#pragma unroll
    for (int q = 0; q < kWlQueues; ++q) {
      uint64_t e = r[q] + (q < 4 ? (ex >> (CB * q)) & CM : ex4);
      uint32_t bits = m[q];
      while (bits) {
        const int j = __ffs(bits) - 1;
        bits &= bits - 1;
        if (e < S) qs.q[q][e] = (uint32_t)(id0 + j);
        ++e;
      }
      r[q] += q < 4 ? (tot >> (CB * q)) & CM : tot4;
    }
    if (clear) wl_clear16<BITS>(f, id0, hi, f16);
  }
}

// ------------------------------------------------------------ modularity
// modularityByOmp (inc/properties.hxx:205-233)
template <class WT>
__global__ __launch_bounds__(kBlock) void modularity_parts_kernel(GView g, const uint32_t* memb, double* cin,
                                                                  double* ctot) {
  const WT* W = (const WT*)g.w;
  const int lane = __lane_id();
  for (uint64_t base = 0; base < g.S; base += (uint64_t)gridDim.x * kBlock) {  // grid-uniform
    const uint64_t u = scrambled_vertex(base);
    const bool live = u < g.S && (!g.ex || g.ex[u]);
    const uint64_t o = live ? g.off[u] : 0;
    const uint32_t d = live ? g.deg[u] : 0;
    const uint32_t c = live ? memb[u] : 0;
    const bool big = live && d > 64;
    if (live && !big) {
      double si = 0, st = 0;
      for (uint32_t i = 0; i < d; ++i) {
        double w = (double)W[o + i];
        if (memb[g.keys[o + i]] == c) si += w;
        st += w;
      }
      if (si != 0) unsafeAtomicAdd(&cin[c], si);
      if (st != 0) unsafeAtomicAdd(&ctot[c], st);
    }
    unsigned long long mask = __ballot(big);
    while (mask) {
      const int j = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const uint64_t oj = __shfl(o, j);
      const uint32_t dj = __shfl(d, j);
      const uint32_t cj = __shfl(c, j);
      double si = 0, st = 0;
      for (uint32_t i0 = 0; i0 < dj; i0 += 4 * 64) {
        double w[4];
        uint32_t k[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t i = i0 + r * 64 + lane;
          w[r] = i < dj ? (double)W[oj + i] : 0.0;
          k[r] = i < dj ? memb[g.keys[oj + i]] : ~cj;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (k[r] == cj) si += w[r];
          st += w[r];
        }
      }
#pragma unroll
      for (int m = 32; m > 0; m >>= 1) { si += __shfl_xor(si, m, 64); st += __shfl_xor(st, m, 64); }
      if (lane == 0) {
        if (si != 0) unsafeAtomicAdd(&cin[cj], si);
        if (st != 0) unsafeAtomicAdd(&ctot[cj], st);
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void modularity_sum_kernel(const double* cin, const double* ctot, uint64_t n,
                                                                double M, double R, double* out) {
  double a = 0;
  for (uint64_t c = blockIdx.x * (uint64_t)kBlock + threadIdx.x; c < n; c += (uint64_t)gridDim.x * kBlock) {
    double x = ctot[c] / (2 * M);
    a += cin[c] / (2 * M) - R * x * x;
  }
  a = block_sum<kBlock>(a);
  if (threadIdx.x == 0) unsafeAtomicAdd(out, a);
}

#define LV_LAUNCH(kernel, grid, block, st, ...) hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, __VA_ARGS__)

}  // namespace

// ------------------------------------------------------------ launchers
__global__ void offset_bound_kernel(const uint64_t* off, uint64_t S, uint64_t share, uint64_t* out) {
  if (blockIdx.x || threadIdx.x) return;
  const uint64_t target = off[S] / share;
  uint64_t lo = 0, hi = S;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (off[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  *out = lo;
}

__global__ void list_caps_kernel(const uint32_t* ids, uint64_t n, const uint32_t* deg, uint64_t* caps) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i == n) {
      caps[i] = 0;
      continue;
    }
    const uint64_t d = deg[ids[i]];
    uint64_t c = 2;
    while (c < 2 * d) c <<= 1;
    caps[i] = c;
  }
}

void k_fill_u8(uint8_t* a, uint8_t v, uint64_t n, hipStream_t st) {
  if (n) LV_LAUNCH(fill_kernel<uint8_t>, grid_for(n, kBlock * 4), kBlock, st, a, v, n);
}
void k_fill_u32(uint32_t* a, uint32_t v, uint64_t n, hipStream_t st) {
  if (n) LV_LAUNCH(fill_kernel<uint32_t>, grid_for(n, kBlock * 4), kBlock, st, a, v, n);
}
void k_fill_f64(double* a, double v, uint64_t n, hipStream_t st) {
  if (n) LV_LAUNCH(fill_kernel<double>, grid_for(n, kBlock * 4), kBlock, st, a, v, n);
}
void k_fill_u64(uint64_t* a, uint64_t v, uint64_t n, hipStream_t st) {
  if (n) LV_LAUNCH(fill_kernel<uint64_t>, grid_for(n, kBlock * 4), kBlock, st, a, v, n);
}
void k_iota_u32(uint32_t* a, uint64_t n, hipStream_t st) {
  if (n) LV_LAUNCH(iota_kernel, grid_for(n, kBlock * 4), kBlock, st, a, n);
}
void k_degrees_from_offsets(const uint64_t* off, uint32_t* deg, uint64_t S, hipStream_t st) {
  if (S) LV_LAUNCH(degrees_kernel, grid_for(S, kBlock * 4), kBlock, st, off, deg, S);
}
void k_vertex_weights(GView g, bool f64, double* vtot, hipStream_t st) {
  if (!g.S) return;
  uint64_t grid = grid_for(g.S, kBlock, 1u << 16);
  if (f64) LV_LAUNCH((row_sum_kernel<double>), grid, kBlock, st, g, vtot);
  else LV_LAUNCH((row_sum_kernel<float>), grid, kBlock, st, g, vtot);
}
void k_edge_weight(GView g, bool f64, double* out, hipStream_t st) {
  if (!g.S) return;
  uint64_t grid = grid_for(g.S, kBlock, 2048);
  if (f64) LV_LAUNCH((edge_weight_kernel<double>), grid, kBlock, st, g, out);
  else LV_LAUNCH((edge_weight_kernel<float>), grid, kBlock, st, g, out);
}
void k_initialize(GView g, uint32_t* vcom, double* ctot, const double* vtot, hipStream_t st) {
  if (g.S) LV_LAUNCH(initialize_kernel, grid_for(g.S, kBlock * 4), kBlock, st, g, vcom, ctot, vtot);
}
void k_update_weights(const uint32_t* du, const float* dw, uint64_t nd, const uint32_t* iu, const float* iw,
                      uint64_t ni, const uint32_t* vcom, double* vtot, double* ctot, hipStream_t st) {
  if (nd + ni)
    LV_LAUNCH(update_weights_kernel, grid_for(nd + ni + 128, kBlock), kBlock, st, du, dw, nd, iu, iw, ni, vcom, vtot,
              ctot);
}
void k_frontier(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* iu, const uint32_t* iv,
                uint64_t ni, const uint32_t* vcom, uint8_t* vaff, hipStream_t st) {
  if (nd + ni) LV_LAUNCH(frontier_kernel, grid_for(nd + ni, kBlock), kBlock, st, du, dv, nd, iu, iv, ni, vcom, vaff);
}
void k_ds_deletions(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* vcom, uint8_t* vertices,
                    uint8_t* neighbors, uint8_t* communities, hipStream_t st) {
  if (nd)
    LV_LAUNCH(ds_deletions_kernel, grid_for(nd, kBlock), kBlock, st, du, dv, nd, vcom, vertices, neighbors,
              communities);
}
void k_ds_insertions(const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni, const uint32_t* vcom,
                     const double* vtot, const double* ctot, double M, double R, uint32_t* tkey, double* tval,
                     uint32_t* tlist, uint8_t* vertices, uint8_t* neighbors, uint8_t* communities, hipStream_t st,
                     const uint32_t* gid, const double* gM) {
  if (ni)
    LV_LAUNCH(ds_insertions_kernel, grid_for(ni, kBlock), kBlock, st, iu, iv, iw, ni, vcom, vtot, ctot, M, R, tkey,
              tval, tlist, vertices, neighbors, communities, gid, gM);
}
void k_ds_sweep(GView g, const uint32_t* vcom, const uint8_t* neighbors, const uint8_t* communities,
                uint8_t* vertices, hipStream_t st) {
  if (g.S) LV_LAUNCH(ds_sweep_kernel, grid_for(g.S, kBlock, 1u << 16), kBlock, st, g, vcom, neighbors, communities, vertices);
}
void k_count_u8(const uint8_t* a, uint8_t v, uint64_t n, unsigned long long* out, hipStream_t st) {
  if (n) LV_LAUNCH(count_u8_kernel, grid_for(n, kBlock * 8, kReduceGrid), kBlock, st, a, v, n, out);
}
void k_check_sorted(const uint32_t* u, const uint32_t* v, uint64_t n, uint32_t S, int* bad, hipStream_t st) {
  if (n) LV_LAUNCH(check_sorted_kernel, grid_for(n, kBlock), kBlock, st, u, v, n, S, bad);
}
void k_check_rows(const uint64_t* off, const uint32_t* deg, const uint32_t* keys, uint64_t S, int* bad,
                  hipStream_t st) {
  if (S) LV_LAUNCH(check_rows_kernel, grid_for(S, kBlock / 64, 1u << 16), kBlock, st, off, deg, keys, S, bad);
}
void k_check_edges(const uint64_t* off, const uint32_t* keys, uint64_t S, uint64_t E, uint32_t* starts, int* bad,
                   hipStream_t st) {
  if (!S || !E) return;
  LV_LAUNCH(fill_kernel<uint32_t>, grid_for(E / 32 + 1, kBlock * 4), kBlock, st, starts, 0u, E / 32 + 1);
  LV_LAUNCH(row_starts_kernel, grid_for(S, kBlock, 1u << 16), kBlock, st, off, S, starts);
  LV_LAUNCH(check_edges_kernel, grid_for(E, kBlock * 8, 1u << 16), kBlock, st, keys, E, S, starts, bad);
}

template <class WT>
static void move_tier(const MoveArgs& a, int tier, hipStream_t st) {
  // max_streams groups (no more than there are chunks); chunks are handed
  // out dynamically, so the grid only sets how many walk at once.
  uint64_t streams = a.max_streams ? a.max_streams : 1;
  uint64_t groups = streams;
  if (!a.n_dev && !a.ctab) {
    uint64_t chunk = a.chunk ? a.chunk : std::max<uint64_t>(1, (a.n + streams - 1) / streams);
    groups = std::min<uint64_t>(streams, (a.n + chunk - 1) / chunk);
  }
  switch (tier) {
    case 0: LV_LAUNCH((move_kernel<256, 8, 16, WT, MOVE_LDS>), grid_for(groups, 32, 1u << 16), 256, st, a); break;
    case 1: LV_LAUNCH((move_kernel<256, 16, 64, WT, MOVE_LDS>), grid_for(groups, 16, 1u << 16), 256, st, a); break;
    case 2: LV_LAUNCH((move_kernel<256, 64, 256, WT, MOVE_LDS>), grid_for(groups, 4, 1u << 16), 256, st, a); break;
    case 3: LV_LAUNCH((move_kernel<256, 256, 2048, WT, MOVE_LDS>), grid_for(groups, 1, 1u << 14), 256, st, a); break;
    case 4: {
      // hubs walk in at most hub_streams streams (blocks)
      MoveArgs h = a;
      uint64_t hs = std::max<uint64_t>(1, std::min<uint64_t>(a.hub_streams ? a.hub_streams : kHubBlocks, kHubBlocks));
      h.max_streams = std::min<uint64_t>(a.max_streams ? a.max_streams : 1, hs);
      uint64_t hg = h.n_dev ? h.max_streams : std::min<uint64_t>(groups, h.max_streams);
      LV_LAUNCH((move_kernel<1024, 1024, 0, WT, MOVE_HBM>), grid_for(hg, 1, kHubBlocks), 1024, st, h);
      break;
    }
    default:
      // Many streams with chunks dealt by position: four one-wave groups per
      // block (a quarter of the workgroups to dispatch; regions for
      // round_up(gblocks, 4) groups), each still given at most one chunk per
      // round, so no more than `streams` walk at once.  Chunks handed out
      // on demand go a block at a time: one wave per block then.
      if (a.validate || a.xids) {
        if (!a.chunk && groups >= 64)
          LV_LAUNCH((move_kernel<256, 64, kHybridLds, WT, MOVE_HYBRID_X>), grid_for(groups, 4, (a.gblocks + 3) / 4),
                    256, st, a);
        else
          LV_LAUNCH((move_kernel<64, 64, kHybridLds, WT, MOVE_HYBRID_X>), grid_for(groups, 1, a.gblocks), 64, st, a);
      } else if (!a.chunk && groups >= 64 && !a.gid) {
        LV_LAUNCH((move_one_kernel<256, 64, kHybridLds, WT, MOVE_HYBRID>), grid_for(groups, 4, (a.gblocks + 3) / 4),
                  256, st, a);
      } else if (!a.chunk && groups >= 64) {
        LV_LAUNCH((move_bt_kernel<256, 64, kHybridLds, WT, MOVE_HYBRID>), grid_for(groups, 4, (a.gblocks + 3) / 4),
                  256, st, a);
      } else {
        LV_LAUNCH((move_kernel<64, 64, kHybridLds, WT, MOVE_HYBRID>), grid_for(groups, 1, a.gblocks), 64, st, a);
      }
      break;
  }
}
void k_offset_bound(const uint64_t* off, uint64_t S, uint64_t share, uint64_t* out, hipStream_t st) {
  LV_LAUNCH(offset_bound_kernel, 1, 64, st, off, S, share, out);
}
void k_list_caps(const uint32_t* ids, uint64_t n, const uint32_t* deg, uint64_t* caps, hipStream_t st) {
  LV_LAUNCH(list_caps_kernel, grid_for(n + 1, kBlock), kBlock, st, ids, n, deg, caps);
}
void k_move(const MoveArgs& a, int tier, bool f64, hipStream_t st) {
  if (f64) move_tier<double>(a, tier, st);
  else move_tier<float>(a, tier, st);
}
void k_hub_batch(const MoveArgs& a, hipStream_t st) {
  const uint32_t grid = 2048;
  LV_LAUNCH(hub_prep_kernel, 1, kPrepBlock, st, a);
  if (a.g.w && a.f64_weights) LV_LAUNCH(hub_insert_kernel<double>, grid, kBlock, st, a);
  else LV_LAUNCH(hub_insert_kernel<float>, grid, kBlock, st, a);
  LV_LAUNCH(hub_best_kernel, grid, kBlock, st, a);
  LV_LAUNCH(hub_apply_kernel, 64, kBlock, st, a);
  if (!a.choose_only) LV_LAUNCH(hub_flag_kernel, grid, kBlock, st, a);
}
void k_hub_mid(const MoveArgs& a, int cls, hipStream_t st) {
  const bool f64 = a.g.w && a.f64_weights;
  if (cls == 0) {
    if (f64) LV_LAUNCH((hub_mid_kernel<kSmallHubBlock, kSmallHubSlots, kSmallHubMax, double>), kSmallHubGrid,
                       kSmallHubBlock, st, a);
    else LV_LAUNCH((hub_mid_kernel<kSmallHubBlock, kSmallHubSlots, kSmallHubMax, float>), kSmallHubGrid,
                   kSmallHubBlock, st, a);
  } else {
    if (f64) LV_LAUNCH((hub_mid_kernel<kMidBlock, kMidSlots, kMidHubMax, double>), kMidGrid, kMidBlock, st, a);
    else LV_LAUNCH((hub_mid_kernel<kMidBlock, kMidSlots, kMidHubMax, float>), kMidGrid, kMidBlock, st, a);
  }
}
// about one wave per tabled hub of a sub-round (their number grows with the
// graph: ~750 at 4M vertices); a launch with few hubs then dispatches few
// blocks
static uint64_t hub_tab_grid(const MoveArgs& a) {
  return std::min<uint64_t>(1024, std::max<uint64_t>(128, a.g.S >> 14));
}
void k_hub_tab(const MoveArgs& a, hipStream_t st) {
  const uint64_t grid = hub_tab_grid(a);
  if (a.g.w && a.f64_weights) LV_LAUNCH(hub_tab_kernel<double>, grid, kBlock, st, a);
  else LV_LAUNCH(hub_tab_kernel<float>, grid, kBlock, st, a);
}
bool k_wl_fused(const MoveArgs& a, const MoveArgs& h, const WlQueues& qs, const uint32_t* lens, bool f64,
                hipStream_t st) {
  // the id-order part as move_tier launches it with spread chunks and four
  // streams a block; other schedules take the separate launches
  const uint64_t streams = a.max_streams ? a.max_streams : 1;
  if (a.chunk || a.validate || a.xids || a.ctab || a.gid || streams < 64) return false;  // (one graph)
  const uint64_t nlo = grid_for(streams, 4, (a.gblocks + 3) / 4);
  const uint64_t grid = nlo + hub_tab_grid(h);
  if (f64) LV_LAUNCH(wl_fused_kernel<double>, grid, kBlock, st, a, h, (uint32_t)nlo, qs, lens);
  else LV_LAUNCH(wl_fused_kernel<float>, grid, kBlock, st, a, h, (uint32_t)nlo, qs, lens);
  return true;
}
__global__ void flag_nonzero_kernel(const uint32_t* a, uint64_t n, uint8_t* flags) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    flags[i] = a[i] != 0;
}
// One wave that waits on the device's 100 MHz wall clock for `us`
// microseconds (bounded: the clock always advances).  Put before the start
// event of a timed launch, it lets the host write the event, the kernel and
// the end event into the queue before the device reaches them, so the pair
// of events brackets the kernel and not the host's launch latency.
__global__ void hold_kernel(uint64_t ticks) {
  const uint64_t t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(16);
}
void k_hold(uint32_t us, hipStream_t st) { LV_LAUNCH(hold_kernel, 1, 64, st, (uint64_t)us * 100); }
void k_flag_nonzero(const uint32_t* a, uint64_t n, uint8_t* flags, hipStream_t st) {
  if (n) LV_LAUNCH(flag_nonzero_kernel, grid_for(n, kBlock * 4), kBlock, st, a, n, flags);
}
void k_moved_flag(const MoveArgs& a, const WlQueues& qs, const uint32_t* lens, bool carry, hipStream_t st) {
  if (a.g.w && a.f64_weights) LV_LAUNCH(moved_flag_kernel<double>, 256, kBlock, st, a, qs, lens, (int)carry);
  else LV_LAUNCH(moved_flag_kernel<float>, 256, kBlock, st, a, qs, lens, (int)carry);
}
void k_ht_new_hubs(const uint32_t* iu, uint64_t ni, const uint32_t* deg, uint32_t* cap, uint64_t* off,
                   unsigned long long* next, uint64_t total, uint32_t* list, uint32_t* nlist, hipStream_t st) {
  if (ni)
    LV_LAUNCH(ht_new_hubs_kernel, grid_for(ni, kBlock), kBlock, st, iu, ni, deg, cap, off, next, total, list, nlist);
}
void k_ht_caps(const uint32_t* deg, uint64_t S, uint32_t* cap, uint64_t* cap64, uint8_t* flags, hipStream_t st) {
  if (S) LV_LAUNCH(ht_caps_kernel, grid_for(S, kBlock * 4), kBlock, st, deg, S, cap, cap64, flags);
}
void k_ht_build(GView g, bool f64, const uint32_t* vcom, const uint32_t* list, const uint32_t* n_dev,
                const uint64_t* off, uint32_t* cap, uint32_t* key, int32_t* cnt, double* w,
                unsigned long long* ndrop, hipStream_t st) {
  if (f64) LV_LAUNCH((ht_build_kernel<double, false>), 4096, kBlock, st, g, vcom, list, n_dev, off, cap, key, cnt, w,
                     nullptr, ndrop);
  else LV_LAUNCH((ht_build_kernel<float, false>), 4096, kBlock, st, g, vcom, list, n_dev, off, cap, key, cnt, w,
                 nullptr, ndrop);
}
void k_ht_check(GView g, bool f64, const uint32_t* vcom, const uint32_t* list, const uint32_t* n_dev,
                const uint64_t* off, uint32_t* cap, uint32_t* key, int32_t* cnt, double* w, unsigned long long* out,
                hipStream_t st) {
  if (f64) LV_LAUNCH((ht_build_kernel<double, true>), 4096, kBlock, st, g, vcom, list, n_dev, off, cap, key, cnt, w,
                     out, nullptr);
  else LV_LAUNCH((ht_build_kernel<float, true>), 4096, kBlock, st, g, vcom, list, n_dev, off, cap, key, cnt, w, out,
                 nullptr);
}
void k_ht_batch(const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd, const uint32_t* iu,
                const uint32_t* iv, const float* iw, uint64_t ni, const uint32_t* vcom, const uint64_t* off,
                uint32_t* cap, uint32_t* key, int32_t* cnt, double* w, unsigned long long* ndrop, hipStream_t st) {
  if (nd + ni)
    LV_LAUNCH(ht_batch_kernel, grid_for(nd + ni, kBlock), kBlock, st, du, dv, dw, nd, iu, iv, iw, ni, vcom, off, cap,
              key, cnt, w, ndrop);
}
void k_exists(GView g, const uint32_t* vcom, uint32_t* flags, hipStream_t st) {
  if (g.S) LV_LAUNCH(exists_kernel, grid_for(g.S, kBlock * 4), kBlock, st, g, vcom, flags);
}
void k_lookup(uint32_t* a, uint64_t n, const uint32_t* map, hipStream_t st) {
  if (n) LV_LAUNCH(lookup_kernel, grid_for(n, kBlock * 4), kBlock, st, a, n, map);
}
void k_member_bounds(const uint32_t* sorted_keys, uint64_t n, uint32_t C, uint32_t* coff, hipStream_t st) {
  LV_LAUNCH(member_bounds_kernel, grid_for(n + 1, kBlock * 4), kBlock, st, sorted_keys, n, C, coff);
}
void k_member_degrees(const uint32_t* cedg, uint64_t n, GView g, uint64_t* mdeg, hipStream_t st) {
  LV_LAUNCH(member_degrees_kernel, grid_for(n + 1, kBlock * 4), kBlock, st, cedg, n, g, mdeg);
}
void k_gather_yoff(const uint32_t* coff, const uint64_t* mscan, uint64_t C, uint64_t* yoff, hipStream_t st) {
  LV_LAUNCH(gather_yoff_kernel, grid_for(C + 1, kBlock * 4), kBlock, st, coff, mscan, C, yoff);
}
void k_tier_count(const uint64_t* off, const uint32_t* deg, const uint8_t* ex, uint64_t n, uint32_t thr,
                  uint32_t* block_counts, uint32_t nblocks, hipStream_t st) {
  LV_LAUNCH(tier_count_kernel, nblocks, kBlock, st, off, deg, ex, n, thr, block_counts, nblocks);
}
void k_tier_scatter(const uint64_t* off, const uint32_t* deg, const uint8_t* ex, uint64_t n, uint32_t thr,
                    const uint32_t* block_base, uint32_t nblocks, uint32_t* ids, hipStream_t st) {
  LV_LAUNCH(tier_scatter_kernel, nblocks, kBlock, st, off, deg, ex, n, thr, block_base, nblocks, ids);
}

template <class WT>
static void agg_tier(const AggArgs& a, int tier, hipStream_t st) {
  switch (tier) {
    case 0: LV_LAUNCH((aggregate_kernel<256, 8, 16, WT>), grid_for(a.n, 32, 1u << 16), 256, st, a); break;
    case 1: LV_LAUNCH((aggregate_kernel<256, 16, 64, WT>), grid_for(a.n, 16, 1u << 16), 256, st, a); break;
    case 2: LV_LAUNCH((aggregate_kernel<256, 64, 256, WT>), grid_for(a.n, 4, 1u << 16), 256, st, a); break;
    case 3: LV_LAUNCH((aggregate_kernel<256, 256, 2048, WT>), grid_for(a.n, 1, 1u << 14), 256, st, a); break;
    default: LV_LAUNCH((aggregate_hbm_kernel<WT>), grid_for(a.nchunks, 1, 1u << 16), kBlock, st, a); break;
  }
}
void k_aggregate(const AggArgs& a, int tier, bool f64, hipStream_t st) {
  if (!a.n) return;
  if (f64) agg_tier<double>(a, tier, st);
  else agg_tier<float>(a, tier, st);
}
void k_agg_mark_hbm(const AggArgs& a, uint64_t total, uint32_t* bm, uint64_t words, uint32_t* cnt, hipStream_t st) {
  if (total) LV_LAUNCH(agg_mark_hbm_kernel, grid_for(total, kBlock, 1u << 16), kBlock, st, a, total, bm);
  LV_LAUNCH(popc_words_kernel, grid_for(words + 1, kBlock, 1u << 16), kBlock, st, bm, words, cnt);
}
void k_agg_emit_hbm(const AggArgs& a, uint64_t total, const uint32_t* bm, const uint32_t* bsc, hipStream_t st) {
  if (total) LV_LAUNCH(agg_emit_hbm_kernel, grid_for(total, kBlock, 1u << 16), kBlock, st, a, total, bm, bsc);
  if (a.n) LV_LAUNCH(agg_deg_hbm_kernel, grid_for(a.n, kBlock, 1u << 16), kBlock, st, a, bsc);
}
void k_member_keys(GView g, const uint32_t* vcom, uint32_t C, uint32_t* key, uint32_t* id, hipStream_t st) {
  if (g.S) LV_LAUNCH(member_keys_kernel, grid_for(g.S, kBlock * 4), kBlock, st, g, vcom, C, key, id);
}
hipError_t sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                          int end_bit, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, st);
}
void k_t4_sizes(const uint64_t* yoff, const uint32_t* list, uint64_t n, uint64_t C, uint64_t chunk, uint64_t* caps,
                uint64_t* chunks, hipStream_t st) {
  LV_LAUNCH(t4_sizes_kernel, grid_for(n + 1, kBlock), kBlock, st, yoff, list, n, C, chunk, caps, chunks);
}
void k_bin_queue(const uint32_t* q, const uint32_t* qlen, uint64_t qcap, const uint32_t* deg, uint32_t thr,
                 uint32_t* tq, uint64_t stride, uint32_t* tlen, hipStream_t st) {
  LV_LAUNCH(bin_queue_kernel, grid_for(qcap, kBlock, 2048), kBlock, st, q, qlen, deg, thr, tq, stride, tlen);
}
void k_max_deg(const uint32_t* ids, uint64_t n, const uint32_t* deg, uint32_t* out, hipStream_t st) {
  if (n) LV_LAUNCH(max_deg_kernel, grid_for(n, kBlock, kReduceGrid), kBlock, st, ids, n, deg, out);
}
void k_max_u32(const uint32_t* a, uint64_t n, uint32_t* out, hipStream_t st) {
  if (n) LV_LAUNCH(max_u32_kernel, grid_for(n, kBlock * 8, kReduceGrid), kBlock, st, a, n, out);
}
void k_gather_u32(const uint32_t* a, const uint32_t* idx, uint64_t n, uint32_t* out, hipStream_t st) {
  if (n) LV_LAUNCH(gather_u32_kernel, grid_for(n, kBlock, 1024), kBlock, st, a, idx, n, out);
}
void k_caps_sum(const uint32_t* deg, uint64_t n, uint32_t thr, unsigned long long* out, hipStream_t st) {
  if (n) LV_LAUNCH(caps_sum_kernel, grid_for(n, kBlock * 8, kReduceGrid), kBlock, st, deg, n, thr, out);
}
// The two compaction kernels with min(512, segment / 16) threads a block:
// a segment of up to 8192 ids is one tile.  (1024 threads would cover 16384
// ids a tile, but the scatter kernel then spills registers.)
template <bool BITS>
static void wl_launch(void* f, bool clear, uint64_t S, const uint32_t* deg, const WlSplit& sp, const WlQueues& qs,
                      uint32_t* lens, uint64_t* bsum, hipStream_t st) {
  if (sp.shift >= 13) {
    LV_LAUNCH((wl_count_kernel<BITS, 512>), kWlScanBlocks, 512, st, f, S, deg, sp, bsum);
    LV_LAUNCH((wl_scatter_kernel<BITS, 512>), kWlScanBlocks, 512, st, f, (int)clear, S, deg, sp, bsum, qs, lens);
  } else {
    LV_LAUNCH((wl_count_kernel<BITS, 256>), kWlScanBlocks, 256, st, f, S, deg, sp, bsum);
    LV_LAUNCH((wl_scatter_kernel<BITS, 256>), kWlScanBlocks, 256, st, f, (int)clear, S, deg, sp, bsum, qs, lens);
  }
}
void k_wl_queue(uint8_t* f, bool clear, uint64_t S, const uint32_t* deg, const WlSplit& sp_in, const WlQueues& qs,
                uint32_t* lens, uint64_t* bsum, hipStream_t st) {
  WlSplit sp = sp_in;
  sp.shift = wl_shift(S);
  wl_launch<false>(f, clear, S, deg, sp, qs, lens, bsum, st);
}
void k_wl_queue_bits(uint32_t* bits, bool clear, uint64_t S, const uint32_t* deg, const WlSplit& sp_in,
                     const WlQueues& qs, uint32_t* lens, uint64_t* bsum, hipStream_t st) {
  WlSplit sp = sp_in;
  sp.shift = wl_shift(S);
  wl_launch<true>(bits, clear, S, deg, sp, qs, lens, bsum, st);
}
void k_modularity_parts(GView g, bool f64, const uint32_t* memb, double* cin, double* ctot, hipStream_t st) {
  if (!g.S) return;
  uint64_t grid = grid_for(g.S, kBlock, 1u << 16);
  if (f64) LV_LAUNCH(modularity_parts_kernel<double>, grid, kBlock, st, g, memb, cin, ctot);
  else LV_LAUNCH(modularity_parts_kernel<float>, grid, kBlock, st, g, memb, cin, ctot);
}
void k_modularity_sum(const double* cin, const double* ctot, uint64_t n, double M, double R, double* out,
                      hipStream_t st) {
  if (n) LV_LAUNCH(modularity_sum_kernel, grid_for(n, kBlock * 8, kReduceGrid), kBlock, st, cin, ctot, n, M, R, out);
}

hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), st);
}
hipError_t scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), st);
}

}  // namespace lv
