// Device-side batch tidy and graph update, and the per-graph helpers of the
// batched form (gfx950).
//
// tidyBatchUpdateU (inc/batch.hxx:202-209): deletions are kept when the edge
// exists, insertions when it does not (filterEdgesByExistenceU, :154-162),
// then each list is sorted by (u, v) (:170-177) and made unique (:185-193).
// The sort here is a stable radix sort, so of repeated (u, v) entries the
// first in input order is kept.
//
// applyBatchUpdateU (inc/batch.hxx:225-236, inc/update.hxx:130-150): the
// deletions leave their rows, the insertions join them, rows stay sorted by
// neighbour id (inc/_bitset.hxx:245-262), and an insertion's endpoints
// become existing vertices (inc/Graph.hxx:318-321).  Rows live in a gapped
// CSR: a touched row is rebuilt in scratch and copied back when its new
// degree fits its capacity; otherwise every row is moved to a new CSR with
// slack.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "lv_internal.hpp"

namespace lv {
namespace {

constexpr int kB = 256;

#define LV_UPD_LAUNCH(kernel, grid, block, st, ...) \
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, __VA_ARGS__)

__host__ __device__ inline uint64_t blocks_for(uint64_t n, uint64_t per, uint64_t cap = 1u << 16) {
  uint64_t g = (n + per - 1) / per;
  return g < 1 ? 1 : (g < cap ? g : cap);
}

// First position in keys[o, o + d) holding a key >= v (rows are sorted).
__device__ __forceinline__ uint64_t row_lower(const uint32_t* keys, uint64_t o, uint64_t d, uint32_t v) {
  uint64_t lo = 0, hi = d;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (keys[o + mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// First index in a[0, n) holding a value >= v (a ascending).
template <class T>
__device__ __forceinline__ uint64_t lower_idx(const T* a, uint64_t n, T v) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Largest r in [0, n) with a[r] <= x (a ascending, a[0] <= x).
__device__ __forceinline__ uint64_t owner_of(const uint64_t* a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint64_t mid = (lo + hi + 1) >> 1;
    if (a[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Sort key of an entry: (u << 32 | v) when it is kept, all ones when the
// edge's existence disqualifies it; ids beyond the span set bad.
__global__ void tidy_keys_kernel(GView g, const uint32_t* u, const uint32_t* v, uint64_t n, int keep_existing,
                                 uint64_t* key, uint32_t* idx, int* bad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t a = u[i], b = v[i];
    uint64_t k = ~0ull;
    if (a >= g.S || b >= g.S) {
      atomicOr(bad, 1);
    } else {
      const uint64_t o = g.off[a], d = g.deg[a];
      const uint64_t p = row_lower(g.keys, o, d, b);
      const bool exists = (!g.ex || g.ex[a]) && p < d && g.keys[o + p] == b;
      if (exists == (keep_existing != 0)) k = ((uint64_t)a << 32) | b;
    }
    key[i] = k;
    idx[i] = (uint32_t)i;
  }
}

// First entry of every run of equal sorted keys, disqualified entries out.
__global__ void tidy_flags_kernel(const uint64_t* key, uint64_t n, uint8_t* flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = key[i] != ~0ull && (i == 0 || key[i] != key[i - 1]);
}

__global__ void tidy_gather_kernel(const uint32_t* sel, uint64_t m, const uint32_t* u, const uint32_t* v,
                                   const float* w, uint32_t* ou, uint32_t* ov, float* ow) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = sel[i];
    ou[i] = u[j];
    ov[i] = v[j];
    ow[i] = w[j];
  }
}

// Degree change per row and the touched-row flags.
__global__ void upd_delta_kernel(const uint32_t* u, uint64_t n, int s, int* delta, uint8_t* touched) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    atomicAdd(&delta[u[i]], s);
    touched[u[i]] = 1;
  }
}

// Touched row r: old live degree, new degree; bit 1 of bad when the new
// degree exceeds the row's capacity.  Entry R of both is 0 (scan totals).
__global__ void upd_sizes_kernel(const uint32_t* T, uint64_t R, const uint64_t* off, const uint32_t* deg,
                                 const int* delta, uint64_t* so, uint64_t* sn, int* bad) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r <= R; r += (uint64_t)gridDim.x * blockDim.x) {
    if (r == R) {
      so[r] = 0;
      sn[r] = 0;
      continue;
    }
    const uint32_t u = T[r];
    const uint64_t d = deg[u], dn = (uint64_t)((int64_t)d + delta[u]);
    if (dn > off[u + 1] - off[u]) atomicOr(bad, 2);
    so[r] = d;
    sn[r] = dn;
  }
}

// keep[bo[r] + p] = 0 for the deleted entry p of touched row r; the
// graph's weight of the deleted edge goes to delw[i] (2M loses it).
__global__ void upd_tomb_kernel(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* T, uint64_t R,
                                const uint64_t* bo, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                                const float* w, uint32_t* keep, float* delw) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nd; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = du[i];
    const uint64_t r = lower_idx(T, R, u);
    const uint64_t p = row_lower(keys, off[u], deg[u], dv[i]);
    keep[bo[r] + p] = 0;
    delw[i] = w[off[u] + p];
  }
}

// Surviving old entries of the touched rows, to their merged positions in
// scratch: rank among the row's survivors + insertions of the row below it.
template <class WT>
__global__ void upd_scatter_old_kernel(const uint32_t* T, uint64_t R, const uint64_t* bo, const uint64_t* bn,
                                       uint64_t total, const uint64_t* off, const uint32_t* keys, const WT* w,
                                       const uint32_t* keep, const uint32_t* rank, const uint32_t* iu,
                                       const uint32_t* iv, uint64_t ni, uint32_t* yk, WT* yw) {
  for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
    if (!keep[x]) continue;
    const uint64_t r = owner_of(bo, R, x);
    const uint32_t u = T[r];
    const uint64_t k = x - bo[r];
    const uint32_t key = keys[off[u] + k];
    const uint64_t is = lower_idx(iu, ni, u), ie = lower_idx(iu, ni, u + 1);
    const uint64_t below = lower_idx(iv + is, ie - is, key);
    const uint64_t pos = (rank[x] - rank[bo[r]]) + below;
    yk[bn[r] + pos] = key;
    yw[bn[r] + pos] = w[off[u] + k];
  }
}

// Insertions to their merged positions: index among the row's insertions +
// survivors of the row below the new key.
template <class WT>
__global__ void upd_scatter_ins_kernel(const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni,
                                       const uint32_t* T, uint64_t R, const uint64_t* bo, const uint64_t* bn,
                                       const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                                       const uint32_t* rank, uint32_t* yk, WT* yw) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ni; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = iu[j];
    const uint64_t r = lower_idx(T, R, u);
    const uint64_t is = lower_idx(iu, ni, u);
    const uint64_t p = row_lower(keys, off[u], deg[u], iv[j]);
    const uint64_t pos = (j - is) + (rank[bo[r] + p] - rank[bo[r]]);
    yk[bn[r] + pos] = iv[j];
    yw[bn[r] + pos] = (WT)iw[j];
  }
}

// Merged rows from scratch into the graph (offsets dst_off).
template <class WT>
__global__ void upd_copy_back_kernel(const uint32_t* T, uint64_t R, const uint64_t* bn, uint64_t total,
                                     const uint64_t* dst_off, const uint32_t* yk, const WT* yw, uint32_t* keys,
                                     WT* w) {
  for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = owner_of(bn, R, x);
    const uint32_t u = T[r];
    const uint64_t k = x - bn[r];
    keys[dst_off[u] + k] = yk[x];
    w[dst_off[u] + k] = yw[x];
  }
}

// New degrees of the touched rows; delta and the touched flags are reset.
__global__ void upd_degrees_kernel(const uint32_t* T, uint64_t R, int* delta, uint32_t* deg, uint8_t* touched) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = T[r];
    deg[u] = (uint32_t)((int64_t)deg[u] + delta[u]);
    delta[u] = 0;
    touched[u] = 0;
  }
}

// Rebuild: capacity of every row = new degree + max(4, new degree / 8).
__global__ void upd_caps_kernel(const uint32_t* deg, const int* delta, uint64_t S, uint64_t* cap) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u <= S; u += (uint64_t)gridDim.x * blockDim.x) {
    if (u == S) {
      cap[u] = 0;
      continue;
    }
    const uint64_t dn = (uint64_t)((int64_t)deg[u] + delta[u]);
    cap[u] = dn + (dn / 8 > 4 ? dn / 8 : 4);
  }
}

// Rebuild: rows no batch entry touches move unchanged (one wave per row).
template <class WT>
__global__ void upd_copy_rows_kernel(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys,
                                     const WT* w, const uint8_t* touched, const uint64_t* noff, uint32_t* nkeys,
                                     WT* nw) {
  const int lane = threadIdx.x % 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t u = blockIdx.x * (uint64_t)(blockDim.x / 64) + threadIdx.x / 64; u < S; u += nwaves) {
    if (touched[u]) continue;
    const uint64_t o = off[u], n = noff[u];
    for (uint64_t i = lane; i < deg[u]; i += 64) {
      nkeys[n + i] = keys[o + i];
      nw[n + i] = w[o + i];
    }
  }
}

// Endpoints of insertions exist afterwards.
__global__ void upd_exists_kernel(const uint32_t* iu, const uint32_t* iv, uint64_t ni, uint8_t* ex) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ni; i += (uint64_t)gridDim.x * blockDim.x) {
    ex[iu[i]] = 1;
    ex[iv[i]] = 1;
  }
}

// out += sum of w (f64 accumulation).
__global__ void upd_wsum_kernel(const float* w, uint64_t n, double sign, double* out) {
  __shared__ double s[kB / 64];
  double acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    acc += (double)w[i];
  for (int m = 32; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
  if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < kB / 64; ++i) t += s[i];
    if (t != 0) unsafeAtomicAdd(out, sign * t);
  }
}

// ---------------------------------------------------------- batched form
// Graph of every vertex: the k with voff[k] <= u < voff[k + 1].
__global__ void gid_fill_kernel(const uint64_t* voff, uint64_t G, uint64_t S, uint32_t* gid) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < S; u += (uint64_t)gridDim.x * blockDim.x)
    gid[u] = (uint32_t)owner_of(voff, G, u);
}

// out[gid[u]] += x[u] (f64), one atomic per run of equal graphs in a wave.
__global__ void group_sum_kernel(const double* x, const uint32_t* gid, uint64_t S, double* out) {
  const int lane = threadIdx.x % 64;
  for (uint64_t b = (blockIdx.x * (uint64_t)blockDim.x) & ~63ull; b < S; b += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = b + (threadIdx.x & ~63u) + lane;
    const bool on = u < S;
    const uint32_t gk = on ? gid[u] : 0xFFFFFFFFu;
    double v = on ? x[u] : 0.0;
    const uint32_t g0 = __shfl(gk, 0, 64);
    if (__all(gk == g0 || !on)) {  // the whole wave in one graph
      for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
      if (lane == 0 && g0 != 0xFFFFFFFFu) unsafeAtomicAdd(&out[g0], v);
    } else if (on) {
      unsafeAtomicAdd(&out[gk], v);
    }
  }
}

// Per-community modularity terms of a union of graphs, in place: community
// c (a vertex id, so in the range of graph gid[c]) gets
// cin[c] / 2M - R (ctot[c] / 2M)^2 with its own graph's M = gM[gid[c]]
// (modularityByOmp, inc/properties.hxx:205-233, per graph).
__global__ void group_modularity_kernel(double* cin, const double* ctot, const uint32_t* gid, const double* gM,
                                        uint64_t S, double R) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < S; c += (uint64_t)gridDim.x * blockDim.x) {
    const double m = gM[gid[c]];
    const double x = m > 0 ? ctot[c] / (2 * m) : 0.0;
    cin[c] = m > 0 ? cin[c] / (2 * m) - R * x * x : 0.0;
  }
}

// Graph of the next level's super-vertices: gid_next[map[vcom[u]]] = gid[u].
__global__ void gid_next_kernel(const uint32_t* vcom, const uint32_t* gid, const uint32_t* map, const uint8_t* ex,
                                uint64_t S, uint32_t* gid_next) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < S; u += (uint64_t)gridDim.x * blockDim.x)
    if (!ex || ex[u]) gid_next[map[vcom[u]]] = gid[u];
}

// Final community ids of the graphs flagged in fin, re-based from their
// range of the current level (lo) to their own vertex range (voff): the
// membership of their original vertices (the level id of u is vcom[ucom[u]],
// or ucom[u] at the first level) and Sigma_tot of their communities
// (ftot[voff + j - lo] = ctot[j] for level id j).
__global__ void finalize_memb_kernel(uint32_t* ucom, const uint32_t* vcom, const uint32_t* gid, const uint8_t* ex,
                                     uint64_t S, const uint8_t* fin, const uint64_t* lo, const uint64_t* voff) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < S; u += (uint64_t)gridDim.x * blockDim.x) {
    if (ex && !ex[u]) continue;
    const uint32_t k = gid[u];
    if (!fin[k]) continue;
    const uint32_t id = vcom ? vcom[ucom[u]] : ucom[u];
    ucom[u] = (uint32_t)(id - lo[k] + voff[k]);
  }
}
__global__ void finalize_tot_kernel(const double* ctot, uint64_t SL, const uint64_t* lo, uint64_t G,
                                    const uint8_t* fin, const uint64_t* voff, double* ftot) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < SL; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = owner_of(lo, G, j);
    if (fin[k]) ftot[voff[k] + (j - lo[k])] = ctot[j];
  }
}

// a[i] = map[a[i]] except for entries of the graphs flagged in skip.
__global__ void lookup_active_kernel(uint32_t* a, uint64_t n, const uint32_t* map, const uint32_t* gid,
                                     const uint8_t* skip) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (!skip[gid[i]]) a[i] = map[a[i]];
}

// Level vertices of the graphs still running (and existing, when ex is set).
__global__ void active_mask_kernel(const uint32_t* gid, uint64_t SL, const uint8_t* ex, const uint8_t* done,
                                   uint8_t* mask) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < SL; v += (uint64_t)gridDim.x * blockDim.x)
    mask[v] = (!ex || ex[v]) && !done[gid[v]];
}

// Flagged entries of each graph range [voff[k], voff[k+1]): block b takes
// slice b % kRangeSlices of graph b / kRangeSlices (every kRangeSlices-th
// kB-entry tile of the range), sums it and adds the sum to out[k] (zeroed
// by the launcher), so a large graph's range is read by many blocks.
constexpr uint64_t kRangeSlices = 32;
__global__ __launch_bounds__(kB) void range_count_kernel(const uint8_t* f, const uint64_t* voff, uint64_t G,
                                                         unsigned long long* out) {
  __shared__ unsigned long long s_part[kB / 64];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  for (uint64_t b = blockIdx.x; b < G * kRangeSlices; b += gridDim.x) {  // block-uniform
    const uint64_t k = b / kRangeSlices, j = b % kRangeSlices;
    const uint64_t lo = voff[k], hi = voff[k + 1];
    unsigned long long n = 0;
    for (uint64_t u = lo + j * kB + threadIdx.x; u < hi; u += kRangeSlices * kB) n += f[u] != 0;
    for (int m = 32; m > 0; m >>= 1) n += __shfl_xor(n, m, 64);
    if (lane == 0) s_part[wv] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int i = 0; i < kB / 64; ++i) t += s_part[i];
      if (t) atomicAdd(&out[k], t);
    }
    __syncthreads();
  }
}

// List entries [e0, e1) of chunk t (t < a.nct) of the batched id-order schedule.
__device__ __forceinline__ void chunk_bounds(const MoveArgs& a, uint64_t t, uint64_t& e0, uint64_t& e1) {
  const uint64_t n = a.n_dev ? (uint64_t)*a.n_dev : a.n;
  const uint64_t k = owner_of(a.gcs, a.ng, t);
  auto bound = [&](uint64_t id) -> uint64_t {  // first entry with id >= `id`
    if (!a.list) return id <= a.base ? 0 : min(n, id - a.base);
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (a.list[mid] < id) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const uint64_t b0 = bound(a.glo[k]), b1 = bound(a.glo[k + 1]);
  const uint64_t len = b1 - b0, s = a.gcs[k + 1] - a.gcs[k], j = t - a.gcs[k];
  const uint64_t chunk = max<uint64_t>(1, (len + s - 1) / s);
  e0 = b0 + min(len, j * chunk);
  e1 = b0 + min(len, (j + 1) * chunk);
}

// Chunk t of the batched id-order schedule: graph k owns chunks
// gcs[k] .. gcs[k+1]-1 and list entries [b0, b1) (ids glo[k] .. glo[k+1]-1;
// an ascending list is searched, an id range is offset by base); its entries
// are cut as the kernel cuts a list alone: ceil(len / streams) per chunk.
// With a.nct_dev the chunks that hold entries take the first places of the
// table, a wave's at a time (whole waves run the kernel: a.nct_dev counts them).
__global__ void graph_chunks_kernel(MoveArgs a) {
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (a.nct_dev) {
    uint64_t e0 = 0, e1 = 0;
    if (t < a.nct) chunk_bounds(a, t, e0, e1);
    const bool has = e1 > e0;
    const unsigned long long m = __ballot(has);
    if (!m) return;
    const int lane = __lane_id(), lead = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == lead) base = atomicAdd(a.nct_dev, (uint32_t)__popcll(m));
    base = __shfl(base, lead, 64);
    if (has) {
      const uint64_t p = base + (uint64_t)__popcll(m & ((1ull << lane) - 1));
      a.ctab[2 * p] = e0;
      a.ctab[2 * p + 1] = e1;
    }
    return;
  }
  if (t >= a.nct) return;
  uint64_t e0 = 0, e1 = 0;
  chunk_bounds(a, t, e0, e1);
  a.ctab[2 * t] = e0;
  a.ctab[2 * t + 1] = e1;
}

}  // namespace

void k_gid_fill(const uint64_t* voff, uint64_t G, uint64_t S, uint32_t* gid, hipStream_t st) {
  if (S) LV_UPD_LAUNCH(gid_fill_kernel, blocks_for(S, kB), kB, st, voff, G, S, gid);
}
void k_group_sum(const double* x, const uint32_t* gid, uint64_t S, double* out, hipStream_t st) {
  if (S) LV_UPD_LAUNCH(group_sum_kernel, blocks_for(S, kB), kB, st, x, gid, S, out);
}
void k_group_modularity(double* cin, const double* ctot, const uint32_t* gid, const double* gM, uint64_t S, double R,
                        hipStream_t st) {
  if (S) LV_UPD_LAUNCH(group_modularity_kernel, blocks_for(S, kB), kB, st, cin, ctot, gid, gM, S, R);
}
void k_gid_next(const uint32_t* vcom, const uint32_t* gid, const uint32_t* map, const uint8_t* ex, uint64_t S,
                uint32_t* gid_next, hipStream_t st) {
  if (S) LV_UPD_LAUNCH(gid_next_kernel, blocks_for(S, kB), kB, st, vcom, gid, map, ex, S, gid_next);
}
void k_finalize(uint32_t* ucom, const uint32_t* vcom, const uint32_t* gid0, const uint8_t* ex, uint64_t S,
                const double* ctot, uint64_t SL, const uint64_t* lo, uint64_t G, const uint8_t* fin,
                const uint64_t* voff, double* ftot, hipStream_t st) {
  if (S) LV_UPD_LAUNCH(finalize_memb_kernel, blocks_for(S, kB), kB, st, ucom, vcom, gid0, ex, S, fin, lo, voff);
  if (SL) LV_UPD_LAUNCH(finalize_tot_kernel, blocks_for(SL, kB), kB, st, ctot, SL, lo, G, fin, voff, ftot);
}
void k_lookup_active(uint32_t* a, uint64_t n, const uint32_t* map, const uint32_t* gid, const uint8_t* skip,
                     hipStream_t st) {
  if (n) LV_UPD_LAUNCH(lookup_active_kernel, blocks_for(n, kB), kB, st, a, n, map, gid, skip);
}
void k_active_mask(const uint32_t* gid, uint64_t SL, const uint8_t* ex, const uint8_t* done, uint8_t* mask,
                   hipStream_t st) {
  if (SL) LV_UPD_LAUNCH(active_mask_kernel, blocks_for(SL, kB), kB, st, gid, SL, ex, done, mask);
}
void k_range_count(const uint8_t* f, const uint64_t* voff, uint64_t G, unsigned long long* out, hipStream_t st) {
  if (!G) return;
  k_fill_u64(reinterpret_cast<uint64_t*>(out), 0, G, st);
  LV_UPD_LAUNCH(range_count_kernel, blocks_for(G * kRangeSlices, 1), kB, st, f, voff, G, out);
}

void k_graph_chunks(const MoveArgs& a, hipStream_t st) {
  if (a.nct) LV_UPD_LAUNCH(graph_chunks_kernel, blocks_for(a.nct, kB), kB, st, a);
}

void k_tidy_keys(GView g, const uint32_t* u, const uint32_t* v, uint64_t n, bool keep_existing, uint64_t* key,
                 uint32_t* idx, int* bad, hipStream_t st) {
  if (n) LV_UPD_LAUNCH(tidy_keys_kernel, blocks_for(n, kB), kB, st, g, u, v, n, keep_existing ? 1 : 0, key, idx, bad);
}
hipError_t sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                          void* tmp, size_t& bytes, hipStream_t st) {
  return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (size_t)n, 0u, 64u, st);
}
void k_tidy_flags(const uint64_t* key, uint64_t n, uint8_t* flag, hipStream_t st) {
  if (n) LV_UPD_LAUNCH(tidy_flags_kernel, blocks_for(n, kB), kB, st, key, n, flag);
}
hipError_t select_flagged_u32(const uint32_t* in, const uint8_t* flags, uint32_t* out, uint32_t* count, uint64_t n,
                              void* tmp, size_t& bytes, hipStream_t st) {
  return rocprim::select(tmp, bytes, in, flags, out, count, (size_t)n, st);
}
void k_tidy_gather(const uint32_t* sel, uint64_t m, const uint32_t* u, const uint32_t* v, const float* w,
                   uint32_t* ou, uint32_t* ov, float* ow, hipStream_t st) {
  if (m) LV_UPD_LAUNCH(tidy_gather_kernel, blocks_for(m, kB), kB, st, sel, m, u, v, w, ou, ov, ow);
}
void k_upd_delta(const uint32_t* u, uint64_t n, int s, int* delta, uint8_t* touched, hipStream_t st) {
  if (n) LV_UPD_LAUNCH(upd_delta_kernel, blocks_for(n, kB), kB, st, u, n, s, delta, touched);
}
void k_upd_sizes(const uint32_t* T, uint64_t R, const uint64_t* off, const uint32_t* deg, const int* delta,
                 uint64_t* so, uint64_t* sn, int* bad, hipStream_t st) {
  LV_UPD_LAUNCH(upd_sizes_kernel, blocks_for(R + 1, kB), kB, st, T, R, off, deg, delta, so, sn, bad);
}
void k_upd_tomb(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* T, uint64_t R,
                const uint64_t* bo, const uint64_t* off, const uint32_t* deg, const uint32_t* keys, const float* w,
                uint32_t* keep, float* delw, hipStream_t st) {
  if (nd)
    LV_UPD_LAUNCH(upd_tomb_kernel, blocks_for(nd, kB), kB, st, du, dv, nd, T, R, bo, off, deg, keys, w, keep, delw);
}
void k_upd_scatter(const uint32_t* T, uint64_t R, const uint64_t* bo, const uint64_t* bn, uint64_t total_old,
                   const uint64_t* off, const uint32_t* deg, const uint32_t* keys, const float* w,
                   const uint32_t* keep, const uint32_t* rank, const uint32_t* iu, const uint32_t* iv,
                   const float* iw, uint64_t ni, uint32_t* yk, float* yw, hipStream_t st) {
  if (total_old)
    LV_UPD_LAUNCH(upd_scatter_old_kernel<float>, blocks_for(total_old, kB), kB, st, T, R, bo, bn, total_old, off,
                  keys, w, keep, rank, iu, iv, ni, yk, yw);
  if (ni)
    LV_UPD_LAUNCH(upd_scatter_ins_kernel<float>, blocks_for(ni, kB), kB, st, iu, iv, iw, ni, T, R, bo, bn, off, deg,
                  keys, rank, yk, yw);
}
void k_upd_copy_back(const uint32_t* T, uint64_t R, const uint64_t* bn, uint64_t total, const uint64_t* dst_off,
                     const uint32_t* yk, const float* yw, uint32_t* keys, float* w, hipStream_t st) {
  if (total)
    LV_UPD_LAUNCH(upd_copy_back_kernel<float>, blocks_for(total, kB), kB, st, T, R, bn, total, dst_off, yk, yw, keys,
                  w);
}
void k_upd_degrees(const uint32_t* T, uint64_t R, int* delta, uint32_t* deg, uint8_t* touched, hipStream_t st) {
  if (R) LV_UPD_LAUNCH(upd_degrees_kernel, blocks_for(R, kB), kB, st, T, R, delta, deg, touched);
}
void k_upd_caps(const uint32_t* deg, const int* delta, uint64_t S, uint64_t* cap, hipStream_t st) {
  LV_UPD_LAUNCH(upd_caps_kernel, blocks_for(S + 1, kB), kB, st, deg, delta, S, cap);
}
void k_upd_copy_rows(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys, const float* w,
                     const uint8_t* touched, const uint64_t* noff, uint32_t* nkeys, float* nw, hipStream_t st) {
  if (S)
    LV_UPD_LAUNCH(upd_copy_rows_kernel<float>, blocks_for(S, kB / 64), kB, st, S, off, deg, keys, w, touched, noff,
                  nkeys, nw);
}
void k_upd_exists(const uint32_t* iu, const uint32_t* iv, uint64_t ni, uint8_t* ex, hipStream_t st) {
  if (ni) LV_UPD_LAUNCH(upd_exists_kernel, blocks_for(ni, kB), kB, st, iu, iv, ni, ex);
}
void k_upd_wsum(const float* w, uint64_t n, double sign, double* out, hipStream_t st) {
  if (n) LV_UPD_LAUNCH(upd_wsum_kernel, blocks_for(n, kB * 4, 1024), kB, st, w, n, sign, out);
}

}  // namespace lv
