// Host side of the MI355X Louvain library: the pass driver (the role of
// louvainInvokeOmp, inc/louvain.hxx:1120-1221), the four approaches
// (inc/louvain.hxx:1285, :1349, :1542, :1674) and the C ABI declared in
// include/louvain_hip.h.  All device work runs on the context's stream;
// exceptions are used internally and turned into lv_status codes at the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "lv_internal.hpp"

using namespace lv;

namespace {

struct LvError {
  int code;
  std::string msg;
};

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess)                                                                      \
      throw LvError{e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation ? LV_ERR_NOMEM \
                                                                                : LV_ERR_HIP,  \
                    std::string(#x) + ": " + hipGetErrorString(e_)};                           \
  } while (0)

void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw LvError{code, msg};
}

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Scratch of lv_dmodularity_noalloc: per-community internal weight and
// total weight (kNoAllocSpan doubles each), then the result.  The library
// holds this array on every GPU it runs on until the process ends.  Host
// code never passes the array's name to a kernel or a copy: it uses the
// device address hipGetSymbolAddress gives for the current device
// (scratch_address), looked up once per device.
constexpr uint64_t kNoAllocSpan = 1ull << 21;
constexpr uint64_t kNoAllocDoubles = 2 * kNoAllocSpan + 32;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of scratch on `device`, which must be current.
hipError_t scratch_address(int device, double** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace

struct lv_context {
  int device = 0;
  hipStream_t st = nullptr;
  double* scratch = nullptr;  // the device's lv_dmodularity_noalloc scratch (its address there)
  std::string err;
  int path = 2;
  // Local-moving streams per launch: clamp(stream_fraction * live vertices of
  // the level, min_streams, max_streams) lane groups, each walking dynamically
  // consecutive list entries sequentially.
  double stream_fraction = 0.002;
  uint64_t min_streams = 1;
  uint64_t max_streams = 1ull << 20;
  uint64_t chunk = 0;           // entries per dynamically assigned chunk (0: static spread)
  // worklist sub-rounds: at most wl_streams streams in the id-order kernel,
  // the queue interleaved over them (lv_context_set_worklist_schedule)
  uint64_t wl_streams = kWlStreams;
  int interleave = 0;
  // schedule tuner: choices by graph shape (lv_autotune), applied before
  // every run while autotune is on and no tuned knob was set by hand
  int autotune = 1;
  bool knobs_hand = false;  // a tuned knob set by hand: the knobs stay as set
  bool tuning = false;      // inside a tuning: trial runs keep the candidate's knobs
  std::map<std::string, lv_tune_choice> tuned;
  // Sweep schedules by pass: the first lead_iters iterations of the first
  // pass run lead_streams streams over ascending chunks of lead_chunk
  // entries; passes after the first run late_streams streams over chunks of
  // late_chunk entries handed out in ascending order (0: spread; no more
  // streams than the level's schedule gives) when late_streams is set.
  // Defaults: Static's first iteration sweeps the id prefix that holds a
  // third of the edges (at most kLeadIdsMax ids) in 16 streams of 2048-id
  // ascending chunks, the reference's own sweep pattern there, and the rest
  // of the first pass runs the wide schedule; later passes walk 2048-id
  // chunks in ascending order in 256 streams.  (Measured on R-MAT 22 Static:
  // spread chunks in the later passes gave 0.0660-0.0679, ascending ones
  // 0.0678-0.0689 over 5 runs each of 64, 256 and 1024 streams; the same
  // schedule over the first pass's later iterations gained nothing and took
  // 15-40% longer.)
  int lead_iters = 1;
  uint64_t lead_streams = 16, lead_chunk = 2048;
  uint64_t lead_ids = 0;  // ids under the lead schedule: 0 auto, ~0 all, else [0, lead_ids)
  uint64_t late_streams = 256, late_chunk = 2048;
  int pick_less_iters = 0;      // first iterations of each pass that only move to lower ids
  // validated moves of the id-order kernel: 0 off (default), 1 sweeps (every
  // pass but a DF / DS worklist first pass), 2 sweeps and worklists
  int validate = 0;
  // sweeps: hubs (degree > kHubMin) all at once before each sweep (0,
  // default) or visited in their id turn with tables of their own (1)
  int sweep_hubs = 0;
  // diagnostics (environment LV_TRACE set when the context is made): worklist
  // runs enqueue launch by launch and print every sub-round's queue lengths
  int trace = 0;
  uint64_t slabs = 1;           // id-order sweeps: consecutive id slabs launched in turn
  uint64_t hub_streams = kHubBlocks;  // hubs processed at once
  uint32_t hub_min = kHubMin;         // sweeps: degrees above run in the hub batch
  uint32_t hub_min_wl = 256;          // worklists: degrees above run in the hub batch
  // worklist hub-table slots of the last graph seen (its degree array, span, threshold)
  uint64_t wl_key_stamp = 0;  // graph stamp wl_slots was computed for (0: none)
  uint64_t wl_slots = 0;
  uint32_t wl_key_thr = 0;
  int tier_order = 2;           // 2: ascending ids (hubs first), 1: degree tiers highest first, 0: lowest first
  uint32_t serial = 0;          // worklist iteration serial (visit stamps)
  uint64_t stamp_span = ~0ull;
  int profiling = 0;             // time every local-moving launch with events
  lv_stats stats{};              // accumulated since lv_stats_reset (the work counters: on the device)
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof;  // (tier, events) of a run
  std::map<std::string, Buf> bufs;
  void* pinned = nullptr;  // small pinned host buffer for scalar readbacks
  unsigned long long* tail = nullptr;  // pinned: a run's closing read-back
  unsigned long long* tail_pin() {
    if (!tail) HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&tail), 64, hipHostMallocDefault));
    return tail;
  }
  size_t pinned_bytes = 0;
  std::vector<hipEvent_t> events;
  size_t ev_used = 0;
  size_t scan_tmp_bytes = 0;

  // Grow-only named device buffer; contents are unspecified (*fresh is set
  // when it was just allocated).  The buffers together may not exceed
  // ws_limit bytes: a growth beyond it fails with LV_ERR_NOMEM and keeps the
  // buffer as it was.
  size_t ws_total = 0;
  size_t ws_limit = ~size_t(0);
  template <class T>
  T* ws(const std::string& name, size_t count, bool* fresh = nullptr) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    Buf& b = bufs[name];
    if (fresh) *fresh = b.bytes < bytes;
    if (b.bytes < bytes) {
      size_t want = bytes + bytes / 8;  // headroom for slightly larger later calls
      if (ws_total - b.bytes + want > ws_limit)
        throw LvError{LV_ERR_NOMEM, "context workspace limit reached (buffer " + name + ")"};
      if (b.p) HIPCHK(hipFree(b.p));
      ws_total -= b.bytes;
      b.p = nullptr;
      b.bytes = 0;
      HIPCHK(hipMalloc(&b.p, want));
      b.bytes = want;
      ws_total += want;
    }
    return reinterpret_cast<T*>(b.p);
  }
  size_t ws_bytes(const std::string& name) { return bufs[name].bytes; }
  // Work counters of the local-moving kernels (kSpread slots of kCntStride),
  // accumulated on the device since the last lv_stats_reset and read by
  // lv_stats_get only, so a run reads none of them back.
  unsigned long long* counters() {
    bool fresh = false;
    auto* p = ws<unsigned long long>("cnt", (size_t)kSpread * kCntStride, &fresh);
    if (fresh) k_fill_u64(reinterpret_cast<uint64_t*>(p), 0, ws_bytes("cnt") / 8, st);
    return p;
  }
  void* scan_tmp(size_t bytes) {
    scan_tmp_bytes = std::max(scan_tmp_bytes, bytes);
    return ws<uint8_t>("scan_tmp", scan_tmp_bytes);
  }
  template <class T>
  void read(T* host, const T* dev, size_t n) {
    size_t bytes = n * sizeof(T);
    if (pinned_bytes < bytes) {
      if (pinned) HIPCHK(hipHostFree(pinned));
      pinned = nullptr;
      HIPCHK(hipHostMalloc(&pinned, std::max<size_t>(bytes, 4096), hipHostMallocDefault));
      pinned_bytes = std::max<size_t>(bytes, 4096);
    }
    HIPCHK(hipMemcpyAsync(pinned, dev, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::memcpy(host, pinned, bytes);
  }
  // n values dev[0], dev[stride], ... with one synchronisation
  template <class T>
  void read_strided(T* host, const T* dev, size_t n, size_t stride) {
    if (pinned_bytes < n * sizeof(T)) {
      T tmp;
      read(&tmp, dev, 1);  // grows the pinned buffer
      if (pinned_bytes < n * sizeof(T)) throw LvError{LV_ERR_NOMEM, "pinned buffer"};
    }
    T* pin = reinterpret_cast<T*>(pinned);
    for (size_t i = 0; i < n; ++i)
      HIPCHK(hipMemcpyAsync(pin + i, dev + i * stride, sizeof(T), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::memcpy(host, pin, n * sizeof(T));
  }
  template <class T>
  T read1(const T* dev) {
    T v;
    read(&v, dev, 1);
    return v;
  }
  hipEvent_t event() {
    if (ev_used == events.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      events.push_back(e);
    }
    return events[ev_used++];
  }
  hipEvent_t record(hipStream_t s = nullptr) {
    hipEvent_t e = event();
    HIPCHK(hipEventRecord(e, s ? s : st));
    return e;
  }
  // The start event of a timed launch: a 40 us hold first, so the device
  // meets the event only once the launch behind it is queued too.
  hipEvent_t record_start(hipStream_t s = nullptr) {
    k_hold(40, s ? s : st);
    return record(s);
  }
  // Side streams of the worklist sub-rounds (created on first use, default
  // priority): the hub kernels run there beside the low-degree kernel on st.
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};
  void make_side() {
    if (side[0]) return;
    for (int k = 0; k < 2; ++k) {
      HIPCHK(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&join_ev[k], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
  }
  // Worklist sub-round groups replayed from captured graphs (one launch
  // instead of some fifty, and no host gaps between them on the device).  A
  // graph is keyed by everything its kernels and their launch shapes depend
  // on.  A new key is captured and, when a graph of its shape (the same
  // launches: sub-rounds per group, hub kernels, tables) exists already,
  // written into the most recent one of them (hipGraphExecUpdate: other
  // arguments only), else instantiated -- the latter only for a key an
  // earlier call used too.  The last kGraphCache graphs are kept.
  static constexpr size_t kGraphCache = 8, kSeenKeys = 32;
  int use_graphs = 1;
  int use_tables = 1;  // hub neighbour-community tables with resident states
  struct GraphEntry {
    std::string key;
    int shape;  // launches of the same shape have the same topology (updatable in place)
    hipGraphExec_t ex;
  };
  std::vector<GraphEntry> graphs;  // most recent last
  std::vector<std::pair<std::string, uint64_t>> seen;          // (key, call)
  uint64_t wl_calls = 0;                                       // worklist calls so far
  // true if the group was launched from a graph; false: enqueue it directly
  template <class F>
  bool replay(const std::string& key, int shape, F enqueue) {
    for (size_t k = 0; k < graphs.size(); ++k) {
      if (graphs[k].key != key) continue;
      auto hit = graphs[k];
      graphs.erase(graphs.begin() + k);
      graphs.push_back(hit);
      HIPCHK(hipGraphLaunch(hit.ex, st));
      return true;
    }
    auto it = std::find_if(seen.begin(), seen.end(), [&](const auto& e) { return e.first == key; });
    const bool again = it != seen.end() && it->second != wl_calls;
    const bool have_shape =
        std::any_of(graphs.begin(), graphs.end(), [&](const GraphEntry& e) { return e.shape == shape; });
    if (!have_shape && !again) {  // nothing to update, and a key not used before
      if (it == seen.end()) {
        if (seen.size() >= kSeenKeys) seen.erase(seen.begin());
        seen.push_back({key, wl_calls});
      }
      return false;
    }
    if (it != seen.end()) seen.erase(it);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st, &g));
    for (size_t k = graphs.size(); k-- > 0;) {  // the most recent graph of this shape, updated in place
      if (graphs[k].shape != shape) continue;
      hipGraphNode_t bad = nullptr;
      hipGraphExecUpdateResult res;
      if (hipGraphExecUpdate(graphs[k].ex, g, &bad, &res) == hipSuccess) {
        (void)hipGraphDestroy(g);
        auto hit = graphs[k];
        hit.key = key;
        graphs.erase(graphs.begin() + k);
        graphs.push_back(hit);
        HIPCHK(hipGraphLaunch(hit.ex, st));
        return true;
      }
      (void)hipGetLastError();  // not updatable: instantiate it
      break;
    }
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    if (graphs.size() >= kGraphCache) {
      (void)hipGraphExecDestroy(graphs.front().ex);
      graphs.erase(graphs.begin());
    }
    graphs.push_back({key, shape, ex});
    HIPCHK(hipGraphLaunch(ex, st));
    return true;
  }
  void drop_graph() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.ex);
    graphs.clear();
    seen.clear();
  }
  // the first n side streams wait for the work enqueued on st so far
  void fork(int n = 2) {
    HIPCHK(hipEventRecord(fork_ev, st));
    for (int k = 0; k < n; ++k) HIPCHK(hipStreamWaitEvent(side[k], fork_ev, 0));
  }
  // st waits for the work enqueued on the first n side streams so far
  void join(int n = 2) {
    for (int k = 0; k < n; ++k) {
      HIPCHK(hipEventRecord(join_ev[k], side[k]));
      HIPCHK(hipStreamWaitEvent(st, join_ev[k], 0));
    }
  }
  void check() { HIPCHK(hipGetLastError()); }
};

struct lv_dgraph {
  lv_context* ctx = nullptr;
  uint64_t S = 0, E = 0, order = 0;  // span, edge capacity (offsets[S]), existing vertices
  uint64_t live = 0;                  // directed edges (sum of degrees)
  uint64_t* off = nullptr;
  uint32_t* deg = nullptr;
  uint32_t* keys = nullptr;
  void* w = nullptr;
  bool f64 = false;
  uint8_t* ex = nullptr;
  double M2 = 0;  // 2M, edgeWeightOmp
  TierLists tiers;  // five degree tiers
  TierLists hubs;   // tier 4 = degree > kHubMin (id-order mode), ascending
  uint32_t maxdeg = ~0u;  // largest degree above kTierMax[2] (0: none), with the tiers
  uint64_t hub_slots = 0;
  uint64_t version = 0;  // counts the updates applied to the graph
  bool broken = false;   // a failed update left the rows inconsistent: the graph is refused
  // process-unique stamp of the current degrees, new at upload and at every
  // update: contexts key what they derive from the degrees on it
  uint64_t stamp = 0;
  // worklist hub-batch table slots: sum of nextpow2(2 deg) over degrees
  // above kMidSlots / 2 (the hubs that may run in the multi-block batch)
  uint64_t slab_slots = 0;
  // Batched form: 2M and existing vertices of each graph range voff, as of
  // `version` (a property of the resident graph, kept between runs)
  struct Groups {
    uint64_t version = ~0ull;
    std::vector<uint64_t> voff;
    std::vector<double> M2;
    std::vector<uint64_t> order;
  };
  mutable Groups groups;
  GView view() const { return GView{S, off, deg, keys, w, ex}; }
};

struct lv_dbatch {
  uint64_t nd = 0, ni = 0;
  uint32_t *du = nullptr, *dv = nullptr, *iu = nullptr, *iv = nullptr;
  float *dw = nullptr, *iw = nullptr;
  // checked once, when the batch is made: every id below id_bound, and
  // check_flags bit 1 set if a list is unsorted or repeats an entry (bit 0:
  // an id of 2^32 - 1); a run compares id_bound with its graph's span
  bool checked = false;
  uint64_t id_bound = 0;
  int check_flags = 0;
  // Made by lv_dgraph_update: the graph's stamp before and after the batch
  // was applied (0: not made by an update), and the stored weight of every
  // deletion as it was in the graph (the batch keeps the caller's weights,
  // which the weight update uses, inc/louvain.hxx:374-388).
  uint64_t from_stamp = 0, to_stamp = 0;
  float* dws = nullptr;
};

struct lv_dstate {
  uint64_t S = 0;
  uint32_t* memb = nullptr;
  double* vtot = nullptr;
  double* ctot = nullptr;
  // Hub neighbour-community tables kept with the state (worklist DF / DS):
  // valid while every membership change since their build went through the
  // worklist first pass, which keeps them exact; ht_stamp is the stamp of the
  // graph they describe: a run carries them forward only when its batch was
  // made by lv_dgraph_update from that graph into the run's graph (or the
  // batch is empty and the graph is that graph).  runs counts the runs since
  // the last upload or copy: tables are built on the second run of a state.
  uint64_t* ht_off = nullptr;
  uint32_t* ht_cap = nullptr;
  uint64_t ht_slots = 0;              // slab slots allocated
  unsigned long long* ht_next = nullptr;  // device: first spare slot, tables dropped since the build
  // some hub may be without a table (tables dropped by the last run's end,
  // or a build not yet run through a worklist pass)
  bool ht_scan = true;
  uint32_t* ht_key = nullptr;
  int32_t* ht_cnt = nullptr;
  double* ht_w = nullptr;
  bool ht_valid = false;
  uint64_t ht_stamp = 0;
  uint64_t runs = 0;
  void ht_reset() {
    ht_valid = false;
    runs = 0;
  }
};

namespace {

template <class T>
T* dalloc(size_t n) {
  T* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256)));
  return p;
}

template <class T>
void h2d(lv_context* c, T* dst, const T* src, size_t n) {
  if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->st));
}

template <class T>
void d2h(lv_context* c, T* dst, const T* src, size_t n) {
  if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->st));
}

template <class T>
void d2d(lv_context* c, T* dst, const T* src, size_t n) {
  if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, c->st));
}

void scan32(lv_context* c, const uint32_t* in, uint32_t* out, uint64_t n) {
  size_t bytes = 0;
  HIPCHK(scan_u32(in, out, n, nullptr, bytes, c->st));
  void* tmp = c->scan_tmp(bytes);
  bytes = c->scan_tmp_bytes;
  HIPCHK(scan_u32(in, out, n, tmp, bytes, c->st));
}

void scan64(lv_context* c, const uint64_t* in, uint64_t* out, uint64_t n) {
  size_t bytes = 0;
  HIPCHK(scan_u64(in, out, n, nullptr, bytes, c->st));
  void* tmp = c->scan_tmp(bytes);
  bytes = c->scan_tmp_bytes;
  HIPCHK(scan_u64(in, out, n, tmp, bytes, c->st));
}

// Degree-tier lists of ids 0..n-1 (live only), ascending within each tier.
void build_tiers(lv_context* c, const std::string& name, const uint64_t* off, const uint32_t* deg,
                 const uint8_t* ex, uint64_t n, TierLists& t, uint32_t* ids_out = nullptr, uint32_t thr = 0) {
  uint32_t nb = (uint32_t)std::max<uint64_t>(1, (n + 255) / 256);
  uint32_t* counts = c->ws<uint32_t>(name + ".cnt", (size_t)kTiers * nb + 1);
  uint32_t* base = c->ws<uint32_t>(name + ".base", (size_t)kTiers * nb + 1);
  t.ids = ids_out ? ids_out : c->ws<uint32_t>(name + ".ids", std::max<uint64_t>(n, 1));
  k_fill_u32(counts, 0, (size_t)kTiers * nb + 1, c->st);
  if (n) k_tier_count(off, deg, ex, n, thr, counts, nb, c->st);
  scan32(c, counts, base, (size_t)kTiers * nb + 1);
  if (n) k_tier_scatter(off, deg, ex, n, thr, base, nb, t.ids, c->st);
  // tier starts: base[t * nb] for t <= kTiers (the last is the total)
  uint32_t st[kTiers + 1];
  c->read_strided(st, base, kTiers + 1, nb);
  for (int k = 0; k <= kTiers; ++k) t.start[k] = st[k];
}

uint32_t max_degree_tier4(lv_context* c, const TierLists& t, const uint32_t* deg, int from = 4) {
  uint64_t n4 = t.start[5] - t.start[from];
  if (!n4) return 0;
  uint32_t* out = c->ws<uint32_t>("maxdeg", 1);
  k_fill_u32(out, 0, 1, c->st);
  k_max_deg(t.ids + t.start[from], n4, deg, out, c->st);
  return c->read1(out);
}

// Sum of nextpow2(2 deg) over the degrees above thr (hub-batch table slots).
uint64_t caps_total(lv_context* c, const uint32_t* deg, uint64_t S, uint32_t thr) {
  unsigned long long* tot = c->ws<unsigned long long>("hb.need", 1);
  k_fill_u64(reinterpret_cast<uint64_t*>(tot), 0, 1, c->st);
  k_caps_sum(deg, S, thr, tot, c->st);
  return c->read1(tot);
}

// ------------------------------------------------------------ local-moving
struct Level {
  GView g;
  bool f64;
  const TierLists* tiers;  // five degree tiers (tier orders 0, 1)
  uint64_t order;          // live vertices
  uint32_t maxdeg4;        // max degree in tier 4 of `tiers`
  const TierLists* hubs;   // degree > kHubMin, ascending (id order)
  uint64_t hub_slots;      // hub-batch table slots for all of them
  uint64_t stamp = 0;      // resident graph's degree stamp (0: a super-graph)
  uint64_t slab_slots = 0; // resident graph: hub-batch slots above kMidSlots / 2 edges
  uint32_t maxdeg = ~0u;   // resident graph: its largest degree above kTierMax[2] (0: none); ~0 unknown
};

// Launch one tier of the local-moving kernel, bracketed by events when the
// context profiles (per-launch durations on the context's own stream).
constexpr int kTierHubBatch = 6;  // launch_move: hubs through k_hub_batch (stats slot 4)
constexpr int kTierMidHub = 7;    // launch_move: worklist mid hubs through k_hub_mid (stats slot 3)
constexpr int kTierSmallHub = 8;  // launch_move: worklist small hubs through k_hub_mid (stats slot 2)
constexpr int kTierTabHub = 9;    // launch_move: worklist hubs with a table through k_hub_tab (stats slot 1)
constexpr int kSubRounds = 4;      // worklist sub-rounds enqueued per host check
constexpr int kShortGroup = 2;     // ... once no more than kShortLeft vertices are queued (the cascade
constexpr uint64_t kShortLeft = 64;  // then ends within a sub-round or two; an empty one costs its launches)
constexpr int kFirstGroup = 8;     // ... in an iteration's first group (a DF step on an R-MAT graph
                                   // typically needs 5-8: one host check instead of two)
constexpr uint64_t kRefChunk = 2048;  // the reference's OpenMP chunk (inc/louvain.hxx:594)
// The automatic lead range: the id prefix holding 1 / kLeadShare of the edge
// slots, at least kLeadIdsMin ids (the whole of a smaller graph) and at most
// kLeadIdsMax.
constexpr uint64_t kLeadIdsMin = 1ull << 18;
constexpr uint64_t kLeadIdsMax = 1ull << 20;
constexpr uint64_t kLeadShare = 3;

void launch_move(lv_context* c, const MoveArgs& a, int tier, bool f64, hipStream_t s = nullptr) {
  if (!s) s = c->st;
  // the chunk counter of chunks handed out in order (the hub batch and
  // spread chunks do not use it)
  if (tier != kTierHubBatch && a.chunk) k_fill_u64(reinterpret_cast<uint64_t*>(a.work), 0, 1, s);
  hipEvent_t e0 = nullptr;
  if (c->profiling) e0 = c->record_start(s);
  if (tier == kTierHubBatch) k_hub_batch(a, s);
  else if (tier == kTierMidHub || tier == kTierSmallHub) k_hub_mid(a, tier == kTierMidHub ? 1 : 0, s);
  else if (tier == kTierTabHub) k_hub_tab(a, s);
  else {
    if (tier == kTierHybrid && a.ctab && a.gcs) k_graph_chunks(a, s);
    k_move(a, tier, f64, s);
  }
  const int slot = tier == kTierHubBatch ? 4
                   : tier == kTierMidHub  ? 3
                   : tier == kTierSmallHub ? 2
                   : tier == kTierTabHub ? 1
                   : (tier == kTierHybrid ? 0 : tier);
  if (c->profiling) c->prof.push_back({slot, {e0, c->record(s)}});
}

// A timed launch into the context's statistics: a local-moving stats slot,
// or kSlotFused (the fused worklist launch, id-order kernel and tabled hubs).
constexpr int kSlotFused = 5;
void add_prof(lv_context* c, int slot, float ms) {
  if (slot == kSlotFused) {
    c->stats.fused_ms += ms;
    c->stats.fused_launches += 1;
  } else {
    c->stats.move_ms[slot] += ms;
    c->stats.move_launches[slot] += 1;
  }
}

// Worklist queues: degrees above hub_min_wl are hubs; up to this one they
// run a block each (LDS table), above it in the multi-block hub batch.
uint32_t wl_big_hub(const lv_context* c) { return std::max<uint32_t>(c->hub_min_wl, kMidHubMax); }

// Hubs that may reach the multi-block batch: above wl_big_hub, or marked
// there after their capped LDS table (hubs above kMidSlots / 2 edges)
// overflowed.  The batch's table slab covers all of them.
uint32_t wl_slab_hub(const lv_context* c) { return std::max<uint32_t>(c->hub_min_wl, kMidSlots / 2); }

// Streams of a level: stream_fraction of its live vertices, clamped.  A
// level of at most one reference chunk (2048 ids, the schedule(dynamic,
// 2048) of inc/louvain.hxx:594) is one thread's work in the reference, so it
// runs as one stream.
uint64_t level_streams(const lv_context* c, uint64_t order) {
  if (order <= kRefChunk) return 1;
  uint64_t want = (uint64_t)(c->stream_fraction * (double)order);
  return std::min(c->max_streams, std::max<uint64_t>(c->min_streams, want));
}

// Per-graph control of the batched form (lv_run_batched): graph of every
// vertex of the level and per graph its M, community-0 id, gain and skip
// flag on the device; tolerance, iteration count and last gain on the host.
struct Batched {
  uint64_t G = 0;
  uint32_t* gid = nullptr;
  double* gM = nullptr;
  uint32_t* gzero = nullptr;
  double* gel = nullptr;
  uint8_t* gmask = nullptr;
  std::vector<double> E, el;
  std::vector<int> it;
  std::vector<uint8_t> skip, done;
  // streams of the level per graph (id-order kernel with spread chunks):
  // level id ranges, chunk starts (G + 1 each) and the chunk table
  uint64_t* glo = nullptr;
  uint64_t* gcs = nullptr;
  uint64_t* ctab = nullptr;
  uint64_t nct = 0;
  std::vector<uint64_t> hlo, hcs;
};

void batched_args(MoveArgs& a, const Batched* bt) {
  if (!bt) return;
  a.gid = bt->gid;
  a.gM = bt->gM;
  a.gzero = bt->gzero;
  a.gel = bt->gel;
  a.gmask = bt->gmask;
}

// Level schedule: graph k walks its own id range in level_streams(its live
// vertices) streams, as it would alone (so a small graph is one stream
// beside a large one); a finished graph walks none.
void batched_level(lv_context* c, Batched* bt, const std::vector<uint64_t>& lo, const std::vector<uint64_t>& GN) {
  const uint64_t G = bt->G;
  bt->hlo = lo;
  bt->hcs.assign(G + 1, 0);
  for (uint64_t k = 0; k < G; ++k) bt->hcs[k + 1] = bt->hcs[k] + (bt->done[k] ? 0 : level_streams(c, GN[k]));
  bt->nct = bt->hcs[G];
  bt->glo = c->ws<uint64_t>("bt.glo", G + 1);
  bt->gcs = c->ws<uint64_t>("bt.gcs", G + 1);
  bt->ctab = c->ws<uint64_t>("bt.ctab", 2 * std::max<uint64_t>(bt->nct, 1));
  h2d(c, bt->glo, bt->hlo.data(), G + 1);
  h2d(c, bt->gcs, bt->hcs.data(), G + 1);
}

void id_order_regions(lv_context* c, MoveArgs& a, const Level& lv);

// The level schedule applies to the id-order kernel when chunks are spread
// (chunk 0); chunks handed out in ascending order run over the union.
void batched_schedule(lv_context* c, MoveArgs& a, const Level& lv, const Batched* bt) {
  if (!bt || a.chunk || c->tier_order != 2 || !bt->nct) return;
  a.ctab = bt->ctab;
  a.nct = bt->nct;
  a.glo = bt->glo;
  a.gcs = bt->gcs;
  a.ng = (uint32_t)bt->G;
  a.max_streams = bt->nct;
  id_order_regions(c, a, lv);
}

// Start of a pass: every unfinished graph runs.
void batched_begin(lv_context* c, Batched* bt) {
  for (uint64_t k = 0; k < bt->G; ++k) {
    bt->skip[k] = bt->done[k];
    bt->it[k] = 0;
    bt->el[k] = 0;
  }
  h2d(c, bt->gmask, bt->skip.data(), bt->G);
}

// After an iteration: each running graph counts it and stops when its gain
// is within its tolerance or it reached L (inc/louvain.hxx:598-606); returns
// whether any graph runs on.
bool batched_step(lv_context* c, Batched* bt, int L) {
  c->read(bt->el.data(), bt->gel, bt->G);
  bool any = false;
  for (uint64_t k = 0; k < bt->G; ++k) {
    if (bt->skip[k]) continue;
    ++bt->it[k];
    if (bt->el[k] <= bt->E[k] || bt->it[k] >= L) bt->skip[k] = 1;
    else any = true;
  }
  if (any) h2d(c, bt->gmask, bt->skip.data(), bt->G);
  return any;
}

// id-order kernel: one HBM region per stream (up to 2^16, rounded up to a
// multiple of the kernel's 4 streams per block) holding nextpow2(2 deg)
// slots for every degree it runs (<= hub_min)
void id_order_regions(lv_context* c, MoveArgs& a, const Level& lv) {
  a.gblocks = std::min<uint64_t>(a.max_streams, 1u << 16);
  // the largest degree the kernel visits with a region: its hub threshold
  // (a worklist's is at most kHubMin), or above kHubMin the level's largest
  // hub (hubs in the sweep have tables of their own)
  const uint64_t hm = a.xids ? kHubMin : a.hub_min;
  uint64_t dmax = std::min<uint64_t>(hm, kHubMin);
  if (hm > kHubMin) dmax = std::max<uint64_t>(kHubMin, std::min<uint64_t>(hm, lv.hubs->maxdeg));
  a.gcap = std::max<uint64_t>(dmax >= kHubMin ? kHybridHbm : kHybridLds, next_pow2(2 * dmax));
  size_t slots = (size_t)a.gcap * ((a.gblocks + 3) & ~3ull);
  a.gkey = c->ws<uint32_t>("g.key", slots);
  a.gpos = c->ws<uint32_t>("g.pos", slots);
  a.gval = c->ws<double>("g.val", slots);
}

MoveArgs base_args(lv_context* c, const Level& lv, uint32_t* vcom, const double* vtot, double* ctot, uint8_t* vaff,
                   const uint8_t* fa, double M, double R, bool worklist = false) {
  MoveArgs a;
  std::memset(&a, 0, sizeof a);  // padding too: the bytes key captured launch graphs
  a.g = lv.g;
  a.vcom = vcom;
  a.vtot = vtot;
  a.ctot = ctot;
  a.vaff = vaff;
  a.fa = fa;
  a.M = M;
  a.R = R;
  a.el = c->ws<double>("el", (size_t)kSpread * kElStride);
  a.cnt = c->counters();
  a.min_chunk = 0;
  a.max_streams = level_streams(c, lv.order);
  // worklist sub-rounds mostly hold a few thousand vertices: no more streams
  // than the id-order kernel keeps resident at once (a launch with fewer
  // entries then costs one round of short-lived waves, not several)
  if (worklist) a.max_streams = std::min<uint64_t>(a.max_streams, c->wl_streams);
  a.chunk = c->chunk;
  a.work = c->ws<unsigned long long>("work", 1);
  a.hub_min = worklist ? c->hub_min_wl : c->hub_min;
  a.hub_streams = c->hub_streams;
  // hub kernel: one HBM region per block, sized for the largest hub
  uint32_t md = c->tier_order == 2 ? lv.hubs->maxdeg : lv.maxdeg4;
  if (md) {
    a.hcap = next_pow2(2 * (uint64_t)md);
    size_t slots = (size_t)a.hcap * kHubBlocks;
    a.hkey = c->ws<uint32_t>("h.key", slots);
    a.hpos = c->ws<uint32_t>("h.pos", slots);
    a.hval = c->ws<double>("h.val", slots);
  }
  uint64_t nh = lv.hubs ? lv.hubs->start[5] - lv.hubs->start[4] : 0;
  uint64_t slots = lv.hub_slots;
  if (worklist && c->tier_order == 2) {
    // the big-hub queue may hold any vertex above its threshold: scratch
    // for all of them, table slots for the sum over them
    nh = lv.g.S;
    const uint32_t thr = wl_slab_hub(c);
    if (thr == kMidSlots / 2 && lv.stamp) {
      c->wl_slots = lv.slab_slots;  // the resident graph's, for the default threshold
      c->wl_key_stamp = lv.stamp;
      c->wl_key_thr = thr;
    } else if (!lv.stamp || c->wl_key_stamp != lv.stamp || c->wl_key_thr != thr) {
      c->wl_slots = caps_total(c, lv.g.deg, lv.g.S, thr);
      c->wl_key_stamp = lv.stamp;
      c->wl_key_thr = thr;
    }
    slots = c->wl_slots;
  }
  if (c->tier_order == 2 && nh) {
    // hub batches: per-hub scratch and a table slab for every hub of the level
    a.f64_weights = lv.f64;
    a.hb_cap = c->ws<uint64_t>("hb.cap", nh + 1);
    a.hb_chunks = c->ws<uint64_t>("hb.chunks", nh + 1);
    a.hb_tot = c->ws<uint64_t>("hb.tot", 2);
    a.hb_cu = c->ws<uint32_t>("hb.cu", nh);
    a.hb_vd = c->ws<double>("hb.vd", nh);
    a.hb_sd = c->ws<double>("hb.sd", nh);
    a.hb_bpos = c->ws<uint32_t>("hb.bpos", nh);
    if (worklist) {
      // capped tables: overflow flags per queue entry, and the vertices that
      // need a full table (zeroed once; a stale mark only costs a full table)
      a.hb_over = c->ws<uint8_t>("hb.over", nh);
      bool fb = false;
      a.hb_big = c->ws<uint8_t>("hb.big", lv.g.S, &fb);
      if (fb) k_fill_u8(a.hb_big, 0, c->ws_bytes("hb.big"), c->st);
    }
    a.hb_off = c->ws<uint64_t>("hb.off", nh);
    a.hb_deg = c->ws<uint32_t>("hb.deg", nh);
    // edge chunks: every table has >= 2 deg slots, so the chunks number at
    // most slots / (2 kHubChunk) + one partial chunk per hub
    a.hb_cown = c->ws<uint32_t>("hb.cown", slots / (2 * kHubChunkEdges) + nh + 1);
    a.hb_town = c->ws<uint32_t>("hb.town", slots / kHubTabChunk + 1);
    a.hb_ce = c->ws<double>("hb.ce", slots / kHubTabChunk + 1);
    a.hb_cp = c->ws<uint32_t>("hb.cp", slots / kHubTabChunk + 1);
    // the batch's best pass empties every slot it used, so a slab is
    // filled once, when it is allocated
    bool fk = false, fp = false, fv = false;
    a.hb_key = c->ws<uint32_t>("hb.key", slots, &fk);
    a.hb_pos = c->ws<uint32_t>("hb.pos", slots, &fp);
    a.hb_val = c->ws<double>("hb.val", slots, &fv);
    if (fk) k_fill_u32(a.hb_key, kEmpty, c->ws_bytes("hb.key") / sizeof(uint32_t), c->st);
    if (fp) k_fill_u32(a.hb_pos, kNoPos, c->ws_bytes("hb.pos") / sizeof(uint32_t), c->st);
    if (fv) k_fill_f64(a.hb_val, 0, c->ws_bytes("hb.val") / sizeof(double), c->st);
  }
  if (c->tier_order == 2) id_order_regions(c, a, lv);
  return a;
}

// One sweep-mode iteration over the tier lists (or, in id-order mode, over
// all ids ascending through the general HBM-table kernel).  core > 0 (id
// order): ids [0, core) first, by core_streams streams over ascending chunks
// of core_chunk ids, then the rest under a's schedule.
void move_sweep(lv_context* c, MoveArgs a, const Level& lv, uint64_t core = 0, uint64_t core_streams = 0,
                uint64_t core_chunk = 0) {
  unsigned long long* cnt0 = a.cnt;
  if (c->tier_order == 2) {
    // hubs (degree > kHubMin, ascending) first, then every other id ascending
    uint64_t nh = lv.hubs->start[5] - lv.hubs->start[4];
    if (nh && !a.xids && c->hub_min < lv.hubs->maxdeg) {
      a.list = lv.hubs->ids + lv.hubs->start[4];
      a.n = nh;
      if (cnt0) a.cnt = cnt0 + (size_t)4 * CNT_N;
      launch_move(c, a, kTierHubBatch, lv.f64);
    }
    // the id range in `slabs` consecutive slabs, one launch each: a slab sees
    // every move of the slabs before it
    if (cnt0) a.cnt = cnt0;
    a.list = nullptr;
    if (core && core < lv.g.S) {
      MoveArgs b = a;
      b.base = 0;
      b.n = core;
      b.max_streams = core_streams;
      b.chunk = core_chunk;
      b.ctab = nullptr;
      launch_move(c, b, kTierHybrid, lv.f64);
      a.base = core;
      a.n = lv.g.S - core;
      a.ctab = nullptr;
      launch_move(c, a, kTierHybrid, lv.f64);
      return;
    }
    uint64_t P = std::max<uint64_t>(1, std::min<uint64_t>(c->slabs, lv.g.S));
    for (uint64_t k = 0; k < P; ++k) {
      uint64_t lo = lv.g.S * k / P, hi = lv.g.S * (k + 1) / P;
      if (hi <= lo) continue;
      a.base = lo;
      a.n = hi - lo;
      launch_move(c, a, kTierHybrid, lv.f64);
    }
    return;
  }
  for (int k = 0; k < kTiers; ++k) {
    int t = c->tier_order ? kTiers - 1 - k : k;
    uint64_t n = lv.tiers->start[t + 1] - lv.tiers->start[t];
    if (!n) continue;
    a.list = lv.tiers->ids + lv.tiers->start[t];
    a.n = n;
    if (cnt0) a.cnt = cnt0 + (size_t)t * CNT_N;
    launch_move(c, a, t, lv.f64);
  }
}

// Zero the spread gain slots of el (kSpread x kElStride doubles).
void clear_el(lv_context* c, double* el) { k_fill_f64(el, 0, (size_t)kSpread * kElStride, c->st); }

// The gain of a launch sequence: its spread slots added in slot order.
double read_el(lv_context* c, const double* el) {
  double h[kSpread * kElStride];
  c->read(h, el, (size_t)kSpread * kElStride);
  double t = 0;
  for (int s = 0; s < kSpread; ++s) t += h[(size_t)s * kElStride];
  return t;
}

// louvainMoveOmpW (inc/louvain.hxx:587-609): iterate until the summed gain
// el <= E or L iterations; returns the iterations, or 0 if the first one
// gained nothing.
// ascending: hand out chunks in ascending order even if the context spreads
// them -- a sparse frontier needs a vertex flagged by a lower id to be reached
// later in the same sweep (inc/louvain.hxx:602 then :598).
// Batched form (bt): every graph iterates under its own tolerance; the
// per-graph results are in bt->it / bt->el.
int local_move(lv_context* c, const Level& lv, uint32_t* vcom, const double* vtot, double* ctot, uint8_t* vaff,
               const uint8_t* fa, double M, double R, int L, double E, bool ascending = false,
               Batched* bt = nullptr, int pass = -1) {
  MoveArgs a = base_args(c, lv, vcom, vtot, ctot, vaff, fa, M, R);
  batched_args(a, bt);
  // the frontier path: 64-id windows, the flagged ids' row data loaded with
  // the window.  (Worklist-path sweeps check one id at a time: with windows
  // the wide Static sweeps after the lead iteration oscillate -- 27 to 29
  // iterations and 2e-3 to 8e-3 less modularity on R-MAT 22.)
  a.sparse = c->path == 1;
  if (ascending && !a.chunk) a.chunk = a.sparse ? 64 : 16;
  batched_schedule(c, a, lv, bt);
  // hubs in the sweep (a Static run's first pass and every later pass): each
  // visited in its id turn by its stream, with a table of its own
  // (nextpow2(2 deg) slots of one slab, hub i's at xoff[i])
  const uint64_t nhub = lv.hubs ? lv.hubs->start[5] - lv.hubs->start[4] : 0;
  if (c->tier_order == 2 && c->sweep_hubs && nhub && pass >= 0) {
    uint64_t* caps = c->ws<uint64_t>("x.caps", nhub + 1);
    uint64_t* xoff = c->ws<uint64_t>("x.off", nhub + 1);
    k_list_caps(lv.hubs->ids + lv.hubs->start[4], nhub, lv.g.deg, caps, c->st);
    scan64(c, caps, xoff, nhub + 1);
    const uint64_t slots = std::max<uint64_t>(lv.hub_slots, 1);
    a.xids = lv.hubs->ids + lv.hubs->start[4];
    a.xoff = xoff;
    a.nx = nhub;
    a.xkey = c->ws<uint32_t>("x.key", slots);
    a.xpos = c->ws<uint32_t>("x.pos", slots);
    a.xval = c->ws<double>("x.val", slots);
    a.hub_min = 0xFFFFFFFFu;
  }
  // per-pass sweep schedules (id-order sweeps of lv_run only): regions for
  // the most streams any iteration of this pass runs
  const uint64_t s0 = a.max_streams, ch0 = a.chunk;
  const bool by_pass = !bt && c->tier_order == 2 && pass >= 0;
  const bool lead = by_pass && pass == 0 && c->lead_iters > 0;
  const bool late = by_pass && pass > 0 && c->late_streams > 0;
  // ids under the lead schedule: the automatic range, all ids, or the set count
  uint64_t lead_core = 0;
  if (lead) {
    if (c->lead_ids == ~0ull) lead_core = 0;
    else if (c->lead_ids) lead_core = std::min<uint64_t>(c->lead_ids, lv.g.S);
    else {
      uint64_t* at = c->ws<uint64_t>("lead.at", 1);
      k_offset_bound(lv.g.off, lv.g.S, kLeadShare, at, c->st);
      lead_core = std::min<uint64_t>(std::max<uint64_t>(c->read1(at), kLeadIdsMin), kLeadIdsMax);
    }
  }
  if (lead || late) {
    a.max_streams = std::max(s0, lead ? c->lead_streams : c->late_streams);
    id_order_regions(c, a, lv);
  }
  unsigned long long* cnt0 = a.cnt;
  int l = 0;
  double el = 0;
  if (bt) batched_begin(c, bt);
  for (; l < L;) {
    if (bt) k_fill_f64(bt->gel, 0, bt->G, c->st);
    else clear_el(c, a.el);
    a.pick_less = l < c->pick_less_iters;
    a.validate = c->validate >= 1;
    uint64_t core = 0;
    if (lead || late) {
      const bool lead_now = lead && l < c->lead_iters;
      if (lead_now && lead_core && lead_core < lv.g.S) {
        core = lead_core;  // the lead schedule over the first ids only
        a.max_streams = s0;
        a.chunk = ch0;
      } else {
        a.max_streams = lead_now ? std::min(c->lead_streams, s0) : late ? std::min(c->late_streams, s0) : s0;
        a.chunk = lead_now ? c->lead_chunk : late ? c->late_chunk : ch0;
      }
    }
    // (never more streams than the context's schedule gives the level)
    move_sweep(c, a, lv, core, std::min(c->lead_streams, s0), c->lead_chunk);
    c->check();
    ++l;
    if (bt) {
      if (!batched_step(c, bt, L)) break;
      continue;
    }
    el = read_el(c, a.el);
    if (el <= E) break;
  }
  return (l > 1 || el) ? l : 0;
}

// Worklist-mode local-moving for the first pass of the frontier-based
// approaches.  Iteration k visits its queue; a vertex flagged 0 -> 1 by a
// lower-id vertex and not yet visited in iteration k joins a follow-up
// sub-round of the same iteration (the reference's ascending sweep reaches it
// later in the same iteration), any other flagged vertex waits for
// iteration k+1.  Sub-rounds run until no vertex is left for iteration k;
// the gain el of iteration k sums over its sub-rounds.  Queues are binned by
// degree tier on the device.  q[0..3]: iteration input / next iteration /
// two sub-round buffers; qlen[0..3] their lengths.
int local_move_worklist(lv_context* c, const Level& lv, uint32_t* vcom, const double* vtot, double* ctot,
                        uint8_t* vaff, const uint8_t* fa, double M, double R, int L, double E,
                        Batched* bt = nullptr, lv_dstate* ht = nullptr) {
  MoveArgs a = base_args(c, lv, vcom, vtot, ctot, vaff, fa, M, R, true);
  if (ht && c->tier_order == 2) {  // hubs with a neighbour-community table
    a.ht_off = ht->ht_off;
    a.ht_cap = ht->ht_cap;
    a.ht_key = ht->ht_key;
    a.ht_cnt = ht->ht_cnt;
    a.ht_w = ht->ht_w;
    a.ht_ndrop = ht->ht_next + 1;
    // moved tabled hubs, for the grid-wide flag pass (count zeroed by every compaction)
    a.mv_cap = 4096;
    a.mv_u = c->ws<uint32_t>("mv.u", a.mv_cap);
    a.mv_from = c->ws<uint32_t>("mv.from", a.mv_cap);
    a.mv_to = c->ws<uint32_t>("mv.to", a.mv_cap);
    a.mv_d = c->ws<uint32_t>("mv.d", a.mv_cap);
    a.mv_o = c->ws<uint64_t>("mv.o", a.mv_cap);
    a.mv_n = c->ws<uint32_t>("mv.n", 3);  // entries, pieces reserved, pieces taken
  }
  batched_args(a, bt);
  batched_schedule(c, a, lv, bt);
  if (bt) batched_begin(c, bt);
  a.sparse = 2;  // queue entries read a window at a time, their flags once
  a.interleave = c->interleave;
  unsigned long long* cnt0 = a.cnt;
  const uint64_t S = lv.g.S, stride = std::max<uint64_t>(S, 1), fwords = (S >> 5) + 2;
  // moved-hub pieces of the fused launch (every descriptor pending between launches)
  uint32_t* mv_piece = nullptr;
  if (a.mv_n) {
    bool fresh_p = false;
    mv_piece = c->ws<uint32_t>("mv.piece", kMvPieces, &fresh_p);
    if (fresh_p) k_fill_u32(mv_piece, kMvPending, kMvPieces, c->st);
  }
  // queues: low degree, small hubs, mid hubs, hub batch
  WlQueues qs;
  for (int q = 0; q < kWlQueues; ++q) qs.q[q] = c->ws<uint32_t>("wl.q" + std::to_string(q), stride);
  // queue lengths, the tables dropped so far (when there are tables), then
  // (u32 6..7) the iteration's gain so far as a double: one read per group
  uint32_t* lens = c->ws<uint32_t>("wl.len", 8);
  uint64_t* bsum = c->ws<uint64_t>("wl.bsum", (size_t)kWlQueues * kWlScanBlocks);
  // next-sub-round flags: all zero between calls (each compaction resets
  // the flags it reads, and a sub-round loop ends on an empty compaction)
  bool fresh = false;
  a.fsame = c->ws<uint32_t>("wl.same", fwords, &fresh);
  if (fresh) k_fill_u32(a.fsame, 0, c->ws_bytes("wl.same") / 4, c->st);
  // per compaction segment: a next-sub-round flag may be set there (all zero
  // between calls too; a compaction clears the marks of the segments it reads)
  a.wl_dirty = c->ws<uint8_t>("wl.dirty", kWlScanBlocks, &fresh);
  if (fresh) k_fill_u8(a.wl_dirty, 0, c->ws_bytes("wl.dirty"), c->st);
  a.wl_shift = wl_shift(S);
  uint32_t* tq = nullptr;
  uint32_t* tlen = nullptr;
  if (c->tier_order != 2) {
    tq = c->ws<uint32_t>("wl.tq", (size_t)kTiers * stride);
    tlen = c->ws<uint32_t>("wl.tlen", kTiers);
  }
  a.stamp = c->ws<uint32_t>("wl.stamp", stride);
  if (c->stamp_span != S) {  // fresh stamps for a new span
    k_fill_u32(a.stamp, 0, stride, c->st);
    c->stamp_span = S;
    c->serial = 0;
  }
  WlSplit sp{};
  if (c->tier_order == 2) {
    const uint32_t thr = c->hub_min_wl;
    sp.t[0] = thr;
    sp.t[1] = std::max(thr, kSmallHubMax);
    sp.t[2] = wl_big_hub(c);
    sp.big = a.hb_big;  // hubs marked for the batch
    sp.tab = a.ht_cap;  // hubs with a table
    sp.reset = a.mv_n;  // the moved-hub list starts empty in every sub-round
    sp.nreset = 3;
    if (a.ctab && !a.mv_n) {  // (the batched form, which keeps no hub tables)
      // the id-order kernel walks the chunks that hold entries, packed
      // first by the chunk-table pass after each compaction
      a.nct_dev = c->ws<uint32_t>("bt.nct", 1);
      sp.reset = a.nct_dev;
      sp.nreset = 1;
    }
    sp.ndrop = a.ht_ndrop;
  }
  if (!bt) sp.el = a.el;  // (the batched form reads its per-graph gains)
  WlSplit sp_same = sp;  // next-sub-round compactions read the marked segments only
  sp_same.dirty = a.wl_dirty;
  // Sub-rounds are enqueued kFirstGroup, then kSubRounds at a time (their kernels read the
  // queue lengths on the device; an empty round costs its launches only),
  // with one host check of the last compaction per group.  Without
  // profiling a group is replayed from a captured graph; the iteration
  // serial is then read on the device.
  const bool graphs = c->use_graphs && !c->profiling;
  ++c->wl_calls;
  uint32_t* serial_dev = c->ws<uint32_t>("wl.serial", 1);
  if (graphs) {
    a.serial_dev = serial_dev;
    if (S) c->make_side();
  }
  // With hub tables and no table dropped by the end of the last run, the
  // edge-scanning hub kernels are not launched.  A table dropped since (by
  // this run's batch or a move) sends its hub to them: they are launched from
  // the next group on, the group after its hub was queued.  Below kHtMinDeg
  // edges hubs have no table.
  bool scan_hubs = !a.ht_cap || ht->ht_scan || c->hub_min_wl < kHtMinDeg;
  // a graph without a vertex above the hub degree has no hub to queue: no
  // hub kernel is launched, nor a side stream forked for one
  if (lv.maxdeg <= c->hub_min_wl) scan_hubs = false;
  // one sub-round: hubs and the rest, then the next queue
  auto sub_round = [&]() {
    if (c->tier_order == 2) {
      // id-order mode: the low-degree vertices through the id-order
      // kernel, hubs with a table (a wave each), and while some hub has no
      // table the other hubs (a block each, mid then small) and the largest
      // (multi-block batch), these on two side streams at once.  Like the
      // reference's threads, each sees the moves the others have made.
      // Without the edge-scanning kernels everything runs on the context's
      // stream: the tabled hubs and the id-order kernel in one launch, then
      // the flags of the tabled hubs that moved (the cross-stream waits of
      // side streams cost more than their overlap gains).
      const int sides = scan_hubs ? 2 : 0;
      hipStream_t hs = sides ? c->side[0] : c->st;
      MoveArgs ah = a;  // the tabled hubs (stats slot 1)
      ah.mv_piece = mv_piece;  // (their moves' rows flagged inside the fused launch)
      ah.mv_pcap = kMvPieces;
      ah.list = qs.q[4];
      ah.n_dev = lens + 4;
      ah.cnt = cnt0 + (size_t)1 * CNT_N;
      a.list = qs.q[0];
      a.n_dev = lens;
      a.cnt = cnt0;
      // Every hub has a table: the tabled hubs and the id-order kernel in one
      // launch, then the flags of the hubs that moved
      hipEvent_t f0 = nullptr;
      if (c->profiling && S && a.ht_cap && !sides) f0 = c->record_start();
      const bool fused = S && a.ht_cap && !sides && k_wl_fused(a, ah, qs, lens, lv.f64, c->st);
      if (f0 && fused) c->prof.push_back({kSlotFused, {f0, c->record()}});
      ah.mv_piece = nullptr;  // (the separate launches: a flag pass reads the entries)
      if (S && !fused) {
        if (sides) {
          c->make_side();
          c->fork(sides);
          hs = c->side[0];
        }
        if (a.ht_cap) {
          launch_move(c, ah, kTierTabHub, lv.f64, hs);
          k_moved_flag(ah, qs, lens, !scan_hubs, hs);  // with the hubs whose tables dropped, kept for later
        }
      }
      if (!fused) launch_move(c, a, kTierHybrid, lv.f64);
      if (S) {
        if (scan_hubs) {
          for (int q = 2; q >= 1; --q) {  // stats slots 3 (mid) and 2 (small)
            a.list = qs.q[q];
            a.n_dev = lens + q;
            a.cnt = cnt0 + (size_t)(q + 1) * CNT_N;
            launch_move(c, a, q == 2 ? kTierMidHub : kTierSmallHub, lv.f64, c->side[0]);
          }
          a.list = qs.q[3];
          a.n_dev = lens + 3;
          a.cnt = cnt0 + (size_t)4 * CNT_N;
          launch_move(c, a, kTierHubBatch, lv.f64, c->side[1]);
        }
        if (sides) c->join(sides);
      }
    } else {
      k_fill_u32(tlen, 0, kTiers, c->st);
      k_bin_queue(qs.q[0], lens, std::min<uint64_t>(S, 1u << 19), lv.g.deg, 0, tq, stride, tlen, c->st);
      for (int k = (c->tier_order ? kTiers - 1 : 0); k >= 0 && k < kTiers; k += (c->tier_order ? -1 : 1)) {
        if (k == 4 && !lv.maxdeg4) continue;
        a.list = tq + (size_t)k * stride;
        a.n_dev = tlen + k;
        a.cnt = cnt0 + (size_t)k * CNT_N;
        launch_move(c, a, k, lv.f64);
      }
    }
    k_wl_queue_bits(a.fsame, true, S, lv.g.deg, sp_same, qs, lens, bsum, c->st);
  };
  int l = 0;
  double el = 0;
  for (; l < L;) {
    if (++c->serial == 0) {  // wrapped: reset the stamps
      k_fill_u32(a.stamp, 0, stride, c->st);
      c->serial = 1;
    }
    a.serial = c->serial;
    a.pick_less = l < c->pick_less_iters;
    a.validate = c->validate >= 2;
    if (bt) k_fill_f64(bt->gel, 0, bt->G, c->st);
    else clear_el(c, a.el);
    // the iteration's queue: every flagged id, ascending
    k_wl_queue(vaff, false, S, lv.g.deg, sp, qs, lens, bsum, c->st);
    if (c->trace) {
      uint32_t t[kWlQueues];
      c->read(t, lens, kWlQueues);
      std::fprintf(stderr, "[lv trace] iteration %d queues %u %u %u %u %u\n", l, t[0], t[1], t[2], t[3], t[4]);
    }
    if (graphs) k_fill_u32(serial_dev, c->serial, 1, c->st);
    // what the captured launches depend on (the serial is read on the device)
    int rounds = kFirstGroup;  // sub-rounds of the next group
    auto group_key = [&]() {
      std::string key;
      MoveArgs ka = a;
      ka.serial = 0;
      ka.list = nullptr;
      ka.n_dev = nullptr;
      ka.cnt = cnt0;
      const int flags[5] = {c->tier_order, lv.f64 ? 1 : 0, (int)lv.maxdeg4, rounds, scan_hubs ? 1 : 0};
      key.append(reinterpret_cast<const char*>(&ka), sizeof ka);
      key.append(reinterpret_cast<const char*>(&qs), sizeof qs);
      key.append(reinterpret_cast<const char*>(&sp), sizeof sp);
      const void* ptrs[5] = {lens, bsum, tq, tlen, c->side[0]};
      key.append(reinterpret_cast<const char*>(ptrs), sizeof ptrs);
      key.append(reinterpret_cast<const char*>(flags), sizeof flags);
      key.append(reinterpret_cast<const char*>(&S), sizeof S);
      return key;
    };
    uint64_t left = 0;
    for (;; rounds = left <= kShortLeft ? kShortGroup : kSubRounds) {
      auto group = [&] {
        for (int r = 0; r < rounds; ++r) {
          sub_round();
          if (c->trace) {
            uint32_t t[kWlQueues];
            c->read(t, lens, kWlQueues);
            std::fprintf(stderr, "[lv trace] iteration %d sub-round: next queues %u %u %u %u %u\n", l, t[0], t[1],
                         t[2], t[3], t[4]);
          }
        }
      };
      const int shape = rounds * 4 + (scan_hubs ? 2 : 0) + (a.ht_cap ? 1 : 0);
      if (!graphs || c->trace || !c->replay(group_key(), shape, group)) group();
      c->check();
      uint32_t nq[8];
      c->read(nq, lens, sp.el ? 8 : sp.ndrop ? kWlQueues + 1 : kWlQueues);
      if (sp.ndrop) ht->ht_scan = nq[kWlQueues] != 0;
      if (sp.el) std::memcpy(&el, nq + 6, sizeof el);
      left = 0;
      for (int q = 0; q < kWlQueues; ++q) left += nq[q];
      if (!left) break;
      if (nq[1] + nq[2] + nq[3]) scan_hubs = true;
    }
    ++l;
    if (bt) {
      if (!batched_step(c, bt, L)) break;
      continue;
    }
    // (el came with the last group's queue lengths)
    if (el <= E) break;
  }
  return (l > 1 || el) ? l : 0;
}

// ------------------------------------------------------------ aggregation
struct SuperGraph {
  const char* name = nullptr;  // its workspace buffers (kept through swaps)
  uint64_t* off = nullptr;
  uint32_t* deg = nullptr;
  uint32_t* keys = nullptr;
  double* w = nullptr;
};

SuperGraph super_graph(lv_context* c, const char* name, uint64_t S, uint64_t E) {
  std::string n(name);
  SuperGraph s;
  s.name = name;
  s.off = c->ws<uint64_t>(n + ".off", S + 1);
  s.deg = c->ws<uint32_t>(n + ".deg", S);
  s.keys = c->ws<uint32_t>(n + ".keys", E);
  s.w = c->ws<double>(n + ".w", E);
  return s;
}

// louvainCommunityExistsOmpW + louvainRenumberCommunitiesOmpW
// (inc/louvain.hxx:668-682, :941-945), up to the lookup: map[0..S'] is the
// old -> new id map; returns the community count.
uint64_t count_communities(lv_context* c, const GView& g, const uint32_t* vcom, uint32_t*& map) {
  uint32_t* flags = c->ws<uint32_t>("rn.flags", g.S + 1);
  map = c->ws<uint32_t>("rn.map", g.S + 1);
  k_fill_u32(flags, 0, g.S + 1, c->st);
  k_exists(g, vcom, flags, c->st);
  scan32(c, flags, map, g.S + 1);
  return c->read1(map + g.S);
}

// louvainCommunityVerticesOmpW + louvainAggregateOmpW (inc/louvain.hxx:800-812,
// :987-992) for a renumbered membership: builds super-graph z and its row
// sums (K of the super-vertices) in ztot.
void aggregate(lv_context* c, const Level& lv, const uint32_t* vcom, uint64_t C, SuperGraph& z, double* ztot) {
  const GView& g = lv.g;
  uint64_t GN = lv.order;
  uint32_t* coff = c->ws<uint32_t>("ag.coff", C + 1);
  uint64_t* mdeg = c->ws<uint64_t>("ag.mdeg", GN + 1);
  uint64_t* mscan = c->ws<uint64_t>("ag.mscan", GN + 1);
  // members of each community in ascending id order: a stable sort of
  // (community, id) over the level's ids, the absent ones (key C) last;
  // the community offsets are read off the sorted keys
  uint32_t* mk = c->ws<uint32_t>("ag.mk", g.S);
  uint32_t* mk2 = c->ws<uint32_t>("ag.mk2", g.S);
  uint32_t* mid = c->ws<uint32_t>("ag.mid", g.S);
  uint32_t* cedg = c->ws<uint32_t>("ag.cedg", g.S);
  k_member_keys(g, vcom, (uint32_t)C, mk, mid, c->st);
  int bits = 1;
  while (bits < 32 && (C >> bits)) ++bits;
  {
    size_t bytes = 0;
    HIPCHK(sort_pairs_u32(mk, mk2, mid, cedg, g.S, bits, nullptr, bytes, c->st));
    void* tmp = c->scan_tmp(bytes);
    bytes = c->scan_tmp_bytes;
    HIPCHK(sort_pairs_u32(mk, mk2, mid, cedg, g.S, bits, tmp, bytes, c->st));
  }
  k_member_bounds(mk2, g.S, (uint32_t)C, coff, c->st);
  k_member_degrees(cedg, GN, g, mdeg, c->st);
  scan64(c, mdeg, mscan, GN + 1);
  k_gather_yoff(coff, mscan, C, z.off, c->st);
  TierLists ct;
  build_tiers(c, "ag.tiers", z.off, nullptr, nullptr, C, ct);
  k_fill_u32(z.deg, 0, C, c->st);
  k_fill_f64(ztot, 0, C, c->st);
  AggArgs a{};
  a.g = g;
  a.vcom = vcom;
  a.coff = coff;
  a.cedg = cedg;
  a.mscan = mscan;
  a.yoff = z.off;
  a.C = C;
  a.zdeg = z.deg;
  a.zkeys = z.keys;
  a.zw = z.w;
  a.ztot = ztot;
  for (int t = 0; t < 4; ++t) {
    a.list = ct.ids + ct.start[t];
    a.n = ct.start[t + 1] - ct.start[t];
    k_aggregate(a, t, lv.f64, c->st);
  }
  uint64_t n4 = ct.start[5] - ct.start[4];
  if (n4) {
    a.list = ct.ids + ct.start[4];
    a.n = n4;
    uint64_t* caps = c->ws<uint64_t>("ag.caps", n4 + 1);
    uint64_t* chunks = c->ws<uint64_t>("ag.chunks", n4 + 1);
    uint64_t* hoff = c->ws<uint64_t>("ag.hoff", n4 + 1);
    uint64_t* choff = c->ws<uint64_t>("ag.choff", n4 + 1);
    k_t4_sizes(z.off, a.list, n4, C, kAggHbmChunk, caps, chunks, c->st);
    scan64(c, caps, hoff, n4 + 1);
    scan64(c, chunks, choff, n4 + 1);
    uint64_t tot[1], nch[1];
    c->read(tot, hoff + n4, 1);
    c->read(nch, choff + n4, 1);
    a.hkey = c->ws<uint32_t>("ag.hkey", tot[0]);
    a.hpos = c->ws<uint32_t>("ag.hpos", tot[0]);
    a.hval = c->ws<double>("ag.hval", tot[0]);
    a.hoff = hoff;
    a.choff = choff;
    a.nchunks = nch[0];
    k_fill_u32(a.hkey, kEmpty, tot[0], c->st);
    k_fill_u32(a.hpos, kNoPos, tot[0], c->st);
    k_fill_f64(a.hval, 0, tot[0], c->st);
    k_aggregate(a, 4, lv.f64, c->st);
    // rows in first-seen order: a bitmap of 32 words per 1024-edge chunk
    static_assert(kAggHbmChunk == 32 * 32, "32 bitmap words per tier-4 chunk");
    const uint64_t words = nch[0] * 32;
    uint32_t* bm = c->ws<uint32_t>("ag.bm", words);
    uint32_t* wcnt = c->ws<uint32_t>("ag.wcnt", words + 1);
    uint32_t* bsc = c->ws<uint32_t>("ag.bsc", words + 1);
    k_fill_u32(bm, 0, words, c->st);
    k_agg_mark_hbm(a, tot[0], bm, words, wcnt, c->st);
    scan32(c, wcnt, bsc, words + 1);
    k_agg_emit_hbm(a, tot[0], bm, bsc, c->st);
  }
  c->check();
}

// ------------------------------------------------------------ the driver
struct RunTimes {
  float time = 0, mark = 0, init = 0, first = 0, lmove = 0, aggr = 0;
};

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

void validate_options(const lv_options* o) {
  require(o != nullptr, LV_ERR_INVALID, "options is NULL");
  require(o->repeat >= 1, LV_ERR_INVALID, "repeat must be >= 1");
  require(o->max_iterations >= 0 && o->max_passes >= 0, LV_ERR_INVALID, "negative iteration/pass limit");
  require(o->resolution > 0, LV_ERR_INVALID, "resolution must be > 0");
  require(o->tolerance_drop > 0, LV_ERR_INVALID, "tolerance_drop must be > 0");
}

// The pass loop of louvainInvokeOmp (inc/louvain.hxx:1153-1218) for one
// repeat.  ucom/utot/ctot are the state (membership, K, Sigma_tot).
// Hub tables of state s for graph x and membership vcom: slots per vertex
// (a scan of the per-vertex slot counts), the hub list, then one wave per hub.
void build_hub_tables(lv_context* c, const lv_dgraph* x, lv_dstate* s, const uint32_t* vcom) {
  const uint64_t S = x->S;
  if (!s->ht_cap) {
    s->ht_cap = dalloc<uint32_t>(S);
    s->ht_off = dalloc<uint64_t>(S + 1);
  }
  uint64_t* cap64 = c->ws<uint64_t>("ht.cap64", S + 1);
  uint8_t* flags = c->ws<uint8_t>("ht.flags", (S + 7) & ~3ull);
  k_fill_u8(flags, 0, (S + 7) & ~3ull, c->st);
  k_ht_caps(x->deg, S, s->ht_cap, cap64, flags, c->st);
  k_fill_u64(cap64 + S, 0, 1, c->st);
  scan64(c, cap64, s->ht_off, S + 1);
  const uint64_t used = c->read1(s->ht_off + S);
  // spare slots for vertices that pass kHtMinDeg edges in later runs
  const uint64_t slots = used + std::max<uint64_t>(used / 4, (uint64_t)1024 * kHtMaxCap);
  if (!s->ht_next) s->ht_next = dalloc<unsigned long long>(2);
  k_fill_u64(reinterpret_cast<uint64_t*>(s->ht_next), used, 1, c->st);
  k_fill_u64(reinterpret_cast<uint64_t*>(s->ht_next) + 1, 0, 1, c->st);
  s->ht_scan = true;
  if (slots > s->ht_slots) {
    (void)hipFree(s->ht_key);
    (void)hipFree(s->ht_cnt);
    (void)hipFree(s->ht_w);
    s->ht_key = nullptr;
    s->ht_cnt = nullptr;
    s->ht_w = nullptr;
    s->ht_slots = 0;
    s->ht_key = dalloc<uint32_t>(slots);
    s->ht_cnt = dalloc<int32_t>(slots);
    s->ht_w = dalloc<double>(slots);
    s->ht_slots = slots;
  }
  uint32_t* list = c->ws<uint32_t>("ht.list", std::max<uint64_t>(S, 1));
  uint32_t* lens = c->ws<uint32_t>("ht.len", kWlQueues);
  uint64_t* bsum = c->ws<uint64_t>("ht.bsum", (size_t)kWlQueues * kWlScanBlocks);
  WlQueues qs{};
  qs.q[0] = list;
  k_wl_queue(flags, true, S, x->deg, WlSplit{}, qs, lens, bsum, c->st);
  k_ht_build(x->view(), x->f64, vcom, list, lens, s->ht_off, s->ht_cap, s->ht_key, s->ht_cnt, s->ht_w,
             s->ht_next + 1, c->st);
}

// Static re-optimises the whole graph from singletons, and its modularity
// falls with the number of streams moving vertices at once: on R-MAT 24
// (33,554 streams by the stream fraction) four runs averaged 0.0600, with
// at most 8192 streams 0.0615 (the reference: 0.0612-0.0614), in 2.1 s
// instead of 1.5 s.  The other approaches keep the context's schedule.
constexpr uint64_t kStaticStreams = 8192;

void invoke(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, lv_dstate* s, const lv_options* o,
            int& l, int& p, RunTimes& rt, uint64_t& affected) {
  struct Keep {
    lv_context* c;
    uint64_t max_streams;
    ~Keep() { c->max_streams = max_streams; }
  } keep{c, c->max_streams};
  if (approach == LV_STATIC) c->max_streams = std::min<uint64_t>(c->max_streams, kStaticStreams);
  const uint64_t S = x->S;
  const double R = o->resolution;
  const int L = o->max_iterations, P = o->max_passes;
  const double M = x->M2 / 2;
  uint32_t* ucom = s->memb;
  double* utot = s->vtot;
  double* ctot = s->ctot;
  // buffers (allocation and resets are outside the timed region, as in the
  // reference); the super-level membership and K (vcom, vtot) are written
  // by each level's aggregation and initialisation before any read
  uint8_t* vaff = c->ws<uint8_t>("vaff", (S + 7) & ~3ull);
  uint32_t* vcom = c->ws<uint32_t>("vcom", S);
  double* vtot = c->ws<double>("vtot", S);
  uint8_t *dsv = nullptr, *dsn = nullptr, *dsc = nullptr;
  if (approach == LV_DELTA_SCREENING) {
    dsv = c->ws<uint8_t>("ds.v", S);
    dsn = c->ws<uint8_t>("ds.n", S);
    dsc = c->ws<uint8_t>("ds.c", S);
  }
  bool worklist = c->path == 2 && (approach == LV_DYNAMIC_FRONTIER || approach == LV_DELTA_SCREENING);
  // the two super-graphs (gapped CSRs with f64 weights) are taken when a
  // pass first aggregates, sized to the level: a run that never aggregates
  // (a DF or DS step at small batch sizes) takes none
  SuperGraph y, z;
  y.name = "y";
  z.name = "z";
  k_fill_u8(vaff, 0, (S + 7) & ~3ull, c->st);
  if (approach == LV_STATIC) {
    k_fill_u32(ucom, 0, S, c->st);
    k_fill_f64(utot, 0, S, c->st);
    k_fill_f64(ctot, 0, S, c->st);
  }
  double E = o->tolerance;
  hipEvent_t e0 = c->record();
  // fi: initialisation
  if (approach == LV_STATIC) {
    k_vertex_weights(x->view(), x->f64, utot, c->st);
    k_initialize(x->view(), ucom, ctot, utot, c->st);
  } else {
    k_update_weights(b->du, b->dw, b->nd, b->iu, b->iw, b->ni, ucom, utot, ctot, c->st);
  }
  // hub neighbour-community tables of the state: brought up to this graph by
  // the batch's edges, or built on the second run of a state
  lv_dstate* ht = nullptr;
  if (worklist && c->tier_order == 2 && c->use_tables && o->repeat == 1 && approach != LV_STATIC) {
    // the tables describe graph ht_stamp: carried forward when this run's
    // graph is that graph (empty batch) or that graph with this batch applied
    // by lv_dgraph_update; anything else rebuilds them
    const bool same = !b->nd && !b->ni && s->ht_stamp == x->stamp;
    const bool chained = b->from_stamp && b->from_stamp == s->ht_stamp && b->to_stamp == x->stamp;
    if (s->ht_valid && (same || chained)) {
      // deletions leave the tables with the weight the edge had in the graph
      k_ht_batch(b->du, b->dv, b->dws, b->nd, b->iu, b->iv, b->iw, b->ni, ucom, s->ht_off, s->ht_cap, s->ht_key,
                 s->ht_cnt, s->ht_w, s->ht_next + 1, c->st);
      // vertices that passed kHtMinDeg edges with this batch: tables from
      // the spare slots, built from this graph (after the batch deltas)
      uint32_t* nl = c->ws<uint32_t>("ht.newlist", std::max<uint64_t>(S, 1));
      uint32_t* nn = c->ws<uint32_t>("ht.newn", 1);
      k_fill_u32(nn, 0, 1, c->st);
      k_ht_new_hubs(b->iu, b->ni, x->deg, s->ht_cap, s->ht_off, s->ht_next, s->ht_slots, nl, nn, c->st);
      k_ht_build(x->view(), x->f64, ucom, nl, nn, s->ht_off, s->ht_cap, s->ht_key, s->ht_cnt, s->ht_w,
                 s->ht_next + 1, c->st);
      ht = s;
    } else if (s->runs >= 1) {
      build_hub_tables(c, x, s, ucom);
      ht = s;
    }
  }
  s->ht_valid = false;  // until this run ends with its tables intact
  hipEvent_t e1 = c->record();
  // fm: marking
  const uint8_t* fa = nullptr;
  if (approach == LV_DYNAMIC_FRONTIER) {
    k_frontier(b->du, b->dv, b->nd, b->iu, b->iv, b->ni, ucom, vaff, c->st);
  } else if (approach == LV_DELTA_SCREENING) {
    k_fill_u8(dsv, 0, S, c->st);
    k_fill_u8(dsn, 0, S, c->st);
    k_fill_u8(dsc, 0, S, c->st);
    k_ds_deletions(b->du, b->dv, b->nd, ucom, dsv, dsn, dsc, c->st);
    if (b->ni) {
      uint32_t* tkey = c->ws<uint32_t>("ds.tkey", 2 * b->ni);
      double* tval = c->ws<double>("ds.tval", 2 * b->ni);
      uint32_t* tlist = c->ws<uint32_t>("ds.tlist", b->ni);
      k_fill_u32(tkey, kEmpty, 2 * b->ni, c->st);
      k_ds_insertions(b->iu, b->iv, b->iw, b->ni, ucom, utot, ctot, M, R, tkey, tval, tlist, dsv, dsn, dsc, c->st);
    }
    k_ds_sweep(x->view(), ucom, dsn, dsc, dsv, c->st);
    d2d(c, vaff, dsv, S);
    fa = dsv;
  } else {
    k_fill_u8(vaff, 1, S, c->st);
  }
  hipEvent_t e2 = c->record();
  hipEvent_t t1 = e2;
  bool t1_set = false;
  Level lv0{x->view(), x->f64, &x->tiers, x->order, x->tiers.maxdeg, &x->hubs, x->hub_slots, x->stamp,
            x->slab_slots, x->maxdeg};
  Level cur = lv0;
  TierLists ytiers, yhubs;
  uint64_t ySpan = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> lm, ag;
  for (l = 0, p = 0; M > 0 && P > 0;) {
    if (p == 1) { t1 = c->record(); t1_set = true; }
    bool first = p == 0;
    int m;
    hipEvent_t a0 = c->record();
    if (first) {
      if (worklist) m = local_move_worklist(c, cur, ucom, utot, ctot, vaff, fa, M, R, L, E, nullptr, ht);
      else m = local_move(c, cur, ucom, utot, ctot, vaff, fa, M, R, L, E,
                          approach == LV_DYNAMIC_FRONTIER || approach == LV_DELTA_SCREENING, nullptr,
                          approach == LV_STATIC ? 0 : -1);
    } else {
      m = local_move(c, cur, vcom, vtot, ctot, vaff, nullptr, M, R, L, E, false, nullptr, p);
    }
    lm.push_back({a0, c->record()});
    l += std::max(m, 1);
    ++p;
    if (m <= 1 || p >= P) break;
    uint32_t* map = nullptr;
    uint32_t* mcur = first ? ucom : vcom;
    uint64_t CN = count_communities(c, cur.g, mcur, map);
    uint64_t GN = cur.order;
    if ((double)CN / (double)GN >= o->aggregation_tolerance) break;
    if (first) k_lookup(ucom, S, map, c->st);
    else {
      k_lookup(vcom, cur.g.S, map, c->st);
      k_lookup(ucom, S, vcom, c->st);
    }
    // (a level's super-graph has at most the input's live edges)
    z = super_graph(c, z.name, CN, x->live);
    hipEvent_t g0 = c->record();
    aggregate(c, cur, mcur, CN, z, vtot);
    ag.push_back({g0, c->record()});
    std::swap(y, z);
    ySpan = CN;
    // next level: vcom[u] = u, Sigma_tot = K, every super-vertex affected
    GView yg{ySpan, y.off, y.deg, y.keys, y.w, nullptr};
    k_initialize(yg, vcom, ctot, vtot, c->st);
    k_fill_u8(vaff, 1, CN, c->st);
    uint64_t yslots = 0;
    if (c->tier_order == 2) {
      build_tiers(c, "y.hubs", y.off, y.deg, nullptr, CN, yhubs, nullptr, kHubMin);
      yhubs.maxdeg = max_degree_tier4(c, yhubs, y.deg);
      yslots = caps_total(c, y.deg, CN, kHubMin);
    } else {
      build_tiers(c, "y.tiers", y.off, y.deg, nullptr, CN, ytiers);
      ytiers.maxdeg = max_degree_tier4(c, ytiers, y.deg);
    }
    cur = Level{yg, true, &ytiers, CN, ytiers.maxdeg, &yhubs, yslots};
    E /= o->tolerance_drop;
  }
  if (p > 1) k_lookup(ucom, S, vcom, c->st);
  // the tables stay valid if the membership kept its ids (one pass)
  if (ht && p == 1) {
    s->ht_valid = true;
    s->ht_stamp = x->stamp;
  }
  ++s->runs;
  hipEvent_t e3 = c->record();
  if (!t1_set) t1 = e3;
  // countValueOmp(vaff) of the result (inc/louvain.hxx:1220), after the
  // timed region, read back with the run's one closing synchronisation
  unsigned long long* dcount = c->ws<unsigned long long>("affected", 1);
  k_fill_u64(reinterpret_cast<uint64_t*>(dcount), 0, 1, c->st);
  k_count_u8(vaff, 1, S, dcount, c->st);
  unsigned long long* pin = c->tail_pin();
  HIPCHK(hipMemcpyAsync(pin, dcount, sizeof *pin, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->check();
  affected = *pin;
  rt.time += elapsed(e0, e3);
  rt.init += elapsed(e0, e1);
  rt.mark += elapsed(e1, e2);
  rt.first += elapsed(e2, t1);
  for (auto& pr : lm) rt.lmove += elapsed(pr.first, pr.second);
  for (auto& pr : ag) rt.aggr += elapsed(pr.first, pr.second);
}

lv_tune_choice tune(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, const lv_dstate* s,
                    const lv_options* o, int reps, const std::vector<uint64_t>* voff);
void apply_tuned(lv_context* c, int approach, const lv_dgraph* x, uint64_t G);
std::string shape_key(const lv_context* c, int approach, const lv_dgraph* x, uint64_t G);

void run(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, lv_dstate* s, const lv_options* o,
         lv_result* res) {
  require(c && x && s, LV_ERR_INVALID, "null context, graph or state");
  require(approach >= LV_STATIC && approach <= LV_DYNAMIC_FRONTIER, LV_ERR_INVALID, "unknown approach");
  require(approach == LV_STATIC || b != nullptr, LV_ERR_INVALID, "dynamic approaches need a batch");
  require(s->S == x->S, LV_ERR_INVALID, "state span differs from graph span");
  require(!x->broken, LV_ERR_INVALID, "graph left inconsistent by a failed update");
  validate_options(o);
  HIPCHK(hipSetDevice(c->device));
  apply_tuned(c, approach, x, 1);
  c->ev_used = 0;
  const uint64_t S = x->S;
  int repeat = o->repeat;
  if (repeat > 1) s->ht_reset();  // each repeat restarts from the caller's state
  // keep the caller's state for every repeat (louvainSetupInitialsW, inc/louvain.hxx:1240-1249)
  uint32_t* bm = nullptr;
  double *bv = nullptr, *bc = nullptr;
  if (repeat > 1 && approach != LV_STATIC) {
    bm = c->ws<uint32_t>("bak.m", S);
    bv = c->ws<double>("bak.v", S);
    bc = c->ws<double>("bak.c", S);
    d2d(c, bm, s->memb, S);
    d2d(c, bv, s->vtot, S);
    d2d(c, bc, s->ctot, S);
  }
  RunTimes rt;
  int l = 0, p = 0;
  uint64_t aff = 0;  // affected vertices after the last repeat
  for (int r = 0; r < repeat; ++r) {
    if (r > 0 && approach != LV_STATIC) {
      d2d(c, s->memb, bm, S);
      d2d(c, s->vtot, bv, S);
      d2d(c, s->ctot, bc, S);
    }
    c->ev_used = 0;
    c->prof.clear();
    invoke(c, approach, x, b, s, o, l, p, rt, aff);
    for (auto& pr : c->prof) add_prof(c, pr.first, elapsed(pr.second.first, pr.second.second));
  }
  if (res) {
    res->iterations = l;
    res->passes = p;
    res->time = rt.time / repeat;
    res->marking_time = rt.mark / repeat;
    res->initialization_time = rt.init / repeat;
    res->first_pass_time = rt.first / repeat;
    res->local_move_time = rt.lmove / repeat;
    res->aggregation_time = rt.aggr / repeat;
    res->affected_vertices = aff;
    if (res->membership) d2h(c, res->membership, s->memb, S);
    if (res->vertex_weight) d2h(c, res->vertex_weight, s->vtot, S);
    if (res->community_weight) d2h(c, res->community_weight, s->ctot, S);
    if (res->membership || res->vertex_weight || res->community_weight) HIPCHK(hipStreamSynchronize(c->st));
  }
  c->stats.runs += 1;
}

// ------------------------------------------------------------ batched form
// louvainInvokeOmp (inc/louvain.hxx:1120-1221) for G graphs laid side by
// side in x (graph k owns vertex ids [voff[k], voff[k+1]), no edges between
// graphs): every launch covers all graphs, and each graph keeps its own M,
// tolerance, iteration count, pass count and aggregation decision, so it
// runs as it would alone; a graph that has finished is skipped while the
// others go on.  Community ids are re-based at the end so that graph k's
// final ids lie in its own vertex range.
void invoke_batched(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, lv_dstate* s,
                    const lv_options* o, const std::vector<uint64_t>& voff, std::vector<int>& l_out,
                    std::vector<int>& p_out, RunTimes& rt) {
  const uint64_t S = x->S, G = voff.size() - 1;
  const double R = o->resolution;
  const int L = o->max_iterations, P = o->max_passes;
  uint32_t* ucom = s->memb;
  double* utot = s->vtot;
  double* ctot = s->ctot;
  uint8_t* vaff = c->ws<uint8_t>("vaff", (S + 7) & ~3ull);
  uint32_t* vcom = c->ws<uint32_t>("vcom", S);
  double* vtot = c->ws<double>("vtot", S);
  bool worklist = c->path == 2 && (approach == LV_DYNAMIC_FRONTIER || approach == LV_DELTA_SCREENING);
  uint8_t *dsv = nullptr, *dsn = nullptr, *dsc = nullptr;
  if (approach == LV_DELTA_SCREENING) {
    dsv = c->ws<uint8_t>("ds.v", S);
    dsn = c->ws<uint8_t>("ds.n", S);
    dsc = c->ws<uint8_t>("ds.c", S);
  }
  SuperGraph y, z;  // taken at the first aggregation, as in invoke()
  y.name = "y";
  z.name = "z";
  // per-graph data: graph of every vertex (level 0 and the current level),
  // M from the graph's own edge weights (edgeWeightOmp of each graph)
  uint64_t* dvoff = c->ws<uint64_t>("bt.voff", G + 1);
  h2d(c, dvoff, voff.data(), G + 1);
  uint32_t* gid0 = c->ws<uint32_t>("bt.gid0", S);
  uint32_t* gidA = c->ws<uint32_t>("bt.gidA", S);
  uint32_t* gidB = c->ws<uint32_t>("bt.gidB", S);
  k_gid_fill(dvoff, G, S, gid0, c->st);
  lv_dgraph::Groups& gr = x->groups;
  if (gr.version != x->version || gr.voff != voff) {
    double* kk = c->ws<double>("bt.k", S);
    double* m2 = c->ws<double>("bt.m2", G);
    k_fill_f64(kk, 0, S, c->st);
    k_fill_f64(m2, 0, G, c->st);
    k_vertex_weights(x->view(), x->f64, kk, c->st);
    k_group_sum(kk, gid0, S, m2, c->st);
    gr.M2.assign(G, 0);
    c->read(gr.M2.data(), m2, G);
    // existing vertices per graph (GN of the first aggregation test)
    gr.order.assign(G, 0);
    if (x->ex) {
      unsigned long long* gn = c->ws<unsigned long long>("bt.gn", G);
      k_range_count(x->ex, dvoff, G, gn, c->st);
      std::vector<unsigned long long> h(G);
      c->read(h.data(), gn, G);
      for (uint64_t k = 0; k < G; ++k) gr.order[k] = h[k];
    } else {
      for (uint64_t k = 0; k < G; ++k) gr.order[k] = voff[k + 1] - voff[k];
    }
    gr.voff = voff;
    gr.version = x->version;
  }
  const std::vector<double>& M2 = gr.M2;
  Batched bt;
  bt.G = G;
  bt.gid = gid0;
  bt.gM = c->ws<double>("bt.gM", G);
  bt.gzero = c->ws<uint32_t>("bt.gzero", G);
  bt.gel = c->ws<double>("bt.gel", G);
  bt.gmask = c->ws<uint8_t>("bt.gmask", G);
  bt.E.assign(G, o->tolerance);
  bt.el.assign(G, 0);
  bt.it.assign(G, 0);
  bt.skip.assign(G, 0);
  bt.done.assign(G, 0);
  std::vector<double> gM(G);
  std::vector<uint32_t> gzero(G);
  std::vector<uint64_t> lo(voff), lo_next(G + 1);
  for (uint64_t k = 0; k < G; ++k) {
    gM[k] = M2[k] / 2;
    gzero[k] = (uint32_t)voff[k];
    bt.done[k] = !(gM[k] > 0 && P > 0);
  }
  h2d(c, bt.gM, gM.data(), G);
  h2d(c, bt.gzero, gzero.data(), G);
  std::vector<uint64_t> GN = gr.order;
  l_out.assign(G, 0);
  p_out.assign(G, 0);
  k_fill_u8(vaff, 0, (S + 7) & ~3ull, c->st);
  k_fill_u32(vcom, 0, S, c->st);
  k_fill_f64(vtot, 0, S, c->st);
  if (approach == LV_STATIC) {
    k_fill_u32(ucom, 0, S, c->st);
    k_fill_f64(utot, 0, S, c->st);
    k_fill_f64(ctot, 0, S, c->st);
  }
  hipEvent_t e0 = c->record();
  if (approach == LV_STATIC) {
    k_vertex_weights(x->view(), x->f64, utot, c->st);
    k_initialize(x->view(), ucom, ctot, utot, c->st);
  } else {
    k_update_weights(b->du, b->dw, b->nd, b->iu, b->iw, b->ni, ucom, utot, ctot, c->st);
  }
  hipEvent_t e1 = c->record();
  // marking as in invoke(); delta-screening's insertion gains take each
  // source's M from its own graph
  const uint8_t* fa = nullptr;
  if (approach == LV_DYNAMIC_FRONTIER) {
    k_frontier(b->du, b->dv, b->nd, b->iu, b->iv, b->ni, ucom, vaff, c->st);
  } else if (approach == LV_DELTA_SCREENING) {
    k_fill_u8(dsv, 0, S, c->st);
    k_fill_u8(dsn, 0, S, c->st);
    k_fill_u8(dsc, 0, S, c->st);
    k_ds_deletions(b->du, b->dv, b->nd, ucom, dsv, dsn, dsc, c->st);
    if (b->ni) {
      uint32_t* tkey = c->ws<uint32_t>("ds.tkey", 2 * b->ni);
      double* tval = c->ws<double>("ds.tval", 2 * b->ni);
      uint32_t* tlist = c->ws<uint32_t>("ds.tlist", b->ni);
      k_fill_u32(tkey, kEmpty, 2 * b->ni, c->st);
      k_ds_insertions(b->iu, b->iv, b->iw, b->ni, ucom, utot, ctot, 0.0, R, tkey, tval, tlist, dsv, dsn, dsc, c->st,
                      gid0, bt.gM);
    }
    k_ds_sweep(x->view(), ucom, dsn, dsc, dsv, c->st);
    d2d(c, vaff, dsv, S);
    fa = dsv;
  } else {
    k_fill_u8(vaff, 1, S, c->st);
  }
  hipEvent_t e2 = c->record();
  hipEvent_t t1 = e2;
  bool t1_set = false;
  Level cur{x->view(), x->f64, &x->tiers, x->order, x->tiers.maxdeg, &x->hubs, x->hub_slots, x->stamp,
            x->slab_slots, x->maxdeg};
  TierLists ytiers, yhubs;
  uint32_t* gidcur = gid0;
  uint64_t levelS = S;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> lm, ag;
  auto all_done = [&] {
    for (uint64_t k = 0; k < G; ++k)
      if (!bt.done[k]) return false;
    return true;
  };
  // A graph is finalised at the level where it finishes: its membership and
  // Sigma_tot go to its own vertex range (Sigma_tot into ftot), and the levels
  // after that hold only the graphs still running.
  double* ftot = c->ws<double>("bt.tot", S);
  k_fill_f64(ftot, 0, S, c->st);
  uint64_t* dlo = c->ws<uint64_t>("bt.lo", G + 1);
  uint8_t* dfin = c->ws<uint8_t>("bt.fin", G);
  uint8_t* dnew = c->ws<uint8_t>("bt.new", G);
  uint8_t* ddone = c->ws<uint8_t>("bt.done", G);
  uint8_t* lmask = c->ws<uint8_t>("bt.mask", S);
  std::vector<uint8_t> fin(G, 0), fresh(G, 0);
  // no graph is finalised yet (the flags of an earlier call are cleared:
  // the lookups below leave a flagged graph's entries unmapped)
  k_fill_u8(dfin, 0, G, c->st);
  auto finalize = [&](bool first_level) {
    bool any_new = false;
    for (uint64_t k = 0; k < G; ++k) {
      fresh[k] = bt.done[k] && !fin[k];
      any_new = any_new || fresh[k];
    }
    if (!any_new) return;
    h2d(c, dnew, fresh.data(), G);
    h2d(c, dlo, lo.data(), G + 1);
    k_finalize(ucom, first_level ? nullptr : vcom, gid0, x->ex, S, ctot, levelS, dlo, G, dnew, dvoff, ftot, c->st);
    for (uint64_t k = 0; k < G; ++k) fin[k] = fin[k] || fresh[k];
    h2d(c, dfin, fin.data(), G);
  };
  if (all_done()) finalize(true);
  int p = 0;
  while (!all_done()) {
    if (p == 1) { t1 = c->record(); t1_set = true; }
    const bool first = p == 0;
    bt.gid = gidcur;
    batched_level(c, &bt, lo, GN);
    hipEvent_t a0 = c->record();
    if (first) {
      if (worklist) local_move_worklist(c, cur, ucom, utot, ctot, vaff, fa, 0, R, L, 0, &bt);
      else local_move(c, cur, ucom, utot, ctot, vaff, fa, 0, R, L, 0,
                      approach == LV_DYNAMIC_FRONTIER || approach == LV_DELTA_SCREENING, &bt);
    } else {
      local_move(c, cur, vcom, vtot, ctot, vaff, nullptr, 0, R, L, 0, false, &bt);
    }
    lm.push_back({a0, c->record()});
    ++p;
    // per graph: l += max(m, 1); stop passing when m <= 1 or p >= P
    // (inc/louvain.hxx:1184-1186)
    for (uint64_t k = 0; k < G; ++k) {
      if (bt.done[k]) continue;
      const int m = (bt.it[k] > 1 || bt.el[k]) ? bt.it[k] : 0;
      l_out[k] += std::max(m, 1);
      p_out[k] += 1;
      if (m <= 1 || p_out[k] >= P) bt.done[k] = 1;
    }
    uint32_t* mcur = first ? ucom : vcom;
    uint32_t* map = nullptr;
    uint32_t* lo32 = c->ws<uint32_t>("bt.lo32", G + 1);
    uint32_t* mlo = c->ws<uint32_t>("bt.mlo", G + 1);
    std::vector<uint32_t> lo_h(G + 1), mlo_h(G + 1);
    for (uint64_t k = 0; k <= G; ++k) lo_h[k] = (uint32_t)lo[k];
    // communities per graph: the renumbering map at the graph range bounds
    auto range_counts = [&] {
      h2d(c, lo32, lo_h.data(), G + 1);
      k_gather_u32(map, lo32, G + 1, mlo, c->st);
      c->read(mlo_h.data(), mlo, G + 1);
    };
    const std::vector<uint64_t> GNlevel = GN;
    if (!all_done()) {
      count_communities(c, cur.g, mcur, map);
      range_counts();
      for (uint64_t k = 0; k < G; ++k) {
        const uint64_t CN = mlo_h[k + 1] - mlo_h[k];
        if (!bt.done[k] && (double)CN / (double)GN[k] >= o->aggregation_tolerance) bt.done[k] = 1;
        GN[k] = CN;
      }
    }
    finalize(first);
    if (all_done()) break;
    // the next level: communities of the graphs still running
    h2d(c, ddone, bt.done.data(), G);
    k_active_mask(gidcur, levelS, first ? x->ex : nullptr, ddone, lmask, c->st);
    GView act = cur.g;
    act.ex = lmask;
    uint64_t live = 0;
    for (uint64_t k = 0; k < G; ++k)
      if (!bt.done[k]) live += GNlevel[k];
    count_communities(c, act, mcur, map);
    range_counts();
    for (uint64_t k = 0; k <= G; ++k) lo_next[k] = mlo_h[k];
    uint32_t* gidnext = gidcur == gidA ? gidB : gidA;
    k_gid_next(mcur, gidcur, map, lmask, levelS, gidnext, c->st);
    if (first) {
      k_lookup_active(ucom, S, map, gid0, dfin, c->st);
    } else {
      k_lookup(vcom, cur.g.S, map, c->st);
      k_lookup_active(ucom, S, vcom, gid0, dfin, c->st);
    }
    const uint64_t CN = lo_next[G];
    z = super_graph(c, z.name, CN, x->live);
    hipEvent_t g0 = c->record();
    Level agl = cur;
    agl.g = act;
    agl.order = live;
    aggregate(c, agl, mcur, CN, z, vtot);
    ag.push_back({g0, c->record()});
    std::swap(y, z);
    GView yg{CN, y.off, y.deg, y.keys, y.w, nullptr};
    k_initialize(yg, vcom, ctot, vtot, c->st);
    k_fill_u8(vaff, 1, CN, c->st);
    uint64_t yslots = 0;
    if (c->tier_order == 2) {
      build_tiers(c, "y.hubs", y.off, y.deg, nullptr, CN, yhubs, nullptr, kHubMin);
      yhubs.maxdeg = max_degree_tier4(c, yhubs, y.deg);
      yslots = caps_total(c, y.deg, CN, kHubMin);
    } else {
      build_tiers(c, "y.tiers", y.off, y.deg, nullptr, CN, ytiers);
      ytiers.maxdeg = max_degree_tier4(c, ytiers, y.deg);
    }
    cur = Level{yg, true, &ytiers, CN, ytiers.maxdeg, &yhubs, yslots};
    gidcur = gidnext;
    levelS = CN;
    lo = lo_next;
    for (uint64_t k = 0; k < G; ++k) {
      gzero[k] = (uint32_t)lo[k];
      if (!bt.done[k]) bt.E[k] /= o->tolerance_drop;
    }
    h2d(c, bt.gzero, gzero.data(), G);
  }
  d2d(c, ctot, ftot, S);
  hipEvent_t e3 = c->record();
  if (!t1_set) t1 = e3;
  HIPCHK(hipEventSynchronize(e3));
  c->check();
  rt.time += elapsed(e0, e3);
  rt.init += elapsed(e0, e1);
  rt.mark += elapsed(e1, e2);
  rt.first += elapsed(e2, t1);
  for (auto& pr : lm) rt.lmove += elapsed(pr.first, pr.second);
  for (auto& pr : ag) rt.aggr += elapsed(pr.first, pr.second);
}

void run_batched(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, lv_dstate* s,
                 const uint64_t* voff, uint64_t G, const lv_options* o, lv_result* res) {
  require(c && x && s && voff && G > 0, LV_ERR_INVALID, "null context, graph, state or graph ranges");
  require(approach >= LV_STATIC && approach <= LV_DYNAMIC_FRONTIER, LV_ERR_INVALID, "unknown approach");
  require(approach == LV_STATIC || b != nullptr, LV_ERR_INVALID, "dynamic approaches need a batch");
  require(s->S == x->S, LV_ERR_INVALID, "state span differs from graph span");
  require(voff[0] == 0 && voff[G] == x->S, LV_ERR_INVALID, "graph ranges must cover the span from 0");
  require(!x->broken, LV_ERR_INVALID, "graph left inconsistent by a failed update");
  for (uint64_t k = 0; k < G; ++k) require(voff[k] <= voff[k + 1], LV_ERR_INVALID, "graph ranges not ascending");
  validate_options(o);
  require(o->repeat == 1, LV_ERR_INVALID, "the batched form runs one repeat");
  s->ht_reset();  // the batched form moves vertices without hub tables
  HIPCHK(hipSetDevice(c->device));
  std::vector<uint64_t> vo(voff, voff + G + 1);
  // a shape seen for the first time is tuned first (on clones of s)
  if (c->autotune && !c->tuning && !c->knobs_hand && approach != LV_STATIC &&
      !c->tuned.count(shape_key(c, approach, x, G)))
    tune(c, approach, x, b, s, o, 3, &vo);
  apply_tuned(c, approach, x, G);
  c->ev_used = 0;
  c->prof.clear();
  RunTimes rt;
  std::vector<int> l, p;
  invoke_batched(c, approach, x, b, s, o, vo, l, p, rt);
  for (auto& pr : c->prof) add_prof(c, pr.first, elapsed(pr.second.first, pr.second.second));
  const uint64_t S = x->S;
  unsigned long long* aff = c->ws<unsigned long long>("bt.aff", G);
  k_range_count(c->ws<uint8_t>("vaff", (S + 7) & ~3ull), c->ws<uint64_t>("bt.voff", G + 1), G, aff, c->st);
  std::vector<unsigned long long> affh(G);
  c->read(affh.data(), aff, G);
  if (res) {
    for (uint64_t k = 0; k < G; ++k) {
      lv_result& r = res[k];
      const uint64_t n = voff[k + 1] - voff[k];
      r.iterations = l[k];
      r.passes = p[k];
      r.time = rt.time;
      r.marking_time = rt.mark;
      r.initialization_time = rt.init;
      r.first_pass_time = rt.first;
      r.local_move_time = rt.lmove;
      r.aggregation_time = rt.aggr;
      r.affected_vertices = affh[k];
      if (r.membership) d2h(c, r.membership, s->memb + voff[k], n);
      if (r.vertex_weight) d2h(c, r.vertex_weight, s->vtot + voff[k], n);
      if (r.community_weight) d2h(c, r.community_weight, s->ctot + voff[k], n);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    // graph-local community ids (absent vertices, whose entries carry no
    // community, read 0)
    for (uint64_t k = 0; k < G; ++k) {
      if (!res[k].membership) continue;
      for (uint64_t u = 0; u < voff[k + 1] - voff[k]; ++u) {
        const uint64_t m = res[k].membership[u];
        res[k].membership[u] = (m >= voff[k] && m < voff[k + 1]) ? (uint32_t)(m - voff[k]) : 0u;
      }
    }
  }
  c->stats.runs += 1;
}

// ------------------------------------------------------------ schedule tuner
// Modularity of s's membership on g (modularityByOmp, inc/properties.hxx:205-233).
double dmodularity(lv_context* c, const lv_dgraph* g, const lv_dstate* s, double R) {
  double* cin = c->ws<double>("mod.cin", g->S);
  double* ct = c->ws<double>("mod.ctot", g->S);
  double* o = c->ws<double>("mod.out", 1);
  k_fill_f64(cin, 0, g->S, c->st);
  k_fill_f64(ct, 0, g->S, c->st);
  k_fill_f64(o, 0, 1, c->st);
  k_modularity_parts(g->view(), g->f64, s->memb, cin, ct, c->st);
  k_modularity_sum(cin, ct, g->S, g->M2 / 2, R, o, c->st);
  return c->read1(o);
}

// Modularity of each graph of a union (graph k on ids [voff[k], voff[k+1]),
// each with its own M, as after run_batched): the per-community terms of
// the union summed per graph.
std::vector<double> dmodularity_groups(lv_context* c, const lv_dgraph* g, const lv_dstate* s, double R,
                                       const std::vector<uint64_t>& voff) {
  const uint64_t S = g->S, G = voff.size() - 1;
  const auto& M2 = g->groups.M2;
  require(g->groups.voff == voff && M2.size() == G, LV_ERR_INVALID, "per-graph weights not computed for these ranges");
  double* cin = c->ws<double>("mod.cin", S);
  double* ct = c->ws<double>("mod.ctot", S);
  uint64_t* dvoff = c->ws<uint64_t>("mod.voff", G + 1);
  uint32_t* gid = c->ws<uint32_t>("mod.gid", S);
  double* gM = c->ws<double>("mod.gM", G);
  double* o = c->ws<double>("mod.gout", G);
  std::vector<double> hM(G);
  for (uint64_t k = 0; k < G; ++k) hM[k] = M2[k] / 2;
  h2d(c, dvoff, voff.data(), G + 1);
  h2d(c, gM, hM.data(), G);
  k_gid_fill(dvoff, G, S, gid, c->st);
  k_fill_f64(cin, 0, S, c->st);
  k_fill_f64(ct, 0, S, c->st);
  k_fill_f64(o, 0, G, c->st);
  k_modularity_parts(g->view(), g->f64, s->memb, cin, ct, c->st);
  k_group_modularity(cin, ct, gid, gM, S, R, c->st);
  k_group_sum(cin, gid, S, o, c->st);
  std::vector<double> q(G);
  c->read(q.data(), o, G);
  return q;
}

// The knobs a tuning chooses: the worklist and sweep hub degrees (the
// degree bucket between the id-order kernel and the hub kernels), the
// worklist id-order streams and queue layout, and the streams per live
// vertex of a level.  The reference's counterpart is its fixed
// schedule(dynamic, 2048) (inc/louvain.hxx:594).
lv_tune_choice default_knobs() {
  lv_tune_choice k{};
  k.worklist_hub_degree = 256;
  k.sweep_hub_degree = kHubMin;
  k.worklist_streams = kWlStreams;
  k.interleave = 0;
  k.stream_fraction = 0.002;
  return k;
}
bool same_knobs(const lv_tune_choice& a, const lv_tune_choice& b) {
  return a.worklist_hub_degree == b.worklist_hub_degree && a.sweep_hub_degree == b.sweep_hub_degree &&
         a.worklist_streams == b.worklist_streams && a.interleave == b.interleave &&
         a.stream_fraction == b.stream_fraction;
}
lv_tune_choice knobs_of(const lv_context* c) {
  lv_tune_choice k = default_knobs();
  k.worklist_hub_degree = c->hub_min_wl;
  k.sweep_hub_degree = c->hub_min;
  k.worklist_streams = c->wl_streams;
  k.interleave = c->interleave;
  k.stream_fraction = c->stream_fraction;
  return k;
}
void apply_knobs(lv_context* c, const lv_tune_choice& k) {
  c->hub_min_wl = k.worklist_hub_degree;
  c->hub_min = k.sweep_hub_degree;
  c->wl_streams = k.worklist_streams;
  c->interleave = k.interleave;
  c->stream_fraction = k.stream_fraction;
}
// A public setter of a tuned knob: from the defaults (leaving automatic
// mode) or from the knobs set by hand so far; back in automatic mode when
// every tuned knob is at its default again.
template <class F>
void set_by_hand(lv_context* c, F set) {
  if (!c->knobs_hand) apply_knobs(c, default_knobs());
  set();
  c->knobs_hand = !same_knobs(knobs_of(c), default_knobs());
}

// Graph shape: approach, floor(log2) of the span and of the live edges, and
// for the batched form the graphs' count (log2 + 1).
std::string shape_key(const lv_context* c, int approach, const lv_dgraph* x, uint64_t G) {
  auto lg = [](uint64_t v) {
    int b = 0;
    while (v >>= 1) ++b;
    return b;
  };
  return std::to_string(approach) + "/" + std::to_string(c->path) + "/" + std::to_string(lg(std::max<uint64_t>(x->S, 1))) +
         "/" + std::to_string(lg(std::max<uint64_t>(x->live, 1))) + "/" + std::to_string(G > 1 ? lg(G) + 1 : 0);
}

// Candidates: single-graph worklist approaches (DF, DS on path 2) the
// worklist hub degree, id-order streams and queue layout; single-graph
// sweeps (ND, or any approach off path 2) the sweep hub degree and the
// streams per live vertex; the batched form the streams per live vertex,
// the worklist hub degree and the queue layout.  The defaults first.
std::vector<lv_tune_choice> tune_candidates(const lv_context* c, int approach, bool batched) {
  std::vector<lv_tune_choice> v{default_knobs()};
  const bool worklist = c->path == 2 && (approach == LV_DYNAMIC_FRONTIER || approach == LV_DELTA_SCREENING);
  auto add = [&](const lv_tune_choice& k) {
    for (const auto& e : v)
      if (same_knobs(e, k)) return;
    v.push_back(k);
  };
  if (batched) {
    for (double f : {0.001, 0.002, 0.004})
      for (uint32_t h : {256u, 512u})
        for (int il : {0, 1}) {
          lv_tune_choice k = default_knobs();
          k.stream_fraction = f;
          k.worklist_hub_degree = h;
          k.interleave = il;
          add(k);
        }
  } else if (worklist) {
    for (uint32_t h : {256u, 512u})
      for (uint64_t st : {1536ull, 3072ull, 6144ull})
        for (int il : {0, 1}) {
          lv_tune_choice k = default_knobs();
          k.worklist_hub_degree = h;
          k.worklist_streams = st;
          k.interleave = il;
          add(k);
        }
  } else {
    for (uint32_t h : {(uint32_t)kHubMin, 16384u})
      for (double f : {0.001, 0.002, 0.004}) {
        lv_tune_choice k = default_knobs();
        k.sweep_hub_degree = h;
        k.stream_fraction = f;
        add(k);
      }
  }
  return v;
}

// s's membership, K, Sigma_tot and (when valid) hub tables into t, which
// takes table buffers as it needs; t then stands where s stands (its next
// run carries or builds the tables as s's would).
void clone_state(lv_context* c, lv_dstate* t, const lv_dstate* s) {
  require(t->S == s->S, LV_ERR_INVALID, "state spans differ");
  const uint64_t S = s->S;
  d2d(c, t->memb, s->memb, S);
  d2d(c, t->vtot, s->vtot, S);
  d2d(c, t->ctot, s->ctot, S);
  t->ht_reset();
  if (s->ht_valid && s->ht_cap && s->ht_next && s->ht_slots) {
    if (!t->ht_cap) {
      t->ht_cap = dalloc<uint32_t>(S);
      t->ht_off = dalloc<uint64_t>(S + 1);
    }
    if (!t->ht_next) t->ht_next = dalloc<unsigned long long>(2);
    if (t->ht_slots < s->ht_slots) {
      (void)hipFree(t->ht_key);
      (void)hipFree(t->ht_cnt);
      (void)hipFree(t->ht_w);
      t->ht_key = nullptr;
      t->ht_cnt = nullptr;
      t->ht_w = nullptr;
      t->ht_slots = 0;
      t->ht_key = dalloc<uint32_t>(s->ht_slots);
      t->ht_cnt = dalloc<int32_t>(s->ht_slots);
      t->ht_w = dalloc<double>(s->ht_slots);
      t->ht_slots = s->ht_slots;
    }
    d2d(c, t->ht_cap, s->ht_cap, S);
    d2d(c, t->ht_off, s->ht_off, S + 1);
    d2d(c, t->ht_next, s->ht_next, 2);
    d2d(c, t->ht_key, s->ht_key, s->ht_slots);
    d2d(c, t->ht_cnt, s->ht_cnt, s->ht_slots);
    d2d(c, t->ht_w, s->ht_w, s->ht_slots);
    t->ht_valid = true;
    t->ht_stamp = s->ht_stamp;
    t->ht_scan = s->ht_scan;
  }
  t->runs = s->runs;
}

void run(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, lv_dstate* s, const lv_options* o,
         lv_result* res);

// Times every candidate schedule of `approach` on graph x with batch b from
// state s (each trial on a clone of s: s is left as it was), `reps` runs
// each, median of the host clock around the run; keeps the fastest whose
// modularity (for the batched form, every graph's) is within 1e-3 of the
// defaults', and the defaults unless it is
// faster than them by more than 3% (run-to-run noise).  The choice is
// cached under the graph's shape and applied to the context.
lv_tune_choice tune(lv_context* c, int approach, const lv_dgraph* x, const lv_dbatch* b, const lv_dstate* s,
                    const lv_options* o, int reps, const std::vector<uint64_t>* voff) {
  struct Restore {  // (a failed tuning leaves the knobs as they were)
    lv_context* c;
    lv_stats st;
    lv_tune_choice knobs;
    bool done = false;
    ~Restore() {
      c->tuning = false;
      c->stats = st;
      if (!done) apply_knobs(c, knobs);
    }
  } restore{c, c->stats, knobs_of(c)};
  c->tuning = true;
  lv_dstate* t = nullptr;
  {
    const int rc = lv_dstate_create(c, s->S, &t);
    require(rc == LV_OK, rc, "tuning state: " + c->err);
  }
  struct Free {
    lv_dstate* t;
    ~Free() { lv_dstate_destroy(t); }
  } free_t{t};
  const auto cands = tune_candidates(c, approach, voff != nullptr);
  // quality of a trial: the modularity of the whole graph, and for the
  // batched form that of each graph
  auto measure = [&](const lv_tune_choice& k, double& q, std::vector<double>& qg) {
    apply_knobs(c, k);
    std::vector<double> ms;
    for (int r = 0; r < reps; ++r) {
      clone_state(c, t, s);
      HIPCHK(hipStreamSynchronize(c->st));
      const auto t0 = std::chrono::steady_clock::now();
      if (voff) run_batched(c, approach, x, b, t, voff->data(), voff->size() - 1, o, nullptr);
      else run(c, approach, x, b, t, o, nullptr);
      ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    q = dmodularity(c, x, t, o->resolution);
    if (voff) qg = dmodularity_groups(c, x, t, o->resolution, *voff);
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
  };
  // within 1e-3 of the defaults: the whole graph, and every graph of a union
  auto close = [](double qa, double qb, const std::vector<double>& ga, const std::vector<double>& gb) {
    if (std::fabs(qa - qb) > 1e-3 || ga.size() != gb.size()) return false;
    for (size_t k = 0; k < ga.size(); ++k)
      if (std::fabs(ga[k] - gb[k]) > 1e-3) return false;
    return true;
  };
  double qd = 0;
  std::vector<double> gd, gk;
  const double md = measure(cands[0], qd, gd);
  lv_tune_choice best = cands[0];
  double bm = md, bq = qd;
  for (size_t k = 1; k < cands.size(); ++k) {
    double qk = 0;
    const double mk = measure(cands[k], qk, gk);
    if (close(qk, qd, gk, gd) && mk < bm) {
      best = cands[k];
      bm = mk;
      bq = qk;
    }
  }
  if (bm > 0.97 * md) {
    best = cands[0];
    bm = md;
    bq = qd;
  }
  best.ms = bm;
  best.default_ms = md;
  best.modularity = bq;
  best.default_modularity = qd;
  best.candidates = (int32_t)cands.size();
  apply_knobs(c, best);
  c->tuned[shape_key(c, approach, x, voff ? voff->size() - 1 : 1)] = best;
  restore.done = true;
  return best;
}

// Before a run: the tuned choice for the graph's shape, or the defaults,
// unless tuning is off or a tuned knob was set by hand.
void apply_tuned(lv_context* c, int approach, const lv_dgraph* x, uint64_t G) {
  if (!c->autotune || c->tuning || c->knobs_hand) return;
  const auto it = c->tuned.find(shape_key(c, approach, x, G));
  apply_knobs(c, it != c->tuned.end() ? it->second : default_knobs());
}

// ------------------------------------------------------------ uploads
void graph_check(const lv_graph* g) {
  require(g != nullptr, LV_ERR_INVALID, "graph is NULL");
  require(g->offsets != nullptr, LV_ERR_INVALID, "graph offsets is NULL");
  require(g->span < 0xFFFFFFFFull, LV_ERR_LIMIT, "span must fit 32-bit vertex ids");
  uint64_t E = g->offsets[g->span];
  require(E == 0 || (g->keys && (g->weights || g->weights_f64)), LV_ERR_INVALID, "graph keys/weights NULL");
  // rows must lie inside the edge arrays: offsets ascending from 0, a gapped
  // row's degree within its capacity (the kernels index keys[off[u] + i])
  require(g->offsets[0] == 0, LV_ERR_INVALID, "graph offsets[0] must be 0");
  bool ascending = true, within = true;
  for (uint64_t u = 0; u < g->span; ++u) {
    ascending &= g->offsets[u] <= g->offsets[u + 1];
    within &= !g->degrees || g->degrees[u] <= g->offsets[u + 1] - g->offsets[u];
  }
  require(ascending, LV_ERR_INVALID, "graph offsets not ascending");
  require(within, LV_ERR_INVALID, "graph degree exceeds its row's capacity");
}

// A stamp no graph of this process has had (never 0).
uint64_t new_graph_stamp() {
  static std::atomic<uint64_t> last{0};
  return ++last;
}

lv_dgraph* upload_graph(lv_context* c, const lv_graph* g) {
  graph_check(g);
  HIPCHK(hipSetDevice(c->device));
  auto* d = new lv_dgraph();
  try {
    d->ctx = c;
    d->stamp = new_graph_stamp();
    d->S = g->span;
    d->E = g->offsets[g->span];
    d->f64 = g->weights_f64 != nullptr;
    d->off = dalloc<uint64_t>(d->S + 1);
    d->deg = dalloc<uint32_t>(d->S);
    d->keys = dalloc<uint32_t>(d->E);
    d->w = d->f64 ? (void*)dalloc<double>(d->E) : (void*)dalloc<float>(d->E);
    h2d(c, d->off, g->offsets, d->S + 1);
    if (g->degrees) h2d(c, d->deg, g->degrees, d->S);
    else k_degrees_from_offsets(d->off, d->deg, d->S, c->st);
    h2d(c, d->keys, g->keys, d->E);
    if (d->f64) h2d(c, (double*)d->w, g->weights_f64, d->E);
    else h2d(c, (float*)d->w, g->weights, d->E);
    d->order = d->S;
    d->live = d->E;
    if (g->degrees) {
      d->live = 0;
      for (uint64_t u = 0; u < d->S; ++u) d->live += g->degrees[u];
    }
    if (g->exists) {
      d->ex = dalloc<uint8_t>(d->S);
      h2d(c, d->ex, g->exists, d->S);
      uint64_t n = 0;
      for (uint64_t u = 0; u < d->S; ++u) n += g->exists[u] != 0;
      d->order = n;
    }
    if (d->E) {
      int* bad = c->ws<int>("rows.bad", 1);
      k_fill_u32(reinterpret_cast<uint32_t*>(bad), 0, 1, c->st);
      if (g->degrees) k_check_rows(d->off, d->deg, d->keys, d->S, bad, c->st);
      else k_check_edges(d->off, d->keys, d->S, d->E, c->ws<uint32_t>("rows.starts", d->E / 32 + 1), bad, c->st);
      int f = c->read1(bad);
      require(!(f & 1), LV_ERR_INVALID, "graph has a neighbour id >= span");
      // input graphs keep DiGraph's sorted rows; super-graph rows (f64
      // weights) are in the aggregation's first-seen order
      require(d->f64 || !(f & 2), LV_ERR_UNSORTED, "graph row not sorted by neighbour id, or a repeated neighbour");
    }
    // 2M (edgeWeightOmp), computed once per graph like the reference's M
    double* m2 = c->ws<double>("m2", 1);
    k_fill_f64(m2, 0, 1, c->st);
    k_edge_weight(d->view(), d->f64, m2, c->st);
    d->M2 = c->read1(m2);
    d->tiers.ids = dalloc<uint32_t>(std::max<uint64_t>(d->S, 1));
    build_tiers(c, "g.tiers", d->off, d->deg, d->ex, d->S, d->tiers, d->tiers.ids);
    d->tiers.maxdeg = max_degree_tier4(c, d->tiers, d->deg);
    d->maxdeg = max_degree_tier4(c, d->tiers, d->deg, 3);
    d->hubs.ids = dalloc<uint32_t>(std::max<uint64_t>(d->S, 1));
    build_tiers(c, "g.hubs", d->off, d->deg, d->ex, d->S, d->hubs, d->hubs.ids, kHubMin);
    d->hubs.maxdeg = max_degree_tier4(c, d->hubs, d->deg);
    d->hub_slots = caps_total(c, d->deg, d->S, kHubMin);
    d->slab_slots = caps_total(c, d->deg, d->S, kMidSlots / 2);
    c->check();
  } catch (...) {
    lv_dgraph_destroy(d);
    throw;
  }
  return d;
}

// Sortedness, uniqueness and the largest id of a device batch, read back
// with one synchronisation.
void check_batch(lv_context* c, lv_dbatch* b) {
  uint32_t* chk = c->ws<uint32_t>("batch.check", 2);  // flags, largest id
  k_fill_u32(chk, 0, 2, c->st);
  int* bad = reinterpret_cast<int*>(chk);
  k_check_sorted(b->du, b->dv, b->nd, 0xFFFFFFFFu, bad, c->st);
  k_check_sorted(b->iu, b->iv, b->ni, 0xFFFFFFFFu, bad, c->st);
  k_max_u32(b->du, b->nd, chk + 1, c->st);
  k_max_u32(b->dv, b->nd, chk + 1, c->st);
  k_max_u32(b->iu, b->ni, chk + 1, c->st);
  k_max_u32(b->iv, b->ni, chk + 1, c->st);
  uint32_t h[2];
  c->read(h, chk, 2);
  b->check_flags = (int)h[0];
  b->id_bound = b->nd + b->ni ? (uint64_t)h[1] + 1 : 0;
  b->checked = true;
}

// The batch's ids lie within span S, its lists sorted and unique (the
// errors lv_run reports).
void require_batch(lv_context* c, lv_dbatch* b, uint64_t S) {
  if (!b->checked) check_batch(c, b);
  require(!(b->check_flags & 1) && b->id_bound <= S, LV_ERR_INVALID, "batch vertex id out of range");
  require(!(b->check_flags & 2), LV_ERR_UNSORTED, "batch not sorted by (u, v) or not unique");
}

lv_dbatch* upload_batch(lv_context* c, const lv_batch* del, const lv_batch* ins, uint64_t S) {
  HIPCHK(hipSetDevice(c->device));
  auto* b = new lv_dbatch();
  try {
    b->nd = del ? del->size : 0;
    b->ni = ins ? ins->size : 0;
    require(!b->nd || (del->u && del->v && del->w), LV_ERR_INVALID, "deletions arrays NULL");
    require(!b->ni || (ins->u && ins->v && ins->w), LV_ERR_INVALID, "insertions arrays NULL");
    b->du = dalloc<uint32_t>(b->nd);
    b->dv = dalloc<uint32_t>(b->nd);
    b->dw = dalloc<float>(b->nd);
    b->iu = dalloc<uint32_t>(b->ni);
    b->iv = dalloc<uint32_t>(b->ni);
    b->iw = dalloc<float>(b->ni);
    if (b->nd) { h2d(c, b->du, del->u, b->nd); h2d(c, b->dv, del->v, b->nd); h2d(c, b->dw, del->w, b->nd); }
    if (b->ni) { h2d(c, b->iu, ins->u, b->ni); h2d(c, b->iv, ins->v, b->ni); h2d(c, b->iw, ins->w, b->ni); }
    check_batch(c, b);  // once: runs compare the largest id with their graph's span
    if (S) require_batch(c, b, S);
  } catch (...) {
    lv_dbatch_destroy(b);
    throw;
  }
  return b;
}

lv_dstate* new_state(uint64_t S) {
  auto* s = new lv_dstate();
  try {
    s->S = S;
    s->memb = dalloc<uint32_t>(S);
    s->vtot = dalloc<double>(S);
    s->ctot = dalloc<double>(S);
  } catch (...) {
    lv_dstate_destroy(s);
    throw;
  }
  return s;
}

void check_membership(const uint32_t* q, uint64_t S) {
  uint32_t mx = 0;
  for (uint64_t u = 0; u < S; ++u) mx = std::max(mx, q[u]);
  require(S == 0 || mx < S, LV_ERR_INVALID, "membership has a community id >= span");
}

// ------------------------------------------------------------ graph update
// One list of a raw batch tidied on the device (tidyBatchUpdateU,
// inc/batch.hxx:202-209): kept by the edge's existence in d, sorted by
// (u, v) (stable), unique; the m kept entries go to new arrays.
uint64_t tidy_list(lv_context* c, const lv_dgraph* d, const lv_batch* b, bool keep_existing, uint32_t*& ou,
                   uint32_t*& ov, float*& ow, const std::string& tag) {
  const uint64_t n = b ? b->size : 0;
  ou = ov = nullptr;
  ow = nullptr;
  if (!n) return 0;
  require(b->u && b->v && b->w, LV_ERR_INVALID, "batch arrays NULL");
  require(n < 0x7FFFFFFFull, LV_ERR_LIMIT, "batch longer than 2^31 - 1 entries");
  uint32_t* u = c->ws<uint32_t>(tag + ".u", n);
  uint32_t* v = c->ws<uint32_t>(tag + ".v", n);
  float* w = c->ws<float>(tag + ".w", n);
  h2d(c, u, b->u, n);
  h2d(c, v, b->v, n);
  h2d(c, w, b->w, n);
  uint64_t* key = c->ws<uint64_t>("up.key", n);
  uint64_t* key2 = c->ws<uint64_t>("up.key2", n);
  uint32_t* idx = c->ws<uint32_t>("up.idx", n);
  uint32_t* idx2 = c->ws<uint32_t>("up.idx2", n);
  uint8_t* flags = c->ws<uint8_t>("up.flags", n);
  uint32_t* sel = c->ws<uint32_t>("up.sel", n);
  uint32_t* cnt = c->ws<uint32_t>("up.cnt", 1);
  int* bad = c->ws<int>("up.bad", 1);
  k_fill_u32(reinterpret_cast<uint32_t*>(bad), 0, 1, c->st);
  k_tidy_keys(d->view(), u, v, n, keep_existing, key, idx, bad, c->st);
  size_t bytes = 0;
  HIPCHK(sort_pairs_u64(key, key2, idx, idx2, n, nullptr, bytes, c->st));
  void* tmp = c->scan_tmp(bytes);
  bytes = c->scan_tmp_bytes;
  HIPCHK(sort_pairs_u64(key, key2, idx, idx2, n, tmp, bytes, c->st));
  k_tidy_flags(key2, n, flags, c->st);
  bytes = 0;
  HIPCHK(select_flagged_u32(idx2, flags, sel, cnt, n, nullptr, bytes, c->st));
  tmp = c->scan_tmp(bytes);
  bytes = c->scan_tmp_bytes;
  HIPCHK(select_flagged_u32(idx2, flags, sel, cnt, n, tmp, bytes, c->st));
  require(c->read1(bad) == 0, LV_ERR_INVALID, "batch vertex id >= span");
  const uint64_t m = c->read1(cnt);
  ou = dalloc<uint32_t>(m);
  ov = dalloc<uint32_t>(m);
  ow = dalloc<float>(m);
  k_tidy_gather(sel, m, u, v, w, ou, ov, ow, c->st);
  return m;
}

// Scratch of build_tiers for n ids, taken ahead (growth can fail).
void reserve_tiers(lv_context* c, const std::string& name, uint64_t n) {
  const uint32_t nb = (uint32_t)std::max<uint64_t>(1, (n + 255) / 256);
  c->ws<uint32_t>(name + ".cnt", (size_t)kTiers * nb + 1);
  c->ws<uint32_t>(name + ".base", (size_t)kTiers * nb + 1);
}

// applyBatchUpdateU (inc/batch.hxx:225-236) of a tidy device batch on d;
// delw receives the weight every deletion had in the graph.  Every scratch
// buffer is taken before the graph changes, so a failure up to that point
// leaves the graph as it was; one after it (a HIP error) marks the graph
// unusable.  Either way the context's zero-between-calls scratch (per-row
// degree changes, touched flags) is cleared before the error goes on, and
// the graph's version and stamp move only when the update completes.
void apply_update(lv_context* c, lv_dgraph* d, const lv_dbatch* b, float* delw) {
  const uint64_t S = d->S, fb = (S + 7) & ~3ull;
  if (!b->nd && !b->ni) return;
  bool fresh = false;
  int* delta = c->ws<int>("up.delta", S, &fresh);  // zero between calls
  if (fresh) k_fill_u32(reinterpret_cast<uint32_t*>(delta), 0, c->ws_bytes("up.delta") / 4, c->st);
  uint8_t* touched = c->ws<uint8_t>("up.touched", fb, &fresh);  // zero between calls
  if (fresh) k_fill_u8(touched, 0, c->ws_bytes("up.touched"), c->st);
  bool mutating = false;
  try {
    k_upd_delta(b->du, b->nd, -1, delta, touched, c->st);
    k_upd_delta(b->iu, b->ni, +1, delta, touched, c->st);
    // touched rows, ascending
    uint32_t* T = c->ws<uint32_t>("up.T", S);
    uint32_t* lens = c->ws<uint32_t>("up.lens", kWlQueues);
    uint64_t* bsum = c->ws<uint64_t>("up.bsum", (size_t)kWlQueues * kWlScanBlocks);
    WlQueues qs{};
    qs.q[0] = T;
    k_wl_queue(touched, false, S, d->deg, WlSplit{}, qs, lens, bsum, c->st);  // no degree split
    const uint64_t R = c->read1(lens);
    uint64_t* so = c->ws<uint64_t>("up.so", R + 1);
    uint64_t* sn = c->ws<uint64_t>("up.sn", R + 1);
    uint64_t* bo = c->ws<uint64_t>("up.bo", R + 1);
    uint64_t* bn = c->ws<uint64_t>("up.bn", R + 1);
    int* bad = c->ws<int>("up.bad", 1);
    k_fill_u32(reinterpret_cast<uint32_t*>(bad), 0, 1, c->st);
    k_upd_sizes(T, R, d->off, d->deg, delta, so, sn, bad, c->st);
    scan64(c, so, bo, R + 1);
    scan64(c, sn, bn, R + 1);
    const uint64_t total_old = c->read1(bo + R), total_new = c->read1(bn + R);
    const bool overflow = (c->read1(bad) & 2) != 0;
    // survivors of the touched rows and their ranks within each row
    uint32_t* keep = c->ws<uint32_t>("up.keep", total_old + 1);
    uint32_t* rank = c->ws<uint32_t>("up.rank", total_old + 1);
    k_fill_u32(keep, 1, total_old, c->st);
    k_fill_u32(keep + total_old, 0, 1, c->st);
    k_upd_tomb(b->du, b->dv, b->nd, T, R, bo, d->off, d->deg, d->keys, static_cast<const float*>(d->w), keep, delw,
               c->st);
    scan32(c, keep, rank, total_old + 1);
    // merged rows in scratch, then into the graph
    uint32_t* yk = c->ws<uint32_t>("up.yk", std::max<uint64_t>(total_new, 1));
    float* yw = c->ws<float>("up.yw", std::max<uint64_t>(total_new, 1));
    k_upd_scatter(T, R, bo, bn, total_old, d->off, d->deg, d->keys, static_cast<const float*>(d->w), keep, rank,
                  b->iu, b->iv, b->iw, b->ni, yk, yw, c->st);
    // what the steps after the first change need, taken now
    unsigned long long* order = c->ws<unsigned long long>("up.order", 1);
    double* m2 = c->ws<double>("up.m2", 1);
    reserve_tiers(c, "g.tiers", S);
    reserve_tiers(c, "g.hubs", S);
    c->ws<unsigned long long>("hb.need", 1);
    c->ws<uint32_t>("maxdeg", 1);
    uint64_t* noff = nullptr;
    uint32_t* nkeys = nullptr;
    float* nw = nullptr;
    if (overflow) {
      // every row to a new CSR with slack: capacity = degree + max(4, degree / 8)
      uint64_t* cap = c->ws<uint64_t>("up.cap", S + 1);
      k_upd_caps(d->deg, delta, S, cap, c->st);
      try {
        noff = dalloc<uint64_t>(S + 1);
        scan64(c, cap, noff, S + 1);
        const uint64_t E2 = c->read1(noff + S);
        nkeys = dalloc<uint32_t>(E2);
        nw = dalloc<float>(E2);
        k_upd_copy_rows(S, d->off, d->deg, d->keys, static_cast<const float*>(d->w), touched, noff, nkeys, nw,
                        c->st);
        k_upd_copy_back(T, R, bn, total_new, noff, yk, yw, nkeys, nw, c->st);
        c->check();
        HIPCHK(hipStreamSynchronize(c->st));
        mutating = true;
        (void)hipFree(d->off);
        (void)hipFree(d->keys);
        (void)hipFree(d->w);
        d->off = noff;
        d->keys = nkeys;
        d->w = nw;
        d->E = E2;
      } catch (...) {
        if (!mutating) {
          (void)hipFree(noff);
          (void)hipFree(nkeys);
          (void)hipFree(nw);
        }
        throw;
      }
    } else {
      mutating = true;
      k_upd_copy_back(T, R, bn, total_new, d->off, yk, yw, d->keys, static_cast<float*>(d->w), c->st);
    }
    k_upd_degrees(T, R, delta, d->deg, touched, c->st);  // (clears delta and touched of the rows)
    if (d->ex) {
      k_upd_exists(b->iu, b->iv, b->ni, d->ex, c->st);
      k_fill_u64(reinterpret_cast<uint64_t*>(order), 0, 1, c->st);
      k_count_u8(d->ex, 1, S, order, c->st);
      d->order = c->read1(order);
    }
    // 2M after the batch: the old total minus the deleted edges' weights plus
    // the insertions'
    k_fill_f64(m2, d->M2, 1, c->st);
    k_upd_wsum(b->iw, b->ni, 1.0, m2, c->st);
    k_upd_wsum(delw, b->nd, -1.0, m2, c->st);
    d->M2 = c->read1(m2);
    d->live = d->live + b->ni - b->nd;
    // degree lists of the new degrees
    build_tiers(c, "g.tiers", d->off, d->deg, d->ex, S, d->tiers, d->tiers.ids);
    d->tiers.maxdeg = max_degree_tier4(c, d->tiers, d->deg);
    d->maxdeg = max_degree_tier4(c, d->tiers, d->deg, 3);
    build_tiers(c, "g.hubs", d->off, d->deg, d->ex, S, d->hubs, d->hubs.ids, kHubMin);
    d->hubs.maxdeg = max_degree_tier4(c, d->hubs, d->deg);
    d->hub_slots = caps_total(c, d->deg, S, kHubMin);
    d->slab_slots = caps_total(c, d->deg, S, kMidSlots / 2);
    c->check();
  } catch (...) {
    k_fill_u32(reinterpret_cast<uint32_t*>(delta), 0, c->ws_bytes("up.delta") / 4, c->st);
    k_fill_u8(touched, 0, c->ws_bytes("up.touched"), c->st);
    (void)hipStreamSynchronize(c->st);
    if (mutating) d->broken = true;
    throw;
  }
  ++d->version;
  d->stamp = new_graph_stamp();  // worklist hub-table sizes depend on the degrees
}

template <class F>
int guard(lv_context* c, F f) {
  try {
    f();
    return LV_OK;
  } catch (const LvError& e) {
    if (c) c->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (c) c->err = "host allocation failed";
    return LV_ERR_NOMEM;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return LV_ERR_INVALID;
  }
}

struct Owned {
  lv_dgraph* g = nullptr;
  lv_dbatch* b = nullptr;
  lv_dstate* s = nullptr;
  ~Owned() {
    lv_dgraph_destroy(g);
    lv_dbatch_destroy(b);
    lv_dstate_destroy(s);
  }
};

int host_entry(lv_context* c, int approach, const lv_graph* y, const lv_batch* del, const lv_batch* ins,
               const uint32_t* q, const double* qvtot, const double* qctot, const lv_options* o, lv_result* res) {
  return guard(c, [&] {
    require(c != nullptr, LV_ERR_INVALID, "context is NULL");
    require(res && res->membership && res->vertex_weight && res->community_weight, LV_ERR_INVALID,
            "result arrays must be set");
    Owned w;
    w.g = upload_graph(c, y);
    uint64_t S = w.g->S;
    if (approach != LV_STATIC) {
      require(q && qvtot && qctot, LV_ERR_INVALID, "previous state (q, qvtot, qctot) is NULL");
      check_membership(q, S);
      w.b = upload_batch(c, del, ins, S);
    }
    w.s = new_state(S);
    if (approach != LV_STATIC) {
      h2d(c, w.s->memb, q, S);
      h2d(c, w.s->vtot, qvtot, S);
      h2d(c, w.s->ctot, qctot, S);
    }
    run(c, approach, w.g, w.b, w.s, o, res);
  });
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

void lv_options_default(lv_options* o) {
  if (!o) return;
  o->repeat = 1;
  o->resolution = 1;
  o->tolerance = 1e-2;
  o->aggregation_tolerance = 1.0;
  o->tolerance_drop = 10;
  o->max_iterations = 20;
  o->max_passes = 10;
}

const char* lv_status_string(int s) {
  switch (s) {
    case LV_OK: return "ok";
    case LV_ERR_INVALID: return "invalid argument";
    case LV_ERR_NOMEM: return "out of device memory";
    case LV_ERR_HIP: return "HIP runtime error";
    case LV_ERR_UNSORTED: return "batch not sorted/unique";
    case LV_ERR_NODEVICE: return "no HIP device";
    case LV_ERR_LIMIT: return "input exceeds a limit";
    default: return "unknown status";
  }
}

int lv_context_create(int device, lv_context** out) {
  if (!out) return LV_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return LV_ERR_NODEVICE;
  if (device < 0 || device >= n) return LV_ERR_INVALID;
  auto* c = new lv_context();
  c->device = device;
  c->trace = std::getenv("LV_TRACE") != nullptr;
  int rc = guard(c, [&] {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPCHK(scratch_address(device, &c->scratch));
  });
  if (rc != LV_OK) {
    lv_context_destroy(c);
    return rc;
  }
  *out = c;
  return LV_OK;
}

void lv_context_destroy(lv_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto& kv : c->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  c->drop_graph();
  for (auto e : c->events) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->tail) (void)hipHostFree(c->tail);
  for (int k = 0; k < 2; ++k) {
    if (c->side[k]) {
      (void)hipStreamSynchronize(c->side[k]);
      (void)hipStreamDestroy(c->side[k]);
    }
    if (c->join_ev[k]) (void)hipEventDestroy(c->join_ev[k]);
  }
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char* lv_last_error(const lv_context* c) { return c ? c->err.c_str() : "no context"; }

int lv_context_set_path(lv_context* c, int path) {
  if (!c || path < 0 || path > 2) return LV_ERR_INVALID;
  c->path = path;
  return LV_OK;
}

int lv_context_set_slabs(lv_context* c, uint64_t slabs) {
  if (!c || slabs < 1) return LV_ERR_INVALID;
  c->slabs = slabs;
  return LV_OK;
}

int lv_context_set_hub_streams(lv_context* c, uint64_t streams) {
  if (!c || streams < 1 || streams > kHubBlocks) return LV_ERR_INVALID;
  c->hub_streams = streams;
  return LV_OK;
}

int lv_context_set_hub_threshold(lv_context* c, uint32_t sweep_degree, uint32_t worklist_degree) {
  if (!c || sweep_degree < kHubMin || worklist_degree < kHybridLds / 2 || worklist_degree > kHubMin)
    return LV_ERR_INVALID;
  set_by_hand(c, [&] {
    c->hub_min = sweep_degree;
    c->hub_min_wl = worklist_degree;
  });
  return LV_OK;
}

int lv_context_set_workspace_limit(lv_context* c, uint64_t bytes) {
  if (!c) return LV_ERR_INVALID;
  c->ws_limit = bytes ? (size_t)bytes : ~size_t(0);
  return LV_OK;
}

uint64_t lv_context_workspace_bytes(const lv_context* c) { return c ? (uint64_t)c->ws_total : 0; }

int lv_context_set_pass_schedule(lv_context* c, int lead_iterations, uint64_t lead_streams, uint64_t lead_chunk,
                                 uint64_t late_streams, uint64_t late_chunk) {
  if (!c || lead_iterations < 0 || (lead_iterations && (!lead_streams || !lead_chunk))) return LV_ERR_INVALID;
  c->lead_iters = lead_iterations;
  c->lead_streams = lead_streams;
  c->lead_chunk = lead_chunk;
  c->late_streams = late_streams;
  c->late_chunk = late_chunk;
  return LV_OK;
}

int lv_context_set_lead_ids(lv_context* c, uint64_t ids) {
  // (0: automatic, ~0: every id)
  if (!c) return LV_ERR_INVALID;
  c->lead_ids = ids;
  return LV_OK;
}

int lv_context_set_sweep_hubs(lv_context* c, int in_sweep) {
  if (!c) return LV_ERR_INVALID;
  c->sweep_hubs = in_sweep ? 1 : 0;
  return LV_OK;
}

int lv_context_set_validate(lv_context* c, int mode) {
  if (!c || mode < 0 || mode > 2) return LV_ERR_INVALID;
  c->validate = mode;
  return LV_OK;
}

int lv_context_set_pick_less(lv_context* c, int iterations) {
  if (!c || iterations < 0) return LV_ERR_INVALID;
  c->pick_less_iters = iterations;
  return LV_OK;
}

int lv_dstate_check_hub_tables(lv_context* c, const lv_dgraph* g, const lv_dstate* s, uint64_t* tables,
                               uint64_t* mismatches) {
  if (!c || !g || !s || !tables || !mismatches || g->S != s->S) return LV_ERR_INVALID;
  return guard(c, [&] {
    *tables = *mismatches = 0;
    if (!s->ht_valid) return;
    const uint64_t S = g->S;
    uint8_t* flags = c->ws<uint8_t>("htc.flags", (S + 7) & ~3ull);
    uint32_t* list = c->ws<uint32_t>("htc.list", std::max<uint64_t>(S, 1));
    uint32_t* lens = c->ws<uint32_t>("htc.len", kWlQueues);
    uint64_t* bsum = c->ws<uint64_t>("htc.bsum", (size_t)kWlQueues * kWlScanBlocks);
    unsigned long long* out = c->ws<unsigned long long>("htc.out", 2);
    k_fill_u64(reinterpret_cast<uint64_t*>(out), 0, 2, c->st);
    k_fill_u8(flags, 0, (S + 7) & ~3ull, c->st);
    k_flag_nonzero(s->ht_cap, S, flags, c->st);
    WlQueues qs{};
    qs.q[0] = list;
    k_wl_queue(flags, true, S, g->deg, WlSplit{}, qs, lens, bsum, c->st);
    k_ht_check(g->view(), g->f64, s->memb, list, lens, s->ht_off, s->ht_cap, s->ht_key, s->ht_cnt, s->ht_w, out,
               c->st);
    unsigned long long h[2];
    c->read(h, out, 2);
    *tables = h[0];
    *mismatches = h[1];
  });
}

int lv_context_set_hub_tables(lv_context* c, int on) {
  if (!c) return LV_ERR_INVALID;
  c->use_tables = on ? 1 : 0;
  return LV_OK;
}

int lv_context_set_graphs(lv_context* c, int on) {
  if (!c) return LV_ERR_INVALID;
  c->use_graphs = on ? 1 : 0;
  if (!on) c->drop_graph();
  return LV_OK;
}

int lv_context_set_profiling(lv_context* c, int on) {
  if (!c) return LV_ERR_INVALID;
  c->profiling = on ? 1 : 0;
  return LV_OK;
}

int lv_stats_get(const lv_context* cc, lv_stats* out) {
  if (!cc || !out) return LV_ERR_INVALID;
  auto* c = const_cast<lv_context*>(cc);  // reads the device counters on the context's stream
  return guard(c, [&] {
    HIPCHK(hipSetDevice(c->device));
    *out = c->stats;
    std::vector<unsigned long long> h((size_t)kSpread * kCntStride);
    c->read(h.data(), c->counters(), h.size());
    for (int sl = 0; sl < kSpread; ++sl)
      for (int t = 0; t < kTiers; ++t) {
        const unsigned long long* v = &h[(size_t)sl * kCntStride + t * CNT_N];
        out->vertices[t] += v[CNT_VERTICES];
        out->edges[t] += v[CNT_EDGES];
        out->candidates[t] += v[CNT_CANDIDATES];
        out->moves[t] += v[CNT_MOVES];
        out->moved_edges[t] += v[CNT_MOVED_EDGES];
      }
  });
}

int lv_stats_reset(lv_context* c) {
  if (!c) return LV_ERR_INVALID;
  return guard(c, [&] {
    HIPCHK(hipSetDevice(c->device));
    c->stats = lv_stats{};
    k_fill_u64(reinterpret_cast<uint64_t*>(c->counters()), 0, (size_t)kSpread * kCntStride, c->st);
  });
}

int lv_context_set_schedule(lv_context* c, double stream_fraction, uint64_t min_streams, uint64_t max_streams,
                            uint64_t chunk, int hubs_first) {
  if (!c || !(stream_fraction > 0) || min_streams < 1 || max_streams < min_streams) return LV_ERR_INVALID;
  if (hubs_first < 0 || hubs_first > 2) return LV_ERR_INVALID;
  c->chunk = chunk;
  set_by_hand(c, [&] { c->stream_fraction = stream_fraction; });
  c->min_streams = min_streams;
  c->max_streams = max_streams;
  c->tier_order = hubs_first;
  return LV_OK;
}

int lv_context_set_worklist_schedule(lv_context* c, uint64_t streams, int interleave) {
  if (!c || streams < 1 || streams > (1u << 16) || interleave < 0 || interleave > 1) return LV_ERR_INVALID;
  set_by_hand(c, [&] {
    c->wl_streams = streams;
    c->interleave = interleave;
  });
  return LV_OK;
}

int lv_context_set_autotune(lv_context* c, int on) {
  if (!c || on < 0 || on > 1) return LV_ERR_INVALID;
  c->autotune = on;
  if (on && c->knobs_hand) {
    apply_knobs(c, default_knobs());
    c->knobs_hand = false;
  }
  return LV_OK;
}

int lv_context_apply_tuning(lv_context* c, const lv_tune_choice* k) {
  if (!c) return LV_ERR_INVALID;
  if (k && (k->worklist_hub_degree < kHybridLds / 2 || k->worklist_hub_degree > kHubMin ||
            k->sweep_hub_degree < kHubMin || k->worklist_streams < 1 || k->worklist_streams > (1u << 16) ||
            k->interleave < 0 || k->interleave > 1 || !(k->stream_fraction > 0)))
    return LV_ERR_INVALID;
  set_by_hand(c, [&] { apply_knobs(c, k ? *k : default_knobs()); });
  return LV_OK;
}

int lv_context_add_tuning(lv_context* c, int approach, const lv_dgraph* g, const lv_tune_choice* k) {
  if (!c || !g || !k || approach < LV_STATIC || approach > LV_DYNAMIC_FRONTIER) return LV_ERR_INVALID;
  c->tuned[shape_key(c, approach, g, 1)] = *k;
  return LV_OK;
}

int lv_context_tuning(lv_context* c, int approach, const lv_dgraph* g, uint64_t graphs, lv_tune_choice* out) {
  if (!c || !g || !out || graphs < 1) return LV_ERR_INVALID;
  const auto it = c->tuned.find(shape_key(c, approach, g, graphs));
  if (it != c->tuned.end()) {
    *out = it->second;
  } else {
    *out = knobs_of(c);
    if (c->autotune && !c->knobs_hand) {
      const lv_tune_choice d = default_knobs();
      *out = d;
    }
    out->candidates = 0;
  }
  return LV_OK;
}

int lv_dstate_clone(lv_context* c, lv_dstate* dst, const lv_dstate* src) {
  if (!c || !dst || !src || dst->S != src->S) return LV_ERR_INVALID;
  return guard(c, [&] {
    clone_state(c, dst, src);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_autotune(lv_context* c, int approach, const lv_dgraph* g, const lv_dbatch* probe, const lv_dstate* s,
                const lv_options* o, int reps, lv_tune_choice* out) {
  if (!c || !g || !s) return LV_ERR_INVALID;
  return guard(c, [&] {
    require(approach >= LV_NAIVE_DYNAMIC && approach <= LV_DYNAMIC_FRONTIER, LV_ERR_INVALID,
            "lv_autotune tunes the dynamic approaches (Static's modularity moves with its schedule)");
    require(probe != nullptr, LV_ERR_INVALID, "dynamic approaches need a batch");
    require(s->S == g->S, LV_ERR_INVALID, "state span differs from graph span");
    require(reps >= 1 && reps <= 64, LV_ERR_INVALID, "reps must be 1..64");
    lv_options od;
    lv_options_default(&od);
    const lv_options* oo = o ? o : &od;
    validate_options(oo);
    require(oo->repeat == 1, LV_ERR_INVALID, "tuning runs one repeat");
    require_batch(c, const_cast<lv_dbatch*>(probe), g->S);
    c->knobs_hand = false;
    const lv_tune_choice k = tune(c, approach, g, probe, s, oo, reps, nullptr);
    if (out) *out = k;
  });
}

int lv_louvain_static(lv_context* c, const lv_graph* x, const lv_options* o, lv_result* res) {
  return host_entry(c, LV_STATIC, x, nullptr, nullptr, nullptr, nullptr, nullptr, o, res);
}
int lv_louvain_naive_dynamic(lv_context* c, const lv_graph* y, const lv_batch* d, const lv_batch* i,
                             const uint32_t* q, const double* qv, const double* qc, const lv_options* o,
                             lv_result* res) {
  return host_entry(c, LV_NAIVE_DYNAMIC, y, d, i, q, qv, qc, o, res);
}
int lv_louvain_delta_screening(lv_context* c, const lv_graph* y, const lv_batch* d, const lv_batch* i,
                               const uint32_t* q, const double* qv, const double* qc, const lv_options* o,
                               lv_result* res) {
  return host_entry(c, LV_DELTA_SCREENING, y, d, i, q, qv, qc, o, res);
}
int lv_louvain_dynamic_frontier(lv_context* c, const lv_graph* y, const lv_batch* d, const lv_batch* i,
                                const uint32_t* q, const double* qv, const double* qc, const lv_options* o,
                                lv_result* res) {
  return host_entry(c, LV_DYNAMIC_FRONTIER, y, d, i, q, qv, qc, o, res);
}

int lv_dgraph_create(lv_context* c, const lv_graph* g, lv_dgraph** out) {
  if (!c || !out) return LV_ERR_INVALID;
  return guard(c, [&] { *out = upload_graph(c, g); });
}
void lv_dgraph_destroy(lv_dgraph* g) {
  if (!g) return;
  (void)hipFree(g->off);
  (void)hipFree(g->deg);
  (void)hipFree(g->keys);
  (void)hipFree(g->w);
  (void)hipFree(g->ex);
  (void)hipFree(g->tiers.ids);
  (void)hipFree(g->hubs.ids);
  delete g;
}
uint64_t lv_dgraph_device_bytes(const lv_dgraph* g) {
  if (!g) return 0;
  const uint64_t w = g->f64 ? 8 : 4;
  return (g->S + 1) * 8 + g->S * 4 + g->E * (4 + w) + (g->ex ? g->S : 0) + 2 * std::max<uint64_t>(g->S, 1) * 4;
}
uint64_t lv_dstate_device_bytes(const lv_dstate* s) {
  if (!s) return 0;
  uint64_t b = s->S * (4 + 8 + 8);
  if (s->ht_cap) b += s->S * 4 + (s->S + 1) * 8;
  b += s->ht_slots * (4 + 4 + 8);
  return b;
}
uint64_t lv_dgraph_span(const lv_dgraph* g) { return g ? g->S : 0; }
uint64_t lv_dgraph_edges(const lv_dgraph* g) { return g ? g->live : 0; }
uint64_t lv_dgraph_capacity(const lv_dgraph* g) { return g ? g->E : 0; }

int lv_dgraph_update(lv_context* c, lv_dgraph* g, const lv_batch* deletions, const lv_batch* insertions,
                     lv_dbatch** tidy) {
  if (!c || !g || !tidy) return LV_ERR_INVALID;
  *tidy = nullptr;
  return guard(c, [&] {
    require(!g->f64, LV_ERR_INVALID, "only input graphs (f32 weights) can be updated");
    HIPCHK(hipSetDevice(c->device));
    require(!g->broken, LV_ERR_INVALID, "graph left inconsistent by a failed update");
    auto* b = new lv_dbatch();
    try {
      b->nd = tidy_list(c, g, deletions, true, b->du, b->dv, b->dw, "up.del");
      b->ni = tidy_list(c, g, insertions, false, b->iu, b->iv, b->iw, "up.ins");
      b->dws = dalloc<float>(b->nd);
      b->from_stamp = g->stamp;
      apply_update(c, g, b, b->dws);
      b->to_stamp = g->stamp;
      // tidy by construction: sorted, unique, ids of g
      b->checked = true;
      b->id_bound = g->S;
      b->check_flags = 0;
    } catch (...) {
      lv_dbatch_destroy(b);
      throw;
    }
    *tidy = b;
  });
}

int lv_dgraph_download(lv_context* c, const lv_dgraph* g, uint64_t* offsets, uint32_t* degrees, uint32_t* keys,
                       float* weights, uint8_t* exists) {
  if (!c || !g || !offsets || !degrees || !keys || !weights) return LV_ERR_INVALID;
  return guard(c, [&] {
    require(!g->f64, LV_ERR_INVALID, "graph has f64 weights");
    require(!g->broken, LV_ERR_INVALID, "graph left inconsistent by a failed update");
    d2h(c, offsets, g->off, g->S + 1);
    d2h(c, degrees, g->deg, g->S);
    d2h(c, keys, g->keys, g->E);
    d2h(c, weights, static_cast<const float*>(g->w), g->E);
    if (exists) {
      if (g->ex) d2h(c, exists, g->ex, g->S);
      else std::memset(exists, 1, g->S);
    }
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_dbatch_sizes(const lv_dbatch* b, uint64_t* deletions, uint64_t* insertions) {
  if (!b || !deletions || !insertions) return LV_ERR_INVALID;
  *deletions = b->nd;
  *insertions = b->ni;
  return LV_OK;
}

int lv_dbatch_download(lv_context* c, const lv_dbatch* b, uint32_t* du, uint32_t* dv, float* dw, uint32_t* iu,
                       uint32_t* iv, float* iw) {
  if (!c || !b || (b->nd && (!du || !dv || !dw)) || (b->ni && (!iu || !iv || !iw))) return LV_ERR_INVALID;
  return guard(c, [&] {
    d2h(c, du, b->du, b->nd);
    d2h(c, dv, b->dv, b->nd);
    d2h(c, dw, b->dw, b->nd);
    d2h(c, iu, b->iu, b->ni);
    d2h(c, iv, b->iv, b->ni);
    d2h(c, iw, b->iw, b->ni);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_dbatch_create(lv_context* c, const lv_batch* d, const lv_batch* i, lv_dbatch** out) {
  if (!c || !out) return LV_ERR_INVALID;
  return guard(c, [&] { *out = upload_batch(c, d, i, 0); });
}
void lv_dbatch_destroy(lv_dbatch* b) {
  if (!b) return;
  (void)hipFree(b->du);
  (void)hipFree(b->dv);
  (void)hipFree(b->dw);
  (void)hipFree(b->iu);
  (void)hipFree(b->iv);
  (void)hipFree(b->iw);
  (void)hipFree(b->dws);
  delete b;
}

int lv_dstate_create(lv_context* c, uint64_t S, lv_dstate** out) {
  if (!c || !out) return LV_ERR_INVALID;
  return guard(c, [&] {
    HIPCHK(hipSetDevice(c->device));
    *out = new_state(S);
  });
}
void lv_dstate_destroy(lv_dstate* s) {
  if (!s) return;
  (void)hipFree(s->memb);
  (void)hipFree(s->vtot);
  (void)hipFree(s->ctot);
  (void)hipFree(s->ht_off);
  (void)hipFree(s->ht_cap);
  (void)hipFree(s->ht_key);
  (void)hipFree(s->ht_cnt);
  (void)hipFree(s->ht_w);
  (void)hipFree(s->ht_next);
  delete s;
}
int lv_dstate_upload(lv_context* c, lv_dstate* s, const uint32_t* q, const double* qv, const double* qc) {
  if (!c || !s || !q || !qv || !qc) return LV_ERR_INVALID;
  return guard(c, [&] {
    check_membership(q, s->S);
    s->ht_reset();
    h2d(c, s->memb, q, s->S);
    h2d(c, s->vtot, qv, s->S);
    h2d(c, s->ctot, qc, s->S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}
int lv_dstate_download(lv_context* c, const lv_dstate* s, uint32_t* q, double* qv, double* qc) {
  if (!c || !s) return LV_ERR_INVALID;
  return guard(c, [&] {
    if (q) d2h(c, q, s->memb, s->S);
    if (qv) d2h(c, qv, s->vtot, s->S);
    if (qc) d2h(c, qc, s->ctot, s->S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}
int lv_dstate_copy(lv_context* c, lv_dstate* dst, const lv_dstate* src) {
  if (!c || !dst || !src || dst->S != src->S) return LV_ERR_INVALID;
  return guard(c, [&] {
    dst->ht_reset();
    d2d(c, dst->memb, src->memb, src->S);
    d2d(c, dst->vtot, src->vtot, src->S);
    d2d(c, dst->ctot, src->ctot, src->S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_run(lv_context* c, int approach, const lv_dgraph* g, const lv_dbatch* b, lv_dstate* s, const lv_options* o,
           lv_result* res) {
  if (!c) return LV_ERR_INVALID;
  return guard(c, [&] {
    if (b && g) require_batch(c, const_cast<lv_dbatch*>(b), g->S);  // checked when the batch was made
    run(c, approach, g, b, s, o, res);
  });
}

int lv_run_batched(lv_context* c, int approach, const lv_dgraph* g, const lv_dbatch* b, lv_dstate* s,
                   const uint64_t* voff, uint64_t num_graphs, const lv_options* o, lv_result* res) {
  if (!c) return LV_ERR_INVALID;
  return guard(c, [&] {
    if (b && g) require_batch(c, const_cast<lv_dbatch*>(b), g->S);
    run_batched(c, approach, g, b, s, voff, num_graphs, o, res);
  });
}

int lv_dmodularity(lv_context* c, const lv_dgraph* g, const lv_dstate* s, double R, double* out) {
  if (!c || !g || !s || !out) return LV_ERR_INVALID;
  return guard(c, [&] {
    require(!g->broken, LV_ERR_INVALID, "graph left inconsistent by a failed update");
    double* cin = c->ws<double>("mod.cin", g->S);
    double* ct = c->ws<double>("mod.ctot", g->S);
    double* o = c->ws<double>("mod.out", 1);
    k_fill_f64(cin, 0, g->S, c->st);
    k_fill_f64(ct, 0, g->S, c->st);
    k_fill_f64(o, 0, 1, c->st);
    k_modularity_parts(g->view(), g->f64, s->memb, cin, ct, c->st);
    k_modularity_sum(cin, ct, g->S, g->M2 / 2, R, o, c->st);
    *out = c->read1(o);
  });
}

// Modularity as lv_dmodularity computes it (the same kernels), with no
// device allocation: the scratch is the library's static device array of
// this GPU.  Calls on one GPU share that array, so calls that overlap on one
// GPU race; each call clears what it uses first.
int lv_dmodularity_noalloc(lv_context* c, const lv_dgraph* g, const lv_dstate* s, double R, double* out) {
  if (!c || !g || !s || !out) return LV_ERR_INVALID;
  return guard(c, [&] {
    require(!g->broken, LV_ERR_INVALID, "graph left inconsistent by a failed update");
    require(s->S == g->S, LV_ERR_INVALID, "state span differs from graph span");
    require(g->S <= kNoAllocSpan, LV_ERR_LIMIT,
            "lv_dmodularity_noalloc: span above its limit of 2^21 vertices (use lv_dmodularity)");
    require(c->scratch != nullptr, LV_ERR_HIP, "no device scratch address");
    HIPCHK(hipSetDevice(c->device));
    double* cin = c->scratch;
    double* ct = cin + kNoAllocSpan;
    double* o = ct + kNoAllocSpan;
    k_fill_f64(cin, 0, g->S, c->st);
    k_fill_f64(ct, 0, g->S, c->st);
    k_fill_f64(o, 0, 1, c->st);
    k_modularity_parts(g->view(), g->f64, s->memb, cin, ct, c->st);
    k_modularity_sum(cin, ct, g->S, g->M2 / 2, R, o, c->st);
    HIPCHK(hipMemcpyAsync(out, o, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->check();
  });
}

// ------------------------------------------------------------ phase entry points
int lv_edge_weight(lv_context* c, const lv_graph* g, double* out) {
  if (!c || !out) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    *out = w.g->M2;
  });
}

int lv_modularity(lv_context* c, const lv_graph* g, const uint32_t* memb, double R, double* out) {
  if (!c || !memb || !out) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    check_membership(memb, w.g->S);
    w.s = new_state(w.g->S);
    h2d(c, w.s->memb, memb, w.g->S);
    int rc = lv_dmodularity(c, w.g, w.s, R, out);
    require(rc == LV_OK, rc, c->err);
  });
}

int lv_vertex_weights(lv_context* c, const lv_graph* g, double* vtot) {
  if (!c || !vtot) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    double* d = c->ws<double>("ph.vtot", w.g->S);
    k_fill_f64(d, 0, w.g->S, c->st);
    k_vertex_weights(w.g->view(), w.g->f64, d, c->st);
    d2h(c, vtot, d, w.g->S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_update_weights(lv_context* c, uint64_t S, const lv_batch* del, const lv_batch* ins, const uint32_t* vcom,
                      double* vtot, double* ctot) {
  if (!c || !vcom || !vtot || !ctot) return LV_ERR_INVALID;
  return guard(c, [&] {
    check_membership(vcom, S);
    Owned w;
    w.b = upload_batch(c, del, ins, S);
    w.s = new_state(S);
    h2d(c, w.s->memb, vcom, S);
    h2d(c, w.s->vtot, vtot, S);
    h2d(c, w.s->ctot, ctot, S);
    k_update_weights(w.b->du, w.b->dw, w.b->nd, w.b->iu, w.b->iw, w.b->ni, w.s->memb, w.s->vtot, w.s->ctot, c->st);
    d2h(c, vtot, w.s->vtot, S);
    d2h(c, ctot, w.s->ctot, S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_affected_frontier(lv_context* c, uint64_t S, const lv_batch* del, const lv_batch* ins, const uint32_t* vcom,
                         uint8_t* vaff) {
  if (!c || !vcom || !vaff) return LV_ERR_INVALID;
  return guard(c, [&] {
    check_membership(vcom, S);
    Owned w;
    w.b = upload_batch(c, del, ins, S);
    w.s = new_state(S);
    h2d(c, w.s->memb, vcom, S);
    uint8_t* a = c->ws<uint8_t>("ph.vaff", (S + 7) & ~3ull);
    k_fill_u8(a, 0, (S + 7) & ~3ull, c->st);
    k_frontier(w.b->du, w.b->dv, w.b->nd, w.b->iu, w.b->iv, w.b->ni, w.s->memb, a, c->st);
    d2h(c, vaff, a, S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_affected_delta_screening(lv_context* c, const lv_graph* g, const lv_batch* del, const lv_batch* ins,
                                const uint32_t* vcom, const double* vtot, const double* ctot, double M, double R,
                                uint8_t* vertices, uint8_t* neighbors, uint8_t* communities) {
  if (!c || !vcom || !vtot || !ctot || !vertices || !neighbors || !communities) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    uint64_t S = w.g->S;
    check_membership(vcom, S);
    w.b = upload_batch(c, del, ins, S);
    w.s = new_state(S);
    h2d(c, w.s->memb, vcom, S);
    h2d(c, w.s->vtot, vtot, S);
    h2d(c, w.s->ctot, ctot, S);
    uint8_t* v = c->ws<uint8_t>("ph.dsv", S);
    uint8_t* n = c->ws<uint8_t>("ph.dsn", S);
    uint8_t* k = c->ws<uint8_t>("ph.dsc", S);
    k_fill_u8(v, 0, S, c->st);
    k_fill_u8(n, 0, S, c->st);
    k_fill_u8(k, 0, S, c->st);
    k_ds_deletions(w.b->du, w.b->dv, w.b->nd, w.s->memb, v, n, k, c->st);
    if (w.b->ni) {
      uint32_t* tkey = c->ws<uint32_t>("ds.tkey", 2 * w.b->ni);
      double* tval = c->ws<double>("ds.tval", 2 * w.b->ni);
      uint32_t* tlist = c->ws<uint32_t>("ds.tlist", w.b->ni);
      k_fill_u32(tkey, kEmpty, 2 * w.b->ni, c->st);
      k_ds_insertions(w.b->iu, w.b->iv, w.b->iw, w.b->ni, w.s->memb, w.s->vtot, w.s->ctot, M, R, tkey, tval, tlist,
                      v, n, k, c->st);
    }
    k_ds_sweep(w.g->view(), w.s->memb, n, k, v, c->st);
    d2h(c, vertices, v, S);
    d2h(c, neighbors, n, S);
    d2h(c, communities, k, S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_choose_community(lv_context* c, const lv_graph* g, const uint32_t* vcom, const double* vtot,
                        const double* ctot, double M, double R, uint32_t* best_c, double* best_e) {
  if (!c || !vcom || !vtot || !ctot || !best_c || !best_e) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    uint64_t S = w.g->S;
    check_membership(vcom, S);
    w.s = new_state(S);
    h2d(c, w.s->memb, vcom, S);
    h2d(c, w.s->vtot, vtot, S);
    h2d(c, w.s->ctot, ctot, S);
    uint32_t* oc = c->ws<uint32_t>("ph.bc", S);
    double* oe = c->ws<double>("ph.be", S);
    k_fill_u32(oc, 0, S, c->st);
    k_fill_f64(oe, 0, S, c->st);
    Level lv{w.g->view(), w.g->f64, &w.g->tiers, w.g->order, w.g->tiers.maxdeg, &w.g->hubs, w.g->hub_slots};
    MoveArgs a = base_args(c, lv, w.s->memb, w.s->vtot, w.s->ctot, nullptr, nullptr, M, R);
    a.choose_only = 1;
    a.chunk = 1;
    a.max_streams = 1ull << 30;
    a.out_c = oc;
    a.out_e = oe;
    a.cnt = nullptr;
    move_sweep(c, a, lv);
    c->check();
    d2h(c, best_c, oc, S);
    d2h(c, best_e, oe, S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_renumber(lv_context* c, const lv_graph* g, uint32_t* vcom, uint64_t n, uint32_t* cext, uint64_t* C) {
  if (!c || !vcom || !cext || !C) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    uint64_t S = w.g->S;
    require(n >= S, LV_ERR_INVALID, "vcom shorter than the graph span");
    check_membership(vcom, n);
    uint32_t* d = c->ws<uint32_t>("ph.vcom", n);
    h2d(c, d, vcom, n);
    uint32_t* map = nullptr;
    *C = count_communities(c, w.g->view(), d, map);
    k_lookup(d, n, map, c->st);
    d2h(c, vcom, d, n);
    d2h(c, cext, map, S);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int lv_aggregate(lv_context* c, const lv_graph* g, const uint32_t* vcom, uint64_t C, uint64_t* zoff, uint32_t* zdeg,
                 uint32_t* zkeys, double* zw) {
  if (!c || !vcom || !zoff || !zdeg || !zkeys || !zw) return LV_ERR_INVALID;
  return guard(c, [&] {
    Owned w;
    w.g = upload_graph(c, g);
    uint64_t S = w.g->S, E = w.g->E;
    check_membership(vcom, S);
    uint32_t mx = 0;
    for (uint64_t u = 0; u < S; ++u)
      if (!g->exists || g->exists[u]) mx = std::max(mx, vcom[u]);
    require(S == 0 || mx < C, LV_ERR_INVALID, "membership not renumbered below num_communities");
    uint32_t* d = c->ws<uint32_t>("ph.vcom", S);
    h2d(c, d, vcom, S);
    SuperGraph z = super_graph(c, "ph.z", std::max<uint64_t>(C, 1), E);
    double* ztot = c->ws<double>("ph.ztot", std::max<uint64_t>(C, 1));
    Level lv{w.g->view(), w.g->f64, &w.g->tiers, w.g->order, w.g->tiers.maxdeg, &w.g->hubs, w.g->hub_slots};
    aggregate(c, lv, d, C, z, ztot);
    d2h(c, zoff, z.off, C + 1);
    d2h(c, zdeg, z.deg, C);
    d2h(c, zkeys, z.keys, E);
    d2h(c, zw, z.w, E);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

}  // extern "C"
