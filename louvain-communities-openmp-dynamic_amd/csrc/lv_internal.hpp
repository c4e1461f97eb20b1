// Internal declarations shared by the kernels (lv_kernels.hip) and the host
// pass driver / C ABI (lv_capi.hip).  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>

#include "louvain_hip.h"

namespace lv {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // empty hashtable key
constexpr uint32_t kNoPos = 0xFFFFFFFFu;  // "not seen" first-position marker
constexpr int kTiers = 5;                 // degree tiers of the group-per-vertex kernels
// Upper degree bound of tiers 0..3 (LDS hashtables); tier 4 spills to HBM.
constexpr uint32_t kTierMax[4] = {8, 32, 128, 1024};
// Id-order kernel (kernel "tier" 5): one wave per stream, LDS table of
// kHybridLds slots when 2 deg fits, else an HBM region of kHybridHbm slots;
// degrees above kHubMin go to the hub kernel (tier 4, at most kHubBlocks blocks).
constexpr int kHybridLds = 512;
constexpr uint64_t kHybridHbm = 8192;
constexpr uint32_t kHubMin = 4096;
constexpr uint64_t kAggHbmChunk = 1024;  // flat member edges per block chunk, tier-4 aggregation
constexpr uint64_t kHubBlocks = 256;
// hub batch: table slots per best-pass chunk (every hub table is a multiple)
constexpr uint64_t kHubTabChunk = 1024;
// hub batch: edges per insert / flag chunk
constexpr uint64_t kHubChunkEdges = 1024;
// worklist hub batches: table slots a hub starts with (a multiple of kHubTabChunk)
constexpr uint64_t kHubTabCap = 8192;
// Worklist hubs run one block per hub with the hub's table in LDS, in two
// classes: up to kSmallHubMax edges (blocks of kSmallHubBlock threads,
// nextpow2(2 deg) <= kSmallHubSlots slots = 32 KB, kSmallHubGrid blocks)
// and up to kMidHubMax edges (kMidBlock threads, at most kMidSlots slots =
// 32 KB, kMidGrid blocks); kScanBatch x block edges are loaded at a time.  A hub's
// neighbours fall in few communities (tens, against thousands of edges, on
// R-MAT after a static run), so the tables are small -- LDS is what limits
// how many hubs run at once.  Larger hubs, and those with more distinct
// neighbour communities than their table holds, go to the multi-block hub
// batch.
constexpr uint32_t kSmallHubMax = 1024;
constexpr uint32_t kSmallHubSlots = 2 * kSmallHubMax;  // never capped
constexpr int kSmallHubBlock = 256;
constexpr uint64_t kSmallHubGrid = 1024;
constexpr uint32_t kMidHubMax = 16384;
constexpr uint32_t kMidSlots = 2048;
constexpr int kMidBlock = 512;
constexpr int kScanBatch = 4;
constexpr uint64_t kMidGrid = 512;
// Hub neighbour-community tables (worklist DF / DS first pass on a resident
// state): for every vertex of more than kHtMinDeg edges when the tables are
// built, its neighbours' weight and count per community, in an open-addressed
// table of min(nextpow2(2 deg), kHtMaxCap) slots (key u32, count i32, weight
// f64).  Batch edges and every move keep them exact, so such a hub is
// evaluated from its table -- tens of communities -- instead of its edges.
// A table that would hold more than kHtMaxDistinct communities is dropped
// (its hub is evaluated from its edges).
constexpr uint32_t kHtMinDeg = 256;
constexpr uint32_t kHtMaxCap = 256;
// moved-hub pieces of the fused worklist launch: edges a piece, pieces a
// sub-round (more: the hub's wave flags its row itself), the descriptor of a
// reserved piece not yet written, a piece without work
constexpr uint32_t kMvChunk = 1024;
constexpr uint32_t kMvPieces = 16384;
constexpr uint32_t kMvPending = 0xFFFFFFFFu;
constexpr uint32_t kMvSkip = 0xFFFFFFFEu;
constexpr uint32_t kHtMaxDistinct = kHtMaxCap - 64;  // room for a slot per lane of a wave past it
// slots of a vertex without a table: 0 never built, kHtDropped dropped
// (too many communities, or no spare slots); a table has kHtMaxCap
constexpr uint32_t kHtDropped = 1;

// worklist queue compaction: blocks (contiguous id segments) per pass
constexpr int kWlScanBlocks = 1024;
// streams of the worklist id-order kernel: the waves it keeps resident
// (256 CUs x 4 SIMDs x 3 waves at its register count)
constexpr uint64_t kWlStreams = 3072;
// worklist queues: low degree, small hubs, mid hubs, hub batch, tabled hubs
constexpr int kWlQueues = 5;
struct WlSplit {
  uint32_t t[3];              // degree bounds of queues 0..2
  const uint8_t* big;         // ids sent to the hub batch (nullable)
  const uint32_t* tab;        // hub table slots per id: hubs with a table go to queue 4 (nullable)
  uint8_t* dirty;             // per segment: a flag may be set (nullable: scan all)
  uint32_t* reset;            // nreset words (at most 3) zeroed by the compaction (nullable)
  uint32_t nreset;
  const unsigned long long* ndrop;  // copied to lens[kWlQueues], saturated to u32 (nullable)
  const double* el;            // gain slots summed in slot order into lens[6..7] as a double (nullable)
  uint32_t shift;             // segment of block b: ids [b << shift, (b + 1) << shift)
};
// Segment shift of the compaction: 2^shift ids per block (a multiple of
// the 16 x 256 ids a block reads per tile), at most kWlScanBlocks segments.
inline uint32_t wl_shift(uint64_t S) {
  uint32_t sh = 12;
  while ((S + (1ull << sh) - 1) >> sh > (uint64_t)kWlScanBlocks) ++sh;
  return sh;
}
struct WlQueues {
  uint32_t* q[kWlQueues];
};
constexpr int kTierHybrid = 5;

inline int tier_of(uint64_t d) {
  for (int t = 0; t < 4; ++t)
    if (d <= kTierMax[t]) return t;
  return 4;
}

inline uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Device view of a (gapped) CSR graph.  Weight type is a template parameter
// of the kernels: float for the input graph, double for super-graphs.
struct GView {
  uint64_t S;              // span
  const uint64_t* off;     // S + 1
  const uint32_t* deg;     // S (always materialised on the device)
  const uint32_t* keys;
  const void* w;           // float* or double*
  const uint8_t* ex;       // S, or nullptr = every id exists
};

// Per-vertex degree-tier lists of a graph level (ascending ids per tier).
struct TierLists {
  uint32_t* ids = nullptr;      // concatenated lists, tier t at [start[t], start[t+1])
  uint64_t start[kTiers + 1] = {0};
  uint32_t maxdeg = 0;          // max degree of the tier-4 vertices
};

// Device counters of the local-moving kernels (for achieved-bandwidth figures).
enum { CNT_VERTICES = 0, CNT_EDGES = 1, CNT_CANDIDATES = 2, CNT_MOVES = 3, CNT_MOVED_EDGES = 4, CNT_N = 5 };

// Device atomics on one address serialise (about 12 ns each on MI355X, so
// one per block of an 8k-block launch costs 0.1 ms).  Launch-wide sums taken
// by one atomic per block -- the local-moving gain el and the counters -- are
// therefore spread over kSpread slots 128 B apart, picked by block index, and
// the host adds the slots up in slot order.
constexpr int kSpread = 64;
constexpr int kElStride = 16;   // doubles per gain slot (128 B)
constexpr int kCntStride = 32;  // counters per slot: kTiers * CNT_N, rounded up to whole 128 B lines
static_assert(kTiers * CNT_N <= kCntStride, "counter slot too small");

// Arguments of the local-moving kernels.
struct MoveArgs {
  GView g;
  const uint32_t* list;        // vertex ids to consider (nullptr = ids base .. base+n-1)
  uint64_t base;               // first id when list is nullptr
  uint64_t n;                  // entries in list
  const uint32_t* n_dev;       // if set, n is read from here on the device
  uint32_t* vcom;
  const double* vtot;
  double* ctot;
  uint8_t* vaff;
  const uint8_t* fa;           // pass-1 restriction (delta-screening), or nullptr
  double M, R;
  int choose_only;             // 1: write best (c, e) per vertex, move nothing
  uint32_t* out_c;
  double* out_e;
  double* el;                  // accumulated gain (l1-norm of delta-Q)
  unsigned long long* cnt;     // CNT_N counters
  // batched form (gid != nullptr): graph of every vertex of the level, and
  // per graph its M, the id of its community 0 at this level, its gain
  // accumulator and a skip flag (converged this pass, or finished)
  const uint32_t* gid;
  const double* gM;
  const uint32_t* gzero;
  double* gel;
  const uint8_t* gmask;
  // batched form, id-order kernel with spread chunks (chunk == 0): each
  // graph's share of the list cut into its own streams, as it would be alone.
  // Chunk t of nct covers list entries [ctab[2t], ctab[2t+1]) of graph k,
  // gcs[k] <= t < gcs[k+1]; graph k holds level ids [glo[k], glo[k+1]).
  // With nct_dev set the chunks that hold entries are packed first, in no
  // set order, and *nct_dev counts them (a word the compaction before the
  // table zeroes): the other chunks would be walked by no one.
  uint64_t* ctab;
  uint64_t nct;
  const uint64_t* glo;
  const uint64_t* gcs;
  uint32_t ng;
  uint32_t* nct_dev;
  uint32_t* fsame;             // worklist mode, bit v: flagged by a lower id and not
                               //   yet visited this iteration -> next sub-round
  uint8_t* wl_dirty;           // worklist mode: per compaction segment, fsame may hold a flag
  uint32_t wl_shift;           //   (segment of id v: v >> wl_shift)
  uint32_t* stamp;             // worklist mode: iteration serial of the last visit
  uint32_t serial;
  const uint32_t* serial_dev;  // if set, the serial is read from here on the device
  int pick_less;               // 1: only moves to a lower community id
  int validate;                // id-order kernel: a vertex about to move re-reads its neighbours'
                               //   communities and decides again if one changed meanwhile
  uint64_t min_chunk;          // unused (kept zero)
  uint64_t max_streams;        // lane groups launched (streams walking at once)
  uint64_t chunk;              // list entries per dynamically assigned chunk
  int sparse;                  // id-order kernel: scan windows of a group's lanes by ballot, visit
                               //   flagged only: 1 reading each later flag again after a visit
                               //   (sweeps), 2 reading the flags once per window (worklist
                               //   queues: an entry is flagged when queued and only its own
                               //   visit clears its flag)
  int interleave;              // id-order kernel, chunk 0, no ctab: group g takes entries g, g + N,
                               //   g + 2N, ... of the list (N groups), in that order
  unsigned long long* work;    // chunk counter, zero before each launch
  // hub kernel (tier 4) HBM hashtables: one region of hcap slots per block
  uint32_t* hkey;
  uint32_t* hpos;
  double* hval;
  uint64_t hcap;
  // id-order kernel (tier 5): one region of gcap slots per block (stream)
  uint32_t* gkey;
  uint32_t* gpos;
  double* gval;
  uint64_t gcap;
  uint64_t gblocks;            // regions allocated = grid cap
  uint32_t hub_min;            // degrees above this go to the hub kernel
  uint64_t hub_streams;        // hub kernel: hubs processed at once (<= kHubBlocks)
  const int* done;             // device convergence flag: skip work when set
  // hub batches (k_hub_batch): per-hub scratch, nh_max entries each
  uint64_t* hb_cap;            // table slots per hub, then exclusive scan (nh_max + 1)
  uint64_t* hb_chunks;         // edge chunks per hub, then exclusive scan (nh_max + 1)
  uint64_t* hb_tot;            // [0] total slots, [1] total chunks of the batch
  uint32_t* hb_cu;             // hub's community
  double* hb_vd;               // hub's weight to its own community
  double* hb_sd;               // Sigma_tot of the hub's community
  uint32_t* hb_bpos;           // first position of the best community (kNoPos: no move)
  double* hb_ce;               // best gain per kHubTabChunk-slot table chunk
  uint32_t* hb_cp;             // its first position (kNoPos: none)
  int f64_weights;             // hub batch: the graph's weights are f64
  uint32_t* hb_key;            // table slab
  uint32_t* hb_pos;
  double* hb_val;
  // worklist batches (fsame set): tables capped at kHubTabCap slots; a hub
  // whose capped table overflowed (hb_over, per list entry) is marked in
  // hb_big (per vertex) and redone with a full table
  uint8_t* hb_over;
  uint8_t* hb_big;
  // hub neighbour-community tables (nullable ht_cap: none)
  const uint64_t* ht_off;      // per vertex: first slot of its table
  uint32_t* ht_cap;            // per vertex: slots (0: no table)
  uint32_t* ht_key;
  int32_t* ht_cnt;
  double* ht_w;
  unsigned long long* ht_ndrop;  // tables dropped since the build (counted once each)
  // hubs with a table that moved, flagged by a grid-wide pass after their
  // kernel (nullable mv_n: each wave flags its own)
  uint32_t* mv_u;
  uint32_t* mv_from;
  uint32_t* mv_to;
  uint32_t* mv_d;              //   (each hub's degree and row offset)
  uint64_t* mv_o;
  // fused worklist launch (nullable: a flag pass after the kernel reads the
  // entries): each moved hub's row in pieces of kMvChunk edges, taken by the
  // launch's blocks when their own work is done; mv_n[1] pieces reserved,
  // mv_n[2] taken; piece k: entry mv_piece[k] & 4095, part mv_piece[k] >> 12
  uint32_t* mv_piece;
  uint32_t mv_pcap;
  uint32_t* mv_n;
  uint32_t mv_cap;
  uint64_t* hb_off;            // hub's row offset and degree
  uint32_t* hb_deg;
  uint32_t* hb_cown;           // list entry owning each edge chunk
  uint32_t* hb_town;           // list entry owning each kHubTabChunk-slot table chunk
  // id-order sweeps with the hubs in the sweep (xids set): a vertex of more
  // than kHubMin edges is visited in its turn like any other, with its own
  // HBM table: hub i of the ascending list xids (nx entries) owns slots
  // [xoff[i], xoff[i + 1]) of xkey / xpos / xval, nextpow2(2 deg) of them
  const uint32_t* xids;
  const uint64_t* xoff;
  uint64_t nx;
  uint32_t* xkey;
  uint32_t* xpos;
  double* xval;
};

// Hub batch: every listed vertex (hub) at once, each with its own HBM table
// and its edges split over many blocks; the reference's per-vertex steps in
// separate device passes (table build, best gain, move, neighbour flags).
void k_hub_batch(const MoveArgs& a, hipStream_t st);
// Worklist hubs with a neighbour-community table, from a.list (length
// *a.n_dev), one wave each.
void k_hub_tab(const MoveArgs& a, hipStream_t st);
// Hub tables: slots per vertex (cap[u] = min(nextpow2(2 deg), kHtMaxCap) for
// degree > kHtMinDeg, else 0), hub flags for the build list; the build from
// the edges and a membership (a table with too many communities is dropped,
// cap[u] = kHtDropped, and counted in *ndrop); the batch's edges into the
// tables (deletions out, insertions in, with the membership before any move).
void k_ht_caps(const uint32_t* deg, uint64_t S, uint32_t* cap, uint64_t* cap64, uint8_t* flags, hipStream_t st);
void k_ht_build(GView g, bool f64, const uint32_t* vcom, const uint32_t* list, const uint32_t* n_dev,
                const uint64_t* off, uint32_t* cap, uint32_t* key, int32_t* cnt, double* w,
                unsigned long long* ndrop, hipStream_t st);
void k_flag_nonzero(const uint32_t* a, uint64_t n, uint8_t* flags, hipStream_t st);
void k_hold(uint32_t us, hipStream_t st);
// Neighbours of the tabled hubs that moved (a.mv_*), flagged over the grid;
// with carry, the entries of worklist queues 1-3 back to the next sub-round.
void k_moved_flag(const MoveArgs& a, const WlQueues& qs, const uint32_t* lens, bool carry, hipStream_t st);
// A worklist sub-round's id-order kernel (a, queue 0) and tabled hubs (h,
// queue 4) in one launch, which also flags the neighbours of the tabled hubs
// that move (h.mv_piece set) and marks the entries of queues 1-3 (qs,
// lens) for the next sub-round; false (nothing launched) for schedules it
// does not cover (chunks handed out in order, validated moves, hubs in the
// sweep, per-graph chunk tables, fewer than 64 streams).
bool k_wl_fused(const MoveArgs& a, const MoveArgs& h, const WlQueues& qs, const uint32_t* lens, bool f64,
                hipStream_t st);
// Sources of a batch's insertions (iu) that now have more than kHtMinDeg
// edges and no table: a table each from the spare slots (bump counter
// next[0], slots up to total), listed for a build; with no spare slot left,
// dropped (counted in next[1]).
void k_ht_new_hubs(const uint32_t* iu, uint64_t ni, const uint32_t* deg, uint32_t* cap, uint64_t* off,
                   unsigned long long* next, uint64_t total, uint32_t* list, uint32_t* nlist, hipStream_t st);
void k_ht_check(GView g, bool f64, const uint32_t* vcom, const uint32_t* list, const uint32_t* n_dev,
                const uint64_t* off, uint32_t* cap, uint32_t* key, int32_t* cnt, double* w, unsigned long long* out,
                hipStream_t st);
void k_ht_batch(const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd, const uint32_t* iu,
                const uint32_t* iv, const float* iw, uint64_t ni, const uint32_t* vcom, const uint64_t* off,
                uint32_t* cap, uint32_t* key, int32_t* cnt, double* w, unsigned long long* ndrop, hipStream_t st);
// Worklist hubs from a.list (length *a.n_dev), one block each with an LDS
// table: class 0 up to kSmallHubMax edges, class 1 up to kMidHubMax.
void k_hub_mid(const MoveArgs& a, int cls, hipStream_t st);

// *out = the first id u with off[u] >= off[S] / share (a binary search over
// the offsets, S + 1 entries).
void k_offset_bound(const uint64_t* off, uint64_t S, uint64_t share, uint64_t* out, hipStream_t st);
// caps[i] = nextpow2(2 deg[ids[i]]) for i < n, caps[n] = 0.
void k_list_caps(const uint32_t* ids, uint64_t n, const uint32_t* deg, uint64_t* caps, hipStream_t st);

// Kernel launchers (lv_kernels.hip).  All run on `st`.
void k_fill_u8(uint8_t* a, uint8_t v, uint64_t n, hipStream_t st);
void k_fill_u32(uint32_t* a, uint32_t v, uint64_t n, hipStream_t st);
void k_fill_f64(double* a, double v, uint64_t n, hipStream_t st);
void k_fill_u64(uint64_t* a, uint64_t v, uint64_t n, hipStream_t st);
void k_iota_u32(uint32_t* a, uint64_t n, hipStream_t st);
void k_degrees_from_offsets(const uint64_t* off, uint32_t* deg, uint64_t S, hipStream_t st);
void k_vertex_weights(GView g, bool f64, double* vtot, hipStream_t st);
void k_edge_weight(GView g, bool f64, double* out, hipStream_t st);
void k_initialize(GView g, uint32_t* vcom, double* ctot, const double* vtot, hipStream_t st);
void k_update_weights(const uint32_t* du, const float* dw, uint64_t nd, const uint32_t* iu, const float* iw,
                      uint64_t ni, const uint32_t* vcom, double* vtot, double* ctot, hipStream_t st);
void k_frontier(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* iu, const uint32_t* iv,
                uint64_t ni, const uint32_t* vcom, uint8_t* vaff, hipStream_t st);
void k_ds_deletions(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* vcom, uint8_t* vertices,
                    uint8_t* neighbors, uint8_t* communities, hipStream_t st);
void k_ds_insertions(const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni, const uint32_t* vcom,
                     const double* vtot, const double* ctot, double M, double R, uint32_t* tkey, double* tval,
                     uint32_t* tlist, uint8_t* vertices, uint8_t* neighbors, uint8_t* communities, hipStream_t st,
                     const uint32_t* gid = nullptr, const double* gM = nullptr);
void k_ds_sweep(GView g, const uint32_t* vcom, const uint8_t* neighbors, const uint8_t* communities,
                uint8_t* vertices, hipStream_t st);
void k_count_u8(const uint8_t* a, uint8_t v, uint64_t n, unsigned long long* out, hipStream_t st);
void k_check_sorted(const uint32_t* u, const uint32_t* v, uint64_t n, uint32_t S, int* bad, hipStream_t st);
void k_check_rows(const uint64_t* off, const uint32_t* deg, const uint32_t* keys, uint64_t S, int* bad,
                  hipStream_t st);
// Ungapped CSR (offsets[S] edges): starts holds E / 32 + 1 words of scratch.
void k_check_edges(const uint64_t* off, const uint32_t* keys, uint64_t S, uint64_t E, uint32_t* starts, int* bad,
                   hipStream_t st);
void k_move(const MoveArgs& a, int tier, bool f64, hipStream_t st);
void k_exists(GView g, const uint32_t* vcom, uint32_t* flags, hipStream_t st);
void k_lookup(uint32_t* a, uint64_t n, const uint32_t* map, hipStream_t st);
// Community offsets from the level's ids sorted by community (keys in [0, C],
// C for an absent id): coff[c] = the number of keys below c, c = 0..C.
void k_member_bounds(const uint32_t* sorted_keys, uint64_t n, uint32_t C, uint32_t* coff, hipStream_t st);
void k_member_degrees(const uint32_t* cedg, uint64_t n, GView g, uint64_t* mdeg, hipStream_t st);
void k_gather_yoff(const uint32_t* coff, const uint64_t* mscan, uint64_t C, uint64_t* yoff, hipStream_t st);
// Degree partition of ids 0..n-1 (live only), ascending within each part:
// thr == 0 the five tiers, thr > 0 two parts (tier 0: degree <= thr, tier 4: > thr).
void k_tier_count(const uint64_t* off, const uint32_t* deg, const uint8_t* ex, uint64_t n, uint32_t thr,
                  uint32_t* block_counts, uint32_t nblocks, hipStream_t st);
void k_tier_scatter(const uint64_t* off, const uint32_t* deg, const uint8_t* ex, uint64_t n, uint32_t thr,
                    const uint32_t* block_base, uint32_t nblocks, uint32_t* ids, hipStream_t st);
struct AggArgs {
  GView g;                   // graph being aggregated
  const uint32_t* vcom;      // renumbered membership
  const uint32_t* coff;      // community -> members offsets (C + 1)
  const uint32_t* cedg;      // members
  const uint64_t* mscan;     // exclusive scan of member degrees (members + 1)
  const uint64_t* yoff;      // super-graph row offsets (C + 1)
  uint64_t C;
  const uint32_t* list;      // communities of this tier
  uint64_t n;
  uint32_t* zdeg;
  uint32_t* zkeys;
  double* zw;
  double* ztot;              // row sums (K of the super-vertices), accumulated
  // tier 4: shared HBM tables, one per community, chunked over blocks; a
  // slot's position is its community's first member edge relative to the row
  uint32_t* hkey;
  uint32_t* hpos;
  double* hval;
  const uint64_t* hoff;      // table offset per tier-4 community (n + 1)
  const uint64_t* choff;     // chunk offsets per tier-4 community (n + 1)
  uint64_t nchunks;
};
void k_aggregate(const AggArgs& a, int tier, bool f64, hipStream_t st);
// Tier-4 rows in first-seen order: mark each slot's first position in the
// bitmap bm (32 words per chunk, zeroed), popcounts of the words into cnt
// (words + 1 entries, the last 0); after an exclusive scan of cnt into bsc,
// emit the rows and their degrees.
void k_agg_mark_hbm(const AggArgs& a, uint64_t total_slots, uint32_t* bm, uint64_t words, uint32_t* cnt,
                    hipStream_t st);
void k_agg_emit_hbm(const AggArgs& a, uint64_t total_slots, const uint32_t* bm, const uint32_t* bsc, hipStream_t st);
// Members in ascending id order per community: sort keys (community, or C
// for a vertex that does not exist) and ids, then a stable radix sort.
void k_member_keys(GView g, const uint32_t* vcom, uint32_t C, uint32_t* key, uint32_t* id, hipStream_t st);
hipError_t sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                          int end_bit, void* tmp, size_t& tmp_bytes, hipStream_t st);
void k_t4_sizes(const uint64_t* yoff, const uint32_t* list, uint64_t n, uint64_t C, uint64_t chunk,
                uint64_t* caps, uint64_t* chunks, hipStream_t st);
void k_bin_queue(const uint32_t* q, const uint32_t* qlen, uint64_t qcap, const uint32_t* deg, uint32_t thr,
                 uint32_t* tq, uint64_t stride, uint32_t* tlen, hipStream_t st);
void k_max_deg(const uint32_t* ids, uint64_t n, const uint32_t* deg, uint32_t* out, hipStream_t st);
void k_max_u32(const uint32_t* a, uint64_t n, uint32_t* out, hipStream_t st);
void k_gather_u32(const uint32_t* a, const uint32_t* idx, uint64_t n, uint32_t* out, hipStream_t st);
void k_caps_sum(const uint32_t* deg, uint64_t n, uint32_t thr, unsigned long long* out, hipStream_t st);
// Ordered worklist queues: ids with f[id] set, ascending, split by degree
// into kWlQueues queues (queue q takes degrees in (t[q-1], t[q]], the last
// one the rest; t[0] == 0 puts every id in queue 0, and the ids of queues 1
// and 2 marked in big go to queue 3), counts to lens[0..3]; clear resets the
// flags read.  f holds S bytes rounded up to a multiple of 4; bsum holds
// kWlQueues * kWlScanBlocks entries.
void k_wl_queue(uint8_t* f, bool clear, uint64_t S, const uint32_t* deg, const WlSplit& sp, const WlQueues& qs,
                uint32_t* lens, uint64_t* bsum, hipStream_t st);
// k_wl_queue over a bitmap (bit v of u32 word v >> 5; (S >> 5) + 2 words).
void k_wl_queue_bits(uint32_t* bits, bool clear, uint64_t S, const uint32_t* deg, const WlSplit& sp,
                     const WlQueues& qs, uint32_t* lens, uint64_t* bsum, hipStream_t st);
void k_modularity_parts(GView g, bool f64, const uint32_t* memb, double* cin, double* ctot, hipStream_t st);
void k_modularity_sum(const double* cin, const double* ctot, uint64_t n, double M, double R, double* out,
                      hipStream_t st);
// Exclusive scans (hipCUB) with grow-only temp storage owned by the caller.
// Batched form helpers (lv_update.hip): graph of every vertex from the
// vertex ranges voff (G + 1 entries), per-graph sums, flag counts per range, graph of
// the next level's super-vertices, the level vertices of the graphs still
// running, lookups that leave finished graphs alone, and the final ids of
// the graphs flagged in fin: re-based from their range lo of the current
// level to their vertex range, Sigma_tot into ftot.
void k_gid_fill(const uint64_t* voff, uint64_t G, uint64_t S, uint32_t* gid, hipStream_t st);
void k_group_sum(const double* x, const uint32_t* gid, uint64_t S, double* out, hipStream_t st);
void k_group_modularity(double* cin, const double* ctot, const uint32_t* gid, const double* gM, uint64_t S, double R,
                        hipStream_t st);
void k_gid_next(const uint32_t* vcom, const uint32_t* gid, const uint32_t* map, const uint8_t* ex, uint64_t S,
                uint32_t* gid_next, hipStream_t st);
void k_active_mask(const uint32_t* gid, uint64_t SL, const uint8_t* ex, const uint8_t* done, uint8_t* mask,
                   hipStream_t st);
void k_lookup_active(uint32_t* a, uint64_t n, const uint32_t* map, const uint32_t* gid, const uint8_t* skip,
                     hipStream_t st);
void k_finalize(uint32_t* ucom, const uint32_t* vcom, const uint32_t* gid0, const uint8_t* ex, uint64_t S,
                const double* ctot, uint64_t SL, const uint64_t* lo, uint64_t G, const uint8_t* fin,
                const uint64_t* voff, double* ftot, hipStream_t st);
void k_range_count(const uint8_t* f, const uint64_t* voff, uint64_t G, unsigned long long* out, hipStream_t st);
// The chunk table a.ctab (a.nct chunks) of the list the id-order kernel is
// about to walk (a.list with a.n_dev / a.n entries, or ids a.base.. when
// a.list is null): graph k's entries in a.gcs[k+1] - a.gcs[k] chunks of
// ceil(len / streams) entries (packed, with a.nct_dev, to the chunks that
// hold entries).
void k_graph_chunks(const MoveArgs& a, hipStream_t st);

// Device batch tidy and graph update (lv_update.hip).
void k_tidy_keys(GView g, const uint32_t* u, const uint32_t* v, uint64_t n, bool keep_existing, uint64_t* key,
                 uint32_t* idx, int* bad, hipStream_t st);
hipError_t sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                          void* tmp, size_t& bytes, hipStream_t st);
void k_tidy_flags(const uint64_t* key, uint64_t n, uint8_t* flag, hipStream_t st);
hipError_t select_flagged_u32(const uint32_t* in, const uint8_t* flags, uint32_t* out, uint32_t* count, uint64_t n,
                              void* tmp, size_t& bytes, hipStream_t st);
void k_tidy_gather(const uint32_t* sel, uint64_t m, const uint32_t* u, const uint32_t* v, const float* w,
                   uint32_t* ou, uint32_t* ov, float* ow, hipStream_t st);
void k_upd_delta(const uint32_t* u, uint64_t n, int s, int* delta, uint8_t* touched, hipStream_t st);
void k_upd_sizes(const uint32_t* T, uint64_t R, const uint64_t* off, const uint32_t* deg, const int* delta,
                 uint64_t* so, uint64_t* sn, int* bad, hipStream_t st);
void k_upd_tomb(const uint32_t* du, const uint32_t* dv, uint64_t nd, const uint32_t* T, uint64_t R,
                const uint64_t* bo, const uint64_t* off, const uint32_t* deg, const uint32_t* keys, const float* w,
                uint32_t* keep, float* delw, hipStream_t st);
void k_upd_scatter(const uint32_t* T, uint64_t R, const uint64_t* bo, const uint64_t* bn, uint64_t total_old,
                   const uint64_t* off, const uint32_t* deg, const uint32_t* keys, const float* w,
                   const uint32_t* keep, const uint32_t* rank, const uint32_t* iu, const uint32_t* iv,
                   const float* iw, uint64_t ni, uint32_t* yk, float* yw, hipStream_t st);
void k_upd_copy_back(const uint32_t* T, uint64_t R, const uint64_t* bn, uint64_t total, const uint64_t* dst_off,
                     const uint32_t* yk, const float* yw, uint32_t* keys, float* w, hipStream_t st);
void k_upd_degrees(const uint32_t* T, uint64_t R, int* delta, uint32_t* deg, uint8_t* touched, hipStream_t st);
void k_upd_caps(const uint32_t* deg, const int* delta, uint64_t S, uint64_t* cap, hipStream_t st);
void k_upd_copy_rows(uint64_t S, const uint64_t* off, const uint32_t* deg, const uint32_t* keys, const float* w,
                     const uint8_t* touched, const uint64_t* noff, uint32_t* nkeys, float* nw, hipStream_t st);
void k_upd_exists(const uint32_t* iu, const uint32_t* iv, uint64_t ni, uint8_t* ex, hipStream_t st);
void k_upd_wsum(const float* w, uint64_t n, double sign, double* out, hipStream_t st);

hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st);
hipError_t scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, void* tmp, size_t& tmp_bytes, hipStream_t st);

}  // namespace lv
