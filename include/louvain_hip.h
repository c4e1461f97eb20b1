/*
 * louvain_hip.h -- C ABI of the MI355X (gfx950) Louvain library
 * (liblouvain_hip.so).
 *
 * Drop-in for the four OpenMP entry points of
 * puzzlef/louvain-communities-openmp-dynamic:
 *
 *   louvainStaticOmp(x, o)                                    inc/louvain.hxx:1285
 *   louvainNaiveDynamicOmp(y, dels, ins, q, qvtot, qctot, o)  inc/louvain.hxx:1349
 *   louvainDynamicDeltaScreeningOmp(...)                      inc/louvain.hxx:1542
 *   louvainDynamicFrontierOmp(...)                            inc/louvain.hxx:1674
 *
 * with LouvainOptions (inc/louvain.hxx:30-63) and LouvainResult
 * (inc/louvain.hxx:77-145) mirrored field for field.  Types are the
 * reference's: vertex ids u32 (main.cxx:148), edge weights f32
 * (main.cxx:19-21), K / Sigma_tot / M in f64 (inc/louvain.hxx:67).
 *
 * Conventions: every function returns 0 (LV_OK) or a negative lv_status;
 * nothing throws across the ABI.  One lv_context per host thread; a context
 * owns one HIP stream and grow-only device buffers.  The caller owns every
 * host buffer.  Graph rows must be sorted by neighbour id with one entry per
 * (u, v) (the order inc/_bitset.hxx:245-262 keeps); batches must be tidy:
 * existence-filtered, sorted by (u, v), unique, both directions present
 * (inc/batch.hxx:202-209).  Unsorted batches are rejected with
 * LV_ERR_UNSORTED.
 */
#ifndef LOUVAIN_HIP_H
#define LOUVAIN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lv_status {
  LV_OK = 0,
  LV_ERR_INVALID = -1,   /* bad argument (null pointer, size mismatch, id out of range) */
  LV_ERR_NOMEM = -2,     /* device allocation failed */
  LV_ERR_HIP = -3,       /* a HIP runtime call failed (see lv_last_error) */
  LV_ERR_UNSORTED = -4,  /* batch not sorted by (u, v) / not unique */
  LV_ERR_NODEVICE = -5,  /* no HIP device */
  LV_ERR_LIMIT = -6,     /* input exceeds a documented limit */
  LV_ERR_PENDING = -7    /* lv_server_poll: the request has not finished yet */
} lv_status;

typedef enum lv_approach {
  LV_STATIC = 0,
  LV_NAIVE_DYNAMIC = 1,
  LV_DELTA_SCREENING = 2,
  LV_DYNAMIC_FRONTIER = 3
} lv_approach;

/* LouvainOptions, inc/louvain.hxx:30-63; defaults (lv_options_default) from :60. */
typedef struct lv_options {
  int32_t repeat;                /* [1]    */
  double resolution;             /* [1]    */
  double tolerance;              /* [1e-2] */
  double aggregation_tolerance;  /* [1.0]  */
  double tolerance_drop;         /* [10]   */
  int32_t max_iterations;        /* [20]   */
  int32_t max_passes;            /* [10]   */
} lv_options;

/* A host graph: CSR (gapped if degrees != NULL) with rows sorted by key.
 * Exactly one of weights / weights_f64 is set (f64 only for super-graphs fed
 * to the phase entry points).  exists == NULL means every id < span exists
 * (DiGraph::hasVertex, inc/Graph.hxx:148-150). */
typedef struct lv_graph {
  uint64_t span;              /* max vertex id + 1 */
  const uint64_t* offsets;    /* span + 1 */
  const uint32_t* degrees;    /* span, or NULL: offsets[u+1] - offsets[u] */
  const uint32_t* keys;       /* offsets[span] */
  const float* weights;       /* offsets[span] */
  const double* weights_f64;  /* offsets[span] (alternative to weights) */
  const uint8_t* exists;      /* span, or NULL */
} lv_graph;

/* A batch of edges as three parallel arrays (the reference's
 * vector<tuple<K,K,V>>, whose in-memory layout is implementation-defined). */
typedef struct lv_batch {
  const uint32_t* u;
  const uint32_t* v;
  const float* w;
  uint64_t size;
} lv_batch;

/* LouvainResult, inc/louvain.hxx:77-145.  The three arrays are caller-owned,
 * span entries each.  Times are milliseconds of the reference's timed region
 * (inc/louvain.hxx:1167-1217), averaged over repeat. */
typedef struct lv_result {
  uint32_t* membership;
  double* vertex_weight;
  double* community_weight;
  int32_t iterations;
  int32_t passes;
  float time;
  float marking_time;
  float initialization_time;
  float first_pass_time;
  float local_move_time;
  float aggregation_time;
  uint64_t affected_vertices;
} lv_result;

typedef struct lv_context lv_context;

/* Library / context ---------------------------------------------------- */
void lv_options_default(lv_options* o);
const char* lv_status_string(int status);
int lv_context_create(int device, lv_context** out);
void lv_context_destroy(lv_context* ctx);
const char* lv_last_error(const lv_context* ctx);
/* Path used by the local-moving phase.  0 (plain): each iteration is one
 * sweep in which every vertex is checked by its wave.  1 (frontier):
 * sweeps read the flags of 64 ids at once and visit only the flagged ones.
 * 2 (worklist, default): as 1, and DF / DS first passes run over ascending
 * queues of the flagged vertices, in sub-rounds until no vertex is flagged
 * by a lower id that it has not been visited after. */
int lv_context_set_path(lv_context* ctx, int path);
/* Scheduling of the local-moving kernels, the role of the reference's
 * schedule(dynamic, 2048) loop (inc/louvain.hxx:594).  Each launch runs
 * streams = clamp(stream_fraction * live vertices of the graph level,
 * min_streams, max_streams) lane groups; chunks of `chunk` consecutive
 * entries of the vertex list are handed out in ascending order, and a group
 * walks its chunk sequentially, so the list is swept in ascending order by a
 * window of about streams x chunk entries.  order: 0 = degree tiers lowest
 * first, 1 = highest first, 2 = no tiers: ascending ids through one kernel
 * (vertices of degree > 4096 run first through the hub kernel); at one
 * stream, order 2 is the reference's sequential visit order.  chunk == 0
 * hands group g the g-th of `streams` equal parts of the list instead, so
 * the vertices in flight lie far apart in the list.
 * Defaults: 0.002, 1, 2^20, 0, 2. */
int lv_context_set_schedule(lv_context* ctx, double stream_fraction, uint64_t min_streams, uint64_t max_streams,
                            uint64_t chunk, int order);
/* Worklist sub-rounds (the dynamic approaches' first pass, path 2): the
 * id-order kernel runs at most `streams` lane groups (1..65536), and with
 * interleave the sub-round's queue (ascending ids) is dealt out round-robin,
 * group g taking entries g, g + N, g + 2N, ... in that order, instead of in
 * N consecutive parts.  A tuning knob with no counterpart in the reference.
 * Defaults: 3072 (the waves the kernel keeps resident on the chip), 0. */
int lv_context_set_worklist_schedule(lv_context* ctx, uint64_t streams, int interleave);

/* The first `iterations` local-moving iterations of every pass only move a
 * vertex to a community with a lower id than its own (no two vertices can
 * then swap communities in one iteration).  Default 0 (off). */
int lv_context_set_pick_less(lv_context* ctx, int iterations);
/* Sweeps of a Static run's first pass and of every later pass (order 2):
 * in_sweep = 0 (default) runs every vertex of more than 4096 edges at once
 * before each sweep (the hub kernel); 1 visits each in its turn of the
 * ascending sweep, as the reference does (inc/louvain.hxx:594), with an HBM
 * table of its own.  At one stream, 1 is the reference's sequential visit
 * order.  With many streams it costs modularity on R-MAT graphs (a hub's
 * visit takes milliseconds while its neighbours move), so it is off. */
int lv_context_set_sweep_hubs(lv_context* ctx, int in_sweep);
/* Validated moves of the id-order kernel (order 2): a vertex about to move
 * reads its neighbours' communities again and, if a concurrent stream moved
 * one of them while the vertex was being decided, decides again from the
 * new view (at most twice).  This keeps the decisions on the neighbour view
 * a sequential sweep would have (inc/louvain.hxx:594-605), which concurrent
 * streams otherwise lose; one stream never sees a change, so at one stream
 * the results are unchanged.  0 off (default), 1 sweeps (every pass except
 * a DF / DS worklist first pass), 2 sweeps and worklists. */
int lv_context_set_validate(lv_context* ctx, int mode);
/* Device workspace of the context (grow-only scratch buffers kept between
 * calls): lv_context_workspace_bytes reports its size; with a limit set
 * (bytes > 0), a call that would grow it beyond the limit fails with
 * LV_ERR_NOMEM and leaves its graph and state as they were.  0 removes the
 * limit.  No counterpart in the reference (host memory, no limit). */
int lv_context_set_workspace_limit(lv_context* ctx, uint64_t bytes);
uint64_t lv_context_workspace_bytes(const lv_context* ctx);
/* Sweep schedules by pass (id-order sweeps, order 2): the first
 * lead_iterations local-moving iterations of a Static run's first pass run
 * lead_streams streams over ascending chunks of lead_chunk ids (over the
 * lead id range, see lv_context_set_lead_ids; the other ids under the wide
 * schedule); every pass after the first runs late_streams streams over
 * chunks of late_chunk ids handed out in ascending order (0: spread; at most
 * the streams the level's schedule gives) when late_streams is not 0.  The
 * reference's schedule(dynamic, 2048) loop (inc/louvain.hxx:594) with 16
 * threads, where its visit order decides the communities that form first.
 * Defaults: 1, 16, 2048, 256, 2048. */
int lv_context_set_pass_schedule(lv_context* ctx, int lead_iterations, uint64_t lead_streams, uint64_t lead_chunk,
                                 uint64_t late_streams, uint64_t late_chunk);
/* The ids a lead iteration sweeps under the lead schedule: [0, ids), then
 * the rest under the context's schedule.  0 (default): the id prefix that
 * holds a third of the edge slots (on R-MAT graphs the dense low-id core,
 * where the largest communities form), at least 2^18 ids (all of a smaller
 * graph) and at most 2^20; ~0: every id. */
int lv_context_set_lead_ids(lv_context* ctx, uint64_t ids);
/* Id-order sweeps (order 2) cut the ids into `slabs` consecutive ranges
 * launched one after another, each seeing every move of the ranges before
 * it; the streams spread within a range.  Default 1. */
int lv_context_set_slabs(lv_context* ctx, uint64_t slabs);
/* Hubs (degree > 4096) are processed by the hub kernel, at most `streams`
 * (1..256) at once, before the other vertices of each sweep.  Default 256. */
int lv_context_set_hub_streams(lv_context* ctx, uint64_t streams);
/* Degree above which a vertex runs in the hub batch (every such vertex at
 * once, edges split over many blocks) instead of in id order: sweep_degree
 * (>= 4096; larger keeps more hubs in the ascending sweep at the cost of
 * larger per-stream HBM tables) for sweeps, worklist_degree (256..4096) for
 * frontier worklists.  Defaults 4096, 256. */
int lv_context_set_hub_threshold(lv_context* ctx, uint32_t sweep_degree, uint32_t worklist_degree);

/* Kernel statistics of the local-moving phase, per degree tier (0..4),
 * accumulated over lv_run / entry-point calls since lv_stats_reset.  The
 * counters are the algorithmic work (processed vertices, visited edges,
 * distinct candidate communities, moves, edges of moved vertices); move_ms
 * and move_launches are filled only while profiling is on (each launch is
 * then bracketed by HIP events on the context's stream). */
typedef struct lv_stats {
  double move_ms[5];
  uint64_t move_launches[5];
  uint64_t vertices[5];
  uint64_t edges[5];
  uint64_t candidates[5];
  uint64_t moves[5];
  uint64_t moved_edges[5];
  uint64_t runs;
  /* worklist sub-rounds whose id-order kernel and tabled hubs ran as one
   * launch: their time (profiling) and count; their work counts in slots 0
   * and 1 */
  double fused_ms;
  uint64_t fused_launches;
} lv_stats;
int lv_context_set_profiling(lv_context* ctx, int on);
/* Worklist sub-round groups (the DF / DS first pass) replayed from a captured
 * HIP graph (on by default; off while profiling, which times every launch).
 * No counterpart in the reference: a launch-overhead switch of this build. */
int lv_context_set_graphs(lv_context* ctx, int on);
/* Hub neighbour-community tables kept with a resident state (on by default):
 * from the second worklist DF / DS run of a state on, every vertex of more
 * than 256 edges is evaluated from its table instead of its edges; batch
 * edges and moves keep the tables exact.  No counterpart in the reference:
 * an incremental index of this build. */
int lv_context_set_hub_tables(lv_context* ctx, int on);
/* Counters since the last reset.  The work counters of local-moving are kept
 * on the device and read here (this waits for the context's stream). */
int lv_stats_get(const lv_context* ctx, lv_stats* out);
int lv_stats_reset(lv_context* ctx);

/* The four entry points (host buffers in, host buffers out) ------------ */
/* louvainStaticOmp, inc/louvain.hxx:1285-1296 */
int lv_louvain_static(lv_context* ctx, const lv_graph* x, const lv_options* o, lv_result* res);
/* louvainNaiveDynamicOmp, inc/louvain.hxx:1349-1366 */
int lv_louvain_naive_dynamic(lv_context* ctx, const lv_graph* y, const lv_batch* deletions,
                             const lv_batch* insertions, const uint32_t* q, const double* qvtot,
                             const double* qctot, const lv_options* o, lv_result* res);
/* louvainDynamicDeltaScreeningOmp, inc/louvain.hxx:1542-1565 */
int lv_louvain_delta_screening(lv_context* ctx, const lv_graph* y, const lv_batch* deletions,
                               const lv_batch* insertions, const uint32_t* q, const double* qvtot,
                               const double* qctot, const lv_options* o, lv_result* res);
/* louvainDynamicFrontierOmp, inc/louvain.hxx:1674-1690 */
int lv_louvain_dynamic_frontier(lv_context* ctx, const lv_graph* y, const lv_batch* deletions,
                                const lv_batch* insertions, const uint32_t* q, const double* qvtot,
                                const double* qctot, const lv_options* o, lv_result* res);

/* Device-resident form: graph, batch and state stay in HBM between calls.
 * lv_run is the same algorithm as the four entry points above (approach
 * selects which); the state holds (membership, K, Sigma_tot) in and out,
 * the carry of main.cxx:127-135.  res->membership / vertex_weight /
 * community_weight may be NULL (nothing is copied back). */
typedef struct lv_dgraph lv_dgraph;
typedef struct lv_dbatch lv_dbatch;
typedef struct lv_dstate lv_dstate;
int lv_dgraph_create(lv_context* ctx, const lv_graph* g, lv_dgraph** out);
void lv_dgraph_destroy(lv_dgraph* g);
uint64_t lv_dgraph_span(const lv_dgraph* g);
/* Directed edges (sum of the row degrees). */
uint64_t lv_dgraph_edges(const lv_dgraph* g);
/* Device bytes a resident graph / state holds (a state's hub
 * neighbour-community tables included). */
uint64_t lv_dgraph_device_bytes(const lv_dgraph* g);
uint64_t lv_dstate_device_bytes(const lv_dstate* s);
/* Edge slots (offsets[span]): rows may carry slack after lv_dgraph_update. */
uint64_t lv_dgraph_capacity(const lv_dgraph* g);
/* Batch tidy and graph update on the device: tidyBatchUpdateU
 * (inc/batch.hxx:202-209) then applyBatchUpdateU (inc/batch.hxx:225-236).
 * Deletions are kept when the edge exists in g, insertions when it does not;
 * each list is sorted by (u, v) and made unique (the first of repeated
 * entries, in input order, is kept).  g is then updated in place: rows stay
 * sorted by neighbour id, an insertion's endpoints become existing vertices,
 * 2M and the degree lists follow.  A row that outgrows its slack moves every
 * row to a new CSR with slack max(4, degree / 8).  *tidy receives the tidy
 * batch, ready for lv_run on the updated graph.  Ids must be < span (the
 * span does not grow: LV_ERR_INVALID); only input graphs (f32 weights). */
int lv_dgraph_update(lv_context* ctx, lv_dgraph* g, const lv_batch* deletions, const lv_batch* insertions,
                     lv_dbatch** tidy);
/* The device graph's CSR: offsets[span + 1], degrees[span], keys and weights
 * of lv_dgraph_capacity entries (row u is keys[offsets[u] ..
 * offsets[u] + degrees[u])), exists[span] (may be NULL). */
int lv_dgraph_download(lv_context* ctx, const lv_dgraph* g, uint64_t* offsets, uint32_t* degrees, uint32_t* keys,
                       float* weights, uint8_t* exists);
int lv_dbatch_create(lv_context* ctx, const lv_batch* deletions, const lv_batch* insertions, lv_dbatch** out);
void lv_dbatch_destroy(lv_dbatch* b);
int lv_dbatch_sizes(const lv_dbatch* b, uint64_t* deletions, uint64_t* insertions);
int lv_dbatch_download(lv_context* ctx, const lv_dbatch* b, uint32_t* du, uint32_t* dv, float* dw, uint32_t* iu,
                       uint32_t* iv, float* iw);
int lv_dstate_create(lv_context* ctx, uint64_t span, lv_dstate** out);
void lv_dstate_destroy(lv_dstate* s);
int lv_dstate_upload(lv_context* ctx, lv_dstate* s, const uint32_t* q, const double* qvtot, const double* qctot);
int lv_dstate_download(lv_context* ctx, const lv_dstate* s, uint32_t* q, double* qvtot, double* qctot);
int lv_dstate_copy(lv_context* ctx, lv_dstate* dst, const lv_dstate* src);
/* dst = src with its hub neighbour-community tables: dst's next run carries
 * or builds them as src's would (lv_dstate_copy leaves dst without tables). */
int lv_dstate_clone(lv_context* ctx, lv_dstate* dst, const lv_dstate* src);
int lv_run(lv_context* ctx, int approach, const lv_dgraph* g, const lv_dbatch* b, lv_dstate* s,
           const lv_options* o, lv_result* res);
/* Batched form: G graphs side by side in one device graph (graph k owns
 * vertex ids [voff[k], voff[k+1]), no edges between graphs), one batch and
 * one state over all of them (graph k's community ids within its range).
 * Static, ND, DS or DF runs on every graph at once -- each kernel launch
 * covers all graphs -- with each graph's own M (DS's insertion gains too),
 * tolerance, iteration count, pass count and aggregation decision, as
 * lv_run on that graph alone would (inc/louvain.hxx:1120-1221 per graph),
 * and in as many concurrent streams as it would get alone.  A graph that
 * finishes leaves the later
 * levels.  res holds num_graphs results with graph-local membership ids
 * (arrays optional); the state keeps the union ids, each graph's re-based
 * into its own range, ready for the next batch.  community_weight holds
 * Sigma_tot at each graph's community ids and 0 at ids that are no
 * community.  No counterpart in the reference, which runs one graph per
 * call. */
int lv_run_batched(lv_context* ctx, int approach, const lv_dgraph* g, const lv_dbatch* b, lv_dstate* s,
                   const uint64_t* voff, uint64_t num_graphs, const lv_options* o, lv_result* res);
/* Modularity of a device state's membership on a device graph
 * (modularityByOmp, inc/properties.hxx:205-233). */
int lv_dmodularity(lv_context* ctx, const lv_dgraph* g, const lv_dstate* s, double resolution, double* out);
/* Schedule tuning -----------------------------------------------------------
 * The degree buckets and schedules of local-moving, chosen by timing: the
 * worklist hub degree (256 or 512: above it a vertex leaves the id-order
 * kernel for the hub kernels), the sweep hub degree (4096 or 16384), the
 * worklist id-order streams (1536, 3072, 6144) and queue layout, and the
 * streams per live vertex of a level (0.001, 0.002, 0.004).  The
 * reference's counterpart is its fixed schedule(dynamic, 2048)
 * (inc/louvain.hxx:594); no counterpart for the tuning itself. */
typedef struct lv_tune_choice {
  uint32_t worklist_hub_degree;   /* lv_context_set_hub_threshold's worklist_degree */
  uint32_t sweep_hub_degree;      /* lv_context_set_hub_threshold's sweep_degree */
  uint64_t worklist_streams;      /* lv_context_set_worklist_schedule */
  int32_t interleave;
  double stream_fraction;         /* lv_context_set_schedule */
  double ms, default_ms;          /* median host-clock ms of a run: this choice, the defaults */
  double modularity, default_modularity;  /* after the probe run */
  int32_t candidates;             /* schedules timed */
} lv_tune_choice;
/* Times the candidate schedules of `approach` (ND, DS or DF; the ones its
 * path uses) on graph g with batch `probe` from state s, `reps` runs each
 * (1..64; median of the host clock around lv_run), every trial on a copy of
 * s (hub tables included: s is left as it is, and a probe made by
 * lv_dgraph_update from the graph of s's tables carries them as a real
 * batch would).  Keeps the fastest schedule whose modularity is within 1e-3
 * of the defaults', and the defaults unless it beats them by more than 3%;
 * caches it under g's shape (approach, log2 span, log2 live edges) and
 * applies it to the context.  o may be NULL (defaults).  The trials count in
 * the device work counters of lv_stats_get. */
int lv_autotune(lv_context* ctx, int approach, const lv_dgraph* g, const lv_dbatch* probe, const lv_dstate* s,
                const lv_options* o, int reps, lv_tune_choice* out);
/* on (default): before every run the knobs take the cached choice for the
 * run's graph shape, or the defaults, and lv_run_batched tunes a shape it
 * has not seen (on copies of its state, before the run).  Setting a tuned
 * knob by hand (lv_context_set_schedule's stream_fraction,
 * lv_context_set_hub_threshold, lv_context_set_worklist_schedule,
 * lv_context_apply_tuning) to other than its default keeps the knobs as set
 * until lv_context_set_autotune(ctx, 1) or every such knob is back at its
 * default.  off: the knobs stay as they are. */
int lv_context_set_autotune(lv_context* ctx, int on);
/* The cached choice for g's shape (graphs: 1, or the batched form's graph
 * count); out->candidates == 0 when none is cached (out then holds the knobs
 * the next run takes). */
int lv_context_tuning(lv_context* ctx, int approach, const lv_dgraph* g, uint64_t graphs, lv_tune_choice* out);
/* Cache a choice (made on another context) for g's shape. */
int lv_context_add_tuning(lv_context* ctx, int approach, const lv_dgraph* g, const lv_tune_choice* k);
/* Set the tuned knobs to k's (NULL: the defaults), by hand. */
int lv_context_apply_tuning(lv_context* ctx, const lv_tune_choice* k);
/* lv_dmodularity with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Spans up to 2^21 vertices (LV_ERR_LIMIT
 * above); its scratch is a static device array of 2^22 + 32 doubles (32 MiB)
 * that the library holds on every GPU it runs on.  Calls that overlap on one
 * GPU share that array and race. */
int lv_dmodularity_noalloc(lv_context* ctx, const lv_dgraph* g, const lv_dstate* s, double resolution, double* out);
/* Diagnostic: rebuild the state's hub neighbour-community tables (see
 * lv_context_set_hub_tables) from graph g and the state's membership and
 * compare; *tables = tables compared, *mismatches = those that differ (0 and 0
 * when the state has no valid tables).  No counterpart in the reference. */
int lv_dstate_check_hub_tables(lv_context* ctx, const lv_dgraph* g, const lv_dstate* s, uint64_t* tables,
                               uint64_t* mismatches);

/* Phase entry points: each hot-path function of inc/louvain.hxx on its own,
 * host buffers in and out.  They run the same kernels as lv_run. */
/* edgeWeightOmp, inc/properties.hxx:96-106 (2M) */
int lv_edge_weight(lv_context* ctx, const lv_graph* g, double* out);
/* modularityByOmp, inc/properties.hxx:205-233 */
int lv_modularity(lv_context* ctx, const lv_graph* g, const uint32_t* membership, double resolution, double* out);
/* louvainVertexWeightsOmpW, inc/louvain.hxx:210-218 */
int lv_vertex_weights(lv_context* ctx, const lv_graph* g, double* vtot);
/* louvainUpdateWeightsFromOmpU, inc/louvain.hxx:374-388 (vtot, ctot in/out) */
int lv_update_weights(lv_context* ctx, uint64_t span, const lv_batch* deletions, const lv_batch* insertions,
                      const uint32_t* vcom, double* vtot, double* ctot);
/* louvainAffectedVerticesFrontierOmpW, inc/louvain.hxx:1606-1624 */
int lv_affected_frontier(lv_context* ctx, uint64_t span, const lv_batch* deletions, const lv_batch* insertions,
                         const uint32_t* vcom, uint8_t* vaff);
/* louvainAffectedVerticesDeltaScreeningOmpW, inc/louvain.hxx:1441-1486 */
int lv_affected_delta_screening(lv_context* ctx, const lv_graph* g, const lv_batch* deletions,
                                const lv_batch* insertions, const uint32_t* vcom, const double* vtot,
                                const double* ctot, double M, double resolution, uint8_t* vertices,
                                uint8_t* neighbors, uint8_t* communities);
/* louvainScanCommunitiesW + louvainChooseCommunity, inc/louvain.hxx:406-464:
 * every vertex's best community against one fixed snapshot, no move. */
int lv_choose_community(lv_context* ctx, const lv_graph* g, const uint32_t* vcom, const double* vtot,
                        const double* ctot, double M, double resolution, uint32_t* best_c, double* best_e);
/* louvainCommunityExistsOmpW + louvainRenumberCommunitiesOmpW,
 * inc/louvain.hxx:668-682, :941-945: vcom (vcom_size entries) is renumbered
 * in place; cext (span entries) receives the old -> new map. */
int lv_renumber(lv_context* ctx, const lv_graph* g, uint32_t* vcom, uint64_t vcom_size, uint32_t* cext,
                uint64_t* num_communities);
/* louvainCommunityVerticesOmpW + louvainAggregateOmpW,
 * inc/louvain.hxx:800-812, :987-992: the super-graph of a renumbered
 * membership as a gapped CSR (zoff[C+1], zdeg[C], zkeys / zw of capacity
 * offsets[span]).  Row order is unspecified (the reference's depends on
 * thread timing); the rows as sets are exact. */
int lv_aggregate(lv_context* ctx, const lv_graph* g, const uint32_t* vcom, uint64_t num_communities,
                 uint64_t* zoff, uint32_t* zdeg, uint32_t* zkeys, double* zw);

/* Serving ---------------------------------------------------------------
 * A request loop over resident graphs: each graph and its state
 * (membership, K, Sigma_tot and hub neighbour-community tables) stay in HBM;
 * a request is a raw update batch for one graph, tidied and applied on the
 * device (lv_dgraph_update), then Dynamic Frontier Louvain runs from the
 * graph's state (lv_run) -- the per-batch step of the reference's driver,
 * main.cxx:101-137, with the state carried as main.cxx:127-135 does.
 * Continuous batching: `slots` worker threads, each with its own lv_context
 * (one HIP stream and workspace), take requests as they free up; requests of
 * one graph run one at a time in submission order, those of different graphs
 * overlap on the device.  Admission by device memory: a request starts only
 * while the resident graphs and states, every slot's workspace and the
 * batches in flight (the request's own included) fit the memory budget, or
 * when nothing else is in flight.  Each slot's workspace is sized for a graph
 * when the graph is added (a warm-up run), so serving does not regrow it.
 * With autotune, adding a graph also tunes its DF schedule (lv_autotune on a
 * copy of the graph and state, with a random probe batch of the requests'
 * kind: 80% insertions of new edges, 20% deletions of existing ones), and
 * every slot runs the graph's requests under that choice.
 * No counterpart in the reference, which runs one batch at a time. */
typedef struct lv_server lv_server;
typedef struct lv_server_options {
  int device;               /* [0] */
  int slots;                /* worker contexts [4] */
  uint64_t memory_budget;   /* device bytes for graphs, states, workspaces and batches in flight; 0 = no limit [0] */
  lv_options options;       /* Louvain options of every run [lv_options_default] */
  int path;                 /* local-moving path of the slots (lv_context_set_path) [2] */
  int modularity;           /* compute each request's modularity (lv_dmodularity) [0] */
  int keep_states;          /* keep each request's resulting state for lv_server_poll [0] */
  int autotune;             /* tune each added graph's DF schedule (lv_autotune) on a probe batch [1] */
  double probe_fraction;    /* probe batch size, a fraction of the graph's undirected edges [1e-4] */
} lv_server_options;
/* A finished request.  result carries the run's scalars; when the server
 * keeps states and the caller sets result.membership / vertex_weight /
 * community_weight (span entries each), lv_server_poll fills them with the
 * graph's state right after the request.  Times are milliseconds on the host
 * clock since lv_server_create. */
typedef struct lv_completion {
  uint64_t ticket;
  uint32_t graph_id;
  int status;               /* LV_OK or the request's error */
  double submitted_ms, started_ms, done_ms;
  lv_result result;
  double modularity;        /* NaN unless options.modularity */
  uint64_t deletions, insertions;  /* sizes of the tidy batch */
} lv_completion;
typedef struct lv_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t resident_bytes;  /* graphs and states */
  uint64_t workspace_bytes; /* all slots */
  uint64_t inflight_bytes;  /* batches of running requests */
  uint64_t peak_bytes;      /* largest sum of the three at an admission */
} lv_server_stats;
void lv_server_options_default(lv_server_options* o);
int lv_server_create(const lv_server_options* o, lv_server** out);
/* Waits for every submitted request, then frees everything. */
void lv_server_destroy(lv_server* s);
/* A resident graph with its starting state: (q, qvtot, qctot), or NULL for
 * a Static run on the device.  Returns its id in *graph_id. */
int lv_server_add_graph(lv_server* s, const lv_graph* g, const uint32_t* q, const double* qvtot, const double* qctot,
                        uint32_t* graph_id);
/* Queue a raw batch (copied) for a graph; returns a ticket. */
int lv_server_submit(lv_server* s, uint32_t graph_id, const lv_batch* deletions, const lv_batch* insertions,
                     uint64_t* ticket);
/* The completion of a ticket, waiting up to timeout_ms (0: do not wait,
 * < 0: until it finishes); LV_ERR_PENDING if it has not finished.  A ticket
 * is reported once. */
int lv_server_poll(lv_server* s, uint64_t ticket, int timeout_ms, lv_completion* out);
/* Wait until every submitted request has finished. */
int lv_server_drain(lv_server* s);
/* The current state of a graph (waits until the graph is idle). */
int lv_server_graph_state(lv_server* s, uint32_t graph_id, uint32_t* q, double* qvtot, double* qctot);
int lv_server_get_stats(lv_server* s, lv_server_stats* out);
/* The DF schedule chosen for a graph when it was added (candidates == 0:
 * not tuned, the defaults). */
int lv_server_graph_tuning(lv_server* s, uint32_t graph_id, lv_tune_choice* out);
const char* lv_server_last_error(const lv_server* s);
/* Milliseconds on the server's clock (the one lv_completion's times use). */
double lv_server_now_ms(const lv_server* s);

#ifdef __cplusplus
}
#endif
#endif /* LOUVAIN_HIP_H */
