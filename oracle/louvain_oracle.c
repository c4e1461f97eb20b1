/*
 * oracle/louvain_oracle.c -- TEST INFRASTRUCTURE, NOT THE PRODUCT.
 *
 * A plain-C + OpenMP restatement of the reference's Static, Naive-dynamic,
 * Delta-screening and Dynamic-Frontier Louvain
 * (puzzlef/louvain-communities-openmp-dynamic, inc/louvain.hxx), written from
 * the reference's semantics.  It is used only as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * (the HIP library) never links, loads or calls it.
 *
 * Types follow the reference: vertex ids u32 (main.cxx:148), input edge
 * weights f32 (main.cxx:19-21), K / Sigma_tot / hashtable values / super-graph
 * weights / M in f64 (louvain.hxx:67, :1121-1123), flags u8 (louvain.hxx:1123).
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *  - every function below is compared bit-for-bit, at 1 thread, against the
 *    golden vectors written by the compiled reference (tests/golden/*.npz,
 *    made by tests/golden/make_golden.py through oracle/_ref/libref.so);
 *  - closed forms / invariants on top: Sigma_tot sums to 2M, K equals row
 *    sums, renumbering is an order-preserving compaction, super-graph total
 *    weight equals 2M, brute-force modularity on tiny graphs.
 * Nothing here is "parity unpinned".
 *
 * The ORACLE_MUTANT_* switches plant a known mistake; only the mutation
 * test (tests/test_oracle.py) builds them, to show the checks catch it.
 *
 * Built with -ffp-contract=off so that deltaModularity evaluates exactly as
 * the reference's x86-64 build does (no fused multiply-add).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t S;             /* span: max vertex id + 1 */
  const uint64_t* off;    /* row offsets, S+1 */
  const uint32_t* deg;    /* row lengths (gapped CSR), NULL = off[u+1]-off[u] */
  const uint32_t* keys;   /* neighbour ids, rows sorted ascending */
  const float* w32;       /* f32 weights (input graph), or */
  const double* w64;      /* f64 weights (super-graph) */
  const uint8_t* ex;      /* vertex exists mask, NULL = all exist */
} oracle_graph;

typedef struct {
  const uint32_t* u;
  const uint32_t* v;
  const float* w;
  uint64_t n;
} oracle_batch;

typedef struct {
  int32_t repeat;
  double resolution, tolerance, aggregation_tolerance, tolerance_drop;
  int32_t max_iterations, max_passes;
} oracle_options;

typedef struct {
  int32_t iterations, passes;
  float time, marking_time, initialization_time, first_pass_time, local_move_time, aggregation_time;
  uint64_t affected_vertices;
} oracle_out;

/* ---- graph access (Graph.hxx:59-169 DiGraph, :383-639 DiGraphCsr) ---- */
static inline int g_has(const oracle_graph* g, uint64_t u) { return u < g->S && (!g->ex || g->ex[u]); }
static inline uint64_t g_deg(const oracle_graph* g, uint64_t u) {
  return g->deg ? g->deg[u] : g->off[u + 1] - g->off[u];
}
static inline double g_w(const oracle_graph* g, uint64_t i) { return g->w64 ? g->w64[i] : (double)g->w32[i]; }
static uint64_t g_order(const oracle_graph* g) {
  if (!g->ex) return g->S;
  uint64_t n = 0;
  for (uint64_t u = 0; u < g->S; ++u) n += g->ex[u] != 0;
  return n;
}

/* ---- per-thread dense scan table (louvain.hxx:160-166, :406-438) ---- */
typedef struct {
  uint32_t* list;
  uint64_t n, cap;
  double* val; /* dense, S entries */
} table_t;

static void tbl_init(table_t* t, uint64_t S) {
  t->cap = 1024;
  t->n = 0;
  t->list = (uint32_t*)malloc(t->cap * sizeof(uint32_t));
  t->val = (double*)calloc(S ? S : 1, sizeof(double));
}
static void tbl_free(table_t* t) { free(t->list); free(t->val); }
static inline void tbl_clear(table_t* t) {
  for (uint64_t i = 0; i < t->n; ++i) t->val[t->list[i]] = 0;
  t->n = 0;
}
/* louvainScanCommunityW (louvain.hxx:405-411): a key is listed when its
 * value is zero before the add, so listing order is first appearance. */
static inline void tbl_add(table_t* t, uint32_t c, double w) {
  if (!t->val[c]) {
    if (t->n == t->cap) { t->cap *= 2; t->list = (uint32_t*)realloc(t->list, t->cap * sizeof(uint32_t)); }
    t->list[t->n++] = c;
  }
  t->val[c] += w;
}
static inline void scan_communities(table_t* t, const oracle_graph* g, uint32_t u, const uint32_t* vcom, int self) {
  uint64_t i = g->off[u], I = i + g_deg(g, u);
  for (; i < I; ++i) {
    uint32_t v = g->keys[i];
    if (!self && u == v) continue;
    tbl_add(t, vcom[v], g_w(g, i));
  }
}

/* deltaModularity (properties.hxx:253-256), same operation order. */
static inline double delta_modularity(double vcout, double vdout, double vtot, double ctot, double dtot, double M, double R) {
  return (vcout - vdout) / M - R * vtot * (vtot + ctot - dtot) / (2 * M * M);
}

/* louvainChooseCommunity (louvain.hxx:454-464): strict '>' from 0, so the
 * first-listed maximum wins; (0, 0) when no positive gain. */
static inline void choose_community(const table_t* t, uint32_t u, const uint32_t* vcom, const double* vtot,
                                    const double* ctot, double M, double R, uint32_t* cbest, double* ebest) {
  uint32_t cmax = 0, d = vcom[u];
  double emax = 0;
  for (uint64_t i = 0; i < t->n; ++i) {
    uint32_t c = t->list[i];
    if (c == d) continue;
    double e = delta_modularity(t->val[c], t->val[d], vtot[u], ctot[c], ctot[d], M, R);
#ifdef ORACLE_MUTANT_TIEBREAK
    if (e >= emax && e > 0) { emax = e; cmax = c; }
#else
    if (e > emax) { emax = e; cmax = c; }
#endif
  }
  *cbest = cmax;
  *ebest = emax;
}

/* ---- weights / init (louvain.hxx:210-218, :285-293) ---- */
static void vertex_weights(double* vtot, const oracle_graph* g) {
  int64_t S = (int64_t)g->S;
#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    uint64_t i = g->off[u], I = i + g_deg(g, (uint64_t)u);
    for (; i < I; ++i) vtot[u] += g_w(g, i);
  }
}
static void initialize(uint32_t* vcom, double* ctot, const oracle_graph* g, const double* vtot) {
  int64_t S = (int64_t)g->S;
#pragma omp parallel for schedule(static, 2048)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    vcom[u] = (uint32_t)u;
    ctot[u] = vtot[u];
  }
}

/* edgeWeightOmp (properties.hxx:96-106): per-vertex sums reduced in double. */
double oracle_edge_weight(const oracle_graph* g) {
  double a = 0;
  int64_t S = (int64_t)g->S;
#pragma omp parallel for schedule(auto) reduction(+ : a)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    double s = 0;
    uint64_t i = g->off[u], I = i + g_deg(g, (uint64_t)u);
    for (; i < I; ++i) s += g_w(g, i);
    a += s;
  }
  return a;
}

/* modularityBy (properties.hxx:178-192), sequential. */
double oracle_modularity(const oracle_graph* g, const uint32_t* memb, double M, double R) {
  uint64_t S = g->S;
  double* cin = (double*)calloc(S ? S : 1, sizeof(double));
  double* ctot = (double*)calloc(S ? S : 1, sizeof(double));
  for (uint64_t u = 0; u < S; ++u) {
    if (!g_has(g, u)) continue;
    uint32_t c = memb[u];
    uint64_t i = g->off[u], I = i + g_deg(g, u);
    for (; i < I; ++i) {
      uint32_t d = memb[g->keys[i]];
      double w = g_w(g, i);
      if (c == d) cin[c] += w;
      ctot[c] += w;
    }
  }
  double a = 0;
  for (uint64_t c = 0; c < S; ++c) a += cin[c] / (2 * M) - R * pow(ctot[c] / (2 * M), 2);
  free(cin);
  free(ctot);
  return a;
}

/* belongsOmp (_openmp.hxx:16-20): keys are owned in chunks of 1024. */
static inline int belongs(uint64_t key, int t, int T) { return (int)((key / 1024) % (uint64_t)T) == t; }

/* louvainUpdateWeightsFromOmpU (louvain.hxx:374-388): every thread walks the
 * whole batch and applies only the keys it owns, so each key sees its
 * updates in batch order, deletions first. */
void oracle_update_weights(double* vtot, double* ctot, const oracle_batch* del, const oracle_batch* ins,
                           const uint32_t* vcom) {
#pragma omp parallel
  {
    int t = omp_get_thread_num(), T = omp_get_num_threads();
    for (uint64_t i = 0; i < del->n; ++i) {
      uint32_t u = del->u[i], c = vcom[u];
      double w = del->w[i];
      if (belongs(u, t, T)) vtot[u] -= w;
      if (belongs(c, t, T)) ctot[c] -= w;
    }
    for (uint64_t i = 0; i < ins->n; ++i) {
      uint32_t u = ins->u[i], c = vcom[u];
      double w = ins->w[i];
      if (belongs(u, t, T)) vtot[u] += w;
      if (belongs(c, t, T)) ctot[c] += w;
    }
  }
}

/* louvainAffectedVerticesFrontierOmpW (louvain.hxx:1605-1624). */
void oracle_frontier(uint8_t* vaff, uint64_t S, const oracle_batch* del, const oracle_batch* ins, const uint32_t* vcom) {
  memset(vaff, 0, S);
  int64_t D = (int64_t)del->n, I = (int64_t)ins->n;
#pragma omp parallel for schedule(auto)
  for (int64_t i = 0; i < D; ++i)
    if (vcom[del->u[i]] == vcom[del->v[i]]) vaff[del->u[i]] = 1;
#pragma omp parallel for schedule(auto)
  for (int64_t i = 0; i < I; ++i)
    if (vcom[ins->u[i]] != vcom[ins->v[i]]) vaff[ins->u[i]] = 1;
}

/* louvainAffectedVerticesDeltaScreeningOmpW (louvain.hxx:1440-1486).  The
 * insertions must be grouped by source (tidy batches are sorted by (u, v));
 * group n is handled by thread n % T, so the flags do not depend on T. */
static void delta_screening(uint8_t* vertices, uint8_t* neighbors, uint8_t* communities, table_t* tbl,
                            const oracle_graph* g, const oracle_batch* del, const oracle_batch* ins,
                            const uint32_t* vcom, const double* vtot, const double* ctot, double M, double R) {
  uint64_t S = g->S;
  memset(vertices, 0, S);
  memset(neighbors, 0, S);
  memset(communities, 0, S);
  int64_t D = (int64_t)del->n;
#pragma omp parallel for schedule(auto)
  for (int64_t i = 0; i < D; ++i) {
    uint32_t u = del->u[i], v = del->v[i];
    if (vcom[u] != vcom[v]) continue;
    vertices[u] = 1;
    neighbors[u] = 1;
    communities[vcom[v]] = 1;
  }
  uint64_t I = ins->n;
#pragma omp parallel
  {
    int T = omp_get_num_threads(), t = omp_get_thread_num();
    table_t* tb = &tbl[t];
    uint32_t u0 = I > 0 ? ins->u[0] : 0;
    for (uint64_t i = 0, n = 0; i < I;) {
      uint32_t u = ins->u[i];
      if (u != u0) { ++n; u0 = u; }
      if ((int)(n % (uint64_t)T) != t) { ++i; continue; }
      tbl_clear(tb);
      for (; i < I && ins->u[i] == u; ++i) {
        uint32_t v = ins->v[i];
        if (vcom[u] == vcom[v]) continue;
        tbl_add(tb, vcom[v], (double)ins->w[i]);
      }
      uint32_t c;
      double e;
      choose_community(tb, u, vcom, vtot, ctot, M, R, &c, &e);
      if (e <= 0) continue;
      vertices[u] = 1;
      neighbors[u] = 1;
      communities[c] = 1;
    }
  }
  int64_t SS = (int64_t)S;
#pragma omp parallel for schedule(auto)
  for (int64_t u = 0; u < SS; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    if (neighbors[u]) {
      uint64_t i = g->off[u], E = i + g_deg(g, (uint64_t)u);
      for (; i < E; ++i) vertices[g->keys[i]] = 1;
    }
    if (communities[vcom[u]]) vertices[u] = 1;
  }
}

static table_t* tables_new(int T, uint64_t S) {
  table_t* t = (table_t*)malloc(sizeof(table_t) * (size_t)T);
  for (int i = 0; i < T; ++i) tbl_init(&t[i], S);
  return t;
}
static void tables_free(table_t* t, int T) {
  for (int i = 0; i < T; ++i) tbl_free(&t[i]);
  free(t);
}

void oracle_delta_screening(uint8_t* vertices, uint8_t* neighbors, uint8_t* communities, const oracle_graph* g,
                            const oracle_batch* del, const oracle_batch* ins, const uint32_t* vcom,
                            const double* vtot, const double* ctot, double M, double R) {
  int T = omp_get_max_threads();
  table_t* tb = tables_new(T, g->S);
  delta_screening(vertices, neighbors, communities, tb, g, del, ins, vcom, vtot, ctot, M, R);
  tables_free(tb, T);
}

/* Decision of every vertex against one fixed snapshot: the clear/scan/choose
 * steps of louvainMoveOmpW (louvain.hxx:599-601) without the move. */
void oracle_choose(uint32_t* best_c, double* best_e, const oracle_graph* g, const uint32_t* vcom,
                   const double* vtot, const double* ctot, double M, double R) {
  int T = omp_get_max_threads();
  table_t* tb = tables_new(T, g->S);
  int64_t S = (int64_t)g->S;
#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t u = 0; u < S; ++u) {
    table_t* t = &tb[omp_get_thread_num()];
    best_c[u] = 0;
    best_e[u] = 0;
    if (!g_has(g, (uint64_t)u)) continue;
    tbl_clear(t);
    scan_communities(t, g, (uint32_t)u, vcom, 0);
    choose_community(t, (uint32_t)u, vcom, vtot, ctot, M, R, &best_c[u], &best_e[u]);
  }
  tables_free(tb, T);
}

/* louvainMoveOmpW (louvain.hxx:587-609).  fa == NULL means every vertex is
 * allowed.  Returns the iteration count, 0 if it converged with no gain. */
static int local_move(uint32_t* vcom, double* ctot, uint8_t* vaff, table_t* tb, const oracle_graph* g,
                      const double* vtot, double M, double R, int L, double E, const uint8_t* fa) {
  int64_t S = (int64_t)g->S;
  int l = 0;
  double el = 0;
  for (; l < L;) {
    el = 0;
#pragma omp parallel for schedule(dynamic, 2048) reduction(+ : el)
    for (int64_t u = 0; u < S; ++u) {
      table_t* t = &tb[omp_get_thread_num()];
      if (!g_has(g, (uint64_t)u)) continue;
      if ((fa && !fa[u]) || !vaff[u]) continue;
      tbl_clear(t);
#ifdef ORACLE_MUTANT_SELFLOOP
      scan_communities(t, g, (uint32_t)u, vcom, 1);
#else
      scan_communities(t, g, (uint32_t)u, vcom, 0);
#endif
      uint32_t c;
      double e;
      choose_community(t, (uint32_t)u, vcom, vtot, ctot, M, R, &c, &e);
#ifdef ORACLE_MUTANT_COMMUNITY0
      if (e > 0) {
#else
      if (c) {
#endif
        /* louvainChangeCommunityOmpW (louvain.hxx:496-503) */
        uint32_t d = vcom[u];
#pragma omp atomic
        ctot[d] -= vtot[u];
#pragma omp atomic
        ctot[c] += vtot[u];
        vcom[u] = c;
        uint64_t i = g->off[u], I = i + g_deg(g, (uint64_t)u);
        for (; i < I; ++i) vaff[g->keys[i]] = 1;
      }
      vaff[u] = 0;
      el += e;
    }
    int conv = el <= E;
    ++l;
    if (conv) break;
  }
  return (l > 1 || el) ? l : 0;
}

/* louvainCommunityExistsOmpW (louvain.hxx:667-682) on a[0..S'). */
static uint64_t community_exists(uint32_t* a, const oracle_graph* g, const uint32_t* vcom) {
  int64_t S = (int64_t)g->S;
  uint64_t C = 0;
  memset(a, 0, g->S * sizeof(uint32_t));
#pragma omp parallel for schedule(static, 2048) reduction(+ : C)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    uint32_t c = vcom[u], m;
#pragma omp atomic capture
    { m = a[c]; a[c] = 1; }
    if (!m) ++C;
  }
  return C;
}

/* Exclusive scan (_vector.hxx:1463-1524); sequential here, which gives the
 * same integers as the reference's per-thread-chunk version. */
static uint64_t exclusive_scan_u32(uint32_t* a, uint64_t N) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < N; ++i) { uint32_t t = a[i]; a[i] = (uint32_t)acc; acc += t; }
  return acc;
}
static uint64_t exclusive_scan_u64(uint64_t* a, uint64_t N) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < N; ++i) { uint64_t t = a[i]; a[i] = acc; acc += t; }
  return acc;
}

/* louvainLookupCommunitiesOmpU (louvain.hxx:838-844) over all N entries. */
static void lookup(uint32_t* a, uint64_t N, const uint32_t* map) {
  int64_t n = (int64_t)N;
#pragma omp parallel for schedule(static, 2048)
  for (int64_t u = 0; u < n; ++u) a[u] = map[a[u]];
}

/* Renumber (louvain.hxx:940-945): cext holds existence flags of g->S
 * communities and becomes the old->new map; vcom is looked up over all of
 * its vcom_n entries.  Returns the community count. */
uint64_t oracle_renumber(uint32_t* vcom, uint64_t vcom_n, uint32_t* cext, const oracle_graph* g) {
  uint64_t CN = community_exists(cext, g, vcom);
  exclusive_scan_u32(cext, g->S);
  lookup(vcom, vcom_n, cext);
  return CN;
}

/* Super-graph scratch: gapped CSR with f64 weights (DiGraphCsr<K,None,W>). */
typedef struct {
  uint64_t* off;
  uint32_t* deg;
  uint32_t* keys;
  double* w;
} sgraph;

/* louvainCommunityVerticesOmpW (louvain.hxx:799-812) then
 * louvainAggregateOmpW (louvain.hxx:986-992).  vcom must be renumbered to
 * [0, CN).  coff/cedg: community -> member CSR (u32 offsets as in the
 * reference's DiGraphCsr<K,None,None,K>). */
static void aggregate(sgraph* z, uint32_t* coff, uint32_t* cdeg, uint32_t* cedg, table_t* tb, const oracle_graph* g,
                      const uint32_t* vcom, uint64_t CN) {
  int64_t S = (int64_t)g->S, C = (int64_t)CN;
  /* count members, scan, scatter with an atomic slot per community */
  memset(coff, 0, (CN + 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static, 2048)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
#pragma omp atomic
    ++coff[vcom[u]];
  }
  coff[CN] = (uint32_t)exclusive_scan_u32(coff, CN);
  memset(cdeg, 0, CN * sizeof(uint32_t));
#pragma omp parallel for schedule(static, 2048)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    uint32_t c = vcom[u], n;
#pragma omp atomic capture
    { n = cdeg[c]; ++cdeg[c]; }
    cedg[coff[c] + n] = (uint32_t)u;
  }
  /* louvainCommunityTotalDegreeOmpW (louvain.hxx:711-722) then scan */
  memset(z->off, 0, (CN + 1) * sizeof(uint64_t));
#pragma omp parallel for schedule(static, 2048)
  for (int64_t u = 0; u < S; ++u) {
    if (!g_has(g, (uint64_t)u)) continue;
    uint64_t d = g_deg(g, (uint64_t)u);
#pragma omp atomic
    z->off[vcom[u]] += d;
  }
  z->off[CN] = exclusive_scan_u64(z->off, CN);
  /* louvainAggregateEdgesOmpW (louvain.hxx:896-912): self-loops included */
  memset(z->deg, 0, CN * sizeof(uint32_t));
#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t c = 0; c < C; ++c) {
    table_t* t = &tb[omp_get_thread_num()];
    uint32_t n = coff[c + 1] - coff[c];
    if (n == 0) continue;
    tbl_clear(t);
    for (uint32_t j = coff[c]; j < coff[c + 1]; ++j) scan_communities(t, g, cedg[j], vcom, 1);
    for (uint64_t k = 0; k < t->n; ++k) {
      uint32_t d = t->list[k];
      uint64_t i = z->off[c] + z->deg[c]++;
      z->keys[i] = d;
      z->w[i] = t->val[d];
    }
  }
}

void oracle_aggregate(uint64_t* zoff, uint32_t* zdeg, uint32_t* zkeys, double* zw, uint32_t* coff, uint32_t* cedg,
                      const oracle_graph* g, const uint32_t* vcom, uint64_t CN) {
  int T = omp_get_max_threads();
  table_t* tb = tables_new(T, g->S);
  uint32_t* cdeg = (uint32_t*)malloc((CN ? CN : 1) * sizeof(uint32_t));
  sgraph z = {zoff, zdeg, zkeys, zw};
  aggregate(&z, coff, cdeg, cedg, tb, g, vcom, CN);
  free(cdeg);
  tables_free(tb, T);
}

/* ---- the pass driver: louvainInvokeOmp (louvain.hxx:1119-1221) ---- */
enum { APPROACH_STATIC = 0, APPROACH_NAIVE = 1, APPROACH_DS = 2, APPROACH_DF = 3 };

int oracle_louvain(int approach, const oracle_graph* x, const oracle_batch* del, const oracle_batch* ins,
                   const uint32_t* q, const double* qvtot, const double* qctot, const oracle_options* o,
                   uint32_t* out_memb, double* out_vtot, double* out_ctot, oracle_out* out) {
  if (approach < 0 || approach > 3) return -1;
  uint64_t S = x->S, X = x->off[S];
  int T = omp_get_max_threads();
  double R = o->resolution;
  int L = o->max_iterations, P = o->max_passes;
  int repeat = o->repeat > 0 ? o->repeat : 1;
  double M = oracle_edge_weight(x) / 2;
  uint64_t GN0 = g_order(x);
  size_t Sb = S ? S : 1;
  uint8_t* vaff = (uint8_t*)calloc(Sb, 1);
  uint8_t *dsv = NULL, *dsn = NULL, *dsc = NULL;
  if (approach == APPROACH_DS) {
    dsv = (uint8_t*)calloc(Sb, 1);
    dsn = (uint8_t*)calloc(Sb, 1);
    dsc = (uint8_t*)calloc(Sb, 1);
  }
  uint32_t* ucom = (uint32_t*)calloc(Sb, 4);
  uint32_t* vcom = (uint32_t*)calloc(Sb, 4);
  double* utot = (double*)calloc(Sb, 8);
  double* vtot = (double*)calloc(Sb, 8);
  double* ctot = (double*)calloc(Sb, 8);
  uint32_t* cext = (uint32_t*)calloc(Sb, 4);  /* cv.degrees: existence flags, then member counts */
  uint32_t* coff = (uint32_t*)calloc(Sb + 1, 4);
  uint32_t* cedg = (uint32_t*)calloc(Sb, 4);
  size_t Xb = X ? X : 1;
  sgraph y = {(uint64_t*)calloc(Sb + 1, 8), (uint32_t*)calloc(Sb, 4), (uint32_t*)calloc(Xb, 4), (double*)calloc(Xb, 8)};
  sgraph z = {(uint64_t*)calloc(Sb + 1, 8), (uint32_t*)calloc(Sb, 4), (uint32_t*)calloc(Xb, 4), (double*)calloc(Xb, 8)};
  table_t* tb = tables_new(T, S);
  int l = 0, p = 0;
  double tt = 0, tm = 0, ti = 0, tp = 0, tl = 0, ta = 0;
  for (int r = 0; r < repeat; ++r) {
    double E = o->tolerance;
    memset(vaff, 0, Sb);
    memset(ucom, 0, Sb * 4);
    memset(vcom, 0, Sb * 4);
    memset(utot, 0, Sb * 8);
    memset(vtot, 0, Sb * 8);
    memset(ctot, 0, Sb * 8);
    uint64_t ySpan = 0; /* span of the current super-graph */
    double t0 = omp_get_wtime();
    /* fi: initialisation (timed as initializationTime) */
    double a0 = omp_get_wtime();
    if (approach == APPROACH_STATIC) {
      vertex_weights(utot, x);
      initialize(ucom, ctot, x, utot);
    } else {
      memcpy(ucom, q, S * 4);
      memcpy(utot, qvtot, S * 8);
      memcpy(ctot, qctot, S * 8);
      oracle_update_weights(utot, ctot, del, ins, ucom);
    }
    double a1 = omp_get_wtime();
    ti += a1 - a0;
    /* fm: marking (timed as markingTime) */
    if (approach == APPROACH_DF) oracle_frontier(vaff, S, del, ins, ucom);
    else if (approach == APPROACH_DS) {
      delta_screening(dsv, dsn, dsc, tb, x, del, ins, ucom, utot, ctot, M, R);
      memcpy(vaff, dsv, S);
    } else memset(vaff, 1, S);
    double a2 = omp_get_wtime();
    tm += a2 - a1;
    double f0 = a2, f1 = f0;
    for (l = 0, p = 0; M > 0 && P > 0;) {
      if (p == 1) f1 = omp_get_wtime();
      int isFirst = p == 0;
      int m;
      oracle_graph yg = {ySpan, y.off, y.deg, y.keys, NULL, y.w, NULL};
      double b0 = omp_get_wtime();
      if (isFirst) m = local_move(ucom, ctot, vaff, tb, x, utot, M, R, L, E, approach == APPROACH_DS ? dsv : NULL);
      else m = local_move(vcom, ctot, vaff, tb, &yg, vtot, M, R, L, E, NULL);
      tl += omp_get_wtime() - b0;
      l += m > 1 ? m : 1;
      ++p;
      if (m <= 1 || p >= P) break;
      const oracle_graph* gcur = isFirst ? x : &yg;
      uint64_t GN = isFirst ? GN0 : ySpan;
      uint64_t CN = community_exists(cext, gcur, isFirst ? ucom : vcom);
      if ((double)CN / (double)GN >= o->aggregation_tolerance) break;
      exclusive_scan_u32(cext, gcur->S);
      if (isFirst) lookup(ucom, S, cext);
      else {
        lookup(vcom, S, cext);
        lookup(ucom, S, vcom);
      }
      double c0 = omp_get_wtime();
      aggregate(&z, coff, cext, cedg, tb, gcur, isFirst ? ucom : vcom, CN);
      ta += omp_get_wtime() - c0;
      sgraph sw = y; y = z; z = sw;
      ySpan = CN;
      memset(vtot, 0, CN * 8);
      memset(vaff, 1, CN);
      oracle_graph yn = {ySpan, y.off, y.deg, y.keys, NULL, y.w, NULL};
      vertex_weights(vtot, &yn);
      initialize(vcom, ctot, &yn, vtot);
      E /= o->tolerance_drop;
    }
    if (p > 1) lookup(ucom, S, vcom);
    if (p <= 1) f1 = omp_get_wtime();
    tp += f1 - f0;
    tt += omp_get_wtime() - t0;
  }
  uint64_t aff = 0;
  for (uint64_t u = 0; u < S; ++u) aff += vaff[u] == 1;
  memcpy(out_memb, ucom, S * 4);
  memcpy(out_vtot, utot, S * 8);
  memcpy(out_ctot, ctot, S * 8);
  out->iterations = l;
  out->passes = p;
  out->time = (float)(tt * 1000 / repeat);
  out->marking_time = (float)(tm * 1000 / repeat);
  out->initialization_time = (float)(ti * 1000 / repeat);
  out->first_pass_time = (float)(tp * 1000 / repeat);
  out->local_move_time = (float)(tl * 1000 / repeat);
  out->aggregation_time = (float)(ta * 1000 / repeat);
  out->affected_vertices = aff;
  tables_free(tb, T);
  free(vaff); free(dsv); free(dsn); free(dsc);
  free(ucom); free(vcom); free(utot); free(vtot); free(ctot);
  free(cext); free(coff); free(cedg);
  free(y.off); free(y.deg); free(y.keys); free(y.w);
  free(z.off); free(z.deg); free(z.keys); free(z.w);
  return 0;
}

void oracle_set_threads(int t) { omp_set_num_threads(t); }
int oracle_max_threads(void) { return omp_get_max_threads(); }
