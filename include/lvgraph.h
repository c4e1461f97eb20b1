/*
 * lvgraph.h -- host-side graph tools (liblvgraph.so, C++17 + OpenMP):
 * generators, batch generation / tidying / application and file readers.
 * They produce the inputs of the Louvain library (louvain_hip.h) in the
 * layout its boundary requires: CSR rows sorted by neighbour id, one entry
 * per (u, v); tidy batches sorted by (u, v), unique, both directions.
 *
 * Reference semantics kept:
 *   random deletions / insertions    inc/batch.hxx:100-140 (retry x5 on deletions)
 *   tidyBatchUpdateU                 inc/batch.hxx:154-209
 *   applyBatchUpdateU                inc/batch.hxx:219-226 (remove, then add)
 *   Matrix Market reader (1-indexed) inc/mtx.hxx:39-54, :152-188
 *   temporal "u v [t]" reader        inc/snap.hxx:41-53, :158-166 (vertices 1..rows)
 *   symmetrizeOmp                    inc/symmetrize.hxx:72-82
 * All random streams are counter-based (splitmix64 of seed and index), so a
 * result depends on the seed only, never on the thread count.
 */
#ifndef LVGRAPH_H
#define LVGRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lvg_graph lvg_graph;
typedef struct lvg_batch lvg_batch;

/* R-MAT (a, b, c, d = 1-a-b-c), 2^scale vertices, edge_factor * 2^scale
 * generated edges, symmetrised, de-duplicated, unit weights, self-loops
 * kept; ids 0-based (every id exists). */
lvg_graph* lvg_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed);
/* Planted partition: n vertices in `blocks` blocks with power-law sizes
 * (exponent), n * avg_degree / 2 undirected edges of which a fraction
 * p_inter has uniform endpoints, the rest stay inside the source's block;
 * no self-loops, unit weights.  one_indexed: ids 1..n (vertex 0 absent). */
lvg_graph* lvg_planted_partition(uint64_t n, uint32_t blocks, double exponent, double avg_degree, double p_inter,
                                 uint64_t seed, int one_indexed);
/* Matrix Market coordinate file (1-indexed; symmetric files add both
 * directions); vertices 1..rows exist. */
lvg_graph* lvg_read_mtx(const char* path);
int lvg_write_mtx(const lvg_graph* g, const char* path);
/* Temporal edge list "u v [t]": the first `lines` lines, vertices 1..rows,
 * symmetric adds both directions (readTemporalIfOmpW). */
lvg_graph* lvg_read_temporal(const char* path, uint64_t rows, uint64_t lines, int symmetric);
/* CSR from directed (u, v, w) triples (rows sorted, last write wins). */
lvg_graph* lvg_from_edges(uint64_t span, const uint32_t* u, const uint32_t* v, const float* w, uint64_t n,
                          const uint8_t* exists);
void lvg_free(lvg_graph* g);
uint64_t lvg_span(const lvg_graph* g);
uint64_t lvg_edges(const lvg_graph* g);
uint64_t lvg_order(const lvg_graph* g);
const uint64_t* lvg_offsets(const lvg_graph* g);
const uint32_t* lvg_keys(const lvg_graph* g);
const float* lvg_weights(const lvg_graph* g);
const uint8_t* lvg_exists(const lvg_graph* g); /* NULL when every id exists */

/* A random batch against g: n_del undirected deletions (a random edge of a
 * uniform vertex, retried up to 5 times on isolated vertices) and n_ins
 * undirected insertions (uniform vertex pairs, weight w), both directions,
 * then tidied against g. */
lvg_batch* lvg_random_batch(const lvg_graph* g, uint64_t n_del, uint64_t n_ins, float w, uint64_t seed);
/* Tidy arbitrary deletions / insertions against g. */
lvg_batch* lvg_tidy_batch(const lvg_graph* g, const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                          const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni);
void lvg_batch_free(lvg_batch* b);
uint64_t lvg_batch_deletions(const lvg_batch* b);
uint64_t lvg_batch_insertions(const lvg_batch* b);
/* which: 0 deletions, 1 insertions */
const uint32_t* lvg_batch_u(const lvg_batch* b, int which);
const uint32_t* lvg_batch_v(const lvg_batch* b, int which);
const float* lvg_batch_w(const lvg_batch* b, int which);
/* New graph = g minus deletions plus insertions (applyBatchUpdateU). */
lvg_graph* lvg_apply_batch(const lvg_graph* g, const lvg_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* LVGRAPH_H */
