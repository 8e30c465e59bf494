/* Seeded synthetic workload generators for group Steiner tree queries.
 *
 * These have no counterpart on the reference's live path: the reference reads
 * its graphs, groups and queries from files (.in / .g / <name><k>.csv, see
 * code/TrimCDP/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:75-286)
 * and its only dataset ships without edge files. They exist so that benchmark
 * and test inputs can be reproduced bit for bit from a seed. Host-only code
 * (libgstgen.so); no GPU is touched.
 */
#ifndef GST_GEN_H
#define GST_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_graph gg_graph;
typedef struct gg_groups gg_groups;

/* R-MAT over 2^scale vertices; edges are drawn until m_target distinct
 * undirected non-loop edges exist (probabilities a, b, c, d = 1-a-b-c). */
gg_graph* gg_rmat(int32_t scale, int64_t m_target, double a, double b, double c, uint64_t seed);
/* Chung-Lu: expected degree of vertex i proportional to (i + i0)^(-1/(beta-1)),
 * scaled to sum 2*m_target with the largest expected degree dmax; endpoint
 * pairs are drawn until m_target distinct undirected non-loop edges exist. */
gg_graph* gg_chung_lu(int64_t n, int64_t m_target, double beta, double dmax, uint64_t seed);
/* Undirected graph from an edge list (duplicates and loops dropped; the
 * first occurrence's weight is kept when w is not NULL). */
gg_graph* gg_from_edges(int64_t n, int64_t m, const int32_t* src, const int32_t* dst, const int32_t* w);
/* Keep the largest connected component, relabelling vertices 0..V'-1 in
 * increasing order of their old ids. */
void gg_keep_lcc(gg_graph* g);
/* Vertices of the largest connected component (ids unchanged, increasing)
 * written to out (may be NULL); returns their count. */
int64_t gg_lcc_vertices(const gg_graph* g, int32_t* out);
/* Weight every undirected edge uniformly in [lo, hi] (edge-list order). */
void gg_set_weights_uniform(gg_graph* g, int32_t lo, int32_t hi, uint64_t seed);
int64_t gg_num_vertices(const gg_graph* g);
int64_t gg_num_edges(const gg_graph* g); /* undirected */
void gg_get_edges(const gg_graph* g, int32_t* src, int32_t* dst, int32_t* w);
/* Symmetric CSR: beg[V+1], col[2E], w[2E], eid[2E] (undirected edge id of
 * each directed entry); rows sorted by neighbour id. w / eid may be NULL. */
void gg_get_csr(const gg_graph* g, int64_t* beg, int32_t* col, int32_t* w, int32_t* eid);
void gg_free(gg_graph* g);

/* Every vertex joins 1 + Poisson(lambda) distinct uniform groups of G. */
gg_groups* gg_groups_poisson(int64_t n, int32_t G, double lambda, uint64_t seed);
/* G groups with Zipf(s) sizes clipped to [smin, smax]; members are distinct
 * uniform vertices, stored sorted. */
gg_groups* gg_groups_zipf(int64_t n, int32_t G, double s, int32_t smin, int32_t smax, uint64_t seed);
/* The same with members drawn from pool[0 .. npool) (distinct vertex ids in
 * [0, n)) instead of from all n vertices; pool = NULL means all vertices. */
gg_groups* gg_groups_zipf_pool(int64_t n, const int32_t* pool, int64_t npool, int32_t G, double s,
                               int32_t smin, int32_t smax, uint64_t seed);
int64_t gg_groups_count(const gg_groups* g);
int64_t gg_groups_total(const gg_groups* g);
void gg_groups_get(const gg_groups* g, int64_t* off, int32_t* mem);
void gg_groups_free(gg_groups* g);

/* nq queries; query q has ks[q] distinct uniform non-empty group ids, written
 * consecutively to out_groups. Returns 0, or -1 if some k is infeasible. */
int32_t gg_queries(int32_t G, const int64_t* goff, int32_t nq, const int32_t* ks, uint64_t seed, int32_t* out_groups);
/* count uniform integers in [lo, hi]. */
void gg_uniform_ints(int64_t count, int32_t lo, int32_t hi, uint64_t seed, int32_t* out);

#ifdef __cplusplus
}
#endif

#endif
