/* gst.h - MI355X (gfx950) group Steiner tree solver: TrimCDP and D-TrimCDP.
 *
 * C ABI of libgst.so. Plain pointers and sizes only; every entry point
 * returns a status code and never exits the process. Host arrays passed in
 * are read during the call only and stay owned by the caller.
 *
 * The reference has no library boundary: each solver is a C++ free function
 * linked into its own executable. The entry points below replace these
 * reference interfaces (paths relative to the reference root):
 *
 *   gst_graph_create   the CSR the reference solvers read:
 *                      CSR_graph / toCSR (code/TrimCDP/include/graph_v_of_v_idealID/
 *                      csr_graph.hpp:14-31, 39-128) and the int64 *_beg_pos/_csr/_weight.bin
 *                      arrays of the work-balanced solvers (code/TrimCDP-WB/include/
 *                      graph.hpp:7-118, code/TrimCDP-WB/src/GSTnonHop.cu:167-188)
 *   gst_groups_set     group_graph, built by read_Group (code/TrimCDP/include/
 *                      graph_hash_of_mixed_weighted_read_for_GSTP.h:190-239)
 *   gst_solve          DP_GPU (code/TrimCDP/include/DPQ.cuh:67, src/DPQ.cu:292) when D < 0;
 *                      DP_gpu (code/D-TrimCDP/include/DPQ.cuh:66, src/DPQ.cu:147) when D >= 0;
 *                      also graph_v_of_v_idealID_PrunedDPPlusPlus
 *                      (code/PrunedDP++/include/CPUNONHOP.h:496) and HOP_cpu
 *                      (code/D-PrunedDP++/include/CPUHOP.h:107), whose tree output it adds
 *   gst_solve_batch    the per-query harness loops exp_GPU_nonHop / exp_GPU_Hop
 *                      (code/TrimCDP/include/exp_GPU_nonHop.h:40-144,
 *                      code/D-TrimCDP/include/exp_GPU_Hop.h:40-126) and the query loop of
 *                      the work-balanced main() (code/TrimCDP-WB/src/GSTnonHop.cu:247-497)
 *   gst_backward       no reference counterpart (the reference is not differentiable);
 *                      backs torch.autograd.Function of the Python wrapper
 *   gst_graph_destroy  the cudaFree sequences (code/TrimCDP/src/DPQ.cu:644-654)
 *
 * Semantics kept from the reference: bit i of a group set is query_groups[i]
 * (code/TrimCDP/src/DPQ.cu:40-52); a vertex is the root of T(v,p) = 0 for every
 * p contained in the query groups it belongs to; the optimum is
 * min over v of T(v, full) (plain) or min over v, d <= D of T(v, full, d) (hop).
 * Costs are exact int32. Weights must be positive.
 */
#ifndef GST_H
#define GST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GST_OK 0
#define GST_ERR_INVALID 1    /* bad argument (sizes, ids, weights <= 0, k out of range) */
#define GST_ERR_OOM 2        /* device allocation failed or a query exceeds the HBM budget */
#define GST_ERR_DEVICE 3     /* HIP runtime error */
#define GST_ERR_CAPACITY 4   /* caller's output buffer too small; sizes were written */
#define GST_ERR_INTERNAL 5   /* consistency check failed inside the solver */

/* per-query status in gst_solve_batch's out_status */
#define GST_Q_OK 0
#define GST_Q_INFEASIBLE 1   /* no tree covers all query groups */
#define GST_Q_RANGE 2        /* no tree of cost below 2^29 was found, and some path or tree
                                cost reached 2^29 while searching: the optimum, if any, lies
                                beyond the int32 cost range; cost = the infeasible cost */
#define GST_Q_NO_TREE 3      /* the cost is exact, but the tree has more edges than the
                                recovery buffer holds (after one retry with room for
                                min(cost, 4V) edges); no edges are returned */

#define GST_MAX_K 7           /* up to 128 group subsets per query */
#define GST_MAX_D 15          /* hop bound */

typedef struct gst_graph gst_graph;

typedef struct gst_stats {
  int64_t queries;
  int64_t batches;            /* admission rounds: lockstep batches, plus (rolling admission)
                                 host checks that admitted queries */
  int64_t sssp_iterations;    /* frontier iterations of the one-label phase */
  int64_t dp_iterations;      /* frontier iterations of the DP phase */
  int64_t sssp_items;         /* one-label states expanded */
  int64_t sssp_relax;         /* one-label edge visits (each relaxes all k labels) */
  int64_t sssp_updates;       /* one-label atomic updates */
  int64_t sssp_items_unbounded; /* one-label items seen before the query had an upper bound */
  int64_t states_expanded;    /* F: DP states whose grow/merge ran */
  int64_t grow_relax;         /* Rg: DP edge relaxations */
  int64_t merge_relax;        /* Rm: DP merge candidates evaluated */
  int64_t improvements;       /* U: successful DP table decreases */
  int64_t lb_terms;           /* Lb: one-label terms of the lower bounds of improved states
                                 (|full \ P| per successful DP table decrease) */
  int64_t items_processed;    /* DP worklist items (chunks included) */
  int64_t max_frontier;       /* largest DP worklist */
  int64_t bytes_alg;          /* algorithmic bytes of the DP expand kernel:
                                 16 F + 12 Rg + 8 Rm + 12 U + 4 Lb */
  double seconds_total;       /* host wall clock of the call */
  double expand_ms;           /* DP expand kernel time (profiling mode only) */
  int64_t expand_launches;    /* DP expand launches (profiling mode only) */
  double sssp_expand_ms;      /* one-label expand kernel time (profiling mode only) */
  double phase_ms[3];         /* device time: one-label phase, bounds + seeds, DP phase */
  int64_t sssp_expand_launches; /* one-label expand launches (profiling mode only) */
  int64_t reruns;             /* batches re-run with full-size lists after a list overflowed */
  int64_t ub_rows;            /* one-label rows read for DP upper bounds (32 B each) */
  int64_t bytes_layout;       /* bytes the DP expand kernel's own layout moves for the same
                                 work: 24 F + 16 Rg + 12 Rm + 16 U + 32 ub_rows (8-byte
                                 state words, 64-bit items) */
  int64_t backpointers_stored; /* 1 if the last batch stored backpointers */
  int64_t bp_fallbacks;       /* tree-recovery steps whose stored backpointer did not match
                                 the final table (a racing write) and were searched instead */
  int64_t rolling_queries;    /* queries solved under rolling admission (GST_OPT_ROLLING) */
  int64_t max_resident;       /* most queries holding DP tables at once under rolling admission */
} gst_stats;

/* options for gst_set_option */
#define GST_OPT_CHUNK 1            /* edges per virtual-split chunk (power of two, 64..65536;
                                      0 = automatic, the default: see gst_autotune) */
#define GST_OPT_GRID 2             /* workgroups of the expand kernels (0 = the autotuner's) */
#define GST_OPT_ITERS_PER_SYNC 3   /* max graph-captured iterations between host checks */
#define GST_OPT_HOP_PRUNE 4        /* 1 = lower-bound pruning in the hop DP (default 1) */
#define GST_OPT_PROFILE 5          /* 1 = time every expand launch with HIP events */
#define GST_OPT_INFEASIBLE_COST 6  /* cost written for infeasible queries (default INT32_MAX;
                                      100000 reproduces the reference's inf) */
#define GST_OPT_MAX_TREE_EDGES 7   /* per-query edge capacity of the backtracking kernel */
#define GST_OPT_USE_GRAPHS 8       /* 1 = capture iteration steps in HIP graphs (default 1) */
#define GST_OPT_SSSP_DELTA 10      /* one-label phase in bands of label value: first band = value,
                                      doubling up to the radius (hop-bounded: until every label
                                      is final); 0 = unordered; -1 (default) = 16 */
#define GST_OPT_RADIUS_PCT 12      /* one-label radius = upper bound * value / 100 + 2; distances
                                      beyond it are clamped lower bounds. 0 (default) =
                                      automatic: 85 for queries of up to 6 groups, 95 for 7 */
#define GST_OPT_MAX_BATCH 14       /* cap on queries per lockstep batch (0 = HBM budget only) */
#define GST_OPT_DP_DELTA_DIV 11    /* plain DP in A* bands of cost + completion bound: first band
                                      = upper bound / value, doubling; 0 = unordered; -1 (default)
                                      = 256 for batches of 32+ plain queries, else unordered */
#define GST_OPT_SSSP_SLOTS 15      /* queries per one-label pass inside a batch (0 = all at once) */
#define GST_OPT_LIST_PERMILLE 16   /* worklist capacity in permille of the batch's states, without
                                      the usual floors (0 = default: 4%, at least one item per
                                      vertex and query); a list that overflows makes the batch
                                      re-run with full-size lists (gst_stats.reruns) */
#define GST_OPT_BACKPOINTERS 19    /* tree recovery: 1 = from backpointers stored during the DP
                                      (one int32 per state, written after every table decrease,
                                      checked against the final table when read), 0 = recomputed
                                      from the cost table alone, -1 (default) = stored when the
                                      whole call fits one admission batch with them, else
                                      recomputed (memory is tight) */
#define GST_OPT_TRACE 20           /* 1 = with GST_OPT_PROFILE, record one gst_trace_row per
                                      expand launch (read with gst_trace_get) */
#define GST_OPT_ROLLING 21         /* 1 (default) = rolling admission for the queries of a mode
                                      whose tables do not fit the HBM budget at once: the one-label
                                      phase runs for all of them first, then DP table regions are
                                      handed to waiting queries as soon as a query converges;
                                      0 = lockstep batches (each waits for its slowest query) */
/* Ablation switches (the reference's TrimCDP-WB / D-TrimCDP-WB ablation variants, README.md:
 * 181-302); results are the same, only the speed changes: */
#define GST_OPT_LANE_APPENDS 17    /* 1 = every lane appends its frontier items with its own
                                      global atomic (no LDS aggregation; variants without the
                                      shared-memory prefix scan); default 0 */
#define GST_OPT_PER_ITEM_EDGES 18  /* 1 = each lane walks its own item's edges in order (per-thread
                                      contiguous chunks, the "no_global_memory_coalescing"
                                      variants) instead of the workgroup's edges balanced and
                                      lane-strided by an LDS prefix scan; default 0 */

/* Creates a graph on the current HIP device from a symmetric CSR:
 * beg_pos[V+1], csr[E], weight[E] with E = beg_pos[V]; every undirected edge
 * appears in both rows. index_bits = 64 reads int64 arrays (the reference's
 * .bin layout), 32 reads int32 arrays. Weights must be >= 1. */
int gst_graph_create(const void* beg_pos, const void* csr, const void* weight, int64_t V,
                     int64_t E, int32_t index_bits, gst_graph** out);
/* Maps each CSR entry to an undirected edge id in [0, num_undirected); used
 * for tree edge ids and by gst_backward. Without it the id is the CSR index. */
int gst_graph_set_edge_ids(gst_graph* g, const int32_t* csr_edge_id, int64_t num_undirected);
/* Replaces the weights (CSR order, int32, >= 1). When few entries change,
 * only their rows are re-sorted and re-uploaded. */
int gst_graph_set_weights(gst_graph* g, const int32_t* weight);
/* Sets weight[csr_index[i]] = w[i] for i < n (the caller's CSR indices; a
 * repeated index takes its last value). Both entries of an undirected edge
 * must change together (GST_ERR_INVALID otherwise, nothing changed). For
 * re-solves under small perturbations (finite differences of the
 * differentiable cost); no counterpart in the reference, whose graphs are
 * read once per run. */
int gst_graph_update_weights(gst_graph* g, const int64_t* csr_index, const int32_t* w, int64_t n);
/* Vertex groups: members of group i are group_members[group_offsets[i] ..
 * group_offsets[i+1]); group ids are 0-based (".g" line "gN" is id N-1). */
int gst_groups_set(gst_graph* g, const int64_t* group_offsets, const int32_t* group_members,
                   int32_t G);
int gst_set_option(gst_graph* g, int32_t option, int64_t value);

/* One query: k (1..GST_MAX_K) distinct group ids (a repeated id, k outside
 * 1..GST_MAX_K, an id outside 0..G-1 or D > GST_MAX_D is GST_ERR_INVALID);
 * D < 0 = plain TrimCDP, D >= 0 = hop-bounded D-TrimCDP. An empty group makes
 * the query infeasible. Writes the optimal cost, the tree as vertex pairs
 * (out_edge_uv[2*i], out_edge_uv[2*i+1]) up to edge_cap edges, and the edge
 * count. Infeasible queries return GST_OK with *out_status = GST_Q_INFEASIBLE
 * and no edges. out_edge_uv / out_stats / out_status may be NULL. */
int gst_solve(gst_graph* g, const int32_t* query_groups, int32_t k, int32_t D, int32_t* out_cost,
              int32_t* out_edge_uv, int32_t edge_cap, int32_t* out_num_edges,
              int32_t* out_status, gst_stats* out_stats);

/* nq queries in one call: query q is q_groups[q_offsets[q] .. q_offsets[q+1]),
 * with hop bound q_D[q] (q_D may be NULL = all plain). Queries run
 * concurrently within hbm_budget_bytes (0 = 80% of the free HBM plus what the
 * graph's per-batch buffers already hold): when all queries of a mode fit, as
 * one lockstep batch; otherwise under rolling admission (GST_OPT_ROLLING): the
 * one-label phase of all of them first, then DP table regions handed to
 * waiting queries as queries converge (after a worklist overflow the resident
 * queries restart, half as many at a time). A lockstep batch whose worklists
 * overflow is solved again in sub-batches with worklists as large as the state
 * count. Trees go to
 * out_edge_uv (pairs) / out_edge_ids
 * (undirected ids, may be NULL), query q at [out_edge_offsets[q],
 * out_edge_offsets[q+1]). If edge_cap is too small GST_ERR_CAPACITY is
 * returned after costs and offsets are written. */
int gst_solve_batch(gst_graph* g, const int64_t* q_offsets, const int32_t* q_groups,
                    const int32_t* q_D, int32_t nq, int64_t hbm_budget_bytes, int32_t* out_costs,
                    int32_t* out_status, int64_t* out_edge_offsets, int32_t* out_edge_uv,
                    int32_t* out_edge_ids, int64_t edge_cap, gst_stats* out_stats);

/* gst_solve without device allocation: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. Every per-query device buffer
 * (state table, stamps, one-label rows, group masks, route tables,
 * worklists, counters of the call, tree output) is carved from a file-scope
 * device array of GST_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; the graph and its groups stay where gst_graph_create put them. A
 * query that needs more scratch is refused with GST_ERR_OOM (the message
 * gives both sizes; gst_noalloc_bytes tells in advance). Trees longer than
 * GST_OPT_MAX_TREE_EDGES are not recovered again (status GST_Q_NO_TREE, exact
 * cost); stored backpointers and captured graphs are not used. All calls on
 * one GPU share its array: calls that overlap on one GPU (from several
 * threads or streams) race and must be serialised by the caller. */
#define GST_NOALLOC_SCRATCH_BYTES (256ull << 20)
int gst_solve_noalloc(gst_graph* g, const int32_t* query_groups, int32_t k, int32_t D,
                      int32_t* out_cost, int32_t* out_edge_uv, int32_t edge_cap,
                      int32_t* out_num_edges, int32_t* out_status, gst_stats* out_stats);
/* Scratch bytes gst_solve_noalloc needs for one query of k groups and hop
 * bound D on this graph (with the one-query settings in force). */
int64_t gst_noalloc_bytes(gst_graph* g, int32_t k, int32_t D);

/* Backward of cost_q with respect to the undirected edge weights: device
 * pointers only. grad_w[e] += sum over queries q of grad_out[q] for every tree
 * edge e of query q (tree q = edge_ids[edge_offsets[q] .. edge_offsets[q+1])).
 * Runs on the graph's stream and waits for it. */
int gst_backward(gst_graph* g, int32_t nq, const int64_t* d_edge_offsets,
                 const int32_t* d_edge_ids, const float* d_grad_out, float* d_grad_w,
                 int64_t num_undirected);

/* Autotuner. The expand kernels' settings (virtual-split chunk, workgroup
 * count, first DP band divisor, first one-label band width) are chosen per
 * size class of a call: one query, 2..31 queries, 32 or more. By default a
 * model of the class picks them (chunk 256 / 1024 / 4096 edges, 8 / 8 / 4
 * workgroups per CU, the band rules of GST_OPT_DP_DELTA_DIV and
 * GST_OPT_SSSP_DELTA); options set with gst_set_option always win.
 * gst_autotune measures instead: on the given sample of queries it times
 * candidate values of every setting the caller has not fixed, one setting at
 * a time (coordinate descent, best of two passes per candidate, stopping
 * after about max_seconds), and keeps the fastest for that size class on
 * this graph; later calls of the class use it. batch_size = 1 tunes the
 * one-query class (the sample is solved one query per call); 0 tunes the
 * class of nq (the sample is solved in one call). max_seconds <= 0 only
 * reports the settings in force for the class. Replaces the reference's
 * compile-time constants: the degree bins SML_MID / MID_LRG and thread/block
 * counts (code/TrimCDP-WB/include/util.h:19-20, 31-33) and the virtual-split
 * threshold and slice (code/TrimCDP-WB/include/graph.hpp:141, 155). */
typedef struct gst_tuning {
  int32_t chunk;            /* edges per virtual-split chunk in force */
  int32_t grid;             /* workgroups of the expand kernels in force */
  int32_t dp_div;           /* first DP band divisor (-1 = the default rule, 0 = unordered) */
  int32_t sssp_band;        /* first one-label band (-1 = the default rule, 0 = unordered) */
  int32_t size_class;       /* 0: one query, 1: 2..31, 2: 32 or more */
  int32_t trials;           /* candidate settings timed */
  double seconds_before;    /* one pass over the sample with the settings before tuning */
  double seconds_after;     /* ... with the chosen settings */
} gst_tuning;
int gst_autotune(gst_graph* g, const int64_t* q_offsets, const int32_t* q_groups,
                 const int32_t* q_D, int32_t nq, int32_t batch_size, double max_seconds,
                 gst_tuning* out);

/* Per-iteration trace of the last gst_solve_batch call made with
 * GST_OPT_PROFILE and GST_OPT_TRACE set: one row per expand launch. Copies
 * min(cap, length) rows and sets *n to the length (GST_ERR_CAPACITY when cap
 * is smaller). No reference counterpart (its harness records one time per
 * query, code/TrimCDP/include/exp_GPU_nonHop.h:128). */
typedef struct gst_trace_row {
  int32_t phase;          /* 0 one-label, 1 plain DP, 2 hop DP */
  int32_t batch;          /* admission batch of the call, from 1 */
  int64_t items;          /* worklist items the launch consumed */
  int32_t active_slots;   /* queries of the batch that expanded at least one item */
  int32_t slots;          /* queries in the batch */
  float ms;               /* the launch's device time */
  int32_t pad;
} gst_trace_row;
int gst_trace_get(gst_graph* g, gst_trace_row* rows, int64_t cap, int64_t* n);

/* Number of device bytes one query of k groups and hop bound D occupies
 * (with the settings in force for the last call). */
int64_t gst_query_footprint(gst_graph* g, int32_t k, int32_t D);
/* Thread-local description of the last error. */
const char* gst_last_error(void);
void gst_graph_destroy(gst_graph* g);
/* Library version, for the loader check. */
int32_t gst_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GST_H */
