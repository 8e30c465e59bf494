// Shared host/device definitions of the gfx950 TrimCDP / D-TrimCDP engine.
#pragma once

#ifdef GST_CPU_EMU  // CPU emulation build for fault finding (emu/hip_cpu_emu.h)
#include "emu/hip_cpu_emu.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace gst {

// "Unreachable / no tree yet". Sums of three such values stay below 2^31,
// so bound checks of the form a + b + c never overflow (the reference's
// inf = 100000, code/TrimCDP/include/DPQ.cuh:50, is not safe for that).
constexpr int32_t INF = 1 << 29;
constexpr int BLOCK = 256;          // threads per workgroup of the expand kernels
constexpr int WAVE = 64;            // gfx950 wavefront
// ints per subset X in the route tables: 8x8 routes, 8 row minima, and at
// RLBMIN the smallest completion bound lb(v, X) over all vertices v
constexpr int RSTRIDE = 73;
constexpr int RLBMIN = 72;
constexpr int RTABLE = 128 * RSTRIDE;  // route ints per query slot (2^7 subsets)
constexpr int CSTAMP = 7;              // word of a distance row holding the one-label stamp
constexpr int CROW = 8;             // one-label distances per vertex (k <= 7, padded to 8)
constexpr int VV = 64;              // 8x8 group-to-group distances per slot

// Worklist item: one frontier state (or one chunk of a high-degree state).
//   bits  0..23  vertex v        bits 24..31  group set p (one-label phase: group index)
//   bits 32..35  hop layer d     bits 36..51  chunk index of v's CSR row
//   bits 52..63  query slot
// The DP frontier alternates two lists: vertex items (p = 0, chunk 0: vertex v
// at depth d has pending states, listed in its pending mask) and state items
// (one per pending state and chunk), made from the vertex items by
// k_dp_unfold at the start of every iteration with the states of one vertex
// and chunk next to each other.
__host__ __device__ inline uint64_t mk_item(uint32_t v, uint32_t p, uint32_t d, uint32_t ch,
                                            uint32_t s) {
  return (uint64_t)v | ((uint64_t)p << 24) | ((uint64_t)d << 32) | ((uint64_t)ch << 36) |
         ((uint64_t)s << 52);
}
__host__ __device__ inline uint32_t it_v(uint64_t x) { return (uint32_t)(x & 0xFFFFFFu); }
__host__ __device__ inline uint32_t it_p(uint64_t x) { return (uint32_t)((x >> 24) & 0xFFu); }
__host__ __device__ inline uint32_t it_d(uint64_t x) { return (uint32_t)((x >> 32) & 0xFu); }
__host__ __device__ inline uint32_t it_ch(uint64_t x) { return (uint32_t)((x >> 36) & 0xFFFFu); }
__host__ __device__ inline uint32_t it_s(uint64_t x) { return (uint32_t)(x >> 52); }
constexpr uint32_t MAX_V = 1u << 24;
constexpr uint32_t MAX_SLOTS = 1u << 12;
constexpr uint32_t MAX_CHUNKS = 1u << 16;

// Best-solution word of a query: (cost << 32) | witness, updated with a
// 64-bit atomicMin so the cost and a state that reproduces it move together.
//   witness bit 31: 0 = the full state T(v, full[, d]) itself,
//                   1 = T(v, p) plus shortest paths to the groups outside p
//   bits 24..30: p (one-label witness) or d (full witness), bits 0..23: v
__host__ __device__ inline uint32_t wit_full(uint32_t v, uint32_t d) { return (d << 24) | v; }
__host__ __device__ inline uint32_t wit_onelabel(uint32_t v, uint32_t p) {
  return 0x80000000u | (p << 24) | v;
}

// Per-query-slot layout; all offsets count elements of their arena.
struct SlotDesc {
  int64_t t_off;   // state words, vertex-major: entry v * nrows + d * nsub + p holds
                   // {cost, completion bound} as two int32, so all states of one vertex
                   // (every merge partner) share a few cache lines; row 0 (p = 0) is
                   // no state and never read. Stored backpointers use the same indices.
  int64_t e_off;   // one int32 per vertex: the one-label phase's far-list key
  int64_t c_off;   // one-label distances c[v * 8 + i]
  int64_t r_off;   // route tables [p][72]
  int32_t k, full, D, nsub;  // D < 0 = plain problem
  int32_t vv_off;  // group distance matrix [8][8] in the vv array
  int32_t nrows;   // rows per vertex: nsub (plain) or nsub * (D + 1) (hop)
  int32_t mw;      // mask words per (vertex, depth): nsub / 32 rounded up, doubled (fin)
  int32_t fin;     // 1: mask words come in pairs {pending i, finite i}: bit p % 32 of
                   // finite word p / 32 says T(v, p, d) < INF (plain slots, and
                   // hop-bounded slots of k <= 5 groups)
  int64_t m_off;   // DP masks: word mask_ix(S, v, d) + p / 32, bit p % 32, says state
                   // (v, p, d) waits for the next iteration; bit 0 of word 0 (p = 0 is no
                   // state) says the vertex item (v, d) is queued. The words of all depths
                   // of a vertex are adjacent.
};

// Word index (in the mask arena) of the masks of (v, d).
__host__ __device__ inline int64_t mask_ix(const SlotDesc& S, int64_t v, int64_t d) {
  return S.m_off + (v * (S.D < 0 ? 1 : S.D + 1) + d) * S.mw;
}

// Index (in int2 entries of T, or int32 entries of BP) of row `row` of vertex v.
__host__ __device__ inline int64_t t_ix(const SlotDesc& S, int64_t row, int64_t v) {
  return S.t_off + v * S.nrows + row;
}

// Counters accumulated on device (uint64 each).
enum Ctr {
  C_SSSP_ITEMS = 0,
  C_SSSP_RELAX,
  C_F,
  C_RG,
  C_RM,
  C_U,
  C_LB,
  C_ITEMS,
  C_MAX_FRONTIER,
  C_DP_ITERS,
  C_SSSP_ITERS,
  C_OVERFLOW,
  C_SSSP_UPD,
  C_SSSP_NOUB,
  C_UBROW,
  C_BP_MISS,
  C_ACTIVE,  // slots that expanded items in the last iteration (trace mode)
  C_NUM
};

// Per-slot frontier ordering (bucketed label correcting): a state is
// expanded only when its cost is below the slot's threshold `thr`; costlier
// states are carried to the next iteration. When an iteration expands
// nothing for a slot, thr jumps to the cheapest carried cost + delta.
// R bounds the one-label phase: distances >= R are clamped to R afterwards
// (still valid lower bounds) and never enter upper bounds.
struct Sched {
  int32_t thr, mindef, active, delta;
  int32_t R;
  int32_t prev;  // banded one-label phase: the previous band's end
  // 1 once a relaxation of a finite value produced a sum >= INF (a path or
  // tree cost the int32 cost range cannot hold): such a query that ends
  // without a solution is reported as out of range, not as infeasible
  int32_t sat;
  // Rolling admission (a call whose tables do not all fit at once): 1 while
  // the slot's query is in the DP phase, 2 once it has converged; `last` = 1
  // if the slot expanded an item in the last iteration (k_advance), `hasfar`
  // = 1 if it still had live far-list states at the last band refill.
  int32_t phase;
  int32_t last;
  int32_t hasfar;
  int32_t pad[2];
};

struct DevCtx {
  const int32_t* rowp;  // [V+1]
  const int32_t* col;   // [E]
  const int32_t* wt;    // [E]
  const int32_t* eid;   // [E] undirected edge id (may be null = CSR index)
  int32_t V, Vp;
  int32_t chunk_shift;  // virtual-split chunk = 1 << chunk_shift edges
  int32_t hop_prune;
  int32_t* T;
  int32_t* E;
  uint32_t* M;                   // DP pending masks (SlotDesc::m_off)
  uint32_t* GM;                  // query-group mask of each vertex [slot][Vp]
  int32_t* C;
  int32_t* R;
  int32_t* vv;
  uint64_t* bestw;
  const SlotDesc* slots;
  int32_t* iter;                 // current iteration number (stamps use iter + 1)
  unsigned long long* ctr;       // Ctr counters
  uint32_t cap;                  // worklist capacity (items)
  Sched* sched;                  // [slots]
  const int32_t* nslots;         // number of slots of the current batch (device)
  // first DP band width = max(1, best / dp_div), for a hop-bounded slot
  // max(1, smallest key / dp_div); doubling per band; 0 = unordered
  int32_t dp_div;
  int32_t radius_pct;            // one-label radius = best * radius_pct / 100 + 2
  // Near/far schedule of the banded DP (far_mode = 1): states whose key
  // cost + bound is at or past their slot's band go to the far list
  // far[*far_sel] (count far_cnt[*far_sel]) and come back only when the near
  // list has drained (k_far_*).
  int32_t far_mode;
  // Banded one-label phase (sssp_far = 1): labels at or past their slot's
  // band end are not relaxed yet; vertices with work past the band wait in
  // the far list (deduplicated per band with key *band_key in stamp row 0 of
  // E, which the DP never uses) and return, flagged as continuations, when
  // the near list has drained and the bands have grown.
  int32_t sssp_far;
  int32_t* band_key;
  // ablation switches: per-lane global appends, per-item edge walks
  int32_t lane_appends;
  int32_t per_item_edges;
  uint64_t* far[2];
  uint32_t* far_cnt;             // [2]
  int32_t* far_sel;
  // Stored backpointers (null = trees are recomputed from the cost table):
  // one word per state, indexed like E, written with a plain store after every
  // successful table decrease (bp_* below). A race can leave the word of an
  // older write; the tree recovery checks every word against the final table
  // and searches for a predecessor when the check fails.
  int32_t* BP;
  int32_t trace;  // 1 = count the slots active in each iteration (C_ACTIVE)
};

// Backpointer words: grow from vertex u (bit 31 clear), merge of subset p1 at
// depth d1 with its complement in the state's subset at depth d2 (bit 31 set),
// or a base state (all ones).
constexpr uint32_t BP_BASE = 0xFFFFFFFFu;
__host__ __device__ inline uint32_t bp_grow(uint32_t u) { return u; }
__host__ __device__ inline uint32_t bp_merge(uint32_t p1, uint32_t d1, uint32_t d2) {
  return 0x80000000u | p1 | (d1 << 7) | (d2 << 11);
}

struct WL {
  const uint64_t* in;
  const uint32_t* in_cnt;
  uint64_t* out;
  uint32_t* out_cnt;
};

// Host-side launchers (gst_kernels.hip).
struct Task {  // one (slot, query group) pair for member-driven kernels
  int32_t slot, bit, group, pad;
};

void launch_fill_i32(int32_t* p, int64_t n, int32_t val, hipStream_t st);
void launch_fill_u64(uint64_t* p, int64_t n, uint64_t val, hipStream_t st);
void launch_groups_setup(const DevCtx& c, const Task* tasks, int ntasks, const int64_t* goff,
                         const int32_t* gmem, uint64_t* wl, uint32_t* cnt, hipStream_t st);
void launch_sssp_expand(const DevCtx& c, const WL& w, int grid, hipStream_t st);
void launch_dp_expand(const DevCtx& c, const WL& w, int grid, bool hop, hipStream_t st);
void launch_advance(const DevCtx& c, uint32_t* cnt_consumed, const uint32_t* cnt_next, bool dp,
                    uint32_t* cnt_states, hipStream_t st);
// Vertex items of the DP frontier -> one state item per pending state and chunk.
void launch_dp_unfold(const DevCtx& c, const uint64_t* vin, const uint32_t* vin_cnt,
                      uint64_t* sout, uint32_t* sout_cnt, int grid, hipStream_t st);
void launch_vv(const DevCtx& c, const Task* tasks, int ntasks, const int64_t* goff,
               const int32_t* gmem, hipStream_t st);
void launch_route(const DevCtx& c, int nslots, hipStream_t st);
void launch_dp_seed(const DevCtx& c, const Task* tasks, int ntasks, const int64_t* goff,
                    const int32_t* gmem, uint64_t* wl, uint32_t* cnt, hipStream_t st);
// Tree recovery for nslots slots: slot sel[i] (or i when sel is null) writes
// its edges to row i of the outputs (max_edges each).
void launch_backtrack(const DevCtx& c, int nslots, int max_edges, int32_t* out_uv,
                      int32_t* out_eid, int32_t* out_n, int32_t* out_err, const int32_t* sel,
                      hipStream_t st);
// Copies the first dp ints of each of `rows` rows (source pitch sp) to rows of pitch dp.
void launch_pack_rows(const int32_t* src, int64_t sp, int32_t* dst, int64_t dp, int rows,
                      hipStream_t st);
void launch_backward(int nq, const int64_t* off, const int32_t* ids, const float* gout,
                     float* gw, int64_t nund, hipStream_t st);
void launch_init_slot_tables(const DevCtx& c, int nslots, int64_t c_total, hipStream_t st);
void launch_clamp(const DevCtx& c, int nslots, hipStream_t st);
// With `sel` (device, n entries) only the slots listed there; else slots 0..n-1.
void launch_lb_table(const DevCtx& c, int n, hipStream_t st, const int32_t* sel = nullptr);
void launch_dp_sched_init(const DevCtx& c, int n, hipStream_t st, const int32_t* sel = nullptr);
// Rolling admission: lists the DP-phase slots whose query has converged in
// out[0 .. *out_n) and marks them converged. mode 0: slots that expanded
// nothing in the last iteration (unordered DP: nothing can reach them any
// more); 1: slots with no live far-list state at the band refill just made
// (banded DP, run only when the near list had drained); 2: every DP-phase
// slot (both lists are empty).
void launch_collect_done(const DevCtx& c, int mode, int32_t* out, uint32_t* out_n,
                         hipStream_t st);
// selective: only slots with no near-list work left (nothing expanded in the
// last iteration) move to their next band.
void launch_far_refill(const DevCtx& c, int nslots, uint32_t n_far, uint64_t* near,
                       uint32_t* near_cnt, hipStream_t st, bool selective = false);
// *out += the DP-phase slots that expanded nothing in the last iteration.
void launch_idle_count(const DevCtx& c, uint32_t* out, hipStream_t st);
void launch_sssp_far_refill(const DevCtx& c, uint32_t n_far, uint64_t* near, uint32_t* near_cnt,
                            hipStream_t st);
void launch_band_start(const DevCtx& c, hipStream_t st);
// This GPU's address of the allocation-free path's scratch array (resolved
// once per device with hipGetSymbolAddress; the current device must be set).
hipError_t noalloc_scratch(void** out);

}  // namespace gst
