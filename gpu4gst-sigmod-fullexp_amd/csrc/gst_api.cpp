// C ABI of the gfx950 TrimCDP / D-TrimCDP engine: graph and group upload,
// HBM-budgeted batching of queries into lockstep slot batches, the
// graph-captured iteration loop, tree extraction and the backward scatter.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gst.h"
#include "gst_device.h"

using namespace gst;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(_e == hipErrorOutOfMemory ? GST_ERR_OOM : GST_ERR_DEVICE,            \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                  \
  } while (0)

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;  // a view into the allocation-free path's scratch: never (re)allocated
  hipError_t ensure(size_t want, bool* grew = nullptr) {
    if (want <= bytes && p) return hipSuccess;
    if (borrowed) return hipErrorOutOfMemory;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t b = want < 256 ? 256 : want;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    if (grew) *grew = true;
    return e;
  }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int ilog2(int64_t x) {
  int s = 0;
  while ((int64_t(1) << s) < x) ++s;
  return s;
}

struct QueryIn {
  int32_t idx;
  int32_t k, D;
  const int32_t* groups;
};

// Expand-kernel settings chosen per size class of a call: -1 / 0 = the rule.
struct Tune {
  int32_t chunk_shift = -1;  // log2 of the virtual-split chunk
  int32_t grid = 0;          // workgroups of the expand kernels
  int32_t dp_div = -1;       // first DP band = upper bound / dp_div (0 = unordered)
  int32_t sssp_band = -1;    // first one-label band width (0 = unordered)
};

// Size classes of a call: one query, a few (2..31), many (32+: the banded
// DP's regime).
constexpr int kSizeClasses = 3;
int size_class(int64_t nq) { return nq <= 1 ? 0 : (nq < 32 ? 1 : 2); }

}  // namespace

struct gst_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t V = 0, Vp = 0;
  int64_t E = 0;
  int64_t n_und = 0;
  int32_t cus = 256;
  std::vector<int32_t> h_rowp;
  // host copies in the caller's CSR order; the device rows are sorted by weight
  std::vector<int32_t> h_col, h_w, h_ids, perm;  // perm[device position] = caller's CSR index
  DBuf rowp, col, wt, eid;
  bool has_eid = false;
  // groups
  int32_t G = 0;
  std::vector<int64_t> h_goff;
  DBuf goff, gmem;
  // options
  // Settings of the expand kernels. `*_opt` hold the caller's explicit
  // choices (-1 / 0 = not set); the values in force for a call come from
  // them, else from the autotuner's measured choice for the call's size
  // class, else from the size-class model (select_settings).
  int32_t chunk_opt = -1;   // log2 of the virtual-split chunk
  int32_t chunk_shift = 12; // in force
  int32_t grid = 0;         // explicit workgroup count of the expand kernels (0 = not set)
  int32_t grid_eff = 0;     // in force
  Tune tuned[kSizeClasses];
  bool tuned_ok[kSizeClasses] = {false, false, false};
  int32_t dp_div_call = -1, sssp_band_call = -1;  // per call: explicit or tuned (-1 = rule)
  // backpointers: 1 = stored, 0 = recomputed from the cost table, -1 = stored
  // when the whole call fits one admission batch with them (memory not tight)
  int32_t bp_mode = -1;
  bool bp_call = false;  // in force for the call being solved
  // per-iteration trace of the profiled path (GST_OPT_TRACE with GST_OPT_PROFILE)
  int32_t trace = 0;
  // inside gst_solve_noalloc: no allocation, no graph capture, no graph
  // destruction (stale captured graphs are replaced lazily by later calls)
  bool noalloc = false;
  int32_t batch_no = 0;
  std::vector<gst_trace_row> trace_rows;
  int32_t iters_per_sync = 32;
  int32_t hop_prune = 1;
  int32_t profile = 0;
  int32_t infeasible_cost = INT32_MAX;
  int32_t max_tree_edges = 4096;
  int32_t use_graphs = 1;
  int32_t sssp_band = -1;      // first one-label band width; 0 = unordered, -1 = automatic
  int32_t sssp_band_eff = 0;   // the value in force for the batch being solved
  int32_t dp_div = -1;     // -1 = automatic (kAutoBandDiv for batches of kAutoBandSlots+)
  int32_t dp_div_eff = 0;  // the value in force for the batch being solved
  int32_t radius_pct = 0;  // 0: 85 for queries of up to 6 groups, 95 for 7
  int32_t max_batch = 0;
  int32_t sssp_slots = 0;  // slots per one-label pass (0 = the whole batch)
  int32_t list_permille = 0;  // worklist capacity override (0 = default sizing)
  int32_t lane_appends = 0;    // ablation: per-lane global appends
  int32_t per_item_edges = 0;  // ablation: per-item edge walks
  int64_t eff_key = -1;        // the effective settings baked into the captured graphs
  int32_t maxw = 1;
  // engine buffers (grow-only)
  // wl0 / wl1: the frontier lists (DP: vertex items); wls: the DP's unfolded
  // state items of the current iteration; Ms: DP pending masks
  DBuf T, Es, BPs, Ms, GMs, far0, far1, C, R, VVb, bestw, slots, tasks, wl0, wl1, wls, cnt, iter,
      ctr, sched;
  DBuf out_uv, out_eid, out_n, out_err;
  DBuf outpack;  // tree rows packed to their longest filled prefix for the copy back
  uint32_t* h_cnt = nullptr;  // pinned
  int32_t h_nslots[1] = {0};  // slots of the batch being solved
  // Rolling admission (GST_OPT_ROLLING): a call whose tables do not fit the
  // HBM budget at once keeps a pool of DP table regions busy, admitting the
  // next query as soon as one converges. `resident` = queries holding a
  // region now (0 outside a rolling run).
  int32_t rolling = 1;
  int32_t resident = 0;
  int32_t out_rows = 0;  // rows of the tree-recovery output buffers
  DBuf rsel, rtasks;  // slot lists / tasks of the queries admitted or collected
  DBuf rbest;         // upper bounds after the one-label phase
  uint32_t wl_cap = 0;
  int64_t gen = 0;  // bumps whenever a pointer or scalar baked into a captured graph changes
  std::map<std::pair<int, int>, std::pair<int64_t, hipGraphExec_t>> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ph[4] = {nullptr, nullptr, nullptr, nullptr};  // phase boundaries
  int64_t nch_total() const {
    int64_t s = 0;
    const int32_t CH = 1 << chunk_shift;
    for (int32_t v = 0; v < V; ++v) {
      int64_t deg = h_rowp[v + 1] - h_rowp[v];
      int64_t n = (deg + CH - 1) / CH;
      s += n ? n : 1;
    }
    return s;
  }
  int64_t nch_cached = -1;
  int32_t nch_shift = -1;
  int64_t nch() {
    if (nch_shift != chunk_shift) {
      nch_cached = nch_total();
      nch_shift = chunk_shift;
    }
    return nch_cached;
  }
  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second.second);
    graphs.clear();
  }
  // device bytes held by the per-batch buffers; a later batch reuses them, so
  // they count as free when the next batch is admitted
  size_t held_bytes() const {
    size_t s = 0;
    for (const DBuf* b : {&T, &Es, &BPs, &Ms, &GMs, &far0, &far1, &C, &R, &VVb, &wl0, &wl1, &wls, &out_uv,
                          &out_eid})
      s += b->bytes;
    return s;
  }
  void release_batch_buffers() {
    drop_graphs();
    for (DBuf* b : {&T, &Es, &BPs, &Ms, &GMs, &far0, &far1, &C, &R, &VVb, &wl0, &wl1, &wls, &out_uv,
                    &out_eid})
      b->release();
    wl_cap = 0;
    ++gen;
  }
};

namespace {

int32_t rows_of(int32_t k, int32_t D) { return (D < 0 ? 1 : D + 1) << k; }

// Worklists are sized to this fraction of all DP states (frontiers observed on
// power-law graphs stay near 1%); a batch that overflows is re-run with
// worklists as large as the state count, which can never overflow.
constexpr double kWorklistFraction = 0.04;
// Largest worklist (items): positions must stay below 2^32 with room for the
// appends of one round past the end.
constexpr int64_t kListClamp = (int64_t)UINT32_MAX - 65536 - 4096;
constexpr int kAutoBandSlots = 32;  // automatic DP bands from this many plain queries per batch
constexpr int kAutoBandDiv = 256;   // first band width = upper bound / 256 (>= 1), doubling
constexpr int kAutoSsspBand = 16;   // first one-label band: labels below 16, doubling per band

// Mask words per (vertex, depth) of a query of k groups: the pending mask, and
// for a plain query or a hop-bounded one of up to 5 groups a finite mask
// interleaved with it word by word (the merge loop enumerates only the
// partners already reached).
bool fin_slot(int32_t k, int32_t D) { return D < 0 || k <= 5; }
int32_t mask_words(int32_t k, int32_t D) {
  const int32_t pend = k <= 5 ? 1 : (1 << k) / 32;
  return fin_slot(k, D) ? 2 * pend : pend;
}

// Pending-mask words of one slot: one mask per (vertex, depth).
int64_t slot_mask_words(const gst_graph* g, int32_t k, int32_t D) {
  return (int64_t)(D < 0 ? 1 : D + 1) * g->Vp * mask_words(k, D);
}

// Table footprint of one slot (int32 elements): state words, one-label far
// keys, pending masks, group masks, one-label distances, route tables.
int64_t slot_ints(const gst_graph* g, int32_t k, int32_t D) {
  const int64_t rows = rows_of(k, D);
  return 2 * rows * g->Vp + g->Vp + slot_mask_words(g, k, D) + g->Vp +
         (int64_t)g->Vp * CROW + RTABLE;
}

int64_t slot_items(gst_graph* g, int32_t k, int32_t D) { return (int64_t)rows_of(k, D) * g->nch(); }

// Device bytes of one slot's stored backpointers (one int32 per state).
int64_t bp_bytes(const gst_graph* g, int32_t k, int32_t D) { return 4 * rows_of(k, D) * g->Vp; }

DevCtx make_ctx(gst_graph* g) {
  DevCtx c;
  c.rowp = g->rowp.as<int32_t>();
  c.col = g->col.as<int32_t>();
  c.wt = g->wt.as<int32_t>();
  c.eid = g->has_eid ? g->eid.as<int32_t>() : nullptr;
  c.V = g->V;
  c.Vp = g->Vp;
  c.chunk_shift = g->chunk_shift;
  c.hop_prune = g->hop_prune;
  c.T = g->T.as<int32_t>();
  c.E = g->Es.as<int32_t>();
  c.M = g->Ms.as<uint32_t>();
  c.GM = g->GMs.as<uint32_t>();
  c.C = g->C.as<int32_t>();
  c.R = g->R.as<int32_t>();
  c.vv = g->VVb.as<int32_t>();
  c.bestw = g->bestw.as<uint64_t>();
  c.slots = g->slots.as<SlotDesc>();
  c.iter = g->iter.as<int32_t>();
  c.ctr = g->ctr.as<unsigned long long>();
  c.cap = g->wl_cap;
  c.sched = g->sched.as<Sched>();
  c.nslots = g->iter.as<int32_t>() + 1;
  c.dp_div = g->dp_div_eff;
  c.radius_pct = g->radius_pct;
  c.far_mode = g->dp_div_eff > 0 ? 1 : 0;
  c.far[0] = g->far0.as<uint64_t>();
  c.far[1] = g->far1.as<uint64_t>();
  c.far_cnt = g->cnt.as<uint32_t>() + 2;
  c.far_sel = g->iter.as<int32_t>() + 2;
  c.sssp_far = g->sssp_band_eff > 0 ? 1 : 0;
  c.band_key = g->iter.as<int32_t>() + 3;
  c.lane_appends = g->lane_appends;
  c.per_item_edges = g->per_item_edges;
  c.BP = g->bp_call ? g->BPs.as<int32_t>() : nullptr;
  c.trace = g->profile && g->trace ? 1 : 0;
  return c;
}

int expand_grid(const gst_graph* g) { return g->grid_eff > 0 ? g->grid_eff : g->cus * 8; }

// The unfold step: one LDS search per state item it writes; the hop-bounded
// DP's vertex items carry many pending states, and eight workgroups per CU
// hide those latencies (hop DP phase 1,655 -> 1,589 ms on the 1.1M-vertex
// workload; no change on the plain 168k-vertex one), whatever the expand
// kernels use, unless GST_OPT_GRID sets fewer.
int unfold_grid(const gst_graph* g) {
  return g->grid > 0 ? std::min(g->grid, g->cus * 8) : g->cus * 8;
}

// Size-class model of the expand settings, from sweeps on the 168k-vertex /
// 6.8M-edge power-law workload (chunk 128..8192, grid 2..16 x CUs): a lone
// query has little parallelism besides its hub rows, so they are split into
// 256-edge chunks (260 queries/s single vs 190 at 4096); a batch of many
// queries fills the device with items and coarse chunks keep the per-item
// overhead low (4096: 2,190 queries/s vs 1,960 at 256); the DP and one-label
// band rules stay with solve_lockstep's defaults. A batch of 32 or more
// queries gets four workgroups per CU, the number the expand kernels keep
// resident (four waves per SIMD): with eight, their grid-stride loops run in
// two rounds and the second round's tail shows (3,790 -> 3,950 queries/s;
// six per CU: 3,450).
Tune model_settings(const gst_graph* g, int cls) {
  Tune t;
  t.chunk_shift = cls == 0 ? 8 : (cls == 1 ? 10 : 12);
  t.grid = cls == 2 ? g->cus * 4 : g->cus * 8;
  return t;
}

// Settings in force for a call of nq queries: the caller's explicit options,
// else the autotuner's choice for the size class, else the model.
void select_settings(gst_graph* g, int64_t nq) {
  const int cls = size_class(nq);
  Tune t = model_settings(g, cls);
  if (g->tuned_ok[cls]) {
    const Tune& m = g->tuned[cls];
    if (m.chunk_shift >= 0) t.chunk_shift = m.chunk_shift;
    if (m.grid > 0) t.grid = m.grid;
    t.dp_div = m.dp_div;
    t.sssp_band = m.sssp_band;
  }
  const int32_t cs = g->chunk_opt >= 0 ? g->chunk_opt : t.chunk_shift;
  const int32_t gr = g->grid > 0 ? g->grid : t.grid;
  g->dp_div_call = g->dp_div >= 0 ? g->dp_div : t.dp_div;
  g->sssp_band_call = g->sssp_band >= 0 ? g->sssp_band : t.sssp_band;
  if (cs != g->chunk_shift || gr != g->grid_eff) {  // both are baked into captured graphs
    g->chunk_shift = cs;
    g->grid_eff = gr;
    ++g->gen;
    if (!g->noalloc) g->drop_graphs();
  }
}

// Counters of the lists (uint32 each): the two frontier lists, the two far
// lists, the DP's unfolded state list.
constexpr int kCntStates = 4;
constexpr int kCntWords = 5;

// The expand step of half-iteration h (list h -> list 1 - h). The DP first
// unfolds the vertex items of list h into the state list.
void launch_expand_step(gst_graph* g, const DevCtx& c, int mode, int h) {
  uint32_t* cnt = g->cnt.as<uint32_t>();
  uint64_t* l[2] = {g->wl0.as<uint64_t>(), g->wl1.as<uint64_t>()};
  const int grid = expand_grid(g);
  if (mode == 0) {
    launch_sssp_expand(c, WL{l[h], cnt + h, l[1 - h], cnt + (1 - h)}, grid, g->stream);
    return;
  }
  launch_dp_unfold(c, l[h], cnt + h, g->wls.as<uint64_t>(), cnt + kCntStates, unfold_grid(g),
                   g->stream);
  launch_dp_expand(c, WL{g->wls.as<uint64_t>(), cnt + kCntStates, l[1 - h], cnt + (1 - h)}, grid,
                   mode == 2, g->stream);
}

void launch_advance_step(gst_graph* g, const DevCtx& c, int mode, int h) {
  uint32_t* cnt = g->cnt.as<uint32_t>();
  launch_advance(c, cnt + h, cnt + (1 - h), mode != 0, mode != 0 ? cnt + kCntStates : nullptr,
                 g->stream);
}

// One frontier iteration pair: list 0 -> list 1, list 1 -> list 0.
void launch_pair(gst_graph* g, const DevCtx& c, int mode) {
  for (int h = 0; h < 2; ++h) {
    launch_expand_step(g, c, mode, h);
    launch_advance_step(g, c, mode, h);
  }
}

// One round of `pairs` frontier iteration pairs: replayed from a captured
// graph, launched one by one, or (profiling mode) launched one by one with
// every expand launch timed (one pair; `pairs` is set to the pairs run).
int phase_round(gst_graph* g, const DevCtx& c, int mode, int& pairs, double* ms_acc,
                int64_t* launches) {
  uint32_t* cnt = g->cnt.as<uint32_t>();
  {
    if (g->profile) {
      const int grid = expand_grid(g);
      uint64_t* l[2] = {g->wl0.as<uint64_t>(), g->wl1.as<uint64_t>()};
      for (int h = 0; h < 2; ++h) {
        // the DP's unfold step runs before the timed expand launch
        if (mode != 0)
          launch_dp_unfold(c, l[h], cnt + h, g->wls.as<uint64_t>(), cnt + kCntStates, unfold_grid(g),
                           g->stream);
        uint32_t items = 0;
        if (g->trace) {
          HIP_TRY(hipStreamSynchronize(g->stream));
          HIP_TRY(hipMemcpy(&items, cnt + (mode != 0 ? kCntStates : h), sizeof(items),
                            hipMemcpyDeviceToHost));
        }
        HIP_TRY(hipEventRecord(g->ev0, g->stream));
        if (mode == 0)
          launch_sssp_expand(c, WL{l[h], cnt + h, l[1 - h], cnt + (1 - h)}, grid, g->stream);
        else
          launch_dp_expand(c, WL{g->wls.as<uint64_t>(), cnt + kCntStates, l[1 - h], cnt + (1 - h)},
                           grid, mode == 2, g->stream);
        HIP_TRY(hipEventRecord(g->ev1, g->stream));
        launch_advance_step(g, c, mode, h);
        HIP_TRY(hipEventSynchronize(g->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        if (ms_acc) *ms_acc += ms;
        if (launches) ++*launches;
        if (g->trace) {
          uint64_t active = 0;
          HIP_TRY(hipStreamSynchronize(g->stream));
          HIP_TRY(hipMemcpy(&active, g->ctr.as<uint64_t>() + C_ACTIVE, sizeof(active),
                            hipMemcpyDeviceToHost));
          gst_trace_row r;
          r.phase = mode;
          r.batch = g->batch_no;
          r.items = items;
          r.active_slots = (int32_t)active;
          r.slots = g->resident > 0 ? g->resident : *g->h_nslots;
          r.ms = ms;
          g->trace_rows.push_back(r);
        }
      }
      pairs = 1;
    } else if (g->use_graphs) {
      auto key = std::make_pair(mode, pairs);
      auto itg = g->graphs.find(key);
      if (itg == g->graphs.end() || itg->second.first != g->gen) {
        if (itg != g->graphs.end()) {
          (void)hipGraphExecDestroy(itg->second.second);
          g->graphs.erase(itg);
        }
        hipGraph_t graph;
        HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < pairs; ++r) launch_pair(g, c, mode);
        HIP_TRY(hipStreamEndCapture(g->stream, &graph));
        hipGraphExec_t exec;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(e);
        g->graphs[key] = std::make_pair(g->gen, exec);
        itg = g->graphs.find(key);
      }
      HIP_TRY(hipGraphLaunch(itg->second.second, g->stream));
    } else {
      for (int r = 0; r < pairs; ++r) launch_pair(g, c, mode);
    }
    HIP_TRY(hipGetLastError());
  }
  return GST_OK;
}

// Runs frontier iterations until the worklist is empty. mode 0 = one-label,
// 1 = plain DP, 2 = hop DP. Kernel timings are accumulated in profiling mode.
int run_phase(gst_graph* g, int mode, double* ms_acc, int64_t* launches) {
  DevCtx c = make_ctx(g);
  int pairs = 1;
  const int max_pairs = std::max(1, g->iters_per_sync / 2);
  int64_t total_pairs = 0;
  uint32_t* cnt = g->cnt.as<uint32_t>();
  while (true) {
    const int prc = phase_round(g, c, mode, pairs, ms_acc, launches);
    if (prc) return prc;
    total_pairs += pairs;
    HIP_TRY(hipMemcpyAsync(g->h_cnt, cnt, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(g->h_cnt + 2, g->ctr.as<uint64_t>() + C_OVERFLOW, sizeof(uint64_t),
                           hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    uint64_t overflow;
    std::memcpy(&overflow, g->h_cnt + 2, sizeof(overflow));
    if (overflow) return 100;  // worklist too small: the batch is re-run with full capacity
    if (g->h_cnt[0] == 0) {
      if (!((mode != 0 && c.far_mode) || (mode == 0 && c.sssp_far))) break;
      // banded schedule: the near list has drained; bring back the far list's next band
      int32_t sel = 0;
      uint32_t fc[2] = {0, 0};
      HIP_TRY(hipMemcpyAsync(fc, cnt + 2, sizeof(fc), hipMemcpyDeviceToHost, g->stream));
      HIP_TRY(hipMemcpyAsync(&sel, g->iter.as<int32_t>() + 2, sizeof(sel), hipMemcpyDeviceToHost,
                             g->stream));
      HIP_TRY(hipStreamSynchronize(g->stream));
      if (fc[sel & 1] == 0) break;
      if (mode == 0)
        launch_sssp_far_refill(c, fc[sel & 1], g->wl0.as<uint64_t>(), cnt, g->stream);
      else
        launch_far_refill(c, *g->h_nslots, fc[sel & 1], g->wl0.as<uint64_t>(), cnt, g->stream);
      HIP_TRY(hipGetLastError());
      ++total_pairs;
      pairs = 1;
      continue;
    }
    if (total_pairs > (int64_t)1 << 17)
      return fail(GST_ERR_INTERNAL, "frontier iteration limit exceeded");
    pairs = std::min(pairs * 2, max_pairs);
  }
  return GST_OK;
}

// Solves one lockstep batch of same-mode queries. Results go to the caller's
// per-query vectors. Returns GST_ERR_CAPACITY-like internal retry signal 100
// when the worklist overflowed.
struct BatchOut {
  std::vector<int32_t> cost, status;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> uv;
  std::vector<std::vector<int32_t>> ids;
};

// Tree recovery and results of the slots `sel` (host list; `d_sel` the same
// list on the device, or null when sel = 0, 1, ..., n-1): output row i of the
// recovery kernel belongs to slot sel[i]. Costs, statuses and trees go to the
// queries' entries of `out`.
int collect_chunk(gst_graph* g, const DevCtx& c, const std::vector<QueryIn>& qs,
                  const std::vector<int32_t>& sel, const int32_t* d_sel, BatchOut& out) {
  const int n = (int)sel.size();
  if (n == 0) return GST_OK;
  const int nslots = (int)qs.size();
  const int maxe = g->max_tree_edges;
  launch_backtrack(c, n, maxe, g->out_uv.as<int32_t>(), g->out_eid.as<int32_t>(),
                   g->out_n.as<int32_t>(), g->out_err.as<int32_t>(), d_sel, g->stream);
  HIP_TRY(hipGetLastError());
  std::vector<uint64_t> bw(nslots);
  std::vector<int32_t> on(n), oerr(n);
  std::vector<Sched> hsched(nslots);
  HIP_TRY(hipMemcpyAsync(bw.data(), g->bestw.p, nslots * 8, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(on.data(), g->out_n.p, n * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(oerr.data(), g->out_err.p, n * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(hsched.data(), g->sched.p, nslots * sizeof(Sched), hipMemcpyDeviceToHost,
                         g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  // only the filled prefix of every row comes back
  int32_t max_on = 0;
  for (int o = 0; o < n; ++o)
    if (oerr[o] == 0) max_on = std::max(max_on, std::min(on[o], maxe));
  std::vector<int32_t> huv((size_t)n * std::max(max_on, 1) * 2),
      hid((size_t)n * std::max(max_on, 1));
  if (max_on > 0 && n > 1 && !g->noalloc) {
    // rows packed to max_on edges on the device, then one copy each (a 2D copy
    // of a thousand short rows from device to pageable memory took 14 ms)
    const size_t w = (size_t)n * max_on;
    HIP_TRY(g->outpack.ensure(w * 12));
    int32_t* puv = g->outpack.as<int32_t>();
    int32_t* pid = puv + 2 * w;
    launch_pack_rows(g->out_uv.as<int32_t>(), 2 * (int64_t)maxe, puv, 2 * (int64_t)max_on, n,
                     g->stream);
    launch_pack_rows(g->out_eid.as<int32_t>(), maxe, pid, max_on, n, g->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(huv.data(), puv, w * 8, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(hid.data(), pid, w * 4, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
  } else if (max_on > 0) {
    HIP_TRY(hipMemcpy2D(huv.data(), (size_t)max_on * 8, g->out_uv.p, (size_t)maxe * 8,
                        (size_t)max_on * 8, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy2D(hid.data(), (size_t)max_on * 4, g->out_eid.p, (size_t)maxe * 4,
                        (size_t)max_on * 4, n, hipMemcpyDeviceToHost));
  }
  // Trees longer than the per-slot buffer (code 2) are recovered again with
  // room for min(cost, 4V) edges: every recovered edge weighs >= 1 and their
  // total is at most the cost.
  std::vector<int32_t> redo;  // slots
  for (int o = 0; o < n && !g->noalloc; ++o)
    if (oerr[o] == 2) redo.push_back(sel[o]);
  std::map<int, std::pair<std::vector<int32_t>, std::vector<int32_t>>> long_trees;  // by slot
  std::map<int, int32_t> redo_err;  // by slot
  if (!redo.empty()) {
    int64_t cap2 = 1;
    for (int s : redo)
      cap2 = std::max<int64_t>(cap2, std::min<int64_t>((int64_t)(bw[s] >> 32) + 1,
                                                        4 * (int64_t)g->V + 64));
    const int nr = (int)redo.size();
    DBuf uv2, id2, n2, e2, rs;
    auto release = [&]() {
      for (DBuf* b : {&uv2, &id2, &n2, &e2, &rs}) b->release();
    };
    hipError_t he = uv2.ensure((size_t)nr * cap2 * 8);
    if (he == hipSuccess) he = id2.ensure((size_t)nr * cap2 * 4);
    if (he == hipSuccess) he = n2.ensure((size_t)nr * 4);
    if (he == hipSuccess) he = e2.ensure((size_t)nr * 4);
    if (he == hipSuccess) he = rs.ensure((size_t)nr * 4);
    if (he == hipSuccess)
      he = hipMemcpyAsync(rs.p, redo.data(), nr * 4, hipMemcpyHostToDevice, g->stream);
    if (he == hipSuccess) {
      launch_backtrack(c, nr, (int)cap2, uv2.as<int32_t>(), id2.as<int32_t>(), n2.as<int32_t>(),
                       e2.as<int32_t>(), rs.as<int32_t>(), g->stream);
      he = hipGetLastError();
    }
    std::vector<int32_t> n2h(nr), e2h(nr), uvh, idh;
    if (he == hipSuccess)
      he = hipMemcpyAsync(n2h.data(), n2.p, nr * 4, hipMemcpyDeviceToHost, g->stream);
    if (he == hipSuccess)
      he = hipMemcpyAsync(e2h.data(), e2.p, nr * 4, hipMemcpyDeviceToHost, g->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(g->stream);
    for (int r = 0; r < nr && he == hipSuccess; ++r) {
      const int s = redo[r];
      redo_err[s] = e2h[r];
      if (e2h[r] != 0) continue;
      uvh.resize((size_t)n2h[r] * 2);
      idh.resize((size_t)n2h[r]);
      he = hipMemcpy(uvh.data(), uv2.as<int32_t>() + (size_t)r * cap2 * 2, uvh.size() * 4,
                     hipMemcpyDeviceToHost);
      if (he == hipSuccess)
        he = hipMemcpy(idh.data(), id2.as<int32_t>() + (size_t)r * cap2, idh.size() * 4,
                       hipMemcpyDeviceToHost);
      long_trees[s] = {uvh, idh};
    }
    release();
    HIP_TRY(he);
  }
  for (int o = 0; o < n; ++o) {
    const int s = sel[o];
    const int qi = qs[s].idx;
    const int32_t cost = (int32_t)(bw[s] >> 32);
    int32_t err = oerr[o];
    if (err == 2 && redo_err.count(s)) err = redo_err[s];
    if (err == 2) {  // still longer than the retry's buffer: cost only
      out.cost[qi] = cost;
      out.status[qi] = GST_Q_NO_TREE;
      out.uv[qi].clear();
      out.ids[qi].clear();
      continue;
    }
    if (err != 0)
      return fail(GST_ERR_INTERNAL, "tree recovery failed for query " + std::to_string(qi) +
                                        " (code " + std::to_string(err) + ")");
    if (cost >= INF) {
      out.cost[qi] = g->infeasible_cost;
      out.status[qi] = hsched[s].sat ? GST_Q_RANGE : GST_Q_INFEASIBLE;
      continue;
    }
    out.cost[qi] = cost;
    out.status[qi] = GST_Q_OK;
    std::vector<std::pair<std::pair<int32_t, int32_t>, int32_t>> es;
    auto lt = long_trees.find(s);
    if (lt != long_trees.end()) {
      const auto& uvs = lt->second.first;
      const auto& ids = lt->second.second;
      for (size_t i = 0; i < ids.size(); ++i)
        es.push_back({{std::min(uvs[2 * i], uvs[2 * i + 1]), std::max(uvs[2 * i], uvs[2 * i + 1])},
                      ids[i]});
    }
    for (int i = 0; lt == long_trees.end() && i < on[o]; ++i) {
      int32_t a = huv[((size_t)o * max_on + i) * 2], b = huv[((size_t)o * max_on + i) * 2 + 1];
      es.push_back({{std::min(a, b), std::max(a, b)}, hid[(size_t)o * max_on + i]});
    }
    std::sort(es.begin(), es.end());
    es.erase(std::unique(es.begin(), es.end(),
                         [](const auto& x, const auto& y) { return x.first == y.first; }),
             es.end());
    out.uv[qi].clear();
    out.ids[qi].clear();
    for (auto& e : es) {
      out.uv[qi].push_back(e.first);
      out.ids[qi].push_back(e.second);
    }
  }
  return GST_OK;
}

// The one-label phase of slots 0..nslots-1 (tasks in slot order), in passes
// of per_pass slots, then the clamped distances, group distances and route
// tables of every slot. Counters and list counts start from zero; ph[1] is
// recorded when the last pass ends.
int onelabel_stage(gst_graph* g, const DevCtx& c, const std::vector<Task>& tasks, int nslots,
                   int per_pass, double* sssp_ms, int64_t* sssp_launches) {
  launch_fill_u64(g->ctr.as<uint64_t>(), C_NUM, 0, g->stream);
  launch_fill_i32(g->cnt.as<int32_t>(), kCntWords, 0, g->stream);
  launch_init_slot_tables(c, nslots, (int64_t)nslots * g->Vp * CROW, g->stream);
  const Task* dt = g->tasks.as<Task>();
  const int nt = (int)tasks.size();
  uint32_t* cnt = g->cnt.as<uint32_t>();
  // The one-label phase runs over groups of per_pass slots at a time (tasks
  // are in slot order): the label rows one pass touches can stay in the
  // last-level cache, and the lists need room for one pass only.
  launch_fill_i32(g->cnt.as<int32_t>() + 2, 2, 0, g->stream);   // far list counts
  launch_fill_i32(g->iter.as<int32_t>() + 2, 1, 0, g->stream);  // far list in use
  launch_band_start(c, g->stream);
  for (int t0 = 0; t0 < nt;) {
    const int s_end = std::min(nslots, tasks[t0].slot + per_pass);
    int t1 = t0;
    while (t1 < nt && tasks[t1].slot < s_end) ++t1;
    launch_groups_setup(c, dt + t0, t1 - t0, g->goff.as<int64_t>(), g->gmem.as<int32_t>(),
                        g->wl0.as<uint64_t>(), cnt, g->stream);
    HIP_TRY(hipGetLastError());
    int rc = run_phase(g, 0, sssp_ms, sssp_launches);
    if (rc) return rc;  // includes 100 = overflow
    t0 = t1;
  }
  HIP_TRY(hipEventRecord(g->ph[1], g->stream));
  launch_clamp(c, nslots, g->stream);
  launch_vv(c, dt, nt, g->goff.as<int64_t>(), g->gmem.as<int32_t>(), g->stream);
  launch_route(c, nslots, g->stream);
  HIP_TRY(hipGetLastError());
  return GST_OK;
}

// collect_chunk over pieces of `sel` that fit the output rows.
int collect_results(gst_graph* g, const DevCtx& c, const std::vector<QueryIn>& qs,
                    const std::vector<int32_t>& sel, const int32_t* d_sel, BatchOut& out) {
  const size_t rows = (size_t)std::max(1, g->out_rows);
  if ((int64_t)sel.size() <= (int64_t)rows) return collect_chunk(g, c, qs, sel, d_sel, out);
  if (!d_sel) return fail(GST_ERR_INTERNAL, "tree recovery rows too few for the batch");
  for (size_t a = 0; a < sel.size(); a += rows) {
    const size_t b = std::min(sel.size(), a + rows);
    const std::vector<int32_t> part(sel.begin() + a, sel.begin() + b);
    const int rc = collect_chunk(g, c, qs, part, d_sel + a, out);
    if (rc) return rc;
  }
  return GST_OK;
}

int solve_lockstep(gst_graph* g, const std::vector<QueryIn>& qs, double wl_frac, BatchOut& out,
                   gst_stats* st) {
  const int nslots = (int)qs.size();
  std::vector<SlotDesc> sd(nslots);
  std::vector<Task> tasks;
  int64_t t_tot = 0, e_tot = 0, m_tot = 0, items = 0, sssp_items = 0;
  for (int s = 0; s < nslots; ++s) {
    const QueryIn& q = qs[s];
    SlotDesc d;
    std::memset(&d, 0, sizeof(d));
    const int64_t rows = rows_of(q.k, q.D);
    d.t_off = t_tot;
    d.e_off = e_tot;
    d.c_off = (int64_t)s * g->Vp * CROW;
    d.r_off = (int64_t)s * RTABLE;
    d.k = q.k;
    d.full = (1 << q.k) - 1;
    d.D = q.D;
    d.nsub = 1 << q.k;
    d.vv_off = s * VV;
    d.nrows = (int32_t)rows;
    d.mw = mask_words(q.k, q.D);
    d.fin = fin_slot(q.k, q.D) ? 1 : 0;
    d.m_off = m_tot;
    t_tot += rows * g->Vp;
    e_tot += g->Vp;
    m_tot += slot_mask_words(g, q.k, q.D);
    items += slot_items(g, q.k, q.D);
    sssp_items += g->nch();  // one-label items are (slot, vertex[, chunk])
    sd[s] = d;
    for (int i = 0; i < q.k; ++i) tasks.push_back(Task{s, i, q.groups[i], 0});
  }
  int64_t cap = std::max((int64_t)((double)items * wl_frac), sssp_items) + 4096;
  if (g->list_permille > 0 && wl_frac < 1.0) cap = items * g->list_permille / 1000 + 64;
  if (cap > kListClamp + 4096) cap = kListClamp + 4096;
  bool grew = false;
  HIP_TRY(g->T.ensure((size_t)t_tot * 8, &grew));
  bool e_grew = false;
  HIP_TRY(g->Es.ensure((size_t)e_tot * 4, &e_grew));
  if (e_grew) HIP_TRY(hipMemsetAsync(g->Es.p, 0, g->Es.bytes, g->stream));
  grew |= e_grew;
  // backpointer words need no initialisation: every state the tree recovery
  // visits was written in this batch, and each word is checked before use
  if (g->bp_call) HIP_TRY(g->BPs.ensure((size_t)t_tot * 4, &grew));
  HIP_TRY(g->Ms.ensure((size_t)m_tot * 4, &grew));
  HIP_TRY(g->C.ensure((size_t)nslots * g->Vp * CROW * 4, &grew));
  HIP_TRY(g->GMs.ensure((size_t)nslots * g->Vp * 4, &grew));
  HIP_TRY(g->R.ensure((size_t)nslots * RTABLE * 4, &grew));
  HIP_TRY(g->VVb.ensure((size_t)nslots * VV * 4, &grew));
  HIP_TRY(g->bestw.ensure((size_t)nslots * 8, &grew));
  HIP_TRY(g->slots.ensure((size_t)nslots * sizeof(SlotDesc), &grew));
  HIP_TRY(g->sched.ensure((size_t)nslots * sizeof(Sched), &grew));
  HIP_TRY(g->tasks.ensure(tasks.size() * sizeof(Task) + 16, &grew));
  if ((uint64_t)cap * 8 > g->wl0.bytes || (uint64_t)cap * 8 > g->wl1.bytes ||
      (uint64_t)cap * 8 > g->wls.bytes) {
    HIP_TRY(g->wl0.ensure((size_t)cap * 8, &grew));
    HIP_TRY(g->wl1.ensure((size_t)cap * 8, &grew));
    HIP_TRY(g->wls.ensure((size_t)cap * 8, &grew));
  }
  // A* bands (near/far lists) pay off when many queries share each iteration;
  // a lone query runs unordered (measured on the 168k-vertex workload with the
  // unordered one-label phase: 1,500 vs 1,360 queries/s batched, 188 vs 232
  // single; divisor 256 vs 64 with the banded one-label phase: 1,981 vs 1,915)
  g->dp_div_eff = g->dp_div_call >= 0
                      ? g->dp_div_call
                      : (nslots >= kAutoBandSlots && qs[0].D < 0 ? kAutoBandDiv : 0);
  // the one-label phase in bands of label value (delta stepping with near/far
  // lists), so that labels past the final radius are never computed and
  // label-correcting churn collapses (measured on the 168k-vertex workload:
  // 208 vs 326 ms of one-label time per batch of 1,024, 3.3 G vs 7.4 G edge
  // visits; hop-bounded queries, which have no radius, on the 978k-vertex
  // workload: 619 vs 720 ms, 8.5 G vs 15.2 G edge visits)
  g->sssp_band_eff = g->sssp_band_call >= 0 ? g->sssp_band_call : kAutoSsspBand;
  {
    // these values are kernel arguments inside the captured graphs
    const int64_t key = ((int64_t)g->dp_div_eff << 32) | ((uint32_t)g->sssp_band_eff << 1) |
                        (g->bp_call ? 1 : 0);
    if (key != g->eff_key) {
      g->eff_key = key;
      grew = true;
    }
  }
  const bool far_lists = g->dp_div_eff > 0 || g->sssp_band_eff > 0;
  if (far_lists &&
      ((uint64_t)cap * 8 > g->far0.bytes || (uint64_t)cap * 8 > g->far1.bytes)) {
    HIP_TRY(g->far0.ensure((size_t)cap * 8, &grew));
    HIP_TRY(g->far1.ensure((size_t)cap * 8, &grew));
  }
  size_t list_bytes = std::min({g->wl0.bytes, g->wl1.bytes, g->wls.bytes});
  if (far_lists) list_bytes = std::min({list_bytes, g->far0.bytes, g->far1.bytes});
  const uint32_t new_cap = (uint32_t)std::min<size_t>(list_bytes / 8, UINT32_MAX - 65536);
  if (new_cap != g->wl_cap) {
    g->wl_cap = new_cap;
    grew = true;
  }
  const int maxe = g->max_tree_edges;
  HIP_TRY(g->out_uv.ensure((size_t)nslots * maxe * 8, &grew));
  HIP_TRY(g->out_eid.ensure((size_t)nslots * maxe * 4, &grew));
  HIP_TRY(g->out_n.ensure((size_t)nslots * 4, &grew));
  HIP_TRY(g->out_err.ensure((size_t)nslots * 4, &grew));
  g->out_rows = nslots;
  if (grew) {
    ++g->gen;
    if (!g->noalloc) g->drop_graphs();
  }
  HIP_TRY(hipMemcpyAsync(g->slots.p, sd.data(), nslots * sizeof(SlotDesc), hipMemcpyHostToDevice,
                         g->stream));
  HIP_TRY(hipMemcpyAsync(g->tasks.p, tasks.data(), tasks.size() * sizeof(Task),
                         hipMemcpyHostToDevice, g->stream));
  DevCtx c = make_ctx(g);
  {
    std::vector<Sched> sch(nslots);
    for (auto& x : sch) {
      std::memset(&x, 0, sizeof(x));
      x.thr = x.delta = g->sssp_band_eff > 0 ? g->sssp_band_eff : INF / 2;
      x.prev = 0;
      x.mindef = INF;
      x.R = INF;
    }
    HIP_TRY(hipMemcpyAsync(g->sched.p, sch.data(), nslots * sizeof(Sched), hipMemcpyHostToDevice,
                           g->stream));
    HIP_TRY(hipMemcpyAsync(g->iter.as<int32_t>() + 1, &nslots, sizeof(int32_t),
                           hipMemcpyHostToDevice, g->stream));
    g->h_nslots[0] = nslots;
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  ++g->batch_no;
  HIP_TRY(hipEventRecord(g->ph[0], g->stream));
  const Task* dt = g->tasks.as<Task>();
  const int nt = (int)tasks.size();
  uint32_t* cnt = g->cnt.as<uint32_t>();
  double sssp_ms = 0, dp_ms = 0;
  int64_t dp_launches = 0, sssp_launches = 0;
  {
    const int rc = onelabel_stage(g, c, tasks, nslots, g->sssp_slots > 0 ? g->sssp_slots : nslots,
                                  &sssp_ms, &sssp_launches);
    if (rc) return rc;  // includes 100 = overflow
  }
  launch_lb_table(c, nslots, g->stream);
  // pending masks start empty (an overflowed earlier run may have left bits)
  launch_fill_i32(reinterpret_cast<int32_t*>(g->Ms.p), m_tot, 0, g->stream);
  launch_dp_seed(c, dt, nt, g->goff.as<int64_t>(), g->gmem.as<int32_t>(), g->wl0.as<uint64_t>(),
                 cnt, g->stream);
  launch_dp_sched_init(c, nslots, g->stream);
  launch_fill_i32(g->cnt.as<int32_t>() + 2, 2, 0, g->stream);   // far list counts
  launch_fill_i32(g->iter.as<int32_t>() + 2, 1, 0, g->stream);  // far list in use
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(g->ph[2], g->stream));
  int rc = run_phase(g, qs[0].D >= 0 ? 2 : 1, &dp_ms, &dp_launches);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(g->ph[3], g->stream));
  std::vector<uint64_t> ctr(C_NUM);
  HIP_TRY(hipMemcpyAsync(ctr.data(), g->ctr.p, C_NUM * 8, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (ctr[C_OVERFLOW]) return 100;  // worklist too small: caller retries with full capacity
  {
    std::vector<int32_t> all(nslots);
    for (int s = 0; s < nslots; ++s) all[s] = s;
    int crc = collect_results(g, c, qs, all, nullptr, out);
    if (crc) return crc;
  }
  // again: the tree recovery counts its backpointer misses
  HIP_TRY(hipMemcpy(ctr.data(), g->ctr.p, C_NUM * 8, hipMemcpyDeviceToHost));
  float ph_ms[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i) HIP_TRY(hipEventElapsedTime(&ph_ms[i], g->ph[i], g->ph[i + 1]));
  if (st) {
    st->batches += 1;
    st->queries += nslots;
    st->sssp_iterations += (int64_t)ctr[C_SSSP_ITERS];
    st->dp_iterations += (int64_t)ctr[C_DP_ITERS];
    st->sssp_items += (int64_t)ctr[C_SSSP_ITEMS];
    st->sssp_relax += (int64_t)ctr[C_SSSP_RELAX];
    st->sssp_updates += (int64_t)ctr[C_SSSP_UPD];
    st->sssp_items_unbounded += (int64_t)ctr[C_SSSP_NOUB];
    st->states_expanded += (int64_t)ctr[C_F];
    st->grow_relax += (int64_t)ctr[C_RG];
    st->merge_relax += (int64_t)ctr[C_RM];
    st->improvements += (int64_t)ctr[C_U];
    st->lb_terms += (int64_t)ctr[C_LB];
    st->ub_rows += (int64_t)ctr[C_UBROW];
    st->bp_fallbacks += (int64_t)ctr[C_BP_MISS];
    st->backpointers_stored = g->bp_call ? 1 : 0;
    st->items_processed += (int64_t)ctr[C_ITEMS];
    st->max_frontier = std::max<int64_t>(st->max_frontier, (int64_t)ctr[C_MAX_FRONTIER]);
    st->expand_ms += dp_ms;
    st->expand_launches += dp_launches;
    st->sssp_expand_launches += sssp_launches;
    st->sssp_expand_ms += sssp_ms;
    st->phase_ms[0] += ph_ms[0];
    st->phase_ms[1] += ph_ms[1];
    st->phase_ms[2] += ph_ms[2];
  }
  return GST_OK;
}

// First-fit allocator over an arena of elements; a freed region merges with
// free neighbours.
struct Arena {
  std::map<int64_t, int64_t> free_;  // offset -> size
  void reset(int64_t n) {
    free_.clear();
    if (n > 0) free_[0] = n;
  }
  int64_t take(int64_t n) {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        const int64_t off = it->first, rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        return off;
      }
    return -1;
  }
  void give(int64_t off, int64_t n) {
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

// Device bytes a query holds for the whole of a rolling run: one-label far
// keys, distances and group masks per vertex, route tables, group distances,
// best word, descriptor, band state, tasks, slot-list entries.
int64_t rolling_query_bytes(const gst_graph* g) {
  return (int64_t)g->Vp * 4 * (2 + CROW) + RTABLE * 4 + VV * 4 + 8 + (int64_t)sizeof(SlotDesc) +
         (int64_t)sizeof(Sched) + 2 * GST_MAX_K * (int64_t)sizeof(Task) + 8;
}

// Rolling admission: the queries of one mode (plain or hop-bounded, at most
// MAX_SLOTS) whose state tables do not fit the budget at once. The one-label
// phase runs for every query first (in passes whose lists fit a share of the
// budget); each query's distances, route tables and upper bound then stay
// resident, about 40 bytes per vertex. The DP runs over a pool of table
// regions cut from one arena: a query is admitted (completion bounds written,
// pending masks cleared, seeds queued into the frontier the running queries
// share) as soon as a region fits it, in call order, and collected (tree
// recovered, region freed) at the first host check after it has converged,
// which lets the next query in. Convergence is found on the device
// (k_collect_done): in the unordered DP a query that expanded nothing in the
// last iteration has nothing queued any more; in the banded DP a query with
// no live far-list state when the near list has drained.
// Returns 100 when a worklist overflowed; the queries not collected yet are
// then in *left for a re-run with full-size lists.
int solve_rolling(gst_graph* g, const std::vector<QueryIn>& qs, int64_t budget, BatchOut& out,
                  gst_stats* st, std::vector<QueryIn>* left) {

  const int nslots = (int)qs.size();
  const int64_t Vp = g->Vp, nch = g->nch();
  const bool hop = qs[0].D >= 0;
  const int maxe = g->max_tree_edges;
  std::vector<SlotDesc> sd(nslots);
  std::vector<Task> tasks;
  std::vector<int64_t> first_task(nslots + 1, 0);
  int64_t need_max = 0, need_all = 0;
  double rmax = 0;  // pending-mask words per state, largest over the queries
  for (int s = 0; s < nslots; ++s) {
    const QueryIn& q = qs[s];
    SlotDesc d;
    std::memset(&d, 0, sizeof(d));
    d.e_off = (int64_t)s * Vp;
    d.c_off = (int64_t)s * Vp * CROW;
    d.r_off = (int64_t)s * RTABLE;
    d.k = q.k;
    d.full = (1 << q.k) - 1;
    d.D = q.D;
    d.nsub = 1 << q.k;
    d.vv_off = s * VV;
    d.nrows = rows_of(q.k, q.D);
    d.mw = mask_words(q.k, q.D);
    d.fin = fin_slot(q.k, q.D) ? 1 : 0;
    sd[s] = d;
    first_task[s] = (int64_t)tasks.size();
    for (int i = 0; i < q.k; ++i) tasks.push_back(Task{s, i, q.groups[i], 0});
    const int64_t need = (int64_t)d.nrows * Vp;
    need_max = std::max(need_max, need);
    need_all += need;
    rmax = std::max(rmax, (double)mask_words(q.k, q.D) / (double)(1 << q.k));
  }
  first_task[nslots] = (int64_t)tasks.size();
  g->dp_div_eff = g->dp_div_call >= 0
                      ? g->dp_div_call
                      : (nslots >= kAutoBandSlots && !hop ? kAutoBandDiv : 0);
  g->sssp_band_eff = g->sssp_band_call >= 0 ? g->sssp_band_call : kAutoSsspBand;
  const bool far_lists = g->dp_div_eff > 0 || g->sssp_band_eff > 0;
  const int nl = far_lists ? 5 : 3;  // lists: two frontier lists, state list, two far lists
  // Budget: resident per-query arrays and tree outputs first, then the
  // one-label lists of a pass (at most 15% of what is left), then the arena
  // and the DP's lists, which hold 4% of the arena's states as in a lockstep
  // batch (a rolling run that overflows them restarts its resident queries,
  // half as many at a time).
  const double wl_frac = kWorklistFraction;
  const int64_t per_query = rolling_query_bytes(g);
  // trees are recovered in pieces of up to 64 queries, whose output rows
  // take at most 5% of the budget
  const int orows = (int)std::max<int64_t>(
      1, std::min<int64_t>(std::min(nslots, 64), budget / 20 / ((int64_t)maxe * 12)));
  const int64_t avail = budget - per_query * nslots - (int64_t)orows * maxe * 12;
  int per_pass = nslots;
  if (g->sssp_slots > 0) per_pass = std::min(per_pass, (int)g->sssp_slots);
  per_pass = (int)std::max<int64_t>(
      1, std::min<int64_t>(per_pass, (int64_t)(0.15 * (double)avail) / (nl * 8 * nch + 1)));
  const int64_t lists1 = nch * per_pass;
  const double ratio = (double)nch / (double)Vp;  // items per state row
  const double per_state = 8.0 + (g->bp_call ? 4.0 : 0.0) + 4.0 * rmax;
  double A = ((double)avail - nl * 8.0 * (double)lists1) / per_state;
  if (wl_frac * A * ratio > (double)lists1)
    A = (double)avail / (per_state + nl * 8.0 * wl_frac * ratio);
  int64_t arena = std::min<int64_t>((int64_t)A, need_all);
  // A state table from an earlier call within 5% of this size is kept as it
  // is: freeing and mapping a table of hundreds of GB again costs seconds.
  {
    const int64_t held = (int64_t)(g->T.bytes / 8);
    if (held >= need_max && (double)held >= 0.95 * (double)arena &&
        (double)held <= 1.05 * (double)arena)
      arena = std::min(held, need_all);
  }
  if (arena < need_max)
    return fail(GST_ERR_OOM, "a single query exceeds the HBM budget");
  int64_t cap = std::max<int64_t>(lists1, (int64_t)(wl_frac * (double)arena * ratio)) + 4096;
  if (g->list_permille > 0)  // no floors, as for a lockstep batch
    cap = (int64_t)((double)arena * ratio) * g->list_permille / 1000 + 64;
  if (cap > kListClamp + 4096) cap = kListClamp + 4096;
  if (g->wl_cap > 0 && (double)g->wl_cap >= 0.95 * (double)cap &&
      (double)g->wl_cap <= 1.05 * (double)cap && g->wl_cap >= lists1 + 4096)
    cap = g->wl_cap;  // lists from an earlier call of about this size are kept
  const int64_t arena_m = (int64_t)(rmax * (double)arena) + 64;

  bool grew = false;
  HIP_TRY(g->T.ensure((size_t)arena * 8, &grew));
  if (g->bp_call) HIP_TRY(g->BPs.ensure((size_t)arena * 4, &grew));
  HIP_TRY(g->Ms.ensure((size_t)arena_m * 4, &grew));
  bool e_grew = false;
  HIP_TRY(g->Es.ensure((size_t)nslots * Vp * 4, &e_grew));
  if (e_grew) HIP_TRY(hipMemsetAsync(g->Es.p, 0, g->Es.bytes, g->stream));
  grew |= e_grew;
  HIP_TRY(g->C.ensure((size_t)nslots * Vp * CROW * 4, &grew));
  HIP_TRY(g->GMs.ensure((size_t)nslots * Vp * 4, &grew));
  HIP_TRY(g->R.ensure((size_t)nslots * RTABLE * 4, &grew));
  HIP_TRY(g->VVb.ensure((size_t)nslots * VV * 4, &grew));
  HIP_TRY(g->bestw.ensure((size_t)nslots * 8, &grew));
  HIP_TRY(g->slots.ensure((size_t)nslots * sizeof(SlotDesc), &grew));
  HIP_TRY(g->sched.ensure((size_t)nslots * sizeof(Sched), &grew));
  HIP_TRY(g->tasks.ensure(tasks.size() * sizeof(Task) + 16, &grew));
  HIP_TRY(g->rsel.ensure((size_t)(2 * nslots + 16) * 4, &grew));
  HIP_TRY(g->rbest.ensure((size_t)nslots * 8, &grew));
  HIP_TRY(g->rtasks.ensure(tasks.size() * sizeof(Task) + 16, &grew));
  if ((uint64_t)cap * 8 > g->wl0.bytes || (uint64_t)cap * 8 > g->wl1.bytes ||
      (uint64_t)cap * 8 > g->wls.bytes) {
    HIP_TRY(g->wl0.ensure((size_t)cap * 8, &grew));
    HIP_TRY(g->wl1.ensure((size_t)cap * 8, &grew));
    HIP_TRY(g->wls.ensure((size_t)cap * 8, &grew));
  }
  {
    const int64_t key = ((int64_t)g->dp_div_eff << 32) | ((uint32_t)g->sssp_band_eff << 1) |
                        (g->bp_call ? 1 : 0);
    if (key != g->eff_key) {
      g->eff_key = key;
      grew = true;
    }
  }
  if (far_lists &&
      ((uint64_t)cap * 8 > g->far0.bytes || (uint64_t)cap * 8 > g->far1.bytes)) {
    HIP_TRY(g->far0.ensure((size_t)cap * 8, &grew));
    HIP_TRY(g->far1.ensure((size_t)cap * 8, &grew));
  }
  size_t list_bytes = std::min({g->wl0.bytes, g->wl1.bytes, g->wls.bytes});
  if (far_lists) list_bytes = std::min({list_bytes, g->far0.bytes, g->far1.bytes});
  const uint32_t new_cap = (uint32_t)std::min<size_t>(list_bytes / 8, UINT32_MAX - 65536);
  if (new_cap != g->wl_cap) {
    g->wl_cap = new_cap;
    grew = true;
  }
  HIP_TRY(g->out_uv.ensure((size_t)orows * maxe * 8, &grew));
  HIP_TRY(g->out_eid.ensure((size_t)orows * maxe * 4, &grew));
  HIP_TRY(g->out_n.ensure((size_t)orows * 4, &grew));
  HIP_TRY(g->out_err.ensure((size_t)orows * 4, &grew));
  g->out_rows = orows;
  if (grew) {
    ++g->gen;
    g->drop_graphs();
  }
  HIP_TRY(hipMemcpyAsync(g->slots.p, sd.data(), nslots * sizeof(SlotDesc), hipMemcpyHostToDevice,
                         g->stream));
  HIP_TRY(hipMemcpyAsync(g->tasks.p, tasks.data(), tasks.size() * sizeof(Task),
                         hipMemcpyHostToDevice, g->stream));
  DevCtx c = make_ctx(g);
  {
    std::vector<Sched> sch(nslots);
    for (auto& x : sch) {
      std::memset(&x, 0, sizeof(x));
      x.thr = x.delta = g->sssp_band_eff > 0 ? g->sssp_band_eff : INF / 2;
      x.mindef = INF;
      x.R = INF;
    }
    HIP_TRY(hipMemcpyAsync(g->sched.p, sch.data(), nslots * sizeof(Sched), hipMemcpyHostToDevice,
                           g->stream));
    HIP_TRY(hipMemcpyAsync(g->iter.as<int32_t>() + 1, &nslots, sizeof(int32_t),
                           hipMemcpyHostToDevice, g->stream));
    g->h_nslots[0] = nslots;
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  ++g->batch_no;
  HIP_TRY(hipEventRecord(g->ph[0], g->stream));
  double sssp_ms = 0, dp_ms = 0;
  int64_t dp_launches = 0, sssp_launches = 0;
  auto all_left = [&]() {
    if (!left) return;
    left->clear();
    for (int s = 0; s < nslots; ++s) left->push_back(qs[s]);
  };
  {
    const int rc = onelabel_stage(g, c, tasks, nslots, per_pass, &sssp_ms, &sssp_launches);
    if (rc == 100) all_left();
    if (rc) return rc;
  }
  // ---- the DP over a rolling pool of table regions
  // upper bounds as the one-label phase left them (a restarted query starts from these)
  HIP_TRY(hipMemcpyAsync(g->rbest.p, g->bestw.p, (size_t)nslots * 8, hipMemcpyDeviceToDevice,
                         g->stream));
  launch_fill_i32(g->cnt.as<int32_t>(), kCntWords, 0, g->stream);
  launch_fill_i32(g->iter.as<int32_t>() + 2, 1, 0, g->stream);  // far list in use
  Arena at, am;
  at.reset(arena);
  am.reset(arena_m);
  std::vector<char> state(nslots, 0);  // 0 waiting, 1 resident, 2 collected
  std::vector<int64_t> mwords(nslots);
  std::deque<int32_t> waiting;  // queries not admitted yet, in admission order
  for (int s = 0; s < nslots; ++s) waiting.push_back(s);
  int32_t* d_sel = g->rsel.as<int32_t>();
  int32_t* d_done = d_sel + nslots;
  uint32_t* d_ndone = reinterpret_cast<uint32_t*>(d_sel + 2 * nslots);
  std::vector<int32_t> sel;
  std::vector<Task> new_tasks;
  int64_t rounds = 0, max_res = 0;
  g->resident = 0;
  int cap_res = g->max_batch > 0 ? g->max_batch : nslots;
  uint32_t* cnt = g->cnt.as<uint32_t>();
  // admits waiting queries, in call order, while their regions fit
  auto admit = [&]() -> int {
    sel.clear();
    new_tasks.clear();
    while (!waiting.empty() && g->resident < cap_res) {
      const int s = waiting.front();
      const int64_t nt = (int64_t)sd[s].nrows * Vp;
      const int64_t nm = slot_mask_words(g, qs[s].k, qs[s].D);
      const int64_t ot = at.take(nt);
      if (ot < 0) break;
      const int64_t om = am.take(nm);
      if (om < 0) {
        at.give(ot, nt);
        break;
      }
      sd[s].t_off = ot;
      sd[s].m_off = om;
      mwords[s] = nm;
      state[s] = 1;
      sel.push_back(s);
      for (int64_t t = first_task[s]; t < first_task[s + 1]; ++t) new_tasks.push_back(tasks[t]);
      ++g->resident;
      waiting.pop_front();
    }
    if (sel.empty()) return GST_OK;
    ++rounds;
    max_res = std::max<int64_t>(max_res, g->resident);
    HIP_TRY(hipMemcpyAsync(g->slots.p, sd.data(), nslots * sizeof(SlotDesc),
                           hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(d_sel, sel.data(), sel.size() * 4, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(g->rtasks.p, new_tasks.data(), new_tasks.size() * sizeof(Task),
                           hipMemcpyHostToDevice, g->stream));
    launch_lb_table(c, (int)sel.size(), g->stream, d_sel);
    for (int s : sel)  // pending masks start empty
      launch_fill_i32(reinterpret_cast<int32_t*>(g->Ms.p) + sd[s].m_off, mwords[s], 0, g->stream);
    launch_dp_seed(c, g->rtasks.as<Task>(), (int)new_tasks.size(), g->goff.as<int64_t>(),
                   g->gmem.as<int32_t>(), g->wl0.as<uint64_t>(), cnt, g->stream);
    launch_dp_sched_init(c, (int)sel.size(), g->stream, d_sel);
    HIP_TRY(hipGetLastError());
    return GST_OK;
  };
  {
    const int rc = admit();
    if (rc) return rc;
  }
  if (g->resident == 0) return fail(GST_ERR_OOM, "a single query exceeds the HBM budget");
  HIP_TRY(hipEventRecord(g->ph[2], g->stream));
  const int mode = hop ? 2 : 1;
  const bool far_mode = c.far_mode != 0;
  int pairs = 1;
  // a host check every second iteration: a converged query leaves, and an
  // idle query moves to its next band, at most one iteration later (a check
  // costs tens of microseconds, an iteration of a pool this large milliseconds)
  const int max_pairs = 1;
  int64_t total_pairs = 0;
  std::vector<int32_t> done;
  while (true) {
    {
      const int prc = phase_round(g, c, mode, pairs, &dp_ms, &dp_launches);
      if (prc) return prc;
    }
    total_pairs += pairs;
    HIP_TRY(hipMemcpyAsync(g->h_cnt, cnt, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(g->h_cnt + 2, g->ctr.as<uint64_t>() + C_OVERFLOW, sizeof(uint64_t),
                           hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    uint64_t overflow;
    std::memcpy(&overflow, g->h_cnt + 2, sizeof(overflow));
    if (overflow && g->resident > 1) {
      // A list overflowed: the resident queries start their DP again, fewer
      // at a time (half as many may hold tables), from their one-label
      // results and their upper bounds as the one-label phase left them.
      if (st) ++st->reruns;
      std::vector<int32_t> back;
      for (int s = 0; s < nslots; ++s)
        if (state[s] == 1) back.push_back(s);
      for (auto it = back.rbegin(); it != back.rend(); ++it) {
        const int s = *it;
        state[s] = 0;
        at.give(sd[s].t_off, (int64_t)sd[s].nrows * Vp);
        am.give(sd[s].m_off, mwords[s]);
        waiting.push_front(s);
        HIP_TRY(hipMemcpyAsync(g->bestw.as<uint64_t>() + s, g->rbest.as<uint64_t>() + s, 8,
                               hipMemcpyDeviceToDevice, g->stream));
        launch_fill_i32(&g->sched.as<Sched>()[s].phase, 1, 0, g->stream);
      }
      cap_res = std::max(1, g->resident / 2);
      g->resident = 0;
      launch_fill_i32(g->cnt.as<int32_t>(), kCntWords, 0, g->stream);
      launch_fill_i32(g->iter.as<int32_t>() + 2, 1, 0, g->stream);  // far list in use
      launch_fill_u64(g->ctr.as<uint64_t>() + C_OVERFLOW, 1, 0, g->stream);
      {
        const int rc = admit();
        if (rc) return rc;
      }
      pairs = 1;
      continue;
    }
    if (overflow) {  // a lone query's frontier: the caller re-runs with full-size lists
      if (left) {
        left->clear();
        for (int s = 0; s < nslots; ++s)
          if (state[s] != 2) left->push_back(qs[s]);
      }
      g->resident = 0;
      return 100;
    }
    // Which converged queries to collect: cmode 0 = those that expanded
    // nothing in the last iteration (nothing can reach them: no far list, or
    // it is empty), 1 = those a band refill just found without live far-list
    // states, 2 = every resident query (both lists are empty).
    int cmode = -1;
    if (!far_mode) {
      cmode = g->h_cnt[0] == 0 ? 2 : 0;
    } else {
      int32_t fsel = 0;
      uint32_t fc[2] = {0, 0};
      HIP_TRY(hipMemcpyAsync(fc, cnt + 2, sizeof(fc), hipMemcpyDeviceToHost, g->stream));
      HIP_TRY(hipMemcpyAsync(&fsel, g->iter.as<int32_t>() + 2, sizeof(fsel),
                             hipMemcpyDeviceToHost, g->stream));
      HIP_TRY(hipStreamSynchronize(g->stream));
      const uint32_t nfar = fc[fsel & 1];
      if (nfar == 0) {
        cmode = g->h_cnt[0] == 0 ? 2 : 0;
      } else {
        // Band refills per query: a query whose near-list work is done moves
        // to its next band now, not when every resident query's band has
        // drained (a query admitted late would otherwise advance one band
        // per drain of the whole pool).
        uint32_t idle = 1;
        if (g->h_cnt[0] != 0) {
          launch_fill_i32(reinterpret_cast<int32_t*>(d_ndone), 1, 0, g->stream);
          launch_idle_count(c, d_ndone, g->stream);
          HIP_TRY(hipMemcpyAsync(&idle, d_ndone, sizeof(idle), hipMemcpyDeviceToHost, g->stream));
          HIP_TRY(hipStreamSynchronize(g->stream));
        }
        if (idle > 0) {
          launch_far_refill(c, nslots, nfar, g->wl0.as<uint64_t>(), cnt, g->stream,
                            g->h_cnt[0] != 0);
          HIP_TRY(hipGetLastError());
          cmode = 1;
        }
      }
    }
    if (cmode >= 0) {
      launch_fill_i32(reinterpret_cast<int32_t*>(d_ndone), 1, 0, g->stream);
      launch_collect_done(c, cmode, d_done, d_ndone, g->stream);
      HIP_TRY(hipGetLastError());
      uint32_t nd = 0;
      HIP_TRY(hipMemcpyAsync(&nd, d_ndone, sizeof(nd), hipMemcpyDeviceToHost, g->stream));
      HIP_TRY(hipStreamSynchronize(g->stream));
      if (nd > 0) {
        done.resize(nd);
        HIP_TRY(hipMemcpy(done.data(), d_done, nd * 4, hipMemcpyDeviceToHost));
        const int crc = collect_results(g, c, qs, done, d_done, out);
        if (crc) return crc;
        for (int s : done) {
          if (state[s] != 1) return fail(GST_ERR_INTERNAL, "rolling admission: slot collected twice");
          state[s] = 2;
          at.give(sd[s].t_off, (int64_t)sd[s].nrows * Vp);
          am.give(sd[s].m_off, mwords[s]);
          --g->resident;
        }
        const int rc = admit();
        if (rc) return rc;
      }
    }
    if (g->resident == 0) {
      if (waiting.empty()) break;
      return fail(GST_ERR_INTERNAL, "rolling admission: no query could be admitted");
    }
    if (total_pairs > (int64_t)1 << 20)
      return fail(GST_ERR_INTERNAL, "frontier iteration limit exceeded");
    pairs = std::min(pairs * 2, max_pairs);
  }
  HIP_TRY(hipEventRecord(g->ph[3], g->stream));
  std::vector<uint64_t> ctr(C_NUM);
  HIP_TRY(hipMemcpyAsync(ctr.data(), g->ctr.p, C_NUM * 8, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->resident = 0;
  float ph_ms[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i) HIP_TRY(hipEventElapsedTime(&ph_ms[i], g->ph[i], g->ph[i + 1]));
  if (st) {
    st->batches += rounds;
    st->rolling_queries += nslots;
    st->max_resident = std::max<int64_t>(st->max_resident, max_res);
    st->queries += nslots;
    st->sssp_iterations += (int64_t)ctr[C_SSSP_ITERS];
    st->dp_iterations += (int64_t)ctr[C_DP_ITERS];
    st->sssp_items += (int64_t)ctr[C_SSSP_ITEMS];
    st->sssp_relax += (int64_t)ctr[C_SSSP_RELAX];
    st->sssp_updates += (int64_t)ctr[C_SSSP_UPD];
    st->sssp_items_unbounded += (int64_t)ctr[C_SSSP_NOUB];
    st->states_expanded += (int64_t)ctr[C_F];
    st->grow_relax += (int64_t)ctr[C_RG];
    st->merge_relax += (int64_t)ctr[C_RM];
    st->improvements += (int64_t)ctr[C_U];
    st->lb_terms += (int64_t)ctr[C_LB];
    st->ub_rows += (int64_t)ctr[C_UBROW];
    st->bp_fallbacks += (int64_t)ctr[C_BP_MISS];
    st->backpointers_stored = g->bp_call ? 1 : 0;
    st->items_processed += (int64_t)ctr[C_ITEMS];
    st->max_frontier = std::max<int64_t>(st->max_frontier, (int64_t)ctr[C_MAX_FRONTIER]);
    st->expand_ms += dp_ms;
    st->expand_launches += dp_launches;
    st->sssp_expand_launches += sssp_launches;
    st->sssp_expand_ms += sssp_ms;
    st->phase_ms[0] += ph_ms[0];
    st->phase_ms[1] += ph_ms[1];
    st->phase_ms[2] += ph_ms[2];
  }
  return GST_OK;
}

int solve_all(gst_graph* g, std::vector<QueryIn>& qs, int64_t budget, BatchOut& out,
              gst_stats* st) {
  if (!g->goff.p) return fail(GST_ERR_INVALID, "groups not set");
  g->resident = 0;  // a rolling run that failed may have left its count
  if (budget <= 0) {
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    budget = (int64_t)((double)(fr + g->held_bytes()) * 0.8);
  }
  const double frac = kWorklistFraction;
  // Stored backpointers cost 4 bytes per state; in automatic mode they are
  // used only when the whole call still fits one admission batch with them.
  if (g->bp_mode == 0) {
    g->bp_call = false;
  } else if (g->bp_mode == 1) {
    g->bp_call = true;
  } else {
    int64_t tot = 0;
    bool mixed = false;
    for (const QueryIn& q : qs) {
      tot += gst_query_footprint(g, q.k, q.D) + bp_bytes(g, q.k, q.D);
      mixed |= (q.D >= 0) != (qs[0].D >= 0);
    }
    const size_t cap_slots =
        g->max_batch > 0 ? std::min<size_t>(g->max_batch, MAX_SLOTS) : MAX_SLOTS;
    g->bp_call = !mixed && qs.size() <= cap_slots && tot <= budget;
  }
  // stable order: plain queries first, then hop queries (one mode per batch)
  std::stable_sort(qs.begin(), qs.end(),
                   [](const QueryIn& a, const QueryIn& b) { return (a.D >= 0) < (b.D >= 0); });
  // A batch whose worklists overflowed is solved again with worklists as
  // large as the state count, which cannot overflow; they are (4 lists x 8
  // bytes per state) larger, so it runs in sub-batches that fit the budget at
  // that size and the 32-bit list capacity.
  auto rerun_full = [&](const std::vector<QueryIn>& batch) -> int {
    if (st) ++st->reruns;
    g->release_batch_buffers();  // the held tables and the larger lists never coexist
    size_t j = 0;
    while (j < batch.size()) {
      std::vector<QueryIn> sub;
      int64_t u = 0, items = 0;
      while (j < batch.size()) {
        const QueryIn& q = batch[j];
        const int64_t it = slot_items(g, q.k, q.D);
        const int64_t f = gst_query_footprint(g, q.k, q.D) + 32 * it +
                          (g->bp_call ? bp_bytes(g, q.k, q.D) : 0);
        if (!sub.empty() && (u + f > budget || items + it > kListClamp)) break;
        u += f;
        items += it;
        sub.push_back(q);
        ++j;
      }
      const int rc = solve_lockstep(g, sub, 1.0, out, st);
      if (rc == 100)  // only a query with more states than the list can address
        return fail(GST_ERR_OOM, "query " + std::to_string(sub[0].idx) +
                                     ": its frontier exceeds the 2^32-item worklist");
      if (rc) return rc;
    }
    return GST_OK;
  };
  const size_t cap_slots = g->max_batch > 0 ? std::min<size_t>(g->max_batch, MAX_SLOTS) : MAX_SLOTS;
  size_t i = 0;
  while (i < qs.size()) {
    const bool hop = qs[i].D >= 0;
    // Rolling admission when the queries of this mode do not fit one batch.
    if (g->rolling && !g->noalloc) {
      size_t j = i;
      int64_t tot = 0, fmax = 0;
      while (j < qs.size() && (qs[j].D >= 0) == hop) {
        const int64_t f = gst_query_footprint(g, qs[j].k, qs[j].D) +
                          (g->bp_call ? bp_bytes(g, qs[j].k, qs[j].D) : 0);
        tot += f;
        fmax = std::max(fmax, f);
        ++j;
      }
      // the queries of a rolling run keep their one-label results resident
      // until their DP ends: at most half of the budget goes to those, and
      // room for two of the largest queries stays
      const int64_t per_run = std::min<int64_t>(
          MAX_SLOTS, std::min(budget / 2, budget - 2 * fmax) / rolling_query_bytes(g));
      if ((tot > budget || j - i > cap_slots) && per_run >= 2) {
        while (i < j) {
          const size_t e = std::min(j, i + (size_t)per_run);
          std::vector<QueryIn> chunk(qs.begin() + i, qs.begin() + e);
          std::vector<QueryIn> left;
          int rc = solve_rolling(g, chunk, budget, out, st, &left);
          if (rc == 100) rc = rerun_full(left);
          if (rc) return rc;
          i = e;
        }
        continue;
      }
    }
    std::vector<QueryIn> batch;
    int64_t used = 0;
    while (i < qs.size() && (qs[i].D >= 0) == hop && batch.size() < cap_slots) {
      const int64_t f = gst_query_footprint(g, qs[i].k, qs[i].D) +
                        (g->bp_call ? bp_bytes(g, qs[i].k, qs[i].D) : 0);
      if (!batch.empty() && used + f > budget) break;
      if (batch.empty() && f > budget)
        return fail(GST_ERR_OOM, "a single query exceeds the HBM budget");
      used += f;
      batch.push_back(qs[i]);
      ++i;
    }
    int rc = solve_lockstep(g, batch, frac, out, st);
    if (rc == 100) rc = rerun_full(batch);
    if (rc == 100) return fail(GST_ERR_INTERNAL, "worklist overflow at full capacity");
    if (rc) return rc;
  }
  return GST_OK;
}

// Device rows are sorted by (weight, neighbour): an expansion whose useful
// edges are those with weight <= some limit scans only a prefix of the row.
// Edge ids follow the entries (the caller's ids, or the caller's CSR index).
int upload_rows(gst_graph* g) {
  const int64_t E = g->E;
  g->perm.resize(E);
  for (int32_t v = 0; v < g->V; ++v) {
    const int32_t b = g->h_rowp[v], e = g->h_rowp[v + 1];
    for (int32_t j = b; j < e; ++j) g->perm[j] = j;
    std::sort(g->perm.begin() + b, g->perm.begin() + e, [&](int32_t x, int32_t y) {
      if (g->h_w[x] != g->h_w[y]) return g->h_w[x] < g->h_w[y];
      if (g->h_col[x] != g->h_col[y]) return g->h_col[x] < g->h_col[y];
      return x < y;
    });
  }
  std::vector<int32_t> col(E), w(E), id(E);
  g->maxw = 1;
  for (int64_t j = 0; j < E; ++j) {
    const int32_t o = g->perm[j];
    col[j] = g->h_col[o];
    w[j] = g->h_w[o];
    id[j] = g->h_ids.empty() ? o : g->h_ids[o];
    g->maxw = std::max(g->maxw, w[j]);
  }
  HIP_TRY(g->col.ensure((size_t)std::max<int64_t>(E, 1) * 4));
  HIP_TRY(g->wt.ensure((size_t)std::max<int64_t>(E, 1) * 4));
  HIP_TRY(g->eid.ensure((size_t)std::max<int64_t>(E, 1) * 4));
  if (E) {
    HIP_TRY(hipMemcpy(g->col.p, col.data(), E * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->wt.p, w.data(), E * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->eid.p, id.data(), E * 4, hipMemcpyHostToDevice));
  }
  g->has_eid = true;
  ++g->gen;
  g->drop_graphs();
  return GST_OK;
}

// Weight changes of a few CSR entries (caller's indices, new values): the
// changed entries must come in symmetric pairs (checked by the same hash as
// at creation, over the changed entries only); only their rows are re-sorted
// and re-uploaded. The device arrays keep their addresses, so captured
// graphs stay valid.
int update_entries(gst_graph* g, std::vector<std::pair<int64_t, int32_t>>& upd) {
  std::sort(upd.begin(), upd.end(),
            [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<std::pair<int64_t, int32_t>> ch;
  for (size_t i = 0; i < upd.size(); ++i) {
    if (i + 1 < upd.size() && upd[i + 1].first == upd[i].first) continue;  // last value wins
    if (g->h_w[upd[i].first] != upd[i].second) ch.push_back(upd[i]);
  }
  if (ch.empty()) return GST_OK;
  uint64_t f_out = 0, f_in = 0;
  std::vector<int32_t> rows;
  for (const auto& c : ch) {
    const int32_t v = (int32_t)(std::upper_bound(g->h_rowp.begin(), g->h_rowp.end(),
                                                 (int32_t)c.first) - g->h_rowp.begin()) - 1;
    const uint64_t u = (uint64_t)g->h_col[c.first], ww = (uint64_t)c.second;
    f_out += mix64((uint64_t)v * 0x9E3779B97F4A7C15ull ^ u * 0xC2B2AE3D27D4EB4Full ^
                   ww * 0x165667B19E3779F9ull);
    f_in += mix64(u * 0x9E3779B97F4A7C15ull ^ (uint64_t)v * 0xC2B2AE3D27D4EB4Full ^
                  ww * 0x165667B19E3779F9ull);
    rows.push_back(v);
  }
  if (f_out != f_in)
    return fail(GST_ERR_INVALID, "weight update is not symmetric: change both entries of an edge");
  for (const auto& c : ch) g->h_w[c.first] = c.second;
  std::sort(rows.begin(), rows.end());
  rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
  std::vector<int32_t> col, w, id;
  for (int32_t v : rows) {
    const int32_t b = g->h_rowp[v], e = g->h_rowp[v + 1];
    for (int32_t j = b; j < e; ++j) g->perm[j] = j;
    std::sort(g->perm.begin() + b, g->perm.begin() + e, [&](int32_t x, int32_t y) {
      if (g->h_w[x] != g->h_w[y]) return g->h_w[x] < g->h_w[y];
      if (g->h_col[x] != g->h_col[y]) return g->h_col[x] < g->h_col[y];
      return x < y;
    });
    col.resize(e - b);
    w.resize(e - b);
    id.resize(e - b);
    for (int32_t j = b; j < e; ++j) {
      const int32_t o = g->perm[j];
      col[j - b] = g->h_col[o];
      w[j - b] = g->h_w[o];
      id[j - b] = g->h_ids.empty() ? o : g->h_ids[o];
    }
    HIP_TRY(hipMemcpy(g->col.as<int32_t>() + b, col.data(), (e - b) * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->wt.as<int32_t>() + b, w.data(), (e - b) * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->eid.as<int32_t>() + b, id.data(), (e - b) * 4, hipMemcpyHostToDevice));
  }
  g->maxw = *std::max_element(g->h_w.begin(), g->h_w.end());
  return GST_OK;
}

template <class I>
int upload_graph(gst_graph* g, const I* beg, const I* csr, const I* wgt, int64_t V, int64_t E) {
  if (V < 1 || V >= (int64_t)MAX_V) return fail(GST_ERR_INVALID, "V out of range (1 .. 2^24-1)");
  if (E < 0 || E >= (int64_t)INT32_MAX) return fail(GST_ERR_INVALID, "E out of range");
  if ((int64_t)beg[0] != 0 || (int64_t)beg[V] != E)
    return fail(GST_ERR_INVALID, "beg_pos must start at 0 and end at E");
  g->h_rowp.resize(V + 1);
  std::vector<int32_t> col(E), w(E);
  uint64_t f_out = 0, f_in = 0;
  for (int64_t v = 0; v < V; ++v) {
    if (beg[v + 1] < beg[v]) return fail(GST_ERR_INVALID, "beg_pos not monotone");
    g->h_rowp[v] = (int32_t)beg[v];
    for (int64_t j = (int64_t)beg[v]; j < (int64_t)beg[v + 1]; ++j) {
      const int64_t u = (int64_t)csr[j], ww = (int64_t)wgt[j];
      if (u < 0 || u >= V) return fail(GST_ERR_INVALID, "neighbour id out of range");
      if (ww < 1 || ww > (1 << 20)) return fail(GST_ERR_INVALID, "weights must be in [1, 2^20]");
      col[j] = (int32_t)u;
      w[j] = (int32_t)ww;
      if (ww > g->maxw) g->maxw = (int32_t)ww;
      f_out += mix64((uint64_t)v * 0x9E3779B97F4A7C15ull ^ (uint64_t)u * 0xC2B2AE3D27D4EB4Full ^
                     (uint64_t)ww * 0x165667B19E3779F9ull);
      f_in += mix64((uint64_t)u * 0x9E3779B97F4A7C15ull ^ (uint64_t)v * 0xC2B2AE3D27D4EB4Full ^
                    (uint64_t)ww * 0x165667B19E3779F9ull);
    }
  }
  g->h_rowp[V] = (int32_t)E;
  if (f_out != f_in)
    return fail(GST_ERR_INVALID, "CSR is not symmetric: every edge must appear in both rows "
                                 "with the same weight");
  g->V = (int32_t)V;
  g->Vp = (int32_t)((V + 63) / 64 * 64);
  g->E = E;
  for (int64_t v = 0; v < V; ++v) {
    const int64_t deg = (int64_t)beg[v + 1] - (int64_t)beg[v];
    if (deg > ((int64_t)MAX_CHUNKS << 6))
      return fail(GST_ERR_INVALID, "vertex degree above 2^22 is not supported");
  }
  g->h_col = std::move(col);
  g->h_w = std::move(w);
  g->n_und = E;
  HIP_TRY(g->rowp.ensure((size_t)(V + 1) * 4));
  HIP_TRY(hipMemcpy(g->rowp.p, g->h_rowp.data(), (V + 1) * 4, hipMemcpyHostToDevice));
  return upload_rows(g);
}

}  // namespace

extern "C" {

int32_t gst_version(void) { return 2; }

const char* gst_last_error(void) { return g_last_error.c_str(); }

int gst_graph_create(const void* beg_pos, const void* csr, const void* weight, int64_t V,
                     int64_t E, int32_t index_bits, gst_graph** out) {
  if (!out || !beg_pos || (!csr && E) || (!weight && E))
    return fail(GST_ERR_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(GST_ERR_DEVICE, "no HIP device");
  gst_graph* g = new gst_graph;
  HIP_TRY(hipGetDevice(&g->device));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, g->device) == hipSuccess && prop.multiProcessorCount > 0)
    g->cus = prop.multiProcessorCount;
  int rc;
  if (index_bits == 64)
    rc = upload_graph(g, (const int64_t*)beg_pos, (const int64_t*)csr, (const int64_t*)weight, V, E);
  else if (index_bits == 32)
    rc = upload_graph(g, (const int32_t*)beg_pos, (const int32_t*)csr, (const int32_t*)weight, V, E);
  else
    rc = fail(GST_ERR_INVALID, "index_bits must be 32 or 64");
  if (rc) {
    gst_graph_destroy(g);
    return rc;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess ||
      hipEventCreate(&g->ph[0]) != hipSuccess || hipEventCreate(&g->ph[1]) != hipSuccess ||
      hipEventCreate(&g->ph[2]) != hipSuccess || hipEventCreate(&g->ph[3]) != hipSuccess ||
      hipHostMalloc((void**)&g->h_cnt, 16, 0) != hipSuccess) {
    gst_graph_destroy(g);
    return fail(GST_ERR_DEVICE, "stream / event / pinned allocation failed");
  }
  if (g->cnt.ensure(kCntWords * 4) != hipSuccess || g->iter.ensure(16) != hipSuccess ||
      g->ctr.ensure(C_NUM * 8) != hipSuccess) {
    gst_graph_destroy(g);
    return fail(GST_ERR_OOM, "counter allocation failed");
  }
  int32_t one[4] = {1, 0, 0, 0};
  if (hipMemcpy(g->iter.p, one, 16, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(g->cnt.p, 0, kCntWords * 4) != hipSuccess) {
    gst_graph_destroy(g);
    return fail(GST_ERR_DEVICE, "counter initialisation failed");
  }
  *out = g;
  return GST_OK;
}

int gst_graph_set_edge_ids(gst_graph* g, const int32_t* ids, int64_t num_undirected) {
  if (!g || !ids || num_undirected < 0) return fail(GST_ERR_INVALID, "bad argument");
  for (int64_t j = 0; j < g->E; ++j)
    if (ids[j] < 0 || ids[j] >= num_undirected) return fail(GST_ERR_INVALID, "edge id out of range");
  g->h_ids.assign(ids, ids + g->E);
  g->n_und = num_undirected;
  return upload_rows(g);
}

int gst_graph_set_weights(gst_graph* g, const int32_t* w) {
  if (!g || !w) return fail(GST_ERR_INVALID, "bad argument");
  int64_t changed = 0;
  for (int64_t j = 0; j < g->E; ++j) {
    if (w[j] < 1 || w[j] > (1 << 20)) return fail(GST_ERR_INVALID, "weights must be in [1, 2^20]");
    changed += w[j] != g->h_w[j];
  }
  HIP_TRY(hipSetDevice(g->device));
  if (changed * 8 < g->E) {  // a few entries: re-sort and re-upload their rows only
    std::vector<std::pair<int64_t, int32_t>> upd;
    upd.reserve(changed);
    for (int64_t j = 0; j < g->E; ++j)
      if (w[j] != g->h_w[j]) upd.push_back({j, w[j]});
    return update_entries(g, upd);
  }
  // weights must stay symmetric: entry (v,u) and (u,v) carry the same value
  uint64_t f_out = 0, f_in = 0;
  for (int32_t v = 0; v < g->V; ++v)
    for (int32_t j = g->h_rowp[v]; j < g->h_rowp[v + 1]; ++j) {
      const uint64_t u = (uint64_t)g->h_col[j], ww = (uint64_t)w[j];
      f_out += mix64((uint64_t)v * 0x9E3779B97F4A7C15ull ^ u * 0xC2B2AE3D27D4EB4Full ^
                     ww * 0x165667B19E3779F9ull);
      f_in += mix64(u * 0x9E3779B97F4A7C15ull ^ (uint64_t)v * 0xC2B2AE3D27D4EB4Full ^
                    ww * 0x165667B19E3779F9ull);
    }
  if (f_out != f_in) return fail(GST_ERR_INVALID, "weights are not symmetric");
  g->h_w.assign(w, w + g->E);
  return upload_rows(g);
}

int gst_graph_update_weights(gst_graph* g, const int64_t* csr_index, const int32_t* w,
                             int64_t n) {
  if (!g || n < 0 || (n && (!csr_index || !w))) return fail(GST_ERR_INVALID, "bad argument");
  std::vector<std::pair<int64_t, int32_t>> upd((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (csr_index[i] < 0 || csr_index[i] >= g->E)
      return fail(GST_ERR_INVALID, "CSR index out of range");
    if (w[i] < 1 || w[i] > (1 << 20)) return fail(GST_ERR_INVALID, "weights must be in [1, 2^20]");
    upd[i] = {csr_index[i], w[i]};
  }
  HIP_TRY(hipSetDevice(g->device));
  return update_entries(g, upd);
}

int gst_groups_set(gst_graph* g, const int64_t* off, const int32_t* mem, int32_t G) {
  if (!g || !off || G < 1) return fail(GST_ERR_INVALID, "bad argument");
  if (off[0] != 0) return fail(GST_ERR_INVALID, "group offsets must start at 0");
  for (int32_t i = 0; i < G; ++i)
    if (off[i + 1] < off[i]) return fail(GST_ERR_INVALID, "group offsets not monotone");
  const int64_t tot = off[G];
  for (int64_t m = 0; m < tot; ++m)
    if (mem[m] < 0 || mem[m] >= g->V) return fail(GST_ERR_INVALID, "group member out of range");
  g->G = G;
  g->h_goff.assign(off, off + G + 1);
  HIP_TRY(g->goff.ensure((size_t)(G + 1) * 8));
  HIP_TRY(g->gmem.ensure((size_t)std::max<int64_t>(tot, 1) * 4));
  HIP_TRY(hipMemcpy(g->goff.p, off, (G + 1) * 8, hipMemcpyHostToDevice));
  if (tot) HIP_TRY(hipMemcpy(g->gmem.p, mem, tot * 4, hipMemcpyHostToDevice));
  ++g->gen;
  return GST_OK;
}

int gst_set_option(gst_graph* g, int32_t opt, int64_t v) {
  if (!g) return fail(GST_ERR_INVALID, "null graph");
  switch (opt) {
    case GST_OPT_CHUNK:
      if (v != 0 && (v < 64 || v > 65536 || (v & (v - 1))))
        return fail(GST_ERR_INVALID, "chunk must be 0 (automatic) or a power of two in [64, 65536]");
      g->chunk_opt = v ? ilog2(v) : -1;
      break;
    case GST_OPT_GRID:
      if (v < 0 || v > 65536) return fail(GST_ERR_INVALID, "grid out of range");
      g->grid = (int32_t)v;
      break;
    case GST_OPT_ITERS_PER_SYNC:
      if (v < 2 || v > 4096) return fail(GST_ERR_INVALID, "iters_per_sync out of range");
      g->iters_per_sync = (int32_t)v;
      break;
    case GST_OPT_HOP_PRUNE:
      g->hop_prune = v ? 1 : 0;
      break;
    case GST_OPT_PROFILE:
      g->profile = v ? 1 : 0;
      break;
    case GST_OPT_INFEASIBLE_COST:
      if (v < 0 || v > INT32_MAX) return fail(GST_ERR_INVALID, "cost out of range");
      g->infeasible_cost = (int32_t)v;
      break;
    case GST_OPT_MAX_TREE_EDGES:
      if (v < 1 || v > (1 << 24)) return fail(GST_ERR_INVALID, "max tree edges out of range");
      g->max_tree_edges = (int32_t)v;
      break;
    case GST_OPT_USE_GRAPHS:
      g->use_graphs = v ? 1 : 0;
      break;
    case GST_OPT_SSSP_DELTA:
      if (v < -1 || v > INF / 2) return fail(GST_ERR_INVALID, "band width out of range");
      g->sssp_band = (int32_t)v;
      break;
    case GST_OPT_LANE_APPENDS:
      if (v < 0 || v > 1) return fail(GST_ERR_INVALID, "lane appends is 0 or 1");
      g->lane_appends = (int32_t)v;
      break;
    case GST_OPT_PER_ITEM_EDGES:
      if (v < 0 || v > 1) return fail(GST_ERR_INVALID, "per-item edges is 0 or 1");
      g->per_item_edges = (int32_t)v;
      break;
    case GST_OPT_LIST_PERMILLE:
      if (v < 0 || v > 1000) return fail(GST_ERR_INVALID, "list permille out of range");
      g->list_permille = (int32_t)v;
      break;
    case GST_OPT_SSSP_SLOTS:
      if (v < 0 || v > (int64_t)MAX_SLOTS) return fail(GST_ERR_INVALID, "slot group out of range");
      g->sssp_slots = (int32_t)v;
      break;
    case GST_OPT_MAX_BATCH:
      if (v < 0 || v > (int64_t)MAX_SLOTS) return fail(GST_ERR_INVALID, "max batch out of range");
      g->max_batch = (int32_t)v;
      break;
    case GST_OPT_RADIUS_PCT:
      if (v < 0 || v > 100000) return fail(GST_ERR_INVALID, "radius percentage out of range");
      g->radius_pct = (int32_t)v;
      break;
    case GST_OPT_TRACE:
      g->trace = v ? 1 : 0;
      break;
    case GST_OPT_BACKPOINTERS:
      if (v < -1 || v > 1) return fail(GST_ERR_INVALID, "backpointer mode is -1, 0 or 1");
      g->bp_mode = (int32_t)v;
      break;
    case GST_OPT_ROLLING:
      if (v < 0 || v > 1) return fail(GST_ERR_INVALID, "rolling admission is 0 or 1");
      g->rolling = (int32_t)v;
      break;
    case GST_OPT_DP_DELTA_DIV:
      if (v < -1 || v > (1 << 20)) return fail(GST_ERR_INVALID, "band divisor out of range");
      g->dp_div = (int32_t)v;
      break;
    default:
      return fail(GST_ERR_INVALID, "unknown option");
  }
  ++g->gen;
  g->drop_graphs();
  return GST_OK;
}

int gst_trace_get(gst_graph* g, gst_trace_row* rows, int64_t cap, int64_t* n) {
  if (!g || !n || cap < 0 || (cap && !rows)) return fail(GST_ERR_INVALID, "bad argument");
  *n = (int64_t)g->trace_rows.size();
  const int64_t m = std::min<int64_t>(cap, *n);
  for (int64_t i = 0; i < m; ++i) rows[i] = g->trace_rows[i];
  return m < *n ? fail(GST_ERR_CAPACITY, "trace longer than the buffer; *n holds its length")
                : GST_OK;
}

int64_t gst_query_footprint(gst_graph* g, int32_t k, int32_t D) {
  if (!g || k < 1 || k > GST_MAX_K) return -1;
  const double wl_items =
      std::max(kWorklistFraction * (double)slot_items(g, k, D), (double)g->nch());
  // two frontier lists, the unfolded state list and the two far lists
  return slot_ints(g, k, D) * 4 + (int64_t)(5 * 8 * wl_items) +
         (int64_t)g->max_tree_edges * 12 + 1024;
}

}  // extern "C"

namespace {

int parse_queries(const gst_graph* g, const int64_t* q_offsets, const int32_t* q_groups,
                  const int32_t* q_D, int32_t nq, std::vector<QueryIn>& qs) {
  qs.resize(nq);
  for (int32_t q = 0; q < nq; ++q) {
    const int64_t k = q_offsets[q + 1] - q_offsets[q];
    const int32_t D = q_D ? q_D[q] : -1;
    if (k < 1 || k > GST_MAX_K)
      return fail(GST_ERR_INVALID, "query " + std::to_string(q) + ": k must be in [1, 7]");
    if (D > GST_MAX_D) return fail(GST_ERR_INVALID, "hop bound above 15");
    const int32_t* gr = q_groups + q_offsets[q];
    for (int64_t i = 0; i < k; ++i) {
      if (gr[i] < 0 || gr[i] >= g->G)
        return fail(GST_ERR_INVALID, "query " + std::to_string(q) + ": group id out of range");
      for (int64_t j = 0; j < i; ++j)
        if (gr[j] == gr[i])
          return fail(GST_ERR_INVALID, "query " + std::to_string(q) + ": repeated group id");
    }
    qs[q] = QueryIn{q, (int32_t)k, D < 0 ? -1 : D, gr};
  }
  return GST_OK;
}

// Seconds of one pass over the sample with the settings in `t` (one call
// with every query, or one call per query when `one_at_a_time`); the best
// of `reps` passes. Returns < 0 on error.
double time_pass(gst_graph* g, const std::vector<QueryIn>& qs, bool one_at_a_time,
                 const Tune& t, int reps, int* rc) {
  const int32_t save_chunk = g->chunk_opt, save_grid = g->grid, save_dp = g->dp_div,
                save_ss = g->sssp_band;
  g->chunk_opt = t.chunk_shift;
  g->grid = t.grid;
  g->dp_div = t.dp_div;
  g->sssp_band = t.sssp_band;
  double best = 1e30;
  BatchOut out;
  out.cost.assign(qs.size(), 0);
  out.status.assign(qs.size(), 0);
  out.uv.resize(qs.size());
  out.ids.resize(qs.size());
  *rc = GST_OK;
  for (int r = 0; r < reps && *rc == GST_OK; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    if (one_at_a_time) {
      for (size_t i = 0; i < qs.size() && *rc == GST_OK; ++i) {
        std::vector<QueryIn> one{qs[i]};
        one[0].idx = (int32_t)i;
        select_settings(g, 1);
        *rc = solve_all(g, one, 0, out, nullptr);
      }
    } else {
      std::vector<QueryIn> all(qs);
      for (size_t i = 0; i < all.size(); ++i) all[i].idx = (int32_t)i;
      select_settings(g, (int64_t)all.size());
      *rc = solve_all(g, all, 0, out, nullptr);
    }
    best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                              .count());
  }
  g->chunk_opt = save_chunk;
  g->grid = save_grid;
  g->dp_div = save_dp;
  g->sssp_band = save_ss;
  return *rc == GST_OK ? best : -1.0;
}

}  // namespace

extern "C" {

int gst_autotune(gst_graph* g, const int64_t* q_offsets, const int32_t* q_groups,
                 const int32_t* q_D, int32_t nq, int32_t batch_size, double max_seconds,
                 gst_tuning* out) {
  if (!g || nq < 0 || (nq && (!q_offsets || !q_groups)) || (batch_size != 0 && batch_size != 1))
    return fail(GST_ERR_INVALID, "bad argument");
  if (!g->goff.p) return fail(GST_ERR_INVALID, "groups not set");
  HIP_TRY(hipSetDevice(g->device));
  std::vector<QueryIn> qs;
  int rc = parse_queries(g, q_offsets, q_groups, q_D, nq, qs);
  if (rc) return rc;
  const bool single = batch_size == 1;
  const int cls = single ? 0 : size_class(nq);
  // start from what is in force for this class: cached choice or model
  Tune cur = model_settings(g, cls);
  if (g->tuned_ok[cls]) cur = g->tuned[cls];
  if (cur.chunk_shift < 0) cur.chunk_shift = model_settings(g, cls).chunk_shift;
  if (cur.grid <= 0) cur.grid = model_settings(g, cls).grid;
  gst_tuning res;
  std::memset(&res, 0, sizeof(res));
  res.size_class = cls;
  if (nq > 0 && max_seconds > 0) {
    const auto t_start = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    };
    double best = time_pass(g, qs, single, cur, 2, &rc);
    if (rc) return rc;
    res.seconds_before = best;
    int32_t maxdeg = 1;
    for (int32_t v = 0; v < g->V; ++v)
      maxdeg = std::max(maxdeg, g->h_rowp[v + 1] - g->h_rowp[v]);
    const bool any_plain =
        std::any_of(qs.begin(), qs.end(), [](const QueryIn& q) { return q.D < 0; });
    // coordinate descent: one parameter at a time, the others at their best
    for (int param = 0; param < 4; ++param) {
      std::vector<int32_t> cand;
      if (param == 0 && g->chunk_opt < 0) {
        for (int sh = 6; sh <= 14 && (sh <= 6 || (1 << (sh - 1)) < maxdeg); ++sh) cand.push_back(sh);
      } else if (param == 1 && g->dp_div < 0 && any_plain) {
        cand = {0, 64, 256, 1024};
      } else if (param == 2 && g->sssp_band < 0) {
        cand = {0, 8, 16, 32, 64};
      } else if (param == 3 && g->grid <= 0) {
        cand = {g->cus * 4, g->cus * 8, g->cus * 16};
      }
      for (int32_t x : cand) {
        if (elapsed() > max_seconds) break;
        Tune t = cur;
        int32_t* f = param == 0 ? &t.chunk_shift
                                : (param == 1 ? &t.dp_div : (param == 2 ? &t.sssp_band : &t.grid));
        const int32_t before = *f;
        *f = x;
        if (x == before) continue;
        const double sec = time_pass(g, qs, single, t, 2, &rc);
        if (rc) return rc;
        ++res.trials;
        if (sec < best) {
          best = sec;
          cur = t;
        }
      }
    }
    res.seconds_after = best;
    g->tuned[cls] = cur;
    g->tuned_ok[cls] = true;
  }
  select_settings(g, single ? 1 : nq);
  res.chunk = 1 << g->chunk_shift;
  res.grid = expand_grid(g);
  res.dp_div = g->dp_div_call;
  res.sssp_band = g->sssp_band_call;
  if (out) *out = res;
  return GST_OK;
}

int gst_solve_batch(gst_graph* g, const int64_t* q_offsets, const int32_t* q_groups,
                    const int32_t* q_D, int32_t nq, int64_t budget, int32_t* out_costs,
                    int32_t* out_status, int64_t* out_edge_offsets, int32_t* out_edge_uv,
                    int32_t* out_edge_ids, int64_t edge_cap, gst_stats* st) {
  auto t0 = std::chrono::steady_clock::now();
  if (!g || nq < 0 || (nq && (!q_offsets || !q_groups || !out_costs)))
    return fail(GST_ERR_INVALID, "bad argument");
  if (!g->goff.p) return fail(GST_ERR_INVALID, "groups not set");
  HIP_TRY(hipSetDevice(g->device));
  std::vector<QueryIn> qs;
  {
    int prc = parse_queries(g, q_offsets, q_groups, q_D, nq, qs);
    if (prc) return prc;
  }
  select_settings(g, nq);
  g->trace_rows.clear();
  g->batch_no = 0;
  BatchOut out;
  out.cost.assign(nq, 0);
  out.status.assign(nq, 0);
  out.uv.resize(nq);
  out.ids.resize(nq);
  gst_stats local;
  std::memset(&local, 0, sizeof(local));
  int rc = nq ? solve_all(g, qs, budget, out, &local) : GST_OK;
  if (rc) return rc;
  int64_t pos = 0;
  bool over = false;
  for (int32_t q = 0; q < nq; ++q) {
    out_costs[q] = out.cost[q];
    if (out_status) out_status[q] = out.status[q];
    if (out_edge_offsets) out_edge_offsets[q] = pos;
    for (size_t i = 0; i < out.uv[q].size(); ++i, ++pos) {
      if (pos >= edge_cap) {
        over = true;
        continue;
      }
      if (out_edge_uv) {
        out_edge_uv[2 * pos] = out.uv[q][i].first;
        out_edge_uv[2 * pos + 1] = out.uv[q][i].second;
      }
      if (out_edge_ids) out_edge_ids[pos] = out.ids[q][i];
    }
  }
  if (out_edge_offsets) out_edge_offsets[nq] = pos;
  local.seconds_total =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  local.bytes_alg = 16 * local.states_expanded + 12 * local.grow_relax + 8 * local.merge_relax +
                    12 * local.improvements + 4 * local.lb_terms;
  local.bytes_layout = 24 * local.states_expanded + 16 * local.grow_relax +
                       12 * local.merge_relax + 16 * local.improvements + 32 * local.ub_rows;
  if (st) *st = local;
  if (over && (out_edge_uv || out_edge_ids))
    return fail(GST_ERR_CAPACITY, "edge_cap too small; offsets hold the needed size");
  return GST_OK;
}

int gst_solve(gst_graph* g, const int32_t* query_groups, int32_t k, int32_t D, int32_t* out_cost,
              int32_t* out_edge_uv, int32_t edge_cap, int32_t* out_num_edges, int32_t* out_status,
              gst_stats* st) {
  if (!g || !query_groups || !out_cost) return fail(GST_ERR_INVALID, "bad argument");
  int64_t off[2] = {0, k};
  int64_t eoff[2] = {0, 0};
  int32_t status = 0;
  int rc = gst_solve_batch(g, off, query_groups, &D, 1, 0, out_cost, &status, eoff, out_edge_uv,
                           nullptr, out_edge_uv ? edge_cap : 0, st);
  if (out_num_edges) *out_num_edges = (int32_t)eoff[1];
  if (out_status) *out_status = status;
  return rc;
}

}  // extern "C"

namespace {

// Carves the allocation-free path's buffers from the scratch array, 256-byte
// aligned, in a fixed order; `base` null only measures.
struct Carve {
  char* base;
  size_t used = 0;
  void take(DBuf& b, size_t bytes) {
    used = (used + 255) & ~(size_t)255;
    b.p = base ? base + used : nullptr;
    b.bytes = bytes;
    b.borrowed = true;
    used += bytes;
  }
};

constexpr int kNoallocBufs = 20;

// The buffers of one slot at full worklist capacity (lists never overflow).
void carve_slot(gst_graph* g, int32_t k, int32_t D, DBuf* bufs[], Carve& cv) {
  const int64_t rows = rows_of(k, D);
  const int64_t items = slot_items(g, k, D);
  const int64_t cap = std::max<int64_t>(items, g->nch()) + 4096;
  const size_t sizes[] = {(size_t)rows * g->Vp * 8,        // T
                          (size_t)g->Vp * 4,               // Es
                          (size_t)slot_mask_words(g, k, D) * 4,  // Ms
                          (size_t)cap * 8,                 // wls
                          (size_t)g->Vp * CROW * 4,        // C
                          (size_t)g->Vp * 4,               // GMs
                          (size_t)RTABLE * 4,              // R
                          (size_t)VV * 4,                  // VVb
                          8,                               // bestw
                          sizeof(SlotDesc),                // slots
                          sizeof(Sched),                   // sched
                          (size_t)k * sizeof(Task) + 16,   // tasks
                          (size_t)cap * 8, (size_t)cap * 8, (size_t)cap * 8, (size_t)cap * 8,
                          (size_t)g->max_tree_edges * 8,   // out_uv
                          (size_t)g->max_tree_edges * 4,   // out_eid
                          4, 4};                           // out_n, out_err
  for (int i = 0; i < kNoallocBufs; ++i) cv.take(*bufs[i], sizes[i]);
}

DBuf* const* slot_bufs(gst_graph* g, DBuf* out[kNoallocBufs]) {
  DBuf* b[kNoallocBufs] = {&g->T, &g->Es, &g->Ms, &g->wls, &g->C, &g->GMs, &g->R, &g->VVb,
                 &g->bestw, &g->slots, &g->sched, &g->tasks, &g->wl0,  &g->wl1,
                 &g->far0,  &g->far1,  &g->out_uv, &g->out_eid, &g->out_n, &g->out_err};
  for (int i = 0; i < kNoallocBufs; ++i) out[i] = b[i];
  return out;
}

}  // namespace

extern "C" {

int64_t gst_noalloc_bytes(gst_graph* g, int32_t k, int32_t D) {
  if (!g || k < 1 || k > GST_MAX_K || D > GST_MAX_D) return -1;
  const bool save = g->noalloc;
  g->noalloc = true;
  select_settings(g, 1);
  g->noalloc = save;
  DBuf tmp[kNoallocBufs];
  DBuf* bufs[kNoallocBufs];
  for (int i = 0; i < kNoallocBufs; ++i) bufs[i] = &tmp[i];
  Carve cv{nullptr};
  carve_slot(g, k, D < 0 ? -1 : D, bufs, cv);
  return (int64_t)cv.used;
}

int gst_solve_noalloc(gst_graph* g, const int32_t* query_groups, int32_t k, int32_t D,
                      int32_t* out_cost, int32_t* out_edge_uv, int32_t edge_cap,
                      int32_t* out_num_edges, int32_t* out_status, gst_stats* st) {
  auto t0 = std::chrono::steady_clock::now();
  if (!g || !query_groups || !out_cost || edge_cap < 0) return fail(GST_ERR_INVALID, "bad argument");
  if (!g->goff.p) return fail(GST_ERR_INVALID, "groups not set");
  HIP_TRY(hipSetDevice(g->device));
  const int64_t off[2] = {0, k};
  std::vector<QueryIn> qs;
  int rc = parse_queries(g, off, query_groups, &D, 1, qs);
  if (rc) return rc;
  void* base = nullptr;
  HIP_TRY(noalloc_scratch(&base));
  g->noalloc = true;
  select_settings(g, 1);
  DBuf* bufs[kNoallocBufs];
  slot_bufs(g, bufs);
  Carve measure{nullptr};
  {
    DBuf tmp[kNoallocBufs];
    DBuf* tb[kNoallocBufs];
    for (int i = 0; i < kNoallocBufs; ++i) tb[i] = &tmp[i];
    carve_slot(g, qs[0].k, qs[0].D, tb, measure);
  }
  if (measure.used > (size_t)GST_NOALLOC_SCRATCH_BYTES) {
    g->noalloc = false;
    return fail(GST_ERR_OOM, "query needs " + std::to_string(measure.used) +
                                 " bytes of scratch; the allocation-free path holds " +
                                 std::to_string((unsigned long long)GST_NOALLOC_SCRATCH_BYTES));
  }
  // the graph's own buffers step aside for the call and come back unchanged
  DBuf saved[kNoallocBufs];
  for (int i = 0; i < kNoallocBufs; ++i) {
    saved[i] = *bufs[i];
    *bufs[i] = DBuf();
  }
  const int32_t save_graphs = g->use_graphs;
  const bool save_bp = g->bp_call;
  const uint32_t save_cap = g->wl_cap;
  g->use_graphs = 0;
  g->bp_call = false;
  Carve cv{static_cast<char*>(base)};
  carve_slot(g, qs[0].k, qs[0].D, bufs, cv);
  // the array keeps the last call's contents: clear the enqueue stamps (the
  // other buffers are written whole before they are read)
  launch_fill_i32(g->Es.as<int32_t>(), (int64_t)(g->Es.bytes / 4), 0, g->stream);
  BatchOut out;
  out.cost.assign(1, 0);
  out.status.assign(1, 0);
  out.uv.resize(1);
  out.ids.resize(1);
  gst_stats local;
  std::memset(&local, 0, sizeof(local));
  rc = solve_lockstep(g, qs, 1.0, out, &local);
  for (int i = 0; i < kNoallocBufs; ++i) *bufs[i] = saved[i];
  g->use_graphs = save_graphs;
  g->bp_call = save_bp;
  g->wl_cap = save_cap;
  ++g->gen;  // the next regular call rebuilds its captured graphs
  g->noalloc = false;
  if (rc == 100) return fail(GST_ERR_OOM, "worklist overflow in the allocation-free path");
  if (rc) return rc;
  *out_cost = out.cost[0];
  if (out_status) *out_status = out.status[0];
  const int32_t ne = (int32_t)out.uv[0].size();
  if (out_num_edges) *out_num_edges = ne;
  if (out_edge_uv)
    for (int32_t i = 0; i < std::min(ne, edge_cap); ++i) {
      out_edge_uv[2 * i] = out.uv[0][i].first;
      out_edge_uv[2 * i + 1] = out.uv[0][i].second;
    }
  local.seconds_total =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  local.bytes_alg = 16 * local.states_expanded + 12 * local.grow_relax + 8 * local.merge_relax +
                    12 * local.improvements + 4 * local.lb_terms;
  if (st) *st = local;
  if (out_edge_uv && ne > edge_cap)
    return fail(GST_ERR_CAPACITY, "edge_cap too small; *out_num_edges holds the tree size");
  return GST_OK;
}

int gst_backward(gst_graph* g, int32_t nq, const int64_t* d_off, const int32_t* d_ids,
                 const float* d_gout, float* d_gw, int64_t nund) {
  if (!g || nq < 0 || nund < 0) return fail(GST_ERR_INVALID, "bad argument");
  if (nq == 0) return GST_OK;
  if (!d_off || !d_ids || !d_gout || !d_gw) return fail(GST_ERR_INVALID, "null pointer");
  HIP_TRY(hipSetDevice(g->device));
  launch_backward(nq, d_off, d_ids, d_gout, d_gw, nund, g->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  return GST_OK;
}

void gst_graph_destroy(gst_graph* g) {
  if (!g) return;
  g->drop_graphs();
  DBuf* bufs[] = {&g->rowp, &g->col, &g->wt, &g->eid, &g->goff, &g->gmem, &g->T, &g->Es, &g->BPs, &g->Ms, &g->wls, &g->GMs, &g->far0, &g->far1, &g->rsel, &g->rtasks, &g->rbest, &g->outpack,
                  &g->C, &g->R, &g->VVb, &g->bestw, &g->slots, &g->tasks, &g->wl0, &g->wl1,
                  &g->cnt, &g->iter, &g->ctr, &g->sched, &g->out_uv, &g->out_eid, &g->out_n, &g->out_err};
  for (DBuf* b : bufs) b->release();
  if (g->h_cnt) (void)hipHostFree(g->h_cnt);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  for (hipEvent_t e : g->ph)
    if (e) (void)hipEventDestroy(e);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // extern "C"
