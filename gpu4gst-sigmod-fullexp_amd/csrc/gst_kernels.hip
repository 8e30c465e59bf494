// gfx950 kernels of the TrimCDP / D-TrimCDP engine.
//
// Phases of a batch of query slots (all slots advance in lockstep):
//   1. k_groups_setup   group mask of every member vertex, one-label seeds
//   2. k_sssp_expand    label-correcting multi-source shortest paths c_i(v)
//                       (reference: dis0_init_1, code/TrimCDP/src/DPQ.cu:260-290,
//                       and hybrid_bin_scan_push_kernel_only,
//                       code/TrimCDP-WB/include/mapper_enactor.cuh:249-366)
//   3. k_vv, k_route    group-to-group distances and route tables for the tour
//                       lower bounds (host code at code/TrimCDP/src/DPQ.cu:426-501)
//   4. k_dp_seed        T(v, p) = 0 for p inside v's groups (DPQ.cu:349-389)
//   5. k_dp_expand      the DP frontier step: upper bound, 1/2 rule, merge over
//                       submask pairs, grow over CSR rows with int32 atomicMin,
//                       lower-bound filtered enqueue (Relax, DPQ.cu:86-191;
//                       hop variant: code/D-TrimCDP/src/DPQ.cu:62-124)
//   6. k_backtrack      tree recovery from the final table
//   7. k_backward       edge-incidence subgradient scatter
//
// Frontier bookkeeping is the same in every expand kernel: a state enters the
// next worklist only if atomicMax of its stamp to (iteration + 1) raised it,
// so each state is queued at most once per iteration without any clearing
// pass, and appends are aggregated per wavefront (ballot / wave scan, one
// atomicAdd per wave). Rows longer than one chunk are queued as one item per
// chunk (virtual split), and each workgroup balances the edges of its 256
// items with an LDS prefix scan, so high- and low-degree rows share the same
// coalesced edge loop.
#include "gst_device.h"

#include <mutex>

#include "../../include/gst.h"

namespace gst {

namespace {

__device__ inline int lane_id() { return __lane_id(); }

#ifndef GST_EPL
#define GST_EPL 3
#endif
// edges per lane per round in the one-label kernel (3: 111 VGPRs; 4 took 123
// and 2% longer at four workgroups per CU)
constexpr int EPL = GST_EPL;
#ifndef GST_DEPL
#define GST_DEPL 2
#endif
// ... and in the DP kernel, whose rows are cut short by its bounds: two keep
// it unspilled at four waves per SIMD (DP 53 -> 51.5 ms on the 168k-vertex
// workload, hop-bounded DP 1,259 -> 1,225 ms on the 1.1M-vertex one)
constexpr int DEPL = GST_DEPL;
#ifndef GST_MPL
#define GST_MPL 2
#endif
constexpr int MPL = GST_MPL;  // merge pairs per lane per round in the DP kernel

// Appends `cnt` chunk items (item | chunk << 36) per lane; every lane of the
// wavefront must call this together.
__device__ inline void wave_append(uint32_t cnt, uint64_t item, uint64_t* out, uint32_t* out_cnt,
                                   uint32_t cap, unsigned long long* ctr) {
  const int lane = lane_id();
  const uint64_t many = __ballot(cnt > 1);
  uint32_t excl, total;
  if (many == 0) {
    const uint64_t m = __ballot(cnt == 1);
    if (m == 0) return;
    total = (uint32_t)__popcll(m);
    excl = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  } else {
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
      uint32_t y = __shfl_up(incl, off, WAVE);
      if (lane >= off) incl += y;
    }
    total = __shfl(incl, WAVE - 1, WAVE);
    excl = incl - cnt;
  }
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(out_cnt, total);
  base = __shfl(base, 0, WAVE);
  for (uint32_t i = 0; i < cnt; ++i) {
    uint32_t pos = base + excl + i;
    if (pos < cap) out[pos] = item | ((uint64_t)i << 36);
    else atomicAdd(&ctr[C_OVERFLOW], 1ull);
  }
}

// Wave-private staging of worklist appends in LDS: items collect in the
// wave's buffer and go to the list with one global atomicAdd per WSTAGE
// items (the list counter is a single word every wave of the grid would
// otherwise hit once per append round). `n` is the wave-uniform fill level.
#ifndef GST_WSTAGE
#define GST_WSTAGE 256
#endif
constexpr int WSTAGE = GST_WSTAGE;

// Ablation (GST_OPT_LANE_APPENDS): the lane reserves its own range with one
// global atomic, no aggregation. Called by the lanes that have items only.
__device__ inline void lane_append(uint32_t cnt, uint64_t item, uint64_t* out, uint32_t* out_cnt,
                                   uint32_t cap, unsigned long long* ctr) {
  const uint32_t base = atomicAdd(out_cnt, cnt);
  for (uint32_t i = 0; i < cnt; ++i) {
    if (base + i < cap) out[base + i] = item | ((uint64_t)i << 36);
    else atomicAdd(&ctr[C_OVERFLOW], 1ull);
  }
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline void stage_flush(uint64_t* stage, uint32_t& n, uint64_t* out,
                                   uint32_t* out_cnt, uint32_t cap, unsigned long long* ctr) {
  if (n == 0) return;
  wave_sync();
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(out_cnt, n);
  base = __shfl(base, 0, WAVE);
  for (uint32_t i = (uint32_t)lane_id(); i < n; i += WAVE) {
    const uint32_t pos = base + i;
    if (pos < cap) out[pos] = stage[i];
    else atomicAdd(&ctr[C_OVERFLOW], 1ull);
  }
  wave_sync();
  n = 0;
}

// Stages `cnt` chunk items (item | chunk << 36) per lane; every lane of the
// wavefront must call this together.
__device__ inline void stage_append(uint64_t* stage, uint32_t& n, uint32_t cnt, uint64_t item,
                                    uint64_t* out, uint32_t* out_cnt, uint32_t cap,
                                    unsigned long long* ctr, bool per_lane = false) {
  if (per_lane) {  // wave-uniform
    if (cnt) lane_append(cnt, item, out, out_cnt, cap, ctr);
    return;
  }
  // most calls append nothing (a relaxation rarely improves its target):
  // one ballot instead of the wave scan
  if (__ballot(cnt != 0) == 0) return;
  const int lane = lane_id();
  uint32_t incl = cnt;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, WAVE);
    if (lane >= off) incl += y;
  }
  const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
  if (total == 0) return;
  const uint32_t excl = incl - cnt;
  if (total > (uint32_t)WSTAGE) {  // more than a buffer: straight to the list
    wave_append(cnt, item, out, out_cnt, cap, ctr);
    return;
  }
  if (n + total > (uint32_t)WSTAGE) stage_flush(stage, n, out, out_cnt, cap, ctr);
  for (uint32_t i = 0; i < cnt; ++i) stage[n + excl + i] = item | ((uint64_t)i << 36);
  n += total;
}

__device__ inline void ub_update(uint64_t* bw, int32_t cost, uint32_t wit) {
  const uint64_t pk = ((uint64_t)(uint32_t)cost << 32) | wit;
  if (pk < __hip_atomic_load(bw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMin((unsigned long long*)bw, (unsigned long long)pk);
}

__device__ inline int32_t best_of(const uint64_t* bw) {
  return (int32_t)(__hip_atomic_load(bw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
}

__device__ inline uint32_t nchunks(const DevCtx& c, uint32_t u) {
  const int32_t deg = c.rowp[u + 1] - c.rowp[u];
  const uint32_t n = (uint32_t)((deg + (1 << c.chunk_shift) - 1) >> c.chunk_shift);
  return n ? n : 1u;
}

// One-label radius of a plain query: distances at or beyond it are never
// propagated and are clamped to it afterwards (a valid lower bound). Any
// radius keeps every bound valid; radius_pct only trades one-label work
// against lower-bound strength. 0 (automatic): 85% of the upper bound for
// queries of up to six groups, 95% for seven, whose DP over 128 subsets
// depends most on the bounds.
__device__ inline int32_t sssp_radius(const DevCtx& c, int32_t k, int32_t best) {
  if (best >= INF - 2) return INF;
  const int32_t pct = c.radius_pct > 0 ? c.radius_pct : (k >= 7 ? 95 : 85);
  return (int32_t)(((int64_t)best * pct) / 100) + 2;
}

// Rows are sorted by weight: first position in [lo, hi) whose weight exceeds
// `limit` (edges past it cannot lead to a useful relaxation).
__device__ inline int32_t row_cut(const int32_t* wt, int32_t lo, int32_t hi, int32_t limit) {
  if (limit >= INF) return hi;
  if (limit < 1 || lo >= hi) return lo;
  if (wt[hi - 1] <= limit) return hi;  // the whole row (common for the DP's wide limits)
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (wt[mid] <= limit) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The same cut found by galloping from the row start, then a binary search
// in the last gap: the one-label phase's band limits are small, so its cuts
// lie near the start and the probes stay in the row's first cache lines.
__device__ inline int32_t row_cut_near(const int32_t* wt, int32_t lo, int32_t hi,
                                       int32_t limit) {
  if (limit >= INF) return hi;
  if (limit < 1 || lo >= hi) return lo;
  for (int32_t step = 1;; step <<= 1) {
    const int32_t probe = lo + step - 1;
    if (probe >= hi) break;
    if (wt[probe] > limit) {
      hi = probe;
      break;
    }
    lo = probe + 1;
  }
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (wt[mid] <= limit) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Lower bound of any tree that contains u and touches every group in U
// (U != 0), from the one-label distances cs = c_i(u): the largest distance,
// and the two tour bounds of PrunedDP++ (code/PrunedDP++/include/
// CPUNONHOP.h:433-493) over exact group-route tables R1[i][j] = cheapest
// group sequence i ... j visiting all of U, R2[i] = min_j R1[i][j].
// Also returns sum_{i in U} c_i(u) for the one-label upper bound (INF when
// some distance is a clamped value >= radius). Returns INF when some group of
// U is unreachable from u; returns early (without the tour terms) once the
// largest distance alone exceeds `limit`. Since every tour bound is at most
// the distance sum, a state that yields an upper bound never returns early.
__device__ inline int32_t lb_eval(const int32_t* crow, const int32_t* R, uint32_t U,
                                  int32_t radius, int32_t limit, int32_t* sumc) {
  const int4 a = *reinterpret_cast<const int4*>(crow);
  const int4 b = *reinterpret_cast<const int4*>(crow + 4);
  const int32_t cs[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  int32_t one = 0, mn = INF;
  int64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if ((U >> i) & 1u) {
      one = max(one, cs[i]);
      mn = min(mn, cs[i]);
      sum += cs[i];
    }
  if (one >= INF || one > limit) {  // unreachable, or already past the caller's margin
    *sumc = INF;
    return one;
  }
  // distances >= radius are clamped lower bounds, not path lengths
  *sumc = (sum >= INF || one >= radius) ? INF : (int32_t)sum;
  int32_t lb1 = INF, lb2 = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (!((U >> i) & 1u)) continue;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if ((U >> j) & 1u) lb1 = min(lb1, (cs[i] + R[i * 8 + j] + cs[j]) >> 1);
    lb2 = max(lb2, (cs[i] + R[64 + i] + mn) >> 1);
  }
  return max(one, max(lb1, lb2));
}

// sum_{i in U} c_i(u), or INF if some distance is unreachable or a clamped
// value >= radius (then it is a bound, not a path length).
__device__ inline int32_t sum_c(const int32_t* crow, uint32_t U, int32_t radius) {
  const int4 a = *reinterpret_cast<const int4*>(crow);
  const int4 b = *reinterpret_cast<const int4*>(crow + 4);
  const int32_t cs[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  int64_t sum = 0;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if ((U >> i) & 1u) {
      ok = ok && cs[i] < radius && cs[i] < INF;
      sum += cs[i];
    }
  return ok && sum < INF ? (int32_t)sum : INF;
}

// Exclusive prefix sum over the 256 threads of a workgroup; s_off[256] = total.
__device__ inline void block_scan(int32_t x, int32_t* s_off, int32_t* s_wsum) {
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
  int32_t incl = x;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    int32_t y = __shfl_up(incl, off, WAVE);
    if (lane >= off) incl += y;
  }
  if (lane == WAVE - 1) s_wsum[wid] = incl;
  __syncthreads();
  int32_t prefix = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / WAVE; ++w)
    if (w < wid) prefix += s_wsum[w];
  s_off[tid] = prefix + incl - x;
  if (tid == BLOCK - 1) s_off[BLOCK] = prefix + incl;
  __syncthreads();
}

// Largest x over the workgroup; every thread must call it.
__device__ inline int32_t block_max(int32_t x, int32_t* s_wsum) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) x = max(x, __shfl_down(x, off, WAVE));
  __syncthreads();  // s_wsum may still be read by an earlier scan
  if (lane_id() == 0) s_wsum[threadIdx.x / WAVE] = x;
  __syncthreads();
  int32_t m = s_wsum[0];
#pragma unroll
  for (int w = 1; w < BLOCK / WAVE; ++w) m = max(m, s_wsum[w]);
  __syncthreads();  // before s_wsum is reused
  return m;
}

__device__ inline int find_item(const int32_t* s_off, int32_t e) {
  int lo = 0, hi = BLOCK;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (s_off[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

// One-label stamps: word 7 of the distance row of (slot, v) (k <= 7 labels
// use words 0..6) is (iteration << 8) | mask, mask = the labels of v that
// changed since v was last queued; it shares the cache line the relaxation
// just read. Rows are re-initialised (stamp 0) for every batch. Marks the
// labels `bits` of v for iteration `it`, starting from the caller's copy
// `old` of the word; returns true when this call queued v for `it` (the
// caller appends it). No atomic is issued when the word already holds `it`
// and the bits.
__device__ inline bool mark_labels(int32_t* word, uint32_t old, uint32_t it, uint32_t bits) {
  uint32_t* w = reinterpret_cast<uint32_t*>(word);
  while (true) {
    const bool queued = (old >> 8) == it;
    if (queued && (old & bits) == bits) return false;
    const uint32_t nw = queued ? (old | bits) : ((it << 8) | bits);
    const uint32_t seen = atomicCAS(w, old, nw);
    if (seen == old) return !queued;
    old = seen;
  }
}

// Labels an item of iteration `it` relaxes, from the stamp word w: the marked
// ones, or all of them when v was marked again for a later iteration before
// this item ran (the marks of `it` are then merged away, so nothing may be
// skipped).
__device__ inline uint32_t pending_labels(uint32_t w, uint32_t it) {
  return (w >> 8) == it ? (w & 0xffu) : 0xffu;
}

// Marks DP state (v, p, d) of slot S pending for the next iteration: sets
// bit p of the vertex's pending mask, and bit 0 of its first word, the
// "vertex queued" flag. Returns true when this call raised the flag: the
// caller then appends the vertex item (v, d) once. A state already pending
// costs one atomic and queues nothing. k_dp_unfold clears the mask when it
// lists the states, before any thread of the next iteration marks again.
// In a fin slot the same 64-bit atomic sets bit p of the finite word beside
// the pending word (a state is marked when its value first drops below INF,
// or set_finite is called where it is not queued).
__device__ inline bool mark_state(const DevCtx& c, const SlotDesc& S, uint32_t v, uint32_t p,
                                  uint32_t d) {
  uint32_t* m = c.M + mask_ix(S, v, d);
  const uint32_t wi = p >> 5, bit = 1u << (p & 31u);
  if (S.fin) {
    unsigned long long* m2 = reinterpret_cast<unsigned long long*>(m);
    const unsigned long long b = (unsigned long long)bit | ((unsigned long long)bit << 32);
    if (wi == 0) return (atomicOr(m2, b | 1ull) & 1ull) == 0;
    if (atomicOr(m2 + wi, b) & (unsigned long long)bit) return false;  // already pending
    return (atomicOr(m, 1u) & 1u) == 0;
  }
  if (wi == 0) return (atomicOr(m, bit | 1u) & 1u) == 0;
  if (atomicOr(m + wi, bit) & bit) return false;  // already pending: its vertex is queued
  return (atomicOr(m, 1u) & 1u) == 0;
}

// Sets the finite bit of a fin slot's state (v, p, d) whose value dropped
// below INF without it being queued (sent to the far list, or kept out of the
// queue by the 2/3 rule): it is still a merge partner.
__device__ inline void set_finite(const DevCtx& c, const SlotDesc& S, uint32_t v, uint32_t p,
                                  uint32_t d) {
  if (S.fin) atomicOr(c.M + mask_ix(S, v, d) + 2 * (p >> 5) + 1, 1u << (p & 31u));
}

// Subsets of `mask` (k <= 7) as a 128-bit set: bit x is set for x a subset of mask.
__device__ inline uint4 subsets128(uint32_t mask) {
  uint64_t lo = 1ull, hi = 0ull;  // {0}
  for (uint32_t b = 0; b < 7; ++b) {
    if (!((mask >> b) & 1u)) continue;
    const uint32_t sh = 1u << b;  // every subset so far, with bit b added
    if (sh == 64) {
      hi |= lo;
    } else {
      hi |= (hi << sh) | (lo >> (64 - sh));
      lo |= lo << sh;
    }
  }
  return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

// Position of the r-th set bit (r < popcount) of a 128-bit set.
__device__ inline uint32_t nth_bit128(const uint4& m, uint32_t r) {
  const uint32_t w[4] = {m.x, m.y, m.z, m.w};
  uint32_t i = 0;
  for (; i < 3; ++i) {
    const uint32_t pc = (uint32_t)__popc(w[i]);
    if (r < pc) break;
    r -= pc;
  }
  uint32_t x = w[i];
  for (; r; --r) x &= x - 1u;  // drop the r lowest set bits
  return 32u * i + (uint32_t)__builtin_ctz(x);
}

// Non-empty subsets of `mask` (k <= 5) as a set: bit x is set for x a subset of mask, x != 0.
__device__ inline uint32_t subsets_of(uint32_t mask) {
  uint32_t r = 0;
  for (uint32_t x = mask; x; x = (x - 1u) & mask) r |= 1u << x;
  return r;
}

// Plain slots: T(v, full) <- t + T(v, full ^ p), the complement merge a state
// gets when it is expanded. A state that will never be expanded (dropped by
// the 1/2 rule, or not queued under the 2/3 rule) gets it when it is dropped
// instead, so two complementary states that both exist are always combined,
// in whatever order the frontier reaches them. (States dropped by their
// completion bound need none: t + lb > best + 2 bounds every such merge.)
__device__ inline void complement_merge(const DevCtx& c, const SlotDesc& S, uint32_t s,
                                        uint32_t v, uint32_t p, int32_t t) {
  const int32_t t2 = c.T[2 * t_ix(S, (uint32_t)S.full ^ p, v)];
  if (t2 >= INF || t >= INF) return;
  const int32_t mcost = t + t2;
  int32_t* dfull = &c.T[2 * t_ix(S, S.full, v)];
  if (mcost < *dfull && mcost < atomicMin(dfull, mcost)) {
    if (c.BP) c.BP[t_ix(S, S.full, v)] = (int32_t)bp_merge(p, 0, 0);
    ub_update(&c.bestw[s], mcost, wit_full(v, 0));
  }
}

__device__ inline void flush(unsigned long long* ctr, int idx, unsigned long long v) {
  // wave-reduce, then one atomic per wave
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
  if (lane_id() == 0 && v) atomicAdd(&ctr[idx], v);
}

}  // namespace

// ---------------------------------------------------------------------------
// fills and per-batch initialisation

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = val;
}

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = val;
}

void launch_fill_i32(int32_t* p, int64_t n, int32_t val, hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_i32, (int)blocks, 256, 0, st, p, n, val);
}

void launch_fill_u64(uint64_t* p, int64_t n, uint64_t val, hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u64, (int)blocks, 256, 0, st, p, n, val);
}

// The state words of rows p >= 1 are written whole by k_lb_table ({INF,
// bound}) before the DP reads them, so the state table needs no clearing.
__global__ void k_init_crows(int32_t* C, int64_t rows) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x) {
    int4* p = reinterpret_cast<int4*>(C + r * CROW);
    p[0] = make_int4(INF, INF, INF, INF);
    p[1] = make_int4(INF, INF, INF, 0);  // word 7: one-label stamp
  }
}

void launch_init_slot_tables(const DevCtx& c, int nslots, int64_t c_total, hipStream_t st) {
  const int64_t rows = c_total / CROW;
  const int64_t blocks = (rows + 255) / 256;
  hipLaunchKernelGGL(k_init_crows, (int)(blocks < 4096 ? blocks : 4096), 256, 0, st, c.C, rows);
  launch_fill_i32(c.vv, (int64_t)nslots * VV, INF, st);
  launch_fill_u64(c.bestw, nslots, ((uint64_t)INF << 32) | 0xFFFFFFFFull, st);
  launch_fill_i32(reinterpret_cast<int32_t*>(c.GM), (int64_t)nslots * c.Vp, 0, st);
}

// ---------------------------------------------------------------------------
// group masks and one-label seeds: c_i(v) = 0 for members of query group i

__global__ __launch_bounds__(BLOCK) void k_groups_setup(DevCtx c, const Task* tasks, int ntasks,
                                                        const int64_t* goff, const int32_t* gmem,
                                                        uint64_t* wl, uint32_t* cnt) {
  const int32_t stamp = *c.iter;
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const Task tk = tasks[t];
    const SlotDesc S = c.slots[tk.slot];
    const int64_t b = goff[tk.group], e = goff[tk.group + 1];
    for (int64_t base = b; base < e; base += BLOCK) {
      const int64_t m = base + threadIdx.x;
      uint32_t want = 0;
      uint32_t v = 0;
      if (m < e) {
        v = (uint32_t)gmem[m];
        atomicOr(&c.GM[(int64_t)tk.slot * c.Vp + v], 1u << tk.bit);
        atomicMin(&c.C[S.c_off + (int64_t)v * CROW + tk.bit], 0);
        int32_t* sw = &c.C[S.c_off + (int64_t)v * CROW + CSTAMP];
        const uint32_t old = __hip_atomic_load(reinterpret_cast<uint32_t*>(sw), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
        if (mark_labels(sw, old, (uint32_t)stamp, 1u << tk.bit)) want = nchunks(c, v);
      }
      wave_append(want, mk_item(v, 0, 0, 0, (uint32_t)tk.slot), wl, cnt, c.cap, c.ctr);
    }
  }
}

void launch_groups_setup(const DevCtx& c, const Task* tasks, int ntasks, const int64_t* goff,
                         const int32_t* gmem, uint64_t* wl, uint32_t* cnt, hipStream_t st) {
  if (ntasks <= 0) return;
  int grid = ntasks < 4096 ? ntasks : 4096;
  hipLaunchKernelGGL(k_groups_setup, grid, BLOCK, 0, st, c, tasks, ntasks, goff, gmem, wl, cnt);
}

// ---------------------------------------------------------------------------
// one-label phase: label-correcting multi-source shortest paths for all k
// groups of a slot at once. An item is (slot, v); its k distances sit in one
// 32-byte row c[v][0..7], so one edge visit relaxes all k labels with a
// single row access at u. Plain queries drop labels at or beyond the running
// upper bound + 2 (see k_clamp).

__global__ __launch_bounds__(BLOCK) void k_sssp_expand(DevCtx c, WL w) {
  __shared__ int32_t s_off[BLOCK + 1];
  __shared__ int32_t s_wsum[BLOCK / WAVE];
  __shared__ int32_t s_beg[BLOCK];
  __shared__ int4 s_cv[BLOCK][2];      // the item's distances (masked to INF above the radius)
  __shared__ int64_t s_cbase[BLOCK];   // C offset of the item's slot
  __shared__ uint32_t s_meta[BLOCK];   // slot << 8 | k << 1 | plain
  __shared__ uint64_t s_stage[BLOCK / WAVE][WSTAGE];   // frontier appends
  __shared__ uint64_t s_fstage[BLOCK / WAVE][WSTAGE];  // far-list appends (banded phase)
  uint64_t* stage = s_stage[threadIdx.x / WAVE];
  uint32_t st_n = 0;
  const uint32_t n = min(*w.in_cnt, c.cap);  // an overflowed list is never read past its end
  const int32_t cur = *c.iter, next = cur + 1;
  const int tid = threadIdx.x;
  const int32_t CH = 1 << c.chunk_shift;
  const bool banded = c.sssp_far != 0;
  const int32_t bkey = banded ? *c.band_key : 0;
  const int far_sel = banded ? *c.far_sel : 0;
  uint64_t* fstage = s_fstage[threadIdx.x / WAVE];
  uint32_t fst_n = 0;
  unsigned long long n_items = 0, n_relax = 0, n_upd = 0, n_noub = 0;
  for (uint32_t tile = blockIdx.x; tile * BLOCK < n; tile += gridDim.x) {
    const uint32_t idx = tile * BLOCK + tid;
    int32_t deg = 0;
    uint32_t far_want = 0;
    uint64_t far_item = 0;
    if (idx < n) {
      const uint64_t it = w.in[idx];
      const uint32_t v = it_v(it), ch = it_ch(it), s = it_s(it);
      const SlotDesc S = c.slots[s];
      if (c.trace && !c.sched[s].active) c.sched[s].active = 1;  // per-iteration trace only
      const int32_t best = S.D < 0 ? best_of(&c.bestw[s]) : INF;
      const int32_t radius = sssp_radius(c, S.k, best);
      n_noub += S.D < 0 && radius >= INF;
      const int32_t* row = &c.C[S.c_off + (int64_t)v * CROW];
      int4 a = *reinterpret_cast<const int4*>(row);
      int4 b = *reinterpret_cast<const int4*>(row + 4);
      const uint32_t pend = pending_labels((uint32_t)b.w, (uint32_t)cur);
      // Banded: relax only below the band end (eff); a continuation (flag
      // 0x80 in the item's subset field, set when the band grew) re-relaxes
      // every label, but for a label already below the previous band end only
      // past its old cut.
      int32_t eff = radius, prev = 0;
      if (banded) {
        eff = min(radius, c.sched[s].thr);
        prev = c.sched[s].prev;
      }
      const bool cont = (it_p(it) & 0x80u) != 0;
      // a continuation's labels were all relaxed in earlier bands unless
      // marked for this iteration (none are: the near list was empty when it
      // was made)
      const uint32_t changed =
          cont ? ((b.w >> 8) == (int32_t)cur ? pend & 0x7fu : 0u) : (pend == 0xffu ? 0x7fu : pend & 0x7fu);
      int32_t cv[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
      int32_t lo = INF, done_lim = INF, lo_all = INF;
      bool beyond = false;  // a label waits for a later band
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const bool live = i < S.k && cv[i] < radius;
        if (live) lo_all = min(lo_all, cv[i]);
        if (live && cv[i] >= eff) beyond = true;
        // unchanged labels were relaxed with these same values already
        const bool relax = live && cv[i] < eff && (((changed >> i) & 1u) || cont);
        if (relax) {
          // edges of weight <= prev - 1 - c were relaxed with this value in the
          // previous band unless the label changed or is new to this band
          const int32_t lim = (((changed >> i) & 1u) || cv[i] >= prev) ? -1 : prev - 1 - cv[i];
          done_lim = min(done_lim, lim);
        }
        if (!relax) cv[i] = INF;
        lo = min(lo, cv[i]);
      }
      const int32_t rb0 = c.rowp[v] + (int32_t)(ch << c.chunk_shift);
      const int32_t rend = min(c.rowp[v + 1], rb0 + CH);
      int32_t re = rb0;
      if (lo < INF) {
        // a relaxation must stay below the band end: weight <= eff - 1 - lo
        const int32_t rb = done_lim < INF ? row_cut_near(c.wt, rb0, rend, done_lim) : rb0;
        re = row_cut_near(c.wt, rb0, rend, eff < INF ? eff - 1 - lo : INF);
        deg = max(0, re - rb);
        s_beg[tid] = rb;
        s_cv[tid][0] = make_int4(cv[0], cv[1], cv[2], cv[3]);
        s_cv[tid][1] = make_int4(cv[4], cv[5], cv[6], cv[7]);
        s_cbase[tid] = S.c_off;
        s_meta[tid] = (s << 8) | ((uint32_t)S.k << 1) | (S.D < 0 ? 1u : 0u);
        ++n_items;
      }
      // work left for a later band: a label not yet below the band end, or
      // edges past this band's cut that the final radius would still use
      if (banded && eff < radius && lo_all < INF) {
        const int32_t cut_all = lo < INF ? re : row_cut_near(c.wt, rb0, rend, eff - 1 - lo_all);
        const bool more = cut_all < rend && c.wt[cut_all] <= radius - 1 - lo_all;
        if ((beyond || more) && atomicMax(&c.E[S.e_off + v], bkey) < bkey) {
          far_want = 1;
          // the chunk field carries the vertex's chunk count less one: the
          // refill makes one continuation per chunk without reading the row
          far_item = mk_item(v, 0, 0, nchunks(c, v) - 1u, s);
        }
      }
    }
    if (banded)
      stage_append(fstage, fst_n, far_want, far_item, c.far[far_sel], c.far_cnt + far_sel, c.cap,
                   c.ctr, c.lane_appends != 0);
    block_scan(deg, s_off, s_wsum);
    const int32_t total = s_off[BLOCK];
    // Each lane takes EPL edges per round (strided by the block size, so
    // the CSR reads stay coalesced) and issues all their loads before using
    // any: the random neighbour-row reads of one lane are then in flight
    // together instead of one after another. (Ablation: each lane walks its
    // own item's edges, EPL per round, as many rounds as the largest degree.)
    const bool per_item = c.per_item_edges != 0;
    const int32_t span = per_item ? block_max(deg, s_wsum) * BLOCK : total;
    for (int32_t base = 0; base < span; base += BLOCK * EPL) {
      int jq[EPL];
      int32_t oq[EPL];  // edge offset inside the item's chunk
      uint32_t uq[EPL];
      int32_t wq[EPL];
      int4 aq[EPL], bq[EPL];
#pragma unroll
      for (int q = 0; q < EPL; ++q) {
        if (per_item) {
          oq[q] = base / BLOCK + q;
          jq[q] = oq[q] < deg ? tid : -1;
        } else {
          const int32_t e = base + q * BLOCK + tid;
          jq[q] = e < total ? find_item(s_off, e) : -1;
          oq[q] = jq[q] >= 0 ? e - s_off[jq[q]] : 0;
        }
      }
#pragma unroll
      for (int q = 0; q < EPL; ++q) {
        uq[q] = 0;
        wq[q] = 0;
        if (jq[q] >= 0) {
          const int32_t ej = s_beg[jq[q]] + oq[q];
          uq[q] = (uint32_t)c.col[ej];
          wq[q] = c.wt[ej];
        }
      }
#pragma unroll
      for (int q = 0; q < EPL; ++q) {
        if (jq[q] >= 0) {
          const int32_t* row = &c.C[s_cbase[jq[q]] + (int64_t)uq[q] * CROW];
          aq[q] = *reinterpret_cast<const int4*>(row);
          bq[q] = *reinterpret_cast<const int4*>(row + 4);
        }
      }
      uint32_t cntq[EPL];
      uint64_t itemq[EPL];
#pragma unroll
      for (int q = 0; q < EPL; ++q) {
        uint32_t cnt = 0;
        uint64_t item = 0;
        if (jq[q] >= 0) {
          const int j = jq[q];
          const uint32_t u = uq[q];
          const int32_t wgt = wq[q];
          const uint32_t meta = s_meta[j], sl = meta >> 8, k = (meta >> 1) & 7u;
          const int4 va = s_cv[j][0], vb = s_cv[j][1];
          const int32_t cv[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
          int32_t* row = &c.C[s_cbase[j] + (int64_t)u * CROW];
          const int4 a = aq[q], b = bq[q];
          int32_t cu[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
          ++n_relax;
          // The atomics return nothing, so the wave does not wait on them.
          // "Improved" is judged against the row just read: when another lane
          // lowered the label in between, u is queued once more than needed,
          // and cand, a real path length, still gives a valid upper bound.
          uint32_t improved = 0;
          bool sat = false;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int32_t cand = cv[i] + wgt;  // cv[i] = INF for idle labels: never improves
            sat = sat || (cv[i] < INF && cand >= INF);
            if (cv[i] < INF && cand < cu[i]) {
              __hip_atomic_fetch_min(&row[i], cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              ++n_upd;
              improved |= 1u << i;
              cu[i] = cand;
            }
          }
          if (sat) c.sched[sl].sat = 1;  // a distance past the int32 cost range
          if (improved) {
            if (meta & 1u) {
              // every finite distance is a real path length: u joined to all groups
              int64_t sum = 0;
              bool fin = true;
#pragma unroll
              for (int i = 0; i < 8; ++i)
                if (i < (int)k) {
                  fin = fin && cu[i] < INF;
                  sum += cu[i];
                }
              if (fin && sum < INF) ub_update(&c.bestw[sl], (int32_t)sum, wit_onelabel(u, 0));
            }
            if (mark_labels(&row[CSTAMP], (uint32_t)cu[CSTAMP], (uint32_t)next, improved)) {
              cnt = nchunks(c, u);
              item = mk_item(u, 0, 0, 0, sl);
            }
          }
        }
        cntq[q] = cnt;
        itemq[q] = item;
      }
      // wave-staged appends: no workgroup barrier per round, one list
      // atomic per WSTAGE items per wave
#pragma unroll
      for (int q = 0; q < EPL; ++q)
        stage_append(stage, st_n, cntq[q], itemq[q], w.out, w.out_cnt, c.cap, c.ctr,
                     c.lane_appends != 0);
    }
    __syncthreads();
  }
  stage_flush(stage, st_n, w.out, w.out_cnt, c.cap, c.ctr);
  if (banded)
    stage_flush(fstage, fst_n, c.far[far_sel], c.far_cnt + far_sel, c.cap, c.ctr);
  flush(c.ctr, C_SSSP_ITEMS, n_items);
  flush(c.ctr, C_SSSP_RELAX, n_relax);
  flush(c.ctr, C_SSSP_UPD, n_upd);
  flush(c.ctr, C_SSSP_NOUB, n_noub);
}

void launch_sssp_expand(const DevCtx& c, const WL& w, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_sssp_expand, grid, BLOCK, 0, st, c, w);
}

// Consumes one iteration: clears the consumed list's counter, advances the
// iteration number, moves each slot's threshold past an exhausted band and
// records frontier statistics.
__global__ void k_advance(DevCtx c, uint32_t* cnt_consumed, const uint32_t* cnt_next, int dp,
                          uint32_t* cnt_states) {
  __shared__ int n_active;
  if (threadIdx.x == 0) n_active = 0;
  __syncthreads();
  const int ns = *c.nslots;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    Sched& x = c.sched[s];
    if (c.trace && x.active) atomicAdd(&n_active, 1);
    // near/far modes move bands only when the near list has drained (k_far_*)
    const bool fixed = dp ? c.far_mode != 0 : c.sssp_far != 0;
    if (!fixed && !x.active && x.mindef < INF) x.thr = x.mindef + x.delta;
    x.last = x.active;
    x.active = 0;
    if (!fixed) x.mindef = INF;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t produced = *cnt_next;
    if (*cnt_consumed > 0) atomicAdd(&c.ctr[dp ? C_DP_ITERS : C_SSSP_ITERS], 1ull);
    if (dp && produced > c.ctr[C_MAX_FRONTIER]) c.ctr[C_MAX_FRONTIER] = produced;
    if (c.trace) c.ctr[C_ACTIVE] = (unsigned long long)n_active;  // slots that expanded
    *cnt_consumed = 0;
    if (cnt_states) *cnt_states = 0;  // the DP's unfolded state list, consumed too
    *c.iter += 1;
  }
}

void launch_advance(const DevCtx& c, uint32_t* cnt_consumed, const uint32_t* cnt_next, bool dp,
                    uint32_t* cnt_states, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, 1, 256, 0, st, c, cnt_consumed, cnt_next, dp ? 1 : 0,
                     cnt_states);
}

// The DP frontier of one iteration: every vertex item (v, d) of slot s turns
// into one state item per pending state p and chunk of v's row, chunk-major
// and state-minor, so that the states of one vertex and chunk are adjacent
// in the list (the expand kernel's lanes then relax the same edges for them
// together, reading each neighbour's state row once for several subsets).
// The pending mask is cleared as it is read: the vertex was queued once
// (mark_state), so this is its only reader, and the expand kernel that marks
// states for the next iteration starts after this kernel ends.
__global__ __launch_bounds__(BLOCK) void k_dp_unfold(DevCtx c, const uint64_t* vin,
                                                     const uint32_t* vin_cnt, uint64_t* sout,
                                                     uint32_t* sout_cnt) {
  __shared__ int32_t s_off[BLOCK + 1];
  __shared__ int32_t s_wsum[BLOCK / WAVE];
  __shared__ uint32_t s_base;
  __shared__ uint4 s_w[BLOCK];       // the item's pending mask (the queued flag cleared)
  __shared__ uint32_t s_pc[BLOCK];   // its number of pending states
  __shared__ uint64_t s_it[BLOCK];   // the vertex item (v, d, slot)
  const uint32_t n = min(*vin_cnt, c.cap);
  const int tid = threadIdx.x;
  for (uint32_t tile = blockIdx.x; tile * BLOCK < n; tile += gridDim.x) {
    const uint32_t idx = tile * BLOCK + tid;
    int32_t cnt = 0;
    if (idx < n) {
      const uint64_t it = vin[idx];
      const uint32_t v = it_v(it), d = it_d(it);
      const SlotDesc S = c.slots[it_s(it)];
      uint32_t* m = c.M + mask_ix(S, v, d);
      uint32_t w[4] = {0u, 0u, 0u, 0u};
      // a fin slot's pending words are every second word; its finite words stay
      const int step = S.fin ? 2 : 1, nw = S.mw / step;
      for (int i = 0; i < nw; ++i) {
        w[i] = m[i * step];
        m[i * step] = 0u;
      }
      w[0] &= ~1u;  // the queued flag is no state
      const uint32_t pc = (uint32_t)(__popc(w[0]) + __popc(w[1]) + __popc(w[2]) + __popc(w[3]));
      s_w[tid] = make_uint4(w[0], w[1], w[2], w[3]);
      s_pc[tid] = pc;
      s_it[tid] = it;
      cnt = (int32_t)(pc * nchunks(c, v));
    }
    block_scan(cnt, s_off, s_wsum);
    const uint32_t total = (uint32_t)s_off[BLOCK];
    if (tid == 0) s_base = total ? atomicAdd(sout_cnt, total) : 0u;
    __syncthreads();
    // output position e of the tile: item j, chunk r / pc, the (r % pc)-th
    // pending state; consecutive lanes write consecutive positions
    for (uint32_t e = (uint32_t)tid; e < total; e += BLOCK) {
      const int j = find_item(s_off, (int32_t)e);
      const uint32_t r = e - (uint32_t)s_off[j], pc = s_pc[j];
      const uint32_t ch = r / pc;
      uint32_t k = r - ch * pc;
      const uint4 w4 = s_w[j];
      const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
      uint32_t p = 0;
      for (uint32_t i = 0; i < 4; ++i) {
        const uint32_t pcw = (uint32_t)__popc(ws[i]);
        if (k < pcw) {
          uint32_t x = ws[i];
          for (; k; --k) x &= x - 1u;  // drop the k lowest set bits
          p = 32u * i + (uint32_t)__builtin_ctz(x);
          break;
        }
        k -= pcw;
      }
      const uint64_t it = s_it[j];
      const uint32_t pos = s_base + e;
      if (pos < c.cap) sout[pos] = mk_item(it_v(it), p, it_d(it), ch, it_s(it));
      else atomicAdd(&c.ctr[C_OVERFLOW], 1ull);
    }
    __syncthreads();  // the LDS arrays are reused by the next tile
  }
}

void launch_dp_unfold(const DevCtx& c, const uint64_t* vin, const uint32_t* vin_cnt,
                      uint64_t* sout, uint32_t* sout_cnt, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_dp_unfold, grid, BLOCK, 0, st, c, vin, vin_cnt, sout, sout_cnt);
}

// After the one-label phase of a plain query: every vertex whose distance to
// group i is below R = best + 2 holds it exactly (states are expanded in
// cost bands and none below R was dropped); every other distance is >= R, so
// R is a valid lower bound for it.
__global__ void k_clamp(DevCtx c) {
  const int s = blockIdx.y;
  const SlotDesc S = c.slots[s];
  int32_t R = INF;
  if (S.D < 0) {
    const int32_t best = best_of(&c.bestw[s]);
    R = sssp_radius(c, S.k, best);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) c.sched[s].R = R;
  if (R >= INF) return;
  // the row's labels as two 16-byte words; only a changed half is written
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < c.V; v += gridDim.x * blockDim.x) {
    int4* row = reinterpret_cast<int4*>(&c.C[S.c_off + (int64_t)v * CROW]);
    int4 a = row[0];
    if (a.x > R || (S.k > 1 && a.y > R) || (S.k > 2 && a.z > R) || (S.k > 3 && a.w > R)) {
      a.x = min(a.x, R);
      if (S.k > 1) a.y = min(a.y, R);
      if (S.k > 2) a.z = min(a.z, R);
      if (S.k > 3) a.w = min(a.w, R);
      row[0] = a;
    }
    if (S.k > 4) {
      int4 b = row[1];  // b.w is the stamp word
      if (b.x > R || (S.k > 5 && b.y > R) || (S.k > 6 && b.z > R)) {
        b.x = min(b.x, R);
        if (S.k > 5) b.y = min(b.y, R);
        if (S.k > 6) b.z = min(b.z, R);
        row[1] = b;
      }
    }
  }
}

void launch_clamp(const DevCtx& c, int nslots, hipStream_t st) {
  if (nslots <= 0) return;
  // a few thousand workgroups in all, each walking its slot's rows
  int per_slot = (c.V + 255) / 256 < 8192 / nslots ? (c.V + 255) / 256 : 8192 / nslots;
  if (per_slot < 1) per_slot = 1;
  dim3 grid(per_slot, nslots);
  hipLaunchKernelGGL(k_clamp, grid, 256, 0, st, c);
}

// DP band width from the upper bound known after the one-label phase; the
// slot enters the DP phase.
__global__ void k_dp_sched_init(DevCtx c, int n, const int32_t* sel) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int s = sel ? sel[i] : i;
    const SlotDesc S = c.slots[s];
    const int32_t best = best_of(&c.bestw[s]);
    int32_t delta = INF / 2;
    if (c.dp_div <= 0) delta = INF;
    else if (S.D < 0 && best < INF) delta = max(1, best / c.dp_div);
    else if (S.D >= 0) delta = 0;  // everything to the far list; the first refill sizes the band
    Sched& x = c.sched[s];
    x.delta = delta;
    x.thr = delta;
    x.active = 0;
    x.mindef = INF;
    x.phase = 1;
    x.last = 1;  // not converged before it has run
    x.hasfar = 1;
  }
}

void launch_dp_sched_init(const DevCtx& c, int n, hipStream_t st, const int32_t* sel) {
  if (n > 0) hipLaunchKernelGGL(k_dp_sched_init, 1, 256, 0, st, c, n, sel);
}

__global__ void k_collect_done(DevCtx c, int mode, int32_t* out, uint32_t* out_n) {
  const int ns = *c.nslots;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    Sched& x = c.sched[s];
    if (x.phase != 1) continue;
    const bool done = mode == 2 || (mode == 0 ? x.last == 0 : x.hasfar == 0);
    if (done) {
      x.phase = 2;
      out[atomicAdd(out_n, 1u)] = s;
    }
  }
}

void launch_collect_done(const DevCtx& c, int mode, int32_t* out, uint32_t* out_n,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_collect_done, 1, 256, 0, st, c, mode, out, out_n);
}

__global__ void k_idle_count(DevCtx c, uint32_t* out) {
  const int ns = *c.nslots;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    const Sched& x = c.sched[s];
    if (x.phase == 1 && x.last == 0) atomicAdd(out, 1u);
  }
}

void launch_idle_count(const DevCtx& c, uint32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_idle_count, 1, 256, 0, st, c, out);
}

// ---------------------------------------------------------------------------
// Near/far refill of the banded DP, run when the near list has drained:
// far items whose state can no longer beat the upper bound are dropped (the
// same rules as k_dp_expand), every slot's band restarts at its smallest
// remaining key, and the items inside the new band move to the near list.

// true if the state of far item `it` is still worth expanding; key in *f
// state row of a DP item: p, or d * 2^k + p for a hop-bounded slot
__device__ inline int64_t item_row(const SlotDesc& S, uint64_t it) {
  return S.D >= 0 ? (int64_t)it_d(it) * S.nsub + it_p(it) : (int64_t)it_p(it);
}

// *half_drop (may be null) is set when a plain state is dropped by the 1/2
// rule alone: the caller owes it its complement merge.
// The state word of far item `it`.
__device__ inline int2 far_word(const DevCtx& c, uint64_t it) {
  const SlotDesc S = c.slots[it_s(it)];
  return *reinterpret_cast<const int2*>(&c.T[2 * t_ix(S, item_row(S, it), it_v(it))]);
}

// `tl` is the item's state word (far_word), read by the caller so that the
// reads of several items are in flight together.
__device__ inline bool far_live(const DevCtx& c, uint64_t it, int2 tl, int32_t* f,
                                bool* half_drop = nullptr) {
  const SlotDesc S = c.slots[it_s(it)];
  const int32_t t = tl.x;
  if (t >= INF) return false;
  const int32_t best = best_of(&c.bestw[it_s(it)]);
  if (S.D >= 0) {
    if (c.hop_prune && best < INF && t + tl.y > best) return false;
  } else if (t + tl.y > best + 2) {
    return false;
  } else if (!(t < best / 2 + 2)) {
    if (half_drop) *half_drop = true;
    return false;
  }
  *f = min(t + tl.y, INF - 1);
  return true;
}

// The refill kernels take FPL far items per lane per round and issue all
// their reads before using any (each item's state word is a random read).
constexpr int FPL = 4;

__global__ __launch_bounds__(BLOCK) void k_far_min(DevCtx c) {
  const int sel = *c.far_sel;
  const uint32_t n = min(c.far_cnt[sel], c.cap);
  const uint64_t* far = c.far[sel];
  for (uint32_t base = blockIdx.x * BLOCK * FPL; base < n; base += gridDim.x * BLOCK * FPL) {
    uint64_t itq[FPL];
    int2 tlq[FPL];
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      const uint32_t i = base + q * BLOCK + threadIdx.x;
      itq[q] = i < n ? far[i] : 0;
    }
#pragma unroll
    for (int q = 0; q < FPL; ++q)
      if (base + q * BLOCK + threadIdx.x < n) tlq[q] = far_word(c, itq[q]);
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      int32_t f;
      if (base + q * BLOCK + threadIdx.x < n && far_live(c, itq[q], tlq[q], &f)) {
        int32_t* m = &c.sched[it_s(itq[q])].mindef;
        if (f < *m) atomicMin(m, f);
      }
    }
  }
}

// Band widths double at every refill: the first, narrow bands find a tight
// upper bound early; the later, wide ones finish in few iterations.
// selective = 1: only slots that expanded nothing in the last iteration (no
// near-list work left) move to their next band; the others keep theirs.
__global__ void k_far_thr(DevCtx c, int selective) {
  const int ns = *c.nslots;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    Sched& x = c.sched[s];
    if (selective && x.last != 0) {
      x.mindef = INF;
      x.hasfar = 1;
      continue;
    }
    x.hasfar = x.mindef < INF ? 1 : 0;
    if (x.mindef < INF) {
      // a hop-bounded slot starts with no upper bound: its first band is
      // sized from the smallest key (a lower bound of the optimum) instead
      x.delta = x.delta > 0 ? min(x.delta, INF / 4) * 2 : max(1, x.mindef / c.dp_div);
      x.thr = min(x.mindef + x.delta, INF);
    }
    x.mindef = INF;
  }
}

// Both outputs of a round take one range per workgroup (one scan of the
// packed counts, near in the low and far in the high 16 bits; at most
// BLOCK * FPL each), reserved with one atomic per list.
__global__ __launch_bounds__(BLOCK) void k_far_move(DevCtx c, uint64_t* near, uint32_t* near_cnt) {
  __shared__ int32_t s_off[BLOCK + 1];
  __shared__ int32_t s_wsum[BLOCK / WAVE];
  __shared__ uint32_t s_base[2];
  const int tid = threadIdx.x;
  const int sel = *c.far_sel;
  const uint32_t n = min(c.far_cnt[sel], c.cap);
  const uint64_t* far = c.far[sel];
  uint64_t* far_out = c.far[1 - sel];
  uint32_t* far_out_cnt = c.far_cnt + (1 - sel);
  for (uint32_t base = blockIdx.x * BLOCK * FPL; base < n; base += gridDim.x * BLOCK * FPL) {
    uint64_t itq[FPL];
    int2 tlq[FPL];
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      const uint32_t i = base + q * BLOCK + tid;
      itq[q] = i < n ? far[i] : 0;
    }
#pragma unroll
    for (int q = 0; q < FPL; ++q)
      if (base + q * BLOCK + tid < n) tlq[q] = far_word(c, itq[q]);
    uint32_t wmask = 0, fmask = 0;  // items of this lane for the near / far list
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      if (base + q * BLOCK + tid >= n) continue;
      const uint64_t it = itq[q];
      int32_t f;
      bool half_drop = false;
      const bool live = far_live(c, it, tlq[q], &f, &half_drop);
      if (half_drop) {  // never expanded now: combine it with its complement first
        const SlotDesc S = c.slots[it_s(it)];
        complement_merge(c, S, it_s(it), it_v(it), it_p(it),
                         c.T[2 * t_ix(S, item_row(S, it), it_v(it))]);
      }
      if (!live) continue;
      const uint32_t v = it_v(it), s = it_s(it);
      const SlotDesc S = c.slots[s];
      if (f < c.sched[s].thr) {
        // back to the frontier; duplicates of a state in the far list collapse here
        if (mark_state(c, S, v, it_p(it), it_d(it))) {
          wmask |= 1u << q;
          itq[q] = mk_item(v, 0, it_d(it), 0, s);
        }
      } else {
        fmask |= 1u << q;
      }
    }
    block_scan(__popc(wmask) | (__popc(fmask) << 16), s_off, s_wsum);
    if (tid == 0) {
      const uint32_t tot = (uint32_t)s_off[BLOCK], tn = tot & 0xFFFFu, tf = tot >> 16;
      s_base[0] = tn ? atomicAdd(near_cnt, tn) : 0u;
      s_base[1] = tf ? atomicAdd(far_out_cnt, tf) : 0u;
    }
    __syncthreads();
    uint32_t pn = s_base[0] + ((uint32_t)s_off[tid] & 0xFFFFu);
    uint32_t pf = s_base[1] + ((uint32_t)s_off[tid] >> 16);
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      if ((wmask >> q) & 1u) {
        if (pn < c.cap) near[pn] = itq[q];
        else atomicAdd(&c.ctr[C_OVERFLOW], 1ull);
        ++pn;
      } else if ((fmask >> q) & 1u) {
        if (pf < c.cap) far_out[pf] = itq[q];
        else atomicAdd(&c.ctr[C_OVERFLOW], 1ull);
        ++pf;
      }
    }
    __syncthreads();  // s_off and s_base are reused by the next round
  }
}

__global__ void k_far_flip(DevCtx c) {
  const int sel = *c.far_sel;
  c.far_cnt[sel] = 0;
  *c.far_sel = 1 - sel;
}

void launch_far_refill(const DevCtx& c, int nslots, uint32_t n_far, uint64_t* near,
                       uint32_t* near_cnt, hipStream_t st, bool selective) {
  if (nslots <= 0 || n_far == 0) return;
  const uint32_t blocks = (n_far + BLOCK * FPL - 1) / (BLOCK * FPL);
  const int grid = (int)(blocks < 8192 ? blocks : 8192);
  hipLaunchKernelGGL(k_far_min, grid, BLOCK, 0, st, c);
  hipLaunchKernelGGL(k_far_thr, 1, 256, 0, st, c, selective ? 1 : 0);
  hipLaunchKernelGGL(k_far_move, grid, BLOCK, 0, st, c, near, near_cnt);
  hipLaunchKernelGGL(k_far_flip, 1, 1, 0, st, c);
}

// ---------------------------------------------------------------------------
// Banded one-label phase: when the near list has drained, every slot's band
// grows (doubling, never past the radius best * pct / 100 + 2, which the
// labels only need to reach; a hop-bounded slot has no radius and grows until
// no vertex waits) and the vertices waiting in the far list come back as
// continuations (flag 0x80 in the item); vertices of slots already at the
// radius are dropped.

__global__ void k_band_start(DevCtx c) { *c.band_key = *c.iter + 1; }

void launch_band_start(const DevCtx& c, hipStream_t st) {
  hipLaunchKernelGGL(k_band_start, 1, 1, 0, st, c);
}

__global__ void k_sssp_band_grow(DevCtx c) {
  const int ns = *c.nslots;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    Sched& x = c.sched[s];
    const int32_t radius = sssp_radius(c, c.slots[s].k, best_of(&c.bestw[s]));
    x.prev = x.thr;
    if (x.thr < radius) x.thr = min(x.thr < INF / 2 ? 2 * x.thr : INF, radius);
  }
  __syncthreads();
  if (threadIdx.x == 0) *c.band_key = *c.iter + 1;  // the next band's first iteration + 1
}

// The continuations of a round (FPL far items per lane) take one output range
// per workgroup, reserved with a single atomic on the list counter: the
// counter is one word that every workgroup hits, and this kernel does nothing
// but append.
__global__ __launch_bounds__(BLOCK) void k_sssp_far_move(DevCtx c, uint64_t* near,
                                                         uint32_t* near_cnt) {
  __shared__ int32_t s_off[BLOCK + 1];
  __shared__ int32_t s_wsum[BLOCK / WAVE];
  __shared__ uint32_t s_base;
  const int tid = threadIdx.x;
  const int sel = *c.far_sel;
  const uint32_t n = min(c.far_cnt[sel], c.cap);
  const uint64_t* far = c.far[sel];
  for (uint32_t base = blockIdx.x * BLOCK * FPL; base < n; base += gridDim.x * BLOCK * FPL) {
    uint64_t itq[FPL];
    uint32_t wq[FPL];
    int32_t tot = 0;
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      const uint32_t i = base + (uint32_t)(tid * FPL + q);  // a lane's items are adjacent
      itq[q] = i < n ? far[i] : 0;
    }
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      wq[q] = 0;
      if (base + (uint32_t)(tid * FPL + q) < n) {
        const uint64_t it = itq[q];
        const Sched& x = c.sched[it_s(it)];
        if (x.prev < x.thr) {  // the slot's band grew: v continues
          wq[q] = it_ch(it) + 1u;
          // continuation flag in the subset field, chunk field cleared for the chunk index
          itq[q] = (it & ~((uint64_t)0xFFFFu << 36)) | ((uint64_t)0x80u << 24);
        }
      }
      tot += (int32_t)wq[q];
    }
    block_scan(tot, s_off, s_wsum);
    if (tid == 0) s_base = s_off[BLOCK] ? atomicAdd(near_cnt, (uint32_t)s_off[BLOCK]) : 0u;
    __syncthreads();
    uint32_t pos = s_base + (uint32_t)s_off[tid];
#pragma unroll
    for (int q = 0; q < FPL; ++q)
      for (uint32_t r = 0; r < wq[q]; ++r, ++pos) {
        if (pos < c.cap) near[pos] = itq[q] | ((uint64_t)r << 36);
        else atomicAdd(&c.ctr[C_OVERFLOW], 1ull);
      }
    __syncthreads();  // s_off and s_base are reused by the next round
  }
}

void launch_sssp_far_refill(const DevCtx& c, uint32_t n_far, uint64_t* near, uint32_t* near_cnt,
                            hipStream_t st) {
  if (n_far == 0) return;
  const uint32_t blocks = (n_far + BLOCK * FPL - 1) / (BLOCK * FPL);
  const int grid = (int)(blocks < 8192 ? blocks : 8192);
  hipLaunchKernelGGL(k_sssp_band_grow, 1, 256, 0, st, c);
  hipLaunchKernelGGL(k_sssp_far_move, grid, BLOCK, 0, st, c, near, near_cnt);
  hipLaunchKernelGGL(k_far_flip, 1, 1, 0, st, c);
}

// ---------------------------------------------------------------------------
// tour-bound tables

// vv[i][j] = min over members v of query group i of c_j(v).
__global__ __launch_bounds__(BLOCK) void k_vv(DevCtx c, const Task* tasks, int ntasks,
                                              const int64_t* goff, const int32_t* gmem) {
  __shared__ int32_t s_min[8];
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const Task tk = tasks[t];
    const SlotDesc S = c.slots[tk.slot];
    if (threadIdx.x < 8) s_min[threadIdx.x] = INF;
    __syncthreads();
    int32_t loc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) loc[j] = INF;
    for (int64_t m = goff[tk.group] + threadIdx.x; m < goff[tk.group + 1]; m += BLOCK) {
      const int32_t* row = &c.C[S.c_off + (int64_t)gmem[m] * CROW];
#pragma unroll
      for (int j = 0; j < 8; ++j) loc[j] = min(loc[j], row[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int32_t x = loc[j];
#pragma unroll
      for (int off = WAVE / 2; off > 0; off >>= 1) x = min(x, __shfl_down(x, off, WAVE));
      if (lane_id() == 0) atomicMin(&s_min[j], x);
    }
    __syncthreads();
    if (threadIdx.x < 8 && (int)threadIdx.x < S.k)
      atomicMin(&c.vv[S.vv_off + tk.bit * 8 + threadIdx.x], s_min[threadIdx.x]);
    __syncthreads();
  }
}

void launch_vv(const DevCtx& c, const Task* tasks, int ntasks, const int64_t* goff,
               const int32_t* gmem, hipStream_t st) {
  if (ntasks <= 0) return;
  int grid = ntasks < 4096 ? ntasks : 4096;
  hipLaunchKernelGGL(k_vv, grid, BLOCK, 0, st, c, tasks, ntasks, goff, gmem);
}

// Held-Karp over the query groups of one slot: rt[i][j][X] = cheapest
// sequence of groups starting at i, ending at j, visiting exactly X (i, j in X,
// i != j unless X = {i}), with vv as the step cost. Then R1[X][i][j] = rt and
// R2[X][i] = min_j rt[i][j][X].
__global__ __launch_bounds__(64) void k_route(DevCtx c) {
  __shared__ int32_t rt[8 * 8 * 128];
  __shared__ int32_t vv[64];
  const SlotDesc S = c.slots[blockIdx.x];
  const int k = S.k, nsub = S.nsub;
  const int tid = threadIdx.x;
  for (int x = tid; x < 8 * 8 * 128; x += 64) rt[x] = INF;
  vv[tid] = c.vv[S.vv_off + tid];
  __syncthreads();
  if (tid < k) rt[(tid * 8 + tid) * 128 + (1 << tid)] = 0;
  __syncthreads();
  const int i = tid >> 3, j = tid & 7;
  for (int X = 1; X < nsub; ++X) {
    if (i < k && j < k && i != j && ((X >> i) & 1) && ((X >> j) & 1)) {
      const int Y = X ^ (1 << j);
      int32_t best = INF;
      for (int m = 0; m < k; ++m)
        if ((Y >> m) & 1) best = min(best, rt[(i * 8 + m) * 128 + Y] + vv[m * 8 + j]);
      rt[(i * 8 + j) * 128 + X] = min(best, INF);
    }
    __syncthreads();
  }
  int32_t* R = c.R + S.r_off;
  for (int X = 0; X < nsub; ++X) {
    R[X * RSTRIDE + tid] = (i < k && j < k) ? rt[(i * 8 + j) * 128 + X] : INF;
    if (tid < 8) {
      int32_t m = INF;
      for (int jj = 0; jj < k; ++jj) m = min(m, rt[(tid * 8 + jj) * 128 + X]);
      R[X * RSTRIDE + 64 + tid] = m;
    }
    if (tid == 0) R[X * RSTRIDE + RLBMIN] = INF;  // lowered by k_lb_table
  }
}

void launch_route(const DevCtx& c, int nslots, hipStream_t st) {
  if (nslots > 0) hipLaunchKernelGGL(k_route, nslots, 64, 0, st, c);
}

// ---------------------------------------------------------------------------
// Completion bounds, written together with the initial cost INF as one 8-byte
// store per state: the second word of every state (v, p[, d]) holds
// lb(v, full ^ p), the lower bound of any tree containing v that touches the
// groups outside p (the reference's lb0 table, DPQ.cu:193-243, evaluated here
// with exact group-route tables). A relaxation then decides from one 8-byte
// read whether its target can still beat the upper bound.
#ifndef GST_LB_TP
#define GST_LB_TP 32
#endif
#ifndef GST_LB_BLOCK
#define GST_LB_BLOCK 512
#endif
constexpr int LBB = GST_LB_BLOCK;  // threads per workgroup of the bound kernel
typedef long long i64x2 __attribute__((vector_size(16)));  // two state words
#ifndef GST_LB_VPL
#define GST_LB_VPL 2
#endif
__global__ __launch_bounds__(LBB) void k_lb_table(DevCtx c, const int32_t* sel) {
  // A workgroup (eight waves) stages its slot's route tables in LDS (entries
  // for groups 0..6 only: 49 pair routes + 7 row minima per subset) and then
  // walks tiles of 128 consecutive vertices, subsets in passes of TP = 32 (a
  // smaller tile than all 128 subsets at once lets three workgroups share a
  // CU). Each wave takes every eighth subset p, so U = full ^ p is the same in all its lanes
  // (two vertices per lane, which each route read serves) and every route
  // read is an LDS broadcast. The loop
  // over the first group i of a pair branches (wave-uniformly) on i in U; the
  // second group j runs without branches: k_route leaves the routes of groups
  // outside U (and of groups >= k) at INF, which can never be the minimum.
  // The bounds go to an LDS tile [vertex][p] (odd stride: conflict-free),
  // which the workgroup writes out as whole {INF, bound} state words,
  // contiguous in the vertex-major table.
  constexpr int VPL = GST_LB_VPL;  // vertices per lane: each route read serves VPL of them
  constexpr int TV = WAVE * VPL;
  constexpr int TP = GST_LB_TP;
  constexpr int TS = TP + 1;
  constexpr int RC = 57;  // compact route row: [i * 7 + j] for i, j < 7, then [49 + i]
  __shared__ int32_t tile[TV * TS];
  const int s = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  const SlotDesc S = c.slots[s];
  __shared__ int32_t s_r[128 * RC];
  __shared__ int32_t s_lbmin[128];  // the workgroup's smallest bound per subset p
  if (threadIdx.x < 128) s_lbmin[threadIdx.x] = INF;
  for (int x = threadIdx.x; x < S.nsub * RC; x += LBB) {
    const int X = x / RC, e = x - X * RC;
    const int src = e < 49 ? (e / 7) * 8 + e % 7 : 64 + (e - 49);
    s_r[x] = e < 56 ? c.R[S.r_off + X * RSTRIDE + src] : INF;
  }
  __syncthreads();
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int nrows = S.nrows;
  const int nd = nrows / S.nsub;  // hop: D + 1 copies of the subsets
  int2* T2 = reinterpret_cast<int2*>(c.T);
  for (int64_t v0 = (int64_t)blockIdx.x * TV; v0 < c.V; v0 += (int64_t)gridDim.x * TV) {
    int32_t cs[VPL][7];
#pragma unroll
    for (int r = 0; r < VPL; ++r) {
      const int64_t v = v0 + lane + WAVE * r;
      const int64_t vr = v < c.V ? v : v0;
      const int4 a = *reinterpret_cast<const int4*>(&c.C[S.c_off + vr * CROW]);
      const int4 b = *reinterpret_cast<const int4*>(&c.C[S.c_off + vr * CROW + 4]);
      cs[r][0] = a.x; cs[r][1] = a.y; cs[r][2] = a.z; cs[r][3] = a.w;
      cs[r][4] = b.x; cs[r][5] = b.y; cs[r][6] = b.z;  // word 7 is the one-label stamp
    }
    const int nv = (int)min<int64_t>(TV, c.V - v0);
    for (int h = 0; h < S.nsub; h += TP) {
      const int np = min(TP, S.nsub - h);  // a power of two
      for (int p = h + wv; p < h + np; p += LBB / WAVE) {
        if (p == 0) {  // p = 0 is no state: its word is written as {INF, 0}
          for (int r = 0; r < VPL; ++r) tile[(lane + WAVE * r) * TS] = 0;
          continue;
        }
        const uint32_t U = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)S.full ^ (uint32_t)p);
        const int32_t* R = s_r + U * RC;
        int32_t one[VPL], mn[VPL], lb1[VPL], lb2[VPL];
#pragma unroll
        for (int r = 0; r < VPL; ++r) {
          one[r] = 0;
          mn[r] = INF;
          lb1[r] = INF;
          lb2[r] = 0;
#pragma unroll
          for (int i = 0; i < 7; ++i) {
            const bool in = (U >> i) & 1u;
            one[r] = max(one[r], in ? cs[r][i] : 0);
            mn[r] = min(mn[r], in ? cs[r][i] : INF);
          }
        }
#pragma unroll
        for (int i = 0; i < 7; ++i) {
          if (!((U >> i) & 1u)) continue;  // wave-uniform
          // (c_i + R_ij + c_j) / 2 is monotone in R_ij + c_j: take the minimum first
          int32_t m[VPL];
#pragma unroll
          for (int r = 0; r < VPL; ++r) m[r] = INF;
#pragma unroll
          for (int j = 0; j < 7; ++j) {
            const int32_t rij = R[i * 7 + j];
#pragma unroll
            for (int r = 0; r < VPL; ++r) m[r] = min(m[r], rij + cs[r][j]);
          }
          const int32_t rmin = R[49 + i];
#pragma unroll
          for (int r = 0; r < VPL; ++r) {
            lb1[r] = min(lb1[r], (cs[r][i] + m[r]) >> 1);
            lb2[r] = max(lb2[r], (cs[r][i] + rmin + mn[r]) >> 1);
          }
        }
#pragma unroll
        for (int r = 0; r < VPL; ++r) {
          const int32_t lb =
              U == 0 ? 0 : (one[r] >= INF ? one[r] : max(one[r], max(lb1[r], lb2[r])));
          tile[(lane + WAVE * r) * TS + (p - h)] = lb;
        }
      }
      __syncthreads();
      // one lane per (vertex, two consecutive subsets): a 16-byte store of two
      // state words (np >= 2, rows and table offsets are even); a hop slot's
      // lane writes the pair for every depth d (rows d * 2^k + p), consecutive
      // lanes consecutive pairs
      const int sh = __builtin_ctz((unsigned)np) - 1;
      const int64_t tb = t_ix(S, h, v0);
      // the lane's pair of subsets is the same in every round (the workgroup
      // size is a multiple of np / 2): it keeps their minima over the tile
      int32_t m0 = INF, m1 = INF;
#pragma unroll 4
      for (int x = threadIdx.x; x < (nv << sh); x += LBB) {
        const int vl = x >> sh, pp = 2 * (x & ((np >> 1) - 1));
        const int32_t lb0 = tile[vl * TS + pp], lb1 = tile[vl * TS + pp + 1];
        m0 = min(m0, lb0);
        m1 = min(m1, lb1);
        const i64x2 pair = {(long long)(((uint64_t)(uint32_t)lb0 << 32) | (uint32_t)INF),
                            (long long)(((uint64_t)(uint32_t)lb1 << 32) | (uint32_t)INF)};
        int64_t ix = tb + (int64_t)vl * nrows + pp;
        // streaming stores: the table is far larger than the caches
        for (int dd = 0; dd < nd; ++dd, ix += S.nsub)
          __builtin_nontemporal_store(pair, reinterpret_cast<i64x2*>(&T2[ix]));
      }
      const int pl = h + 2 * (threadIdx.x & ((np >> 1) - 1));
      if (m0 < s_lbmin[pl]) atomicMin(&s_lbmin[pl], m0);
      if (m1 < s_lbmin[pl + 1]) atomicMin(&s_lbmin[pl + 1], m1);
      __syncthreads();
    }
  }
  // the slot's smallest bound per U (k_route set it to INF): a grown state of
  // subset p needs cost + lb <= best + 2, so the DP cuts its edges at
  // best + 2 - t - min_v lb(v, full ^ p)
  if (threadIdx.x < S.nsub && s_lbmin[threadIdx.x] < INF)
    atomicMin(&c.R[S.r_off + (int64_t)((uint32_t)S.full ^ threadIdx.x) * RSTRIDE + RLBMIN],
              s_lbmin[threadIdx.x]);
}

void launch_lb_table(const DevCtx& c, int nslots, hipStream_t st, const int32_t* sel) {
  if (nslots <= 0) return;  // must run for every slot: it initialises the state words
  // enough workgroups to fill the device for one slot as for a thousand
  const int tiles = (c.V + WAVE * GST_LB_VPL - 1) / (WAVE * GST_LB_VPL);
  int per_slot = (8192 + nslots - 1) / nslots;
  if (per_slot < 4) per_slot = 4;
  dim3 grid(tiles < per_slot ? tiles : per_slot, nslots);
  hipLaunchKernelGGL(k_lb_table, grid, LBB, 0, st, c, sel);
}

// ---------------------------------------------------------------------------
// DP seeds: T(v, p[, 0]) = 0 for every non-empty p inside v's query groups.
// Each vertex is seeded from its lowest query group only.

__global__ __launch_bounds__(BLOCK) void k_dp_seed(DevCtx c, const Task* tasks, int ntasks,
                                                   const int64_t* goff, const int32_t* gmem,
                                                   uint64_t* wl, uint32_t* cnt) {
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const Task tk = tasks[t];
    const SlotDesc S = c.slots[tk.slot];
    const bool hop = S.D >= 0;
    const int64_t b = goff[tk.group], e = goff[tk.group + 1];
    for (int64_t base = b; base < e; base += BLOCK) {
      const int64_t m = base + threadIdx.x;
      uint32_t v = 0, gsv = 0;
      if (m < e) {
        v = (uint32_t)gmem[m];
        gsv = c.GM[(int64_t)tk.slot * c.Vp + v];
        if ((uint32_t)__builtin_ctz(gsv | 0x100u) != (uint32_t)tk.bit) gsv = 0;
      }
      uint32_t p = gsv;  // enumerate non-empty submasks of gsv
      bool live = p != 0;
      while (__any(live)) {
        uint32_t want = 0;
        uint64_t item = 0;
        if (live) {
          atomicMin(&c.T[2 * t_ix(S, p, v)], 0);
          if (c.BP) c.BP[t_ix(S, p, v)] = (int32_t)BP_BASE;
          if ((int32_t)p == S.full) {
            ub_update(&c.bestw[tk.slot], 0, wit_full(v, 0));
          } else {
            if (!hop) {
              const int32_t sumc = sum_c(&c.C[S.c_off + (int64_t)v * CROW], (uint32_t)S.full ^ p,
                                         c.sched[tk.slot].R);
              if (sumc < INF) ub_update(&c.bestw[tk.slot], sumc, wit_onelabel(v, p));
            }
            if (mark_state(c, S, v, p, 0)) {
              want = 1;
              item = mk_item(v, 0, 0, 0, (uint32_t)tk.slot);
            }
          }
          p = (p - 1) & gsv;
          live = p != 0;
        }
        wave_append(want, item, wl, cnt, c.cap, c.ctr);
      }
    }
  }
}

void launch_dp_seed(const DevCtx& c, const Task* tasks, int ntasks, const int64_t* goff,
                    const int32_t* gmem, uint64_t* wl, uint32_t* cnt, hipStream_t st) {
  if (ntasks <= 0) return;
  int grid = ntasks < 4096 ? ntasks : 4096;
  hipLaunchKernelGGL(k_dp_seed, grid, BLOCK, 0, st, c, tasks, ntasks, goff, gmem, wl, cnt);
}

// ---------------------------------------------------------------------------
// The DP frontier step.
//
// Plain (TrimCDP) item (v, p) with t = T(v, p), best = the query's upper bound:
//   * complement: T(v, full) <- t + T(v, full ^ p); best <- that   (DPQ.cu:95-108)
//   * 1/2 rule: expand only if t < best/2 + 2                       (DPQ.cu:127;
//     the +2 safeguard of code/TrimCDP-multi-core-CPU/include/CPUNONHOP.h:741)
//   * merge with every finite T(v, p2), p2 a proper submask of full ^ p; an
//     improved T(v, P) is queued if 1000c <= 667 best + 2000 and c + lb <= best + 2
//   * grow over the CSR chunk: T(u, p) <- t + w(v, u); queued if c + lb <= best + 2
//   * every improvement also offers c + sum_{i not in P} c_i(u) as an upper bound
// Hop (D-TrimCDP) item (v, p, d):
//   * merge with T(v, p2, d2), p2 a non-empty submask of full ^ p, d2 <= D - d,
//     into (v, p | p2, max(d, d2)); a full result updates best
//   * grow (d < D) into T(u, p, d + 1)
//   * with hop pruning on, an improvement is queued only if c + lb <= best
//     (lb from hop-free distances is a valid bound for hop-bounded trees)

template <bool HOP>
__global__ __launch_bounds__(BLOCK, 4) void k_dp_expand(DevCtx c, WL w) {
  __shared__ int32_t s_off[BLOCK + 1];
  __shared__ int32_t s_wsum[BLOCK / WAVE];
  __shared__ int32_t s_beg[BLOCK];
  __shared__ int32_t s_cost[BLOCK];
  __shared__ int64_t s_dst[BLOCK];     // T offset of the destination row of vertex 0
  __shared__ int32_t s_rows[BLOCK];    // rows per vertex of the item's slot
  __shared__ uint32_t s_meta[BLOCK];   // |full ^ p| << 24 | slot << 12 | d << 8 | p
  __shared__ uint32_t s_v[BLOCK];      // merge side: the item's vertex
  __shared__ int32_t s_mlim[BLOCK];    //   largest useful partner cost
  __shared__ uint32_t s_mm[BLOCK];     //   mask | partner depths << 8
  __shared__ int64_t s_tb[BLOCK];      //   T offset of the vertex's row 0
  __shared__ uint4 s_cv[BLOCK];        //   one-label distances, 16 bits each
  __shared__ uint4 s_cand[BLOCK];      //   plain: the reached partner subsets
  __shared__ uint64_t s_stage[BLOCK / WAVE][WSTAGE];
  __shared__ uint64_t s_fstage[BLOCK / WAVE][WSTAGE];  // far-list appends (banded DP)
  uint64_t* stage = s_stage[threadIdx.x / WAVE];
  uint64_t* fstage = s_fstage[threadIdx.x / WAVE];
  uint32_t st_n = 0, fst_n = 0;
  const bool far_mode = c.far_mode != 0;
  const bool lane_ap = c.lane_appends != 0;
  const int far_sel = far_mode ? *c.far_sel : 0;
  uint64_t* far_out = c.far[far_sel];
  uint32_t* far_out_cnt = c.far_cnt + far_sel;
  const uint32_t n = min(*w.in_cnt, c.cap);  // an overflowed list is never read past its end
  const int tid = threadIdx.x;
  const int32_t CH = 1 << c.chunk_shift;
  // nLb: one-label terms of the completion bounds of improved states
  // (|full ^ P| per successful decrease, where the reference evaluates its
  // lower bound); nUb: one-label rows read for upper bounds
  unsigned long long nF = 0, nRg = 0, nRm = 0, nU = 0, nLb = 0, nUb = 0, nItems = 0;

  for (uint32_t tile = blockIdx.x; tile * BLOCK < n; tile += gridDim.x) {
    const uint32_t idx = tile * BLOCK + tid;
    int32_t deg = 0;
    uint32_t v = 0, p = 0, d = 0, s = 0, mask = 0;
    int32_t t = INF, Dq = 0;
    uint32_t carry_cnt = 0, far_cnt_l = 0;
    uint64_t citem = 0, far_item_l = 0;
    SlotDesc S;
    S.t_off = S.e_off = S.c_off = S.r_off = 0;
    S.k = S.full = S.D = S.nsub = S.vv_off = 0;
    if (idx < n) {
      ++nItems;
      const uint64_t it = w.in[idx];
      v = it_v(it);
      p = it_p(it);
      d = it_d(it);
      s = it_s(it);
      const uint32_t ch = it_ch(it);
      S = c.slots[s];
      Dq = S.D;
      const int64_t row = HOP ? (int64_t)d * S.nsub + p : (int64_t)p;
      const int2 tl = *reinterpret_cast<const int2*>(&c.T[2 * t_ix(S, row, v)]);
      t = tl.x;
      // A* key of the state: its cost plus its completion bound. With bands
      // (GST_OPT_DP_DELTA_DIV > 0) states are expanded in increasing key, so
      // states that can lie on the optimum come first and the upper bound
      // drops early. Capped below INF so every key can open a band.
      const int32_t f = min(t + tl.y, INF - 1);
      Sched* sc = &c.sched[s];
      bool expand = t < INF;
      const int32_t best = best_of(&c.bestw[s]);
      // the state's own completion bound: useless once t + lb exceeds best + 2 (hop: best)
      if (!HOP && t + tl.y > best + 2) expand = false;
      else if (HOP && c.hop_prune && best < INF && t + tl.y > best) expand = false;
      else if (!HOP) {
        const bool half = t < best / 2 + 2;  // 1/2 rule: costlier states are never expanded
        if (ch == 0 && (!half || f < sc->thr)) complement_merge(c, S, s, v, p, t);
        if (!half) expand = false;
      }
      if (expand && f >= sc->thr) {  // not in the current band yet
        expand = false;
        if (far_mode) {
          // to the far list, once per state (chunk items are rebuilt on return)
          if (ch == 0) {
            far_cnt_l = 1;
            far_item_l = mk_item(v, p, d, 0, s);
          }
        } else {
          if (f < sc->mindef) atomicMin(&sc->mindef, f);
          if (ch == 0 && mark_state(c, S, v, p, d)) {
            carry_cnt = 1;
            citem = mk_item(v, 0, d, 0, s);
          }
        }
      }
      if (expand) {
        if (!sc->active) sc->active = 1;
        nF += ch == 0;
        // a grown state whose cost plus completion bound exceeds best + 2
        // (hop: best) is never written: its bound lb(u, full ^ p) is at least
        // the slot's smallest for that subset, so edges heavier than
        // best + 2 - t - that minimum are skipped
        const int32_t bnow = best_of(&c.bestw[s]);
        int32_t limit = INF;
        if (bnow < INF && (!HOP || c.hop_prune)) {
          const int32_t lbmin = c.R[S.r_off + (int64_t)((uint32_t)S.full ^ p) * RSTRIDE + RLBMIN];
          limit = (HOP ? bnow : bnow + 2) - t - min(lbmin, INF);
        }
        const int32_t rb = c.rowp[v] + (int32_t)(ch << c.chunk_shift);
        const int32_t re = row_cut(c.wt, rb, min(c.rowp[v + 1], rb + CH), limit);
        deg = (HOP && (int32_t)d >= Dq) ? 0 : max(0, re - rb);
        s_beg[tid] = rb;
        s_cost[tid] = t;
        s_v[tid] = v;
        s_dst[tid] = S.t_off + (HOP ? ((int64_t)(d + 1) * S.nsub + p) : (int64_t)p);
        s_rows[tid] = S.nrows;
        s_meta[tid] = ((uint32_t)__popc((uint32_t)S.full ^ p) << 24) | (s << 12) | (d << 8) | p;
        if (ch == 0) mask = (uint32_t)S.full ^ p;
      }
    }

    stage_append(stage, st_n, carry_cnt, citem, w.out, w.out_cnt, c.cap, c.ctr, lane_ap);
    if (far_mode)
      stage_append(fstage, fst_n, far_cnt_l, far_item_l, far_out, far_out_cnt, c.cap, c.ctr,
                   lane_ap);

    // The merge of item (vj, pj, dj) of slot sl (cost tj, its vertex's state row at tb,
    // mk = full ^ pj) with partner (vj, p2, d2) of cost t2 (t2 < INF, within the
    // margin): writes T(vj, pj | p2, max(dj, d2)) when the merged tree is cheaper and
    // can still beat the upper bound, and queues it.
    auto merge_pair = [&](uint32_t sl, uint32_t vj, uint32_t pj, uint32_t dj, uint32_t mk,
                          uint32_t p2, uint32_t d2, int32_t tj, int32_t t2, int64_t tb,
                          uint32_t& want, uint64_t& item, uint32_t& fwant, uint64_t& fitem) {
      const uint32_t P = pj | p2;
      const int32_t mc = tj + t2;
      if (mc >= INF) c.sched[sl].sat = 1;  // a tree cost past the int32 cost range
      const uint32_t nd = HOP ? max(dj, d2) : 0u;
      const int64_t drow = HOP ? (int64_t)nd * (int64_t)((mk | pj) + 1) + P : (int64_t)P;
      int32_t* dst = &c.T[2 * (tb + drow)];
      const int2 dl = *reinterpret_cast<const int2*>(dst);
      if (!(mc < dl.x)) return;
      if (HOP && p2 == mk) {  // a full tree
        if (mc < atomicMin(dst, mc)) {
          ++nU;
          if (c.BP) c.BP[tb + drow] = (int32_t)bp_merge(pj, dj, d2);
          ub_update(&c.bestw[sl], mc, wit_full(vj, nd));
        }
        return;
      }
      // Completion bound first: a state that cannot beat best + 2 (hop: best)
      // lies on no better solution, so its entry is left alone and no atomic
      // is issued.
      int32_t best = INF;
      bool keep = true;
      if (!HOP || c.hop_prune) {
        best = best_of(&c.bestw[sl]);
        keep = dl.y < INF && (best >= INF || mc + dl.y <= (HOP ? best : best + 2));
      }
      if (!(keep && mc < atomicMin(dst, mc))) return;
      ++nU;
      nLb += (uint32_t)__popc(mk ^ p2);  // the improved state's bound terms
      const SlotDesc Sj = c.slots[sl];
      if (c.BP) c.BP[tb + drow] = (int32_t)bp_merge(pj, dj, d2);
      if (!HOP) {
        ++nUb;
        const int32_t sumc = sum_c(&c.C[Sj.c_off + (int64_t)vj * CROW], mk ^ p2, c.sched[sl].R);
        if (sumc < INF && mc + sumc < best) {
          ub_update(&c.bestw[sl], mc + sumc, wit_onelabel(vj, P));
          best = mc + sumc;
        }
      }
      // 2/3 rule (Conditional Tree Merging) for the queue only; a state left
      // out of the queue is combined with its complement now
      const bool enq = HOP || (int64_t)1000 * mc <= (int64_t)667 * best + 2000;
      if (!enq) {
        set_finite(c, Sj, vj, P, nd);
        complement_merge(c, Sj, sl, vj, P, mc);
        return;
      }
      if (far_mode && min(mc + dl.y, INF - 1) >= c.sched[sl].thr) {
        fwant = 1;  // past the band: wait in the far list
        fitem = mk_item(vj, P, nd, 0, sl);
        set_finite(c, Sj, vj, P, nd);
      } else if (mark_state(c, Sj, vj, P, nd)) {
        want = 1;
        item = mk_item(vj, 0, nd, 0, sl);
      }
    };
    bool fin_item = false;  // the item's merges walk its finite masks (fin slots)
    uint32_t fin_mask = 0, fin_nd2 = 0;
    int64_t fin_base = 0;   // finite word of (v, 0); depth d2 at + 2 d2
    // ---- merges, balanced over the workgroup like the grow edges: item j
    // owns pairs s_off[j] .. s_off[j + 1] - 1; its pair r merges with subset
    // number r / nd2 + 1 of its mask (the bits of that number deposited into
    // the mask's set bits) at partner depth r % nd2 (hop: nd2 = D - d + 1;
    // plain: nd2 = 1 and the mask itself is left out, the complement having
    // been merged above).
    {
      int32_t mcnt = 0;
      if (mask) {
        // A merged tree costlier than best + 2 (hop: best) is useless, and
        // T(v, p2) >= max_{i in p2} c_i(v), so partners whose one-label bound
        // already exceeds the margin are skipped without reading T. The
        // labels are kept in LDS capped at 65535 (still lower bounds).
        int32_t mlim = INF;
        const int32_t bnow = best_of(&c.bestw[s]);
        if (bnow < INF) {
          if (!HOP) mlim = bnow + 2 - t;
          else if (c.hop_prune) mlim = bnow - t;
        }
        uint4 cv = make_uint4(0u, 0u, 0u, 0u);
        if (mlim < INF) {
          const int32_t* crow = &c.C[S.c_off + (int64_t)v * CROW];
          const int4 a = *reinterpret_cast<const int4*>(crow);
          const int4 b = *reinterpret_cast<const int4*>(crow + 4);
          auto h = [](int32_t x) { return (uint32_t)min(x, 65535); };
          cv = make_uint4(h(a.x) | h(a.y) << 16, h(a.z) | h(a.w) << 16, h(b.x) | h(b.y) << 16,
                          h(b.z));
        }
        const int32_t nd2 = HOP ? Dq - (int32_t)d + 1 : 1;
        if (HOP && S.fin) {
          fin_item = true;
          fin_mask = mask;
          fin_nd2 = (uint32_t)nd2;
          fin_base = mask_ix(S, v, 0) + 1;
        } else if (!HOP && S.fin) {
          // the partners already reached: non-empty proper subsets of the mask
          // (its complement was merged above) whose finite bit is set, as one
          // snapshot the pairs below are numbered in
          const uint32_t* fw = c.M + mask_ix(S, v, 0);
          const uint4 subs = subsets128(mask);
          uint32_t cw[4] = {fw[1] & subs.x, 0u, 0u, 0u};
          if (S.mw > 2) cw[1] = fw[3] & subs.y;
          if (S.mw > 4) {
            cw[2] = fw[5] & subs.z;
            cw[3] = fw[7] & subs.w;
          }
          cw[0] &= ~1u;
          cw[mask >> 5] &= ~(1u << (mask & 31u));
          s_cand[tid] = make_uint4(cw[0], cw[1], cw[2], cw[3]);
          mcnt = __popc(cw[0]) + __popc(cw[1]) + __popc(cw[2]) + __popc(cw[3]);
        } else {
          mcnt = ((1 << __popc(mask)) - (HOP ? 1 : 2)) * nd2;
        }
        s_v[tid] = v;
        s_mlim[tid] = mlim;
        s_mm[tid] = mask | ((uint32_t)nd2 << 8);
        s_tb[tid] = S.t_off + (int64_t)v * S.nrows;
        s_cv[tid] = cv;
      }
      block_scan(mcnt, s_off, s_wsum);
    }
    const int32_t mtotal = s_off[BLOCK];
    for (int32_t base = 0; base < mtotal; base += BLOCK * MPL) {
      int jq[MPL];
      uint32_t p2q[MPL], d2q[MPL];
      int32_t t2q[MPL];
#pragma unroll
      for (int q = 0; q < MPL; ++q) {
        const int32_t e = base + q * BLOCK + tid;
        jq[q] = e < mtotal ? find_item(s_off, e) : -1;
      }
#pragma unroll
      for (int q = 0; q < MPL; ++q) {
        p2q[q] = 0;
        d2q[q] = 0;
        t2q[q] = INF;
        if (jq[q] >= 0) {
          const int j = jq[q];
          const uint32_t r = (uint32_t)(base + q * BLOCK + tid - s_off[j]);
          const uint32_t mm = s_mm[j], mk = mm & 0xFFu;
          uint32_t idx = r + 1;
          if (HOP) {
            const uint32_t nd2 = mm >> 8;
            idx = r / nd2 + 1;
            d2q[q] = r - (idx - 1) * nd2;
          }
          uint32_t p2 = 0;
          if (HOP) {
            for (uint32_t rest = mk; rest; rest &= rest - 1u, idx >>= 1)
              if (idx & 1u) p2 |= rest & (0u - rest);
          } else {
            p2 = nth_bit128(s_cand[j], r);
          }
          p2q[q] = p2;
          const uint4 cv = s_cv[j];
          const uint32_t w4[4] = {cv.x, cv.y, cv.z, cv.w};
          int32_t lbp = 0;
#pragma unroll
          for (int i = 0; i < 7; ++i)
            if ((p2 >> i) & 1u) lbp = max(lbp, (int32_t)((w4[i >> 1] >> (16 * (i & 1))) & 0xFFFFu));
          if (lbp <= s_mlim[j]) {
            ++nRm;
            const int32_t nsub = (int32_t)(mk | (s_meta[j] & 0xFFu)) + 1;
            t2q[q] = c.T[2 * (s_tb[j] + (HOP ? (int64_t)d2q[q] * nsub : 0) + p2)];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < MPL; ++q) {
        uint32_t want = 0, fwant = 0;
        uint64_t item = 0, fitem = 0;
        const int j = jq[q];
        if (j >= 0 && t2q[q] < INF && t2q[q] <= s_mlim[j]) {
          const uint32_t meta = s_meta[j];
          merge_pair((meta >> 12) & 0xFFFu, s_v[j], meta & 0xFFu, (meta >> 8) & 0xFu,
                     s_mm[j] & 0xFFu, p2q[q], d2q[q], s_cost[j], t2q[q], s_tb[j], want, item,
                     fwant, fitem);
        }
        stage_append(stage, st_n, want, item, w.out, w.out_cnt, c.cap, c.ctr, lane_ap);
        if (far_mode)
          stage_append(fstage, fst_n, fwant, fitem, far_out, far_out_cnt, c.cap, c.ctr, lane_ap);
      }
    }
    // Hop-bounded slots of up to 5 groups (fin): each lane walks the partners
    // of its own item that are already reached, depth by depth, from the
    // vertex's finite masks (one cache line); the partners never reached,
    // nearly all of those the dense enumeration reads, are not visited. A
    // partner reached after its word was read is merged when it is expanded.
    if (HOP) {
      bool act = fin_item;
      uint32_t d2 = 0, cw = 0;
      const uint32_t subs = act ? subsets_of(fin_mask) : 0u;
      if (act) cw = c.M[fin_base] & subs;
      while (__any(act)) {
        uint32_t want = 0, fwant = 0;
        uint64_t item = 0, fitem = 0;
        if (act) {
          if (cw) {
            const uint32_t p2 = (uint32_t)__builtin_ctz(cw);
            cw &= cw - 1u;
            const uint4 cv = s_cv[tid];
            const uint32_t w4[4] = {cv.x, cv.y, cv.z, cv.w};
            int32_t lbp = 0;
#pragma unroll
            for (int i = 0; i < 7; ++i)
              if ((p2 >> i) & 1u)
                lbp = max(lbp, (int32_t)((w4[i >> 1] >> (16 * (i & 1))) & 0xFFFFu));
            const int32_t mlim = s_mlim[tid];
            if (lbp <= mlim) {
              ++nRm;
              const int64_t tb = s_tb[tid];
              const int32_t t2 = c.T[2 * (tb + (int64_t)d2 * (int32_t)(fin_mask | p) + d2 + p2)];
              if (t2 < INF && t2 <= mlim)
                merge_pair(s, v, p, d, fin_mask, p2, d2, t, t2, tb, want, item, fwant, fitem);
            }
          } else if (++d2 < fin_nd2) {
            cw = c.M[fin_base + 2 * (int64_t)d2] & subs;
          } else {
            act = false;
          }
        }
        stage_append(stage, st_n, want, item, w.out, w.out_cnt, c.cap, c.ctr, lane_ap);
        if (far_mode)
          stage_append(fstage, fst_n, fwant, fitem, far_out, far_out_cnt, c.cap, c.ctr, lane_ap);
      }
    }
    __syncthreads();  // s_off is reused by the grow scan

    // ---- grow: the workgroup's edges, balanced by an LDS prefix scan
    // (ablation: each lane walks its own item's edges, as in the SSSP kernel)
    block_scan(deg, s_off, s_wsum);
    const int32_t total = s_off[BLOCK];
    const bool per_item = c.per_item_edges != 0;
    const int32_t span = per_item ? block_max(deg, s_wsum) * BLOCK : total;
    // DEPL edges per lane per round, all loads issued before any is used; the
    // next round's CSR reads are issued before this round's state words are
    // used, so each round waits on one memory latency, not two
    int jq[DEPL];
    uint32_t uq[DEPL];
    int32_t wq[DEPL];
    // item, neighbour and weight of the DEPL edges of the round at `b`
    auto fetch = [&](int32_t b, int* jo, uint32_t* uo, int32_t* wo) {
#pragma unroll
      for (int q = 0; q < DEPL; ++q) {
        int32_t o;
        if (per_item) {
          o = b / BLOCK + q;
          jo[q] = o < deg ? tid : -1;
        } else {
          const int32_t e = b + q * BLOCK + tid;
          jo[q] = e < total ? find_item(s_off, e) : -1;
          o = jo[q] >= 0 ? e - s_off[jo[q]] : 0;
        }
        uo[q] = 0;
        wo[q] = 0;
        if (jo[q] >= 0) {
          const int32_t ej = s_beg[jo[q]] + o;
          uo[q] = (uint32_t)c.col[ej];
          wo[q] = c.wt[ej];
        }
      }
    };
    if (span > 0) fetch(0, jq, uq, wq);
    for (int32_t base = 0; base < span; base += BLOCK * DEPL) {
      int2 dq[DEPL];
#pragma unroll
      for (int q = 0; q < DEPL; ++q)
        if (jq[q] >= 0)
          dq[q] = *reinterpret_cast<const int2*>(
              &c.T[2 * (s_dst[jq[q]] + (int64_t)uq[q] * s_rows[jq[q]])]);
      int jn[DEPL];
      uint32_t un[DEPL];
      int32_t wn[DEPL];
      if (base + BLOCK * DEPL < span) {
        fetch(base + BLOCK * DEPL, jn, un, wn);
      } else {
#pragma unroll
        for (int q = 0; q < DEPL; ++q) {
          jn[q] = -1;
          un[q] = 0;
          wn[q] = 0;
        }
      }
#pragma unroll
      for (int q = 0; q < DEPL; ++q) {
        uint32_t want = 0, fwant = 0;
        uint64_t item = 0, fitem = 0;
        if (jq[q] >= 0) {
          const int j = jq[q];
          const uint32_t u = uq[q];
          const int32_t gc = s_cost[j] + wq[q];
          const int64_t dbase = s_dst[j];
          const int64_t uoff = (int64_t)u * s_rows[j];
          int32_t* dst = &c.T[2 * (dbase + uoff)];
          const int2 dl = dq[q];
          ++nRg;
          if (gc >= INF) c.sched[(s_meta[j] >> 12) & 0xFFFu].sat = 1;  // past the cost range
          if (gc < dl.x) {
            const uint32_t meta = s_meta[j];
            const uint32_t sl = (meta >> 12) & 0xFFFu, pd = meta & 0xFFu, dd = (meta >> 8) & 0xFu;
            int32_t best = INF;
            bool keep = true;
            if (!HOP || c.hop_prune) {  // completion bound before the atomic, as for merges
              best = best_of(&c.bestw[sl]);
              keep = dl.y < INF && (best >= INF || gc + dl.y <= (HOP ? best : best + 2));
            }
            if (keep && gc < atomicMin(dst, gc)) {
              ++nU;
              nLb += meta >> 24;  // the improved state's bound terms
              const SlotDesc Sj = c.slots[sl];
              if (c.BP) c.BP[dbase + uoff] = (int32_t)bp_grow(s_v[j]);
              if (!HOP) {
                ++nUb;
                const int32_t sumc = sum_c(&c.C[Sj.c_off + (int64_t)u * CROW],
                                           (uint32_t)Sj.full ^ pd, c.sched[sl].R);
                if (sumc < INF && gc + sumc < best)
                  ub_update(&c.bestw[sl], gc + sumc, wit_onelabel(u, pd));
              }
              if (far_mode && min(gc + dl.y, INF - 1) >= c.sched[sl].thr) {
                fwant = 1;  // past the band: wait in the far list
                fitem = mk_item(u, pd, HOP ? dd + 1 : 0, 0, sl);
                set_finite(c, Sj, u, pd, HOP ? dd + 1 : 0);
              } else if (mark_state(c, Sj, u, pd, HOP ? dd + 1 : 0)) {
                want = 1;
                item = mk_item(u, 0, HOP ? dd + 1 : 0, 0, sl);
              }
            }
          }
        }
        stage_append(stage, st_n, want, item, w.out, w.out_cnt, c.cap, c.ctr, lane_ap);
        if (far_mode)
          stage_append(fstage, fst_n, fwant, fitem, far_out, far_out_cnt, c.cap, c.ctr, lane_ap);
      }
#pragma unroll
      for (int q = 0; q < DEPL; ++q) {
        jq[q] = jn[q];
        uq[q] = un[q];
        wq[q] = wn[q];
      }
    }
    __syncthreads();
  }
  stage_flush(stage, st_n, w.out, w.out_cnt, c.cap, c.ctr);
  stage_flush(fstage, fst_n, far_out, far_out_cnt, c.cap, c.ctr);
  flush(c.ctr, C_F, nF);
  flush(c.ctr, C_RG, nRg);
  flush(c.ctr, C_RM, nRm);
  flush(c.ctr, C_U, nU);
  flush(c.ctr, C_LB, nLb);
  flush(c.ctr, C_UBROW, nUb);
  flush(c.ctr, C_ITEMS, nItems);
}

void launch_dp_expand(const DevCtx& c, const WL& w, int grid, bool hop, hipStream_t st) {
  if (hop) hipLaunchKernelGGL(k_dp_expand<true>, grid, BLOCK, 0, st, c, w);
  else hipLaunchKernelGGL(k_dp_expand<false>, grid, BLOCK, 0, st, c, w);
}

// ---------------------------------------------------------------------------
// Tree recovery. One wavefront per slot walks back from the best word's
// witness. A state (v, p[, d]) with budget b (its current table value) is a
// base state when p lies inside v's groups; otherwise some merge split with
// T(v,p1[,d1]) + T(v,p2[,d2]) <= b or some neighbour u with
// T(u,p[,d-1]) + w(u,v) <= b exists, because every table value was written
// from predecessor values that can only have decreased since. Recursing on
// such predecessors yields an edge set of weight <= cost; at the optimum it
// is an optimal tree. Among candidates the lowest index is taken.

namespace {
constexpr int STACK = 2048;

__device__ inline uint32_t deposit(uint32_t r, uint32_t mask) {  // r-th submask of mask
  uint32_t out = 0;
  for (uint32_t b = 0; mask; ++b) {
    const uint32_t low = mask & (0u - mask);
    if ((r >> b) & 1u) out |= low;
    mask ^= low;
  }
  return out;
}
}  // namespace

__global__ __launch_bounds__(64) void k_backtrack(DevCtx c, int max_edges, int32_t* out_uv,
                                                  int32_t* out_eid, int32_t* out_n,
                                                  int32_t* out_err, const int32_t* sel) {
  __shared__ int32_t st_v[STACK], st_p[STACK], st_d[STACK], st_b[STACK];
  __shared__ int32_t s_sp, s_ne, s_err;
  __shared__ int32_t s_pick;
  const int s = sel ? sel[blockIdx.x] : (int)blockIdx.x;
  const int o = blockIdx.x;  // output row
  const int lane = threadIdx.x;
  const SlotDesc S = c.slots[s];
  const uint64_t bw = c.bestw[s];
  const int32_t cost = (int32_t)(bw >> 32);
  const uint32_t wit = (uint32_t)bw;
  int32_t* ouv = out_uv + (int64_t)o * max_edges * 2;
  int32_t* oid = out_eid + (int64_t)o * max_edges;
  const bool hop = S.D >= 0;
  if (lane == 0) {
    s_sp = 0;
    s_ne = 0;
    s_err = 0;
  }
  __syncthreads();
  if (cost >= INF) {
    if (lane == 0) {
      out_n[o] = 0;
      out_err[o] = 0;
    }
    return;
  }
  auto Tat = [&](uint32_t v, uint32_t p, uint32_t d) -> int32_t {
    return c.T[2 * t_ix(S, (int64_t)d * S.nsub + p, v)];
  };
  auto emit = [&](uint32_t a, uint32_t b2, int32_t ej) {  // lane 0 only
    if (s_ne >= max_edges) {
      s_err = 2;
      return;
    }
    ouv[2 * s_ne] = (int32_t)a;
    ouv[2 * s_ne + 1] = (int32_t)b2;
    oid[s_ne] = c.eid ? c.eid[ej] : ej;
    ++s_ne;
  };
  const uint32_t rv = wit & 0xFFFFFFu, rx = (wit >> 24) & 0x7Fu;
  const bool onelabel = (wit >> 31) != 0;
  if (lane == 0) {
    if (onelabel) {
      st_v[0] = (int32_t)rv;
      st_p[0] = (int32_t)rx;
      st_d[0] = 0;
      st_b[0] = Tat(rv, rx, 0);
    } else {
      const uint32_t d = hop ? rx : 0u;
      st_v[0] = (int32_t)rv;
      st_p[0] = S.full;
      st_d[0] = (int32_t)d;
      st_b[0] = Tat(rv, (uint32_t)S.full, d);
    }
    s_sp = 1;
  }
  __syncthreads();
  // shortest paths from the witness vertex to the groups outside p
  if (onelabel) {
    const uint32_t U = (uint32_t)S.full ^ rx;
    for (int i = 0; i < S.k; ++i) {
      if (!((U >> i) & 1u)) continue;
      uint32_t x = rv;
      for (int guard = 0; guard < c.V && s_err == 0; ++guard) {
        const int32_t cx = c.C[S.c_off + (int64_t)x * CROW + i];
        if (cx == 0) break;
        const int32_t rb = c.rowp[x], re = c.rowp[x + 1];
        int32_t pick = -1;
        for (int32_t jb = rb; jb < re && pick < 0; jb += 64) {
          const int32_t j = jb + lane;
          bool ok = false;
          if (j < re) {
            const uint32_t y = (uint32_t)c.col[j];
            ok = c.C[S.c_off + (int64_t)y * CROW + i] + c.wt[j] <= cx;
          }
          const uint64_t m = __ballot(ok);
          if (m) pick = jb + (int32_t)__ffsll((unsigned long long)m) - 1;
        }
        if (pick < 0) {
          if (lane == 0) s_err = 4;
          __syncthreads();
          break;
        }
        const uint32_t y = (uint32_t)c.col[pick];
        if (lane == 0) emit(x, y, pick);
        __syncthreads();
        x = y;
      }
    }
  }
  for (int guard = 0;; ++guard) {
    __syncthreads();
    if (s_sp == 0 || s_err) break;
    if (guard > 8 * max_edges + 4096) {  // cannot happen for a consistent table
      if (lane == 0) s_err = 5;
      __syncthreads();
      break;
    }
    const int top = s_sp - 1;
    const uint32_t v = (uint32_t)st_v[top], p = (uint32_t)st_p[top], d = (uint32_t)st_d[top];
    const int32_t b = st_b[top];
    __syncthreads();
    if (lane == 0) s_sp = top;
    const uint32_t gs = c.GM[(int64_t)s * c.Vp + v];
    __syncthreads();
    if ((p & ~gs) == 0) continue;  // p inside v's groups: a single vertex of cost 0
    if (c.BP) {
      // the stored backpointer, used when it checks out against the final table
      const uint32_t bp = (uint32_t)c.BP[t_ix(S, (int64_t)d * S.nsub + p, v)];
      bool used = false;
      if (bp != BP_BASE && (bp >> 31)) {
        const uint32_t p1 = bp & 0x7Fu, d1 = (bp >> 7) & 0xFu, d2 = (bp >> 11) & 0xFu;
        const uint32_t p2 = p ^ p1;
        const bool shape = p1 != 0 && (p1 & ~p) == 0 && p2 != 0 &&
                           (hop ? (max(d1, d2) == d && (int32_t)(d1 + d2) <= S.D)
                                : (d1 == 0 && d2 == 0));
        if (shape) {
          const int32_t a1 = Tat(v, p1, d1), a2 = Tat(v, p2, d2);
          if (a1 < INF && a2 < INF && a1 + a2 <= b) {
            used = true;
            if (lane == 0) {
              if (s_sp + 2 > STACK) {
                s_err = 3;
              } else {
                st_v[s_sp] = (int32_t)v;
                st_p[s_sp] = (int32_t)p1;
                st_d[s_sp] = (int32_t)d1;
                st_b[s_sp] = a1;
                st_v[s_sp + 1] = (int32_t)v;
                st_p[s_sp + 1] = (int32_t)p2;
                st_d[s_sp + 1] = (int32_t)d2;
                st_b[s_sp + 1] = a2;
                s_sp += 2;
              }
            }
          }
        }
      } else if (bp != BP_BASE && bp < (uint32_t)c.V && (!hop || d >= 1)) {
        const uint32_t u = bp, pd = hop ? d - 1 : 0;
        const int32_t tu = Tat(u, p, pd);
        int32_t pick = -1;
        if (tu < INF) {
          // the CSR entry of edge (v, u): rows are sorted by weight, so scan
          const int32_t rb = c.rowp[v], re = c.rowp[v + 1];
          for (int32_t jb = rb; jb < re && pick < 0; jb += 64) {
            const int32_t j = jb + lane;
            const bool ok = j < re && (uint32_t)c.col[j] == u && tu + c.wt[j] <= b;
            const uint64_t m = __ballot(ok);
            if (m) pick = jb + (int32_t)__ffsll((unsigned long long)m) - 1;
          }
        }
        if (pick >= 0) {
          used = true;
          if (lane == 0) {
            emit(v, u, pick);
            if (s_sp + 1 > STACK) {
              s_err = 3;
            } else {
              st_v[s_sp] = (int32_t)u;
              st_p[s_sp] = (int32_t)p;
              st_d[s_sp] = (int32_t)pd;
              st_b[s_sp] = tu;
              s_sp += 1;
            }
          }
        }
      }
      if (used) continue;
      if (lane == 0) atomicAdd(&c.ctr[C_BP_MISS], 1ull);  // stale word: search instead
    }
    // merge splits: proper non-empty submasks p1 (hop: with layers d1, d2,
    // max(d1, d2) = d, d1 + d2 <= D)
    const uint32_t pc = (uint32_t)__popc(p);
    const uint32_t nsplit = (1u << pc) - 2u;
    const uint32_t npair = hop ? 2u * d + 1u : 1u;
    const uint32_t ncand = nsplit * npair;
    int32_t pick = -1;
    for (uint32_t cb = 0; cb < ncand && pick < 0; cb += 64) {
      const uint32_t r = cb + lane;
      bool ok = false;
      if (r < ncand) {
        const uint32_t p1 = deposit(r / npair + 1u, p), p2 = p ^ p1;
        uint32_t d1 = 0, d2 = 0;
        bool legal = true;
        if (hop) {
          const uint32_t q = r % npair;  // (d, 0..d) then (0..d-1, d)
          if (q <= d) {
            d1 = d;
            d2 = q;
          } else {
            d1 = q - d - 1;
            d2 = d;
          }
          legal = (int32_t)(d1 + d2) <= S.D;
        }
        if (legal) {
          const int32_t a1 = Tat(v, p1, d1), a2 = Tat(v, p2, d2);
          ok = a1 < INF && a2 < INF && a1 + a2 <= b;
        }
      }
      const uint64_t m = __ballot(ok);
      if (m) pick = (int32_t)(cb + __ffsll((unsigned long long)m) - 1);
    }
    if (pick >= 0) {
      if (lane == 0) {
        const uint32_t r = (uint32_t)pick;
        const uint32_t p1 = deposit(r / npair + 1u, p), p2 = p ^ p1;
        uint32_t d1 = 0, d2 = 0;
        if (hop) {
          const uint32_t q = r % npair;
          if (q <= d) {
            d1 = d;
            d2 = q;
          } else {
            d1 = q - d - 1;
            d2 = d;
          }
        }
        if (s_sp + 2 > STACK) {
          s_err = 3;
        } else {
          st_v[s_sp] = (int32_t)v;
          st_p[s_sp] = (int32_t)p1;
          st_d[s_sp] = (int32_t)d1;
          st_b[s_sp] = Tat(v, p1, d1);
          st_v[s_sp + 1] = (int32_t)v;
          st_p[s_sp + 1] = (int32_t)p2;
          st_d[s_sp + 1] = (int32_t)d2;
          st_b[s_sp + 1] = Tat(v, p2, d2);
          s_sp += 2;
        }
      }
      continue;
    }
    // grow: a neighbour u with T(u, p[, d-1]) + w <= b
    if (!hop || d >= 1) {
      const uint32_t pd = hop ? d - 1 : 0;
      const int32_t rb = c.rowp[v], re = c.rowp[v + 1];
      for (int32_t jb = rb; jb < re && pick < 0; jb += 64) {
        const int32_t j = jb + lane;
        bool ok = false;
        if (j < re) {
          const int32_t tu = Tat((uint32_t)c.col[j], p, pd);
          ok = tu < INF && tu + c.wt[j] <= b;
        }
        const uint64_t m = __ballot(ok);
        if (m) pick = jb + (int32_t)__ffsll((unsigned long long)m) - 1;
      }
      if (pick >= 0) {
        if (lane == 0) {
          const uint32_t u = (uint32_t)c.col[pick];
          emit(v, u, pick);
          if (s_sp + 1 > STACK) {
            s_err = 3;
          } else {
            st_v[s_sp] = (int32_t)u;
            st_p[s_sp] = (int32_t)p;
            st_d[s_sp] = (int32_t)pd;
            st_b[s_sp] = Tat(u, p, pd);
            s_sp += 1;
          }
        }
        continue;
      }
    }
    if (lane == 0) s_err = 1;  // no predecessor: table inconsistent
  }
  __syncthreads();
  if (lane == 0) {
    out_n[o] = s_ne;
    out_err[o] = s_err;
  }
}

void launch_backtrack(const DevCtx& c, int nslots, int max_edges, int32_t* out_uv,
                      int32_t* out_eid, int32_t* out_n, int32_t* out_err, const int32_t* sel,
                      hipStream_t st) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_backtrack, nslots, 64, 0, st, c, max_edges, out_uv, out_eid, out_n,
                       out_err, sel);
}

// ---------------------------------------------------------------------------
// backward: grad_w[e] += grad_out[q] for every tree edge e of query q

__global__ void k_backward(int nq, const int64_t* off, const int32_t* ids, const float* gout,
                           float* gw, int64_t nund) {
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const float g = gout[q];
    for (int64_t i = off[q] + threadIdx.x; i < off[q + 1]; i += blockDim.x) {
      const int32_t e = ids[i];
      if (e >= 0 && e < nund) atomicAdd(&gw[e], g);
    }
  }
}

// rows x cols ints from a source of pitch sp to a destination of pitch dp (cols = dp <= sp)
__global__ void k_pack_rows(const int32_t* src, int64_t sp, int32_t* dst, int64_t dp, int rows) {
  for (int r = blockIdx.x; r < rows; r += gridDim.x)
    for (int64_t i = threadIdx.x; i < dp; i += blockDim.x) dst[r * dp + i] = src[r * sp + i];
}

void launch_pack_rows(const int32_t* src, int64_t sp, int32_t* dst, int64_t dp, int rows,
                      hipStream_t st) {
  if (rows <= 0 || dp <= 0) return;
  hipLaunchKernelGGL(k_pack_rows, rows < 4096 ? rows : 4096, 128, 0, st, src, sp, dst, dp, rows);
}

void launch_backward(int nq, const int64_t* off, const int32_t* ids, const float* gout,
                     float* gw, int64_t nund, hipStream_t st) {
  if (nq <= 0) return;
  int grid = nq < 4096 ? nq : 4096;
  hipLaunchKernelGGL(k_backward, grid, 256, 0, st, nq, off, ids, gout, gw, nund);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (gst_solve_noalloc): every
// per-query device buffer of that path is carved from this array. Each GPU
// holds its own copy from module load until the process ends. Host code
// never passes the name itself to a kernel or copy: noalloc_scratch()
// resolves the current GPU's address once and keeps it per device.
static __device__ __align__(256) unsigned char scratch[GST_NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(void** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *out = addr[dev];
  return hipSuccess;
}

}  // namespace gst
