// CPU emulation of the small HIP subset the engine uses, for fault finding
// and logic tests without a GPU: every workgroup runs as one fiber per
// work-item on the calling thread, __syncthreads and the wavefront
// collectives (__ballot, __any, __shfl*) are barriers across fibers, and
// device memory is plain heap memory, so a build with -fsanitize=address
// reports any out-of-bounds access of a kernel with its source line.
// Wavefront collectives must be reached by every live lane of the wave with
// the same operation; anything else aborts with a message, since on the GPU
// it would be undefined.
#pragma once

#include <ucontext.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>
#include <vector>

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct int4 {
  int x, y, z, w;
};
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
struct alignas(16) uint4 {
  unsigned x, y, z, w;
};
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) {
  return uint4{a, b, c, d};
}
struct alignas(8) int2 {
  int x, y;
};
inline int2 make_int2(int a, int b) { return int2{a, b}; }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __launch_bounds__(...)
#define __forceinline__ inline
#define __align__(n) __attribute__((aligned(n)))
#define __HIP_MEMORY_SCOPE_AGENT 0

typedef int hipError_t;
enum {
  hipSuccess = 0,
  hipErrorOutOfMemory = 2,
  hipErrorInvalidValue = 1,
};
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef int hipMemcpyKind;
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1 };
typedef int hipStreamCaptureMode;
enum { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
struct hipDeviceProp_t {
  int multiProcessorCount;
};

namespace emu {

constexpr int WAVE = 64;
enum Op { NONE = 0, SYNC, BALLOT, ANY, SHFL, SHFL_UP, SHFL_DOWN };

struct Fiber {
  ucontext_t ctx;
  dim3 tid;
  bool done = false;
  Op op = NONE;      // barrier / collective the fiber waits in
  uint64_t val = 0;  // operand of a collective
  int arg = 0, width = WAVE;
  uint64_t result = 0;
};

inline std::vector<std::vector<char>>& stacks() {  // reused across workgroups
  static std::vector<std::vector<char>> s;
  return s;
}

struct Block {
  dim3 bid, bdim, gdim;
  std::vector<Fiber> fibers;
  ucontext_t sched;
  int cur = 0;
  std::function<void()> entry;
};

inline Block*& blk() {
  static Block* b = nullptr;
  return b;
}
inline Fiber* cur() { return &blk()->fibers[blk()->cur]; }

inline void trampoline() {
  blk()->entry();
  Fiber* f = cur();
  f->done = true;
  swapcontext(&f->ctx, &blk()->sched);
}

inline void die(const char* msg) {
  std::fprintf(stderr, "hip cpu emulation: %s (block %u,%u thread %u)\n", msg, blk()->bid.x,
               blk()->bid.y, cur()->tid.x);
  std::abort();
}

inline uint64_t wait(Op op, uint64_t val, int arg = 0, int width = WAVE) {
  Fiber* f = cur();
  f->op = op;
  f->val = val;
  f->arg = arg;
  f->width = width;
  swapcontext(&f->ctx, &blk()->sched);
  return f->result;
}

// Releases a complete barrier or collective; returns true if one was released.
inline bool release(Block& b) {
  const int n = (int)b.fibers.size();
  bool all_sync = true, any_live = false;
  for (auto& f : b.fibers)
    if (!f.done) {
      any_live = true;
      if (f.op != SYNC) all_sync = false;
    }
  if (any_live && all_sync) {
    for (auto& f : b.fibers) f.op = NONE;
    return true;
  }
  bool released = false;
  for (int w0 = 0; w0 < n; w0 += WAVE) {
    const int w1 = std::min(n, w0 + WAVE);
    Op op = NONE;
    bool ready = true, live = false;
    for (int i = w0; i < w1; ++i) {
      Fiber& f = b.fibers[i];
      if (f.done) continue;
      live = true;
      if (f.op == NONE || f.op == SYNC) {
        ready = false;
        break;
      }
      if (op == NONE) op = f.op;
      else if (op != f.op) die("lanes of one wavefront reached different collectives");
    }
    if (!live || !ready) continue;
    uint64_t mask = 0;
    for (int i = w0; i < w1; ++i)
      if (!b.fibers[i].done && b.fibers[i].val) mask |= 1ull << (i - w0);
    for (int i = w0; i < w1; ++i) {
      Fiber& f = b.fibers[i];
      if (f.done) continue;
      const int lane = i - w0;
      int src = lane;
      switch (op) {
        case BALLOT: f.result = mask; break;
        case ANY: f.result = mask != 0; break;
        case SHFL: src = (lane / f.width) * f.width + (f.arg % f.width); break;
        case SHFL_UP: src = lane - f.arg >= (lane / f.width) * f.width ? lane - f.arg : lane; break;
        case SHFL_DOWN:
          src = lane + f.arg < (lane / f.width + 1) * f.width ? lane + f.arg : lane;
          break;
        default: break;
      }
      if (op == SHFL || op == SHFL_UP || op == SHFL_DOWN) {
        if (w0 + src >= w1 || b.fibers[w0 + src].done) die("shuffle from an inactive lane");
        f.result = b.fibers[w0 + src].val;
      }
    }
    for (int i = w0; i < w1; ++i) b.fibers[i].op = NONE;
    released = true;
  }
  return released;
}

inline size_t& stack_bytes() {
  static size_t s = 1 << 17;
  return s;
}

inline void run_block(Block& b) {
  const int n = (int)(b.bdim.x * b.bdim.y * b.bdim.z);
  b.fibers.clear();
  b.fibers.resize(n);
  for (int i = 0; i < n; ++i) {
    Fiber& f = b.fibers[i];
    f.tid = dim3(i % b.bdim.x, (i / b.bdim.x) % b.bdim.y, i / (b.bdim.x * b.bdim.y));
    if ((int)stacks().size() <= i) stacks().emplace_back(stack_bytes());
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = stacks()[i].data();
    f.ctx.uc_stack.ss_size = stacks()[i].size();
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, trampoline, 0);
  }
  blk() = &b;
  while (true) {
    bool ran = false, live = false;
    for (int i = 0; i < n; ++i) {
      Fiber& f = b.fibers[i];
      if (f.done) continue;
      live = true;
      if (f.op != NONE) continue;
      b.cur = i;
      swapcontext(&b.sched, &f.ctx);
      ran = true;
    }
    if (!live) break;
    if (!release(b) && !ran) {
      std::fprintf(stderr, "hip cpu emulation: workgroup deadlock in block %u\n", b.bid.x);
      std::abort();
    }
  }
  blk() = nullptr;
}

// ---- graph capture: launches are recorded instead of run
struct Graph {
  std::vector<std::function<void()>> steps;
};
inline Graph*& capturing() {
  static Graph* g = nullptr;
  return g;
}

template <class F, class... A>
void launch(dim3 grid, dim3 block, F kernel, A... args) {
  auto run = [=]() {
    for (uint32_t bz = 0; bz < grid.z; ++bz)
    for (uint32_t by = 0; by < grid.y; ++by)
      for (uint32_t bx = 0; bx < grid.x; ++bx) {
        Block b;
        b.bid = dim3(bx, by, bz);
        b.bdim = block;
        b.gdim = grid;
        b.entry = [=]() { kernel(args...); };
        run_block(b);
      }
  };
  if (capturing()) capturing()->steps.push_back(run);
  else run();
}

}  // namespace emu

#define threadIdx (emu::cur()->tid)
#define blockIdx (emu::blk()->bid)
#define blockDim (emu::blk()->bdim)
#define gridDim (emu::blk()->gdim)

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  emu::launch(dim3(grid), dim3(block), kernel, __VA_ARGS__)

inline void __syncthreads() { emu::wait(emu::SYNC, 0); }
inline int __lane_id() { return (int)(threadIdx.x % emu::WAVE); }
// callers pass wave-uniform values only
template <class T>
inline T __builtin_amdgcn_readfirstlane(T x) { return x; }
// fibers of a wave run one at a time: a fence orders nothing more, and the
// wave barrier is a collective like the others
inline void __builtin_amdgcn_fence(int, const char*) {}
// a streaming store is a store
template <class T>
inline void __builtin_nontemporal_store(T v, T* p) { *p = v; }
inline void __builtin_amdgcn_wave_barrier() { (void)emu::wait(emu::BALLOT, 1); }
inline uint64_t __ballot(int p) { return emu::wait(emu::BALLOT, p != 0); }
inline int __any(int p) { return (int)emu::wait(emu::ANY, p != 0); }

template <class T>
inline uint64_t emu_bits(T v) {
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T emu_from(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}
template <class T>
inline T __shfl(T v, int src, int width = emu::WAVE) {
  return emu_from<T>(emu::wait(emu::SHFL, emu_bits(v), src, width));
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = emu::WAVE) {
  return emu_from<T>(emu::wait(emu::SHFL_UP, emu_bits(v), (int)d, width));
}
template <class T>
inline T __shfl_down(T v, unsigned d, int width = emu::WAVE) {
  return emu_from<T>(emu::wait(emu::SHFL_DOWN, emu_bits(v), (int)d, width));
}

inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
template <class T>
inline T min(T a, T b) {
  return b < a ? b : a;
}
template <class T>
inline T max(T a, T b) {
  return a < b ? b : a;
}

template <class T>
inline T atomicAdd(T* p, T v) {
  T o = *p;
  *p = o + v;
  return o;
}
template <class T>
inline T atomicMin(T* p, T v) {
  T o = *p;
  if (v < o) *p = v;
  return o;
}
template <class T>
inline T atomicMax(T* p, T v) {
  T o = *p;
  if (o < v) *p = v;
  return o;
}
template <class T>
inline T atomicOr(T* p, T v) {
  T o = *p;
  *p = o | v;
  return o;
}
template <class T>
inline T atomicCAS(T* p, T cmp, T v) {
  T o = *p;
  if (o == cmp) *p = v;
  return o;
}
template <class T>
inline T __hip_atomic_load(const T* p, int, int) {
  return *p;
}
template <class T>
inline T __hip_atomic_fetch_min(T* p, T v, int, int) {
  T o = *p;
  if (v < o) *p = v;
  return o;
}

// ---- runtime API
inline const char* hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  p->multiProcessorCount = 2;
  return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t* fr, size_t* tot) {
  *fr = (size_t)4 << 30;
  *tot = (size_t)8 << 30;
  return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n);
  if (!*p) return hipErrorOutOfMemory;
  std::memset(*p, 0xA5, n);  // device memory starts undefined
  return hipSuccess;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemcpy2D(void* d, size_t dpitch, const void* s, size_t spitch, size_t width,
                              size_t height, hipMemcpyKind) {
  if (width > dpitch || width > spitch) return hipErrorInvalidValue;
  for (size_t r = 0; r < height; ++r)
    std::memcpy(static_cast<char*>(d) + r * dpitch, static_cast<const char*>(s) + r * spitch,
                width);
  return hipSuccess;
}
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
inline hipError_t hipMemset(void* d, int v, size_t n) {
  std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  if (emu::capturing()) {
    std::fprintf(stderr, "hip cpu emulation: memset inside a captured graph\n");
    std::abort();
  }
  return hipMemset(d, v, n);
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = (void*)0x1;
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = (void*)0x1;
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0;
  return hipSuccess;
}
typedef emu::Graph* hipGraph_t;
typedef emu::Graph* hipGraphExec_t;
typedef void* hipGraphNode_t;
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  emu::capturing() = new emu::Graph;
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = emu::capturing();
  emu::capturing() = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*,
                                      size_t) {
  *e = new emu::Graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& s : e->steps) s();
  return hipSuccess;
}
