// CPU build of the product's HIP sources, for bounds checking only.
//
// TEST INFRASTRUCTURE.  This header stands in for <hip/hip_runtime.h> when
// tests/emul/Makefile compiles psgpu_amd/csrc/*.hip with g++.  Every kernel
// launch runs on the host: workgroups are spread over OpenMP threads, and the
// threads of one workgroup are fibers (ucontext) on one host thread that
// switch at every barrier (__syncthreads, wavefront barrier), so the LDS
// exchange order of the device is kept.  The dynamic LDS of each workgroup
// and every device buffer are separate heap allocations, so AddressSanitizer
// reports any out-of-bounds access of a kernel or of the host code.  Streams
// are synchronous; a stream capture records launches and a graph launch
// replays them.  Nothing here is linked into the product library.
#pragma once
#include <omp.h>
#include <sanitizer/common_interface_defs.h>
#include <ucontext.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <cmath>
#include <functional>
#include <memory>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define PSG_WAVES_PER_EU(...)
#define __shared__ static thread_local
#define PSG_DYN_SMEM(name) char *name = ::emu::dyn_smem
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
inline thread_local char *dyn_smem = nullptr;
inline thread_local dim3 tidx, bidx;
inline dim3 gdim, bdim;
inline void tidx_set(int t) { tidx = dim3(t % bdim.x, (t / bdim.x) % bdim.y, t / (bdim.x * bdim.y)); }
}  // namespace emu
#define threadIdx (::emu::tidx)
#define blockIdx (::emu::bidx)
#define gridDim (::emu::gdim)
#define blockDim (::emu::bdim)

namespace emu {
// fiber scheduler of the workgroup running on this host thread
struct Fiber {
  ucontext_t ctx;
  char *stack = nullptr;
  bool done = false;
  void *fake = nullptr;
};
struct Group {
  ucontext_t sched;
  std::vector<Fiber> f;
  int cur = 0;
  const void *sbot = nullptr;   // scheduler stack, for the sanitizer's fiber switches
  size_t ssize = 0;
  int yielded = 0, finished = 0;
};
inline thread_local Group *grp = nullptr;
constexpr size_t kStack = 256 * 1024;
inline void yield_to_sched() {
  Group *g = grp;
  Fiber &me = g->f[g->cur];
  g->yielded++;
  __sanitizer_start_switch_fiber(&me.fake, g->sbot, g->ssize);
  swapcontext(&me.ctx, &g->sched);
  __sanitizer_finish_switch_fiber(me.fake, nullptr, nullptr);
}
}  // namespace emu

// every thread of the group reaches the barrier before any continues
inline void __syncthreads() { ::emu::yield_to_sched(); }
inline void __builtin_amdgcn_wave_barrier() { ::emu::yield_to_sched(); }
#define __builtin_amdgcn_fence(order, scope) ((void)0)

inline double atomicAdd(double *a, double v) {   // fibers of one group never run concurrently
  const double old = *a;
  *a += v;
  return old;
}

// ------------------------------------------------------------------ runtime
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorMemoryAllocation = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
enum { hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };

namespace emu {
struct Stream {
  bool capturing = false;
  std::vector<std::function<void()>> rec;
};
struct Graph {
  std::vector<std::function<void()>> ops;
};
}  // namespace emu
typedef emu::Stream *hipStream_t;
typedef emu::Graph *hipGraph_t;
typedef emu::Graph *hipGraphExec_t;
typedef double *hipEvent_t;

inline const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorMemoryAllocation;
}
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
// the emulated device reports 16 GiB, all of it free
inline hipError_t hipMemGetInfo(size_t *fr, size_t *tot) {
  *fr = *tot = (size_t)16 << 30;
  return hipSuccess;
}
// a file-scope device array is ordinary host memory here: its address is the symbol's
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st = nullptr) {
  if (st && st->capturing) {
    st->rec.push_back([=]() { memmove(d, s, n); });
    return hipSuccess;
  }
  memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr) {
  if (st && st->capturing) {   // memset nodes in captured graphs are not allowed in the product
    fprintf(stderr, "emulator: hipMemsetAsync recorded into a stream capture\n");
    abort();
  }
  memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = new emu::Stream(); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t s) { return (s && s->capturing) ? hipErrorInvalidValue : hipSuccess; }
inline double emu_now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new double(0.0); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) { *e = emu_now_ms(); return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(*b - *a); return hipSuccess; }
inline hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  s->capturing = true;
  s->rec.clear();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  s->capturing = false;
  *g = new emu::Graph{std::move(s->rec)};
  s->rec.clear();
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t) {
  *e = new emu::Graph{g->ops};
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &op : e->ops) op();
  return hipSuccess;
}

namespace emu {
template <typename F> struct Entry {
  static inline thread_local F *fn = nullptr;
  static void run(int idx) {
    Group *g = grp;
    __sanitizer_finish_switch_fiber(nullptr, &g->sbot, &g->ssize);
    (*fn)(idx);
    g->f[idx].done = true;
    g->finished++;
    __sanitizer_start_switch_fiber(nullptr, g->sbot, g->ssize);   // this fiber ends
    setcontext(&g->sched);
  }
};
// fiber stacks of this host thread, reused by its workgroups and freed when
// the thread ends (the OpenMP runtime may retire workers before exit)
struct StackPool {
  std::vector<char *> v;
  ~StackPool() {
    for (char *p : v) free(p);
  }
};
inline thread_local StackPool stack_pool;
// Run one workgroup: in each round every unfinished fiber runs to its next
// barrier or to its end.  A round in which some fibers stop at a barrier and
// others end is a barrier that not every thread reaches (a hang on the device):
// reported and fatal.
template <typename F> void run_group(int nt, F &&body) {
  Group g;
  g.f.resize(nt);
  Group *prev = grp;
  grp = &g;
  using FD = std::decay_t<F>;
  Entry<FD>::fn = &body;
  while ((int)stack_pool.v.size() < nt) stack_pool.v.push_back((char *)malloc(kStack));
  for (int i = 0; i < nt; i++) {
    Fiber &f = g.f[i];
    f.stack = stack_pool.v[i];
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, (void (*)())(void (*)(int))Entry<FD>::run, 1, i);
  }
  int alive = nt;
  while (alive > 0) {
    g.yielded = g.finished = 0;
    for (int i = 0; i < nt; i++) {
      if (g.f[i].done) continue;
      g.cur = i;
      tidx_set(i);
      void *fs = nullptr;
      __sanitizer_start_switch_fiber(&fs, g.f[i].stack, kStack);
      swapcontext(&g.sched, &g.f[i].ctx);
      __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
      if (g.f[i].done) alive--;
    }
    if (g.yielded && g.finished) {
      fprintf(stderr, "emulator: barrier not reached by every thread of a workgroup (%d at a barrier, %d ended)\n",
              g.yielded, g.finished);
      abort();
    }
  }
  grp = prev;
}
template <typename... KArgs, typename... Args>
void run_kernel(void (*k)(KArgs...), dim3 grid, dim3 block, size_t shmem, Args... args) {
  const int nt = (int)(block.x * block.y * block.z);
  const long nb = (long)grid.x * grid.y * grid.z;
  gdim = grid;
  bdim = block;
#pragma omp parallel for schedule(dynamic, 1)
  for (long b = 0; b < nb; b++) {
    char *smem = shmem ? (char *)malloc(shmem) : nullptr;
    if (smem) memset(smem, 0xA5, shmem);   // garbage-like LDS contents
    const dim3 bi((unsigned)(b % grid.x), (unsigned)((b / grid.x) % grid.y), (unsigned)(b / ((long)grid.x * grid.y)));
    auto body = [&](int t) {
      bidx = bi;
      dyn_smem = smem;
      k(args...);
    };
    run_group(nt, body);
    free(smem);
  }
}
template <typename... KArgs, typename... Args>
void launch(hipStream_t st, void (*k)(KArgs...), dim3 grid, dim3 block, size_t shmem, Args... args) {
  if (st && st->capturing) {
    st->rec.push_back([=]() { run_kernel(k, grid, block, shmem, args...); });
    return;
  }
  run_kernel(k, grid, block, shmem, args...);
}
}  // namespace emu
#define hipLaunchKernelGGL(k, grid, block, shmem, stream, ...) \
  ::emu::launch(stream, k, dim3(grid), dim3(block), (size_t)(shmem), __VA_ARGS__)
