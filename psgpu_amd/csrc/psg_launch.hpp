// Launch helpers: compile-time shapes per pencil length and runtime dispatch.
#pragma once
#include <mutex>
#include <unordered_map>
#include "psg_passes.hpp"

namespace psg {

// column tile: T kx columns x N/8 threads per column; at least 8 columns
// (128 B of FP64 per pencil element row), and in FP32 at least TMIN32
// columns, capped at 1024 threads
#ifndef PSG_COLT32
#define PSG_COLT32 8
#endif
template <int N, typename P = double> struct Col {
  static constexpr int TP = N / EPT;
  static constexpr int T0 = 256 / TP;
  static constexpr int TMIN = sizeof(P) == 4 ? PSG_COLT32 : 8;
  static constexpr int T1 = T0 < TMIN ? TMIN : (T0 > 32 ? 32 : T0);
  static constexpr int T = T1 * TP > 1024 ? 1024 / TP : T1;
  static constexpr int THREADS = T * TP;
};
// y passes: the same tile with their own FP32 minimum width
#ifndef PSG_COLT32Y
#define PSG_COLT32Y PSG_COLT32
#endif
template <int N, typename P = double> struct ColY {
  static constexpr int TP = N / EPT;
  static constexpr int T0 = 256 / TP;
  static constexpr int TMIN = sizeof(P) == 4 ? PSG_COLT32Y : 8;
  static constexpr int T1 = T0 < TMIN ? TMIN : (T0 > 32 ? 32 : T0);
  static constexpr int T = T1 * TP > 1024 ? 1024 / TP : T1;
  static constexpr int THREADS = T * TP;
};
// x-pass row teams
template <int N> struct Row {
  static constexpr int TPX = N / EPT;
  static constexpr int TEAMS = (256 / TPX) < 1 ? 1 : 256 / TPX;
  static constexpr int THREADS = TEAMS * TPX;
};

template <typename P> inline const cpx<P> *twp(const Ctx *c, int axis);
template <> inline const cpx<double> *twp<double>(const Ctx *c, int axis) { return c->tw64[axis]; }
template <> inline const cpx<float> *twp<float>(const Ctx *c, int axis) { return c->tw32[axis]; }

// raise the dynamic-LDS limit of a kernel when it needs more than 64 KiB; each
// kernel is raised once per process (the call is then skipped, so it never
// runs inside a stream capture after the kernel's first direct launch)
inline hipError_t lds_attr_raw(const void *kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  static std::mutex mu;
  static std::unordered_map<const void *, size_t> done;
  std::lock_guard<std::mutex> lk(mu);
  auto it = done.find(kernel);
  if (it != done.end() && it->second >= bytes) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) done[kernel] = bytes;
  return e;
}
template <typename F> inline hipError_t lds_attr(F *kernel, size_t bytes) {
  return lds_attr_raw(reinterpret_cast<const void *>(kernel), bytes);
}

#define PSG_CASE_N(NV, ...) \
  case NV: {                \
    constexpr int N = NV;   \
    __VA_ARGS__;            \
  } break;

#define PSG_DISPATCH_N(n, ...)                                   \
  switch (n) {                                                   \
    PSG_CASE_N(16, __VA_ARGS__)                                  \
    PSG_CASE_N(32, __VA_ARGS__)                                  \
    PSG_CASE_N(64, __VA_ARGS__)                                  \
    PSG_CASE_N(128, __VA_ARGS__)                                 \
    PSG_CASE_N(256, __VA_ARGS__)                                 \
    PSG_CASE_N(512, __VA_ARGS__)                                 \
    PSG_CASE_N(1024, __VA_ARGS__)                                \
    default:                                                     \
      ::psg::set_error("unsupported transform length %d", (int)(n)); \
      return PSG_EINVAL;                                         \
  }

inline bool supported_len(int n) { return n >= 16 && n <= 1024 && (n & (n - 1)) == 0; }

}  // namespace psg
