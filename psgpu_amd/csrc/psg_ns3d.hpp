// Shared pieces of the 3D Navier-Stokes right-hand side: forcing, curl,
// and the launch helpers of the five pencil passes.  Included by the stepper
// and by the tangent-linear/adjoint translation unit.
#pragma once
#include <type_traits>
#include "psg_launch.hpp"

namespace psg {

struct Force {
  int mode;
  double kf, a;
};

// forcing on the curl result: 3D_NavierStokes/src/kernels.h:342-343 (DNS),
// src_gmres/kernels.h:234-238 (GMRES), Kolmogorov f=(sin kf y,0,0) as
// 3D_Boussinesq/src/kernels.h:382-384
__device__ __forceinline__ void add_forcing(const Force &F, double kx, double ky, double kz, double &xiX,
                                            double &etY, double &ztX, double &ztY) {
  if (F.mode == PSG_FORCE_DNS) {
    if (kx == 0.0 && kz == F.kf && ky == F.kf) xiX += F.a;
    if (kx == 0.0 && kz == -F.kf && ky == F.kf) xiX -= F.a;
  } else if (F.mode == PSG_FORCE_GMRES) {
    if (kx == 0.0 && kz == F.kf && ky == F.kf) { etY += F.a; ztY -= F.a; }
    if (kx == 0.0 && kz == -F.kf && ky == F.kf) { etY += F.a; ztY += F.a; }
  } else if (F.mode == PSG_FORCE_KOLMO) {
    if (kx == 0.0 && kz == 0.0 && ky == F.kf) ztX += F.a;
  }
}

// curl of the transformed product (the last part of convol, kernels.h:326-334)
#define PSG_CURL                                                      \
  const double UKX = N[0].x, UKY = N[0].y, VKX = N[1].x, VKY = N[1].y; \
  const double WKX = N[2].x, WKY = N[2].y;                            \
  double xiX = ky * WKY - kz * VKY, xiY = kz * VKX - ky * WKX;        \
  double etX = kz * UKY - kx * WKY, etY = kx * WKX - kz * UKX;        \
  double ztX = kx * VKY - ky * UKY, ztY = ky * UKX - kx * VKX;


// ------------------------------------------------------------------ launches
template <typename P> int launch_yinv_g(Ctx *c, const cpx<P> *in, cpx<P> *out, int nmf, int z0 = 0, int nzs = 0);
template <typename P> int launch_yfwd_g(Ctx *c, const cpx<P> *in, cpx<P> *out, int nmf, int z0 = 0, int nzs = 0);

// ky rows per block of the FP32 z-inverse and z-forward (the next row's loads
// overlap this row's transforms)
#ifndef PSG_ZPIPE32
#define PSG_ZPIPE32 1
#endif
#ifndef PSG_ZFPIPE32
#define PSG_ZFPIPE32 1
#endif
// LAY: layout of the state st (LayNKT: the context's state and caller
// buffers; LayW: the stepper's working copy, with the mask in LayW order)
template <typename P, typename S = double, class LAY = LayNKT>
static int launch_zinv(Ctx *c, const cpx<S> *st, long st_ms, cpx<P> *out, long out_ms, bool bs, int umask,
                       int nm) {
  const Grid &g = c->g;
  constexpr bool W = std::is_same<LAY, LayW>::value;
  const long st_fs = W ? 8 : g.nkt;
  const int8_t *mask = W ? c->wmask : c->mask;
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N, P>::T;
    constexpr int PIPE = sizeof(P) == 4 ? PSG_ZPIPE32 : 1;
    dim3 grid((g.iktx + T - 1) / T, (g.ikty + PIPE - 1) / PIPE, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    if (bs) {
      auto k = k_zinv<P, N, T, true, S, LAY, PIPE>;
      PSG_HIP(lds_attr(k, lds));
      hipLaunchKernelGGL(k, grid, dim3(Col<N, P>::THREADS), lds, c->stream, st, st_ms, st_fs, out, out_ms, g,
                         twp<P>(c, 2), mask, umask);
    } else {
      auto k = k_zinv<P, N, T, false, S, LAY, PIPE>;
      PSG_HIP(lds_attr(k, lds));
      hipLaunchKernelGGL(k, grid, dim3(Col<N, P>::THREADS), lds, c->stream, st, st_ms, st_fs, out, out_ms, g,
                         twp<P>(c, 2), mask, umask);
    }
  });
  PSG_HIP(hipGetLastError());
  return 0;
}


// TEAMS_MAX caps the rows per block (0: as many as fit in 256 threads)
template <typename P, int FI, int FO, class PROD, int E = EPT, bool FRONT = true, bool LAZY = false,
          int TEAMS_MAX = 0, int WPE = 0>
// nrow_m rows per member from plane zs0 on (nrow_m = NY x planes)
static int launch_xpass(Ctx *c, const cpx<P> *in, cpx<P> *out, long nrow_m, int nm, PROD prod, int zs0 = 0) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nx, {
    if (g.iktx != xband<N>()) {
      set_error("x-pass: IKTX %d is not NX/3+1", g.iktx);
      return PSG_EINVAL;
    }
    constexpr int TPX = N / E;
    constexpr int TE0 = (256 / TPX) < 1 ? 1 : 256 / TPX;
    constexpr int TE = (TEAMS_MAX > 0 && TEAMS_MAX < TE0) ? TEAMS_MAX : TE0;
    long nrows = nrow_m * nm;
    if (nrows >= (1L << 31)) {
      set_error("x-pass: %ld rows exceed the 32-bit row index", nrows);
      return PSG_EINVAL;
    }
    dim3 grid((unsigned)((nrows + TE - 1) / TE));
    size_t lds = (size_t)(row_len<N>() * TE + N) * sizeof(cpx<P>);
    auto k = k_xpass<P, N, TE, FI, FO, PROD, E, FRONT, LAZY, WPE>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(TE * TPX), lds, c->stream, in, out, nrow_m, nrows, g, twp<P>(c, 0), prod,
                       zs0);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P, int FO, class UPD, bool FRONT = (FO <= 4)>
static int launch_zfwd(Ctx *c, const cpx<P> *in, UPD upd, int nm) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N, P>::T;
    constexpr int PIPE = (sizeof(P) == 4 && FRONT) ? PSG_ZFPIPE32 : 1;
    dim3 grid((g.iktx + T - 1) / T, (g.ikty + PIPE - 1) / PIPE, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k_zfwd<P, N, T, FO, UPD, FRONT, PIPE>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N, P>::THREADS), lds, c->stream, in, g, twp<P>(c, 2), upd);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

#define PSG_TRY(x)          \
  do {                      \
    int r_ = (x);           \
    if (r_) return r_;      \
  } while (0)

// bytes of the five passes, per member, for the profiler (algorithmic:
// every buffer element read or written once)
inline void pass_bytes(const Ctx *c, int fi, int fo, size_t b, long out[5], size_t sb = 16) {
  const Grid &g = c->g;
  long Z = (long)g.iktx * g.ikty * g.nz, Y = (long)g.iktx * g.ny * g.nz;
  out[0] = 3 * g.nkt * (long)sb + fi * Z * (long)b;
  out[1] = fi * (Z + Y) * (long)b;
  out[2] = (fi + fo) * Y * (long)b;
  out[3] = fo * (Y + Z) * (long)b;
  out[4] = fo * Z * (long)b;
}


}  // namespace psg
