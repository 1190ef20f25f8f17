// 3D Navier-Stokes (Heun + Crank-Nicolson) and 3D Euler (Heun) steppers.
//
// Scheme (3D_NavierStokes/src/cuda.cu:60-112, kernels.h:309-464;
//         3D_Euler/src/cuda.cu:49-101, kernels.h:296-440):
//   predictor  RH = C(w^n)+F ;  w* = L NU1 (NUZN w^n + RH)          [Euler: RH = L C, w* = w^n + dt C]
//   corrector  w^{n+1} = L (w* + NU1/2 (C(w*)+F - RH))              [Euler: L (w* + dt/2 (C* - RH))]
// with C(w) = curl(u x w) assembled from the forward transform of -(u x w)
// and the host-built tables NUZN = (1-NU dt/2)/dt, NU1 = dt/(1+NU dt/2),
// NU = 10^(log10 nu + 2 log10|k|) (3D_NavierStokes/src/init.F90:194-198).
// The reference stores the velocity UK,VK,WK after every update kernel; here
// it is recomputed from the vorticity inside the first transform pass, which
// is the same Biot-Savart expression and saves 6 NKT complex of traffic.
#include <type_traits>
#include "psg_ns3d.hpp"

namespace psg {

// R: storage precision of the predictor's right-hand side RH, which the
// corrector subtracts (the pass precision: in FP32-bulk it only carries
// FP32-accurate products).  S: storage type of the state (double; float for
// the working state of an FP32-bulk run).  The update is evaluated in FP64.
// LAY: layout of st, rh and the tables (LayNKT: ik is the mode's index;
// LayW: the working copy, indexed from the mode's wavenumbers through g)
template <int PHASE, bool EULER, typename R = double, typename S = double, class LAY = LayNKT> struct UpdHeun {
  cpx<S> *st;
  cpx<R> *rh;
  long ms, fs;  // member stride, field stride (same for st and rh)
  const double *nuzn, *nu1;   // FP64 always: rounded tables would bias the damping every step
  const int8_t *L;
  Force F;
  double dt;
  Grid g;
  __device__ __forceinline__ void operator()(int m, long ik_nkt, int ikx, double kx, double ky, double kz,
                                             cpx<double> (&N)[3]) const {
    PSG_CURL
    long ik = ik_nkt, it = ik_nkt;
    if constexpr (std::is_same<LAY, LayW>::value) {
      const int iky = (int)ky + g.kty, ikz = (int)kz + g.ktz;
      ik = LayW::at(g, ikx, iky, ikz);
      it = LayW::tab(g, ikx, iky, ikz);
    }
    cpx<S> *s = st + (long)m * ms + ik;
    cpx<R> *r = rh + (long)m * ms + ik;
    const double ll = (double)L[it];
    if constexpr (PHASE == 0) {
      if constexpr (EULER) {
        r[0] = to<R>(cpx<double>{ll * xiX, ll * xiY});
        r[fs] = to<R>(cpx<double>{ll * etX, ll * etY});
        r[2 * fs] = to<R>(cpx<double>{ll * ztX, ll * ztY});
        const cpx<double> a = up(s[0]), b = up(s[fs]), c = up(s[2 * fs]);
        s[0] = to<S>(cpx<double>{a.x + dt * xiX, a.y + dt * xiY});
        s[fs] = to<S>(cpx<double>{b.x + dt * etX, b.y + dt * etY});
        s[2 * fs] = to<S>(cpx<double>{c.x + dt * ztX, c.y + dt * ztY});
      } else {
        const double NUZN = nuzn[it];
        const double NU1 = ll * nu1[it];
        add_forcing(F, kx, ky, kz, xiX, etY, ztX, ztY);
        r[0] = to<R>(cpx<double>{xiX, xiY});
        r[fs] = to<R>(cpx<double>{etX, etY});
        r[2 * fs] = to<R>(cpx<double>{ztX, ztY});
        const cpx<double> a = up(s[0]), b = up(s[fs]), c = up(s[2 * fs]);
        s[0] = to<S>(cpx<double>{NU1 * (NUZN * a.x + xiX), NU1 * (NUZN * a.y + xiY)});
        s[fs] = to<S>(cpx<double>{NU1 * (NUZN * b.x + etX), NU1 * (NUZN * b.y + etY)});
        s[2 * fs] = to<S>(cpx<double>{NU1 * (NUZN * c.x + ztX), NU1 * (NUZN * c.y + ztY)});
      }
    } else {
      double NU2;
      if constexpr (EULER) NU2 = 0.5 * dt;
      else {
        NU2 = 0.5 * nu1[it];
        add_forcing(F, kx, ky, kz, xiX, etY, ztX, ztY);
      }
      const cpx<double> a = up(s[0]), b = up(s[fs]), c = up(s[2 * fs]);
      const cpx<double> ra = up(r[0]), rb = up(r[fs]), rc = up(r[2 * fs]);
      s[0] = to<S>(cpx<double>{ll * (a.x + NU2 * (xiX - ra.x)), ll * (a.y + NU2 * (xiY - ra.y))});
      s[fs] = to<S>(cpx<double>{ll * (b.x + NU2 * (etX - rb.x)), ll * (b.y + NU2 * (etY - rb.y))});
      s[2 * fs] = to<S>(cpx<double>{ll * (c.x + NU2 * (ztX - rc.x)), ll * (c.y + NU2 * (ztY - rc.y))});
    }
  }
};

// masked C(w)+F into out (nonlinear right-hand side, no viscous term)
struct UpdStoreNL {
  cpx<double> *out;
  long ms, fs;
  const int8_t *L;
  Force F;
  int forced;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky, double kz,
                                             cpx<double> (&N)[3]) const {
    PSG_CURL
    if (forced) add_forcing(F, kx, ky, kz, xiX, etY, ztX, ztY);
    const double ll = (double)L[ik];
    cpx<double> *o = out + (long)m * ms + ik;
    o[0] = cpx<double>{ll * xiX, ll * xiY};
    o[fs] = cpx<double>{ll * etX, ll * etY};
    o[2 * fs] = cpx<double>{ll * ztX, ll * ztY};
  }
};

// generic column passes (also used by the Boussinesq and tangent-linear units)
// planes z0 .. z0+nzs-1 of every field (nzs = 0: all planes)
template <typename P> int launch_yinv_g(Ctx *c, const cpx<P> *in, cpx<P> *out, int nmf, int z0, int nzs) {
  const Grid &g = c->g;
  if (nzs <= 0) z0 = 0, nzs = g.nz;
  if (z0 < 0 || z0 + nzs > g.nz) { set_error("y-inverse: planes %d+%d outside %d", z0, nzs, g.nz); return PSG_EINVAL; }
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = ColY<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, nzs, nmf);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k_yinv<P, N, T>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(ColY<N, P>::THREADS), lds, c->stream, in, out, g, twp<P>(c, 1), z0);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P> int launch_yfwd_g(Ctx *c, const cpx<P> *in, cpx<P> *out, int nmf, int z0, int nzs) {
  const Grid &g = c->g;
  if (nzs <= 0) z0 = 0, nzs = g.nz;
  if (z0 < 0 || z0 + nzs > g.nz) { set_error("y-forward: planes %d+%d outside %d", z0, nzs, g.nz); return PSG_EINVAL; }
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = ColY<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, nzs, nmf);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k_yfwd<P, N, T>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(ColY<N, P>::THREADS), lds, c->stream, in, out, g, twp<P>(c, 1), z0);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template int launch_yinv_g<float>(Ctx *, const cpx<float> *, cpx<float> *, int, int, int);
template int launch_yinv_g<double>(Ctx *, const cpx<double> *, cpx<double> *, int, int, int);
template int launch_yfwd_g<float>(Ctx *, const cpx<float> *, cpx<float> *, int, int, int);
template int launch_yfwd_g<double>(Ctx *, const cpx<double> *, cpx<double> *, int, int, int);

// x-pass shapes of the Navier-Stokes product, chosen per context (psg_autotune):
// points per thread, input front-loading, per-pair product re-evaluation
template <typename P> static int xpass_ns(Ctx *c, const cpx<P> *B, cpx<P> *Cb, int nm, int z0, int nzs) {
  const long nr = (long)c->g.ny * nzs;
  switch (c->xvar) {
    case 1: return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, true>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 2: return launch_xpass<P, 6, 3, ProdNS<P>, 8, false, true>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 3: return launch_xpass<P, 6, 3, ProdNS<P>, 4, true, false>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 4: return launch_xpass<P, 6, 3, ProdNS<P>, 4, false, true>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 5: return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, true, 2>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 6: return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, true, 4>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 7: return launch_xpass<P, 6, 3, ProdNS<P>, 4, true, true, 1>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    case 8:
    case 9:
      // 16 points per thread (radix-16 stages, one LDS exchange per 256-point
      // transform): FP32 only, the FP64 registers would not fit
      if constexpr (std::is_same<P, float>::value) {
        if (c->xvar == 8) return launch_xpass<P, 6, 3, ProdNS<P>, 16, false, true>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
        return launch_xpass<P, 6, 3, ProdNS<P>, 16, false, true, 4>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
      }
      return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, true>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
    // register caps for occupancy: FP64 4 or 5 waves per SIMD, FP32 5 or 6
    case 10:
      return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, true, 0, sizeof(P) == 8 ? 4 : 5>(c, B, Cb, nr, nm,
                                                                                       ProdNS<P>{}, z0);
    case 11:
      return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, true, 0, sizeof(P) == 8 ? 5 : 6>(c, B, Cb, nr, nm,
                                                                                       ProdNS<P>{}, z0);
    default: return launch_xpass<P, 6, 3, ProdNS<P>, 8, true, false>(c, B, Cb, nr, nm, ProdNS<P>{}, z0);
  }
}

// One right-hand side evaluation for all members, ending in functor upd.
template <typename P, class UPD, typename S = double, class LAY = LayNKT>
static int rhs3(Ctx *c, const cpx<S> *st, long st_ms, int umask, UPD upd, long state_bytes, int nm) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->bufA, *B = (cpx<P> *)c->bufB, *Cb = (cpx<P> *)c->bufC;
  long pb[5];
  pass_bytes(c, 6, 3, sizeof(cpx<P>), pb, sizeof(cpx<S>));
  prof_begin(c, "zinv", pb[0] * nm);
  PSG_TRY((launch_zinv<P, S, LAY>(c, st, st_ms, A, 6 * g.zsz, true, umask, nm)));
  prof_end(c);
  // passes 2-4 run slab by slab of z planes when the context asks for it, so
  // that the slab's Y and C blocks are read back while they are still held
  // by the device's last-level cache; A' of a slab only overwrites A planes
  // of the same slab, which its y-inverse has already read
  const int SL = slab_planes(c);
  for (int z0 = 0; z0 < g.nz; z0 += SL) {
    const long fr = SL;   // bytes of a slab: the planes' share of the pass
    prof_begin(c, "yinv", pb[1] * nm / g.nz * fr);
    PSG_TRY(launch_yinv_g<P>(c, A, B, 6 * nm, z0, SL));
    prof_end(c);
    prof_begin(c, "xpass", pb[2] * nm / g.nz * fr);
    PSG_TRY(xpass_ns<P>(c, B, Cb, nm, z0, SL));
    prof_end(c);
    prof_begin(c, "yfwd", pb[3] * nm / g.nz * fr);
    PSG_TRY(launch_yfwd_g<P>(c, Cb, A, 3 * nm, z0, SL));
    prof_end(c);
  }
  prof_begin(c, "zfwd_update", (pb[4] + state_bytes) * nm);
  if (c->zvar == 1) PSG_TRY((launch_zfwd<P, 3, UPD, false>(c, A, upd, nm)));
  else PSG_TRY((launch_zfwd<P, 3, UPD, true>(c, A, upd, nm)));
  prof_end(c);
  return 0;
}

// Heun steps of the state st (storage type S) with the passes in precision P
// Heun steps of the working copy st (layout LayW, storage type S; RH in the
// same layout) with the passes in precision P
template <typename P, typename S> static int run_t(Ctx *c, cpx<S> *st, int nsteps, bool first) {
  const Grid &g = c->g;
  const bool euler = c->cfg.flavour == PSG_EULER3D;
  Force F{c->cfg.force, c->cfg.kf, c->cfg.ampfor};
  const long ms = 3 * c->nktw, fs = 8;
  // state + RH (pass precision) traffic per mode of each half-step
  const long sb = (6 * (long)sizeof(cpx<S>) + 3 * (long)sizeof(cpx<P>)) * g.nkt;
  cpx<P> *rh = (cpx<P> *)c->rh;
  const double *nuzn = c->wnuzn, *nu1 = c->wnu1;
  const int8_t *L = c->wmask;
  for (int nt = 0; nt < nsteps; nt++) {
    const int umask = (first && nt == 0) ? 1 : 0;
    if (euler) {
      UpdHeun<0, true, P, S, LayW> up{st, rh, ms, fs, nuzn, nu1, L, F, c->cfg.dt, g};
      PSG_TRY((rhs3<P, decltype(up), S, LayW>(c, st, ms, umask, up, sb, c->batch)));
      UpdHeun<1, true, P, S, LayW> uc{st, rh, ms, fs, nuzn, nu1, L, F, c->cfg.dt, g};
      PSG_TRY((rhs3<P, decltype(uc), S, LayW>(c, st, ms, 0, uc, sb, c->batch)));
    } else {
      UpdHeun<0, false, P, S, LayW> up{st, rh, ms, fs, nuzn, nu1, L, F, c->cfg.dt, g};
      PSG_TRY((rhs3<P, decltype(up), S, LayW>(c, st, ms, umask, up, sb, c->batch)));
      UpdHeun<1, false, P, S, LayW> uc{st, rh, ms, fs, nuzn, nu1, L, F, c->cfg.dt, g};
      PSG_TRY((rhs3<P, decltype(uc), S, LayW>(c, st, ms, 0, uc, sb, c->batch)));
    }
  }
  return 0;
}

// State <-> working copy: every entry of the LayW copy of `batch` members
// (padding columns past IKTX are zero), with the precision change of an
// FP32-bulk run
template <typename A, typename B, bool TO_W>
__global__ void k_work(const cpx<A> *__restrict__ a, cpx<B> *__restrict__ b, Grid g, long nktw, long total) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long m = i / (3 * nktw), w = i - m * 3 * nktw;
    const int c = (int)(w & 7), f = (int)((w >> 3) % 3);
    const long run = w / 24;
    const int ikz = (int)(run % g.iktz);
    const long t = run / g.iktz;
    const int iky = (int)(t % g.ikty), ikx = (int)(t / g.ikty) * 8 + c;
    const long n = m * 3 * g.nkt + f * g.nkt + LayNKT::at(g, ikx < g.iktx ? ikx : 0, iky, ikz);
    if constexpr (TO_W) {
      const cpx<A> v = ikx < g.iktx ? a[n] : cpx<A>{0, 0};
      b[i] = cpx<B>{(B)v.x, (B)v.y};
    } else if (ikx < g.iktx) {
      b[n] = cpx<B>{(B)a[i].x, (B)a[i].y};
    }
  }
}

template <typename A, typename B, bool TO_W> static int work_copy(Ctx *c, const cpx<A> *a, cpx<B> *b) {
  const long total = 3 * c->nktw * c->batch;
  hipLaunchKernelGGL((k_work<A, B, TO_W>), dim3((unsigned)std::min<long>(8192, (total + 255) / 256)), dim3(256), 0,
                     c->stream, a, b, c->g, c->nktw, total);
  PSG_HIP(hipGetLastError());
  return 0;
}

// CN tables and mask in the working layout (zero at padding columns)
__global__ void k_tables_w(const double *__restrict__ nuzn, const double *__restrict__ nu1,
                           const int8_t *__restrict__ L, double *__restrict__ wnuzn, double *__restrict__ wnu1,
                           int8_t *__restrict__ wL, Grid g, long nktw) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nktw; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i & 7);
    const long run = i >> 3;
    const int ikz = (int)(run % g.iktz);
    const long t = run / g.iktz;
    const int iky = (int)(t % g.ikty), ikx = (int)(t / g.ikty) * 8 + c;
    const bool in = ikx < g.iktx;
    const long n = LayNKT::at(g, in ? ikx : 0, iky, ikz);
    wnuzn[i] = in ? nuzn[n] : 0.0;
    wnu1[i] = in ? nu1[n] : 0.0;
    wL[i] = in ? L[n] : (int8_t)0;
  }
}

int ns3d_tables_w(Ctx *c) {
  if (!c->wnuzn) return 0;
  hipLaunchKernelGGL(k_tables_w, dim3((unsigned)std::min<long>(4096, (c->nktw + 255) / 256)), dim3(256), 0, c->stream,
                     c->nuzn, c->nu1, c->mask, c->wnuzn, c->wnu1, c->wmask, c->g, c->nktw);
  PSG_HIP(hipGetLastError());
  return 0;
}

// A run steps a working copy of the state in the LayW layout (FP32 in
// FP32-bulk, where a run starts from the state rounded to FP32): the FP64
// state in the reference order is kept between runs, where Newton's
// residual, the diagnostics and the API read it.  Between ns3d_work_begin
// and ns3d_work_end the copy is live and ns3d_run steps it without
// converting (run_steps brackets a call's eager steps and graph replays
// this way).
int ns3d_work_begin(Ctx *c) {
  if (c->work_live) return 0;
  if (c->st32) PSG_TRY((work_copy<double, float, true>(c, c->state, c->st32)));
  else PSG_TRY((work_copy<double, double, true>(c, c->state, c->stw)));
  c->work_live = true;
  return 0;
}

int ns3d_work_end(Ctx *c) {
  if (!c->work_live) return 0;
  c->work_live = false;
  if (c->st32) return work_copy<float, double, false>(c, c->st32, c->state);
  return work_copy<double, double, false>(c, c->stw, c->state);
}

int ns3d_run(Ctx *c, int nsteps, bool first) {
  const bool own = !c->work_live;
  if (own) PSG_TRY(ns3d_work_begin(c));
  const int r = c->st32 ? run_t<float, float>(c, c->st32, nsteps, first) : run_t<double, double>(c, c->stw, nsteps, first);
  if (!own) return r;
  const int r2 = ns3d_work_end(c);
  return r ? r : r2;
}

// One Navier-Stokes step of member 0 that records w^n (before the step) and
// the predictor state w* into wn / wstar (device, 3*nkt complex each, either
// may be null): the trajectory the adjoint sweep of a checkpoint segment needs.
template <typename P> static int step_rec_t(Ctx *c, int umask, cpx<double> *wn, cpx<double> *wstar) {
  const Grid &g = c->g;
  Force F{c->cfg.force, c->cfg.kf, c->cfg.ampfor};
  const long ms = 3 * g.nkt, fs = g.nkt;
  const size_t bytes = sizeof(cpx<double>) * 3 * g.nkt;
  if (wn) PSG_HIP(hipMemcpyAsync(wn, c->state, bytes, hipMemcpyDeviceToDevice, c->stream));
  const long sb = (6 * 16 + 3 * (long)sizeof(cpx<P>)) * g.nkt;
  UpdHeun<0, false, P> up{c->state, (cpx<P> *)c->rh, ms, fs, c->nuzn, c->nu1, c->mask, F, c->cfg.dt, c->g};
  PSG_TRY(rhs3<P>(c, c->state, ms, umask, up, sb, 1));
  if (wstar) PSG_HIP(hipMemcpyAsync(wstar, c->state, bytes, hipMemcpyDeviceToDevice, c->stream));
  UpdHeun<1, false, P> uc{c->state, (cpx<P> *)c->rh, ms, fs, c->nuzn, c->nu1, c->mask, F, c->cfg.dt, c->g};
  PSG_TRY(rhs3<P>(c, c->state, ms, 0, uc, sb, 1));
  return 0;
}

int ns3d_step_rec(Ctx *c, int umask, cpx<double> *wn, cpx<double> *wstar) {
  return c->cfg.precision == PSG_FP32BULK ? step_rec_t<float>(c, umask, wn, wstar)
                                          : step_rec_t<double>(c, umask, wn, wstar);
}

template <typename P> static int nonlin_t(Ctx *c, const cpx<double> *in, cpx<double> *out) {
  const Grid &g = c->g;
  Force F{c->cfg.force, c->cfg.kf, c->cfg.ampfor};
  UpdStoreNL u{out, 3 * g.nkt, g.nkt, c->mask, F, c->cfg.flavour == PSG_EULER3D ? 0 : 1};
  return rhs3<P>(c, in, 3 * g.nkt, 1, u, 6 * g.nkt * 16, 1);
}

int ns3d_nonlin(Ctx *c, int, const cpx<double> *in_dev, cpx<double> *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? nonlin_t<float>(c, in_dev, out_dev)
                                          : nonlin_t<double>(c, in_dev, out_dev);
}

// physical fields (w then u): passes 1,2 then a C2R-only row pass
template <typename P, int N, int TE>
__global__ void __launch_bounds__(TE *(N / EPT))
k_xphys(const cpx<P> *__restrict__ in, double *__restrict__ out, int nf, long nrow, Grid g,
        const cpx<P> *__restrict__ tw) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TPX = N / EPT;
  const int team = threadIdx.x / TPX, j = threadIdx.x % TPX;
  const long row = (long)blockIdx.x * TE + team;
  const bool valid = row < nrow;
  const int f = blockIdx.y;
  const cpx<P> *ra = in + (long)f * g.ysz + yrow<P>(g, valid ? row : 0);
  cpx<P> v[EPT];
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) {
    const int k = j + TPX * mm;
    cpx<P> a{0, 0};
    if (valid) {
      if (k < g.iktx) { a = ra[ycol<P>(g, k)]; if (k == 0) a.y = 0; }
      else if (k > N - g.iktx) { a = ra[ycol<P>(g, N - k)]; a.y = -a.y; }
    }
    v[mm] = a;
  }
  fft_pencil<P, N, EPT, true>(v, lds + team * N, j, 1, 0, tw);
  if (valid) {
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) out[((long)f * nrow + row) * N + j + TPX * mm] = (double)v[mm].x;
  }
}

// C2R row pass over nf fields already through the z and y inverse passes
template <typename P> int launch_xphys(Ctx *c, const cpx<P> *B, double *out, int nf) {
  const Grid &g = c->g;
  long nrow = (long)g.ny * g.nz;
  PSG_DISPATCH_N(g.nx, {
    constexpr int TE = Row<N>::TEAMS;
    dim3 grid((unsigned)((nrow + TE - 1) / TE), nf);
    size_t lds = (size_t)(N * TE + N) * sizeof(cpx<P>);
    auto k = k_xphys<P, N, TE>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Row<N>::THREADS), lds, c->stream, B, out, nf, nrow, g, twp<P>(c, 0));
  });
  PSG_HIP(hipGetLastError());
  return 0;
}
template int launch_xphys<float>(Ctx *, const cpx<float> *, double *, int);
template int launch_xphys<double>(Ctx *, const cpx<double> *, double *, int);

template <typename P> static int physical_t(Ctx *c, const cpx<double> *st, double *out) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->bufA, *B = (cpx<P> *)c->bufB;
  PSG_TRY(launch_zinv<P>(c, st, 3 * g.nkt, A, 6 * g.zsz, true, 1, 1));
  PSG_TRY(launch_yinv_g<P>(c, A, B, 6));
  return launch_xphys<P>(c, B, out, 6);
}

int ns3d_physical_state(Ctx *c, const cpx<double> *st, double *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? physical_t<float>(c, st, out_dev) : physical_t<double>(c, st, out_dev);
}

int ns3d_physical(Ctx *c, int member, double *out_dev) {
  return ns3d_physical_state(c, c->state + (long)member * 3 * c->g.nkt, out_dev);
}

int ns3d_init(Ctx *c) {
  const Grid &g = c->g;
  size_t b = c->cfg.precision == PSG_FP32BULK ? sizeof(cpx<float>) : sizeof(cpx<double>);
  c->szA = (size_t)6 * g.zsz * c->batch * b;
  c->szB = (size_t)6 * g.ysz * c->batch * b;
  c->szC = (size_t)3 * g.ysz * c->batch * b;
  PSG_HIP(hipMalloc(&c->bufA, c->szA));
  PSG_HIP(hipMalloc(&c->bufB, c->szB));
  PSG_HIP(hipMalloc(&c->bufC, c->szC));
  // RH serves the working-layout stepper and the reference-order steps of
  // the adjoint's recorded trajectory: the larger of the two sizes
  c->nktw = (long)((g.iktx + 7) / 8) * 8 * g.ikty * g.iktz;
  const long nrh = 3 * std::max<long>(c->nktw, g.nkt) * c->batch;
  PSG_HIP(hipMalloc(&c->rh, sizeof(cpx<double>) * nrh));
  PSG_HIP(hipMemsetAsync(c->rh, 0, sizeof(cpx<double>) * nrh, c->stream));
  if (c->cfg.precision == PSG_FP32BULK) PSG_HIP(hipMalloc(&c->st32, sizeof(cpx<float>) * 3 * c->nktw * c->batch));
  else PSG_HIP(hipMalloc(&c->stw, sizeof(cpx<double>) * 3 * c->nktw * c->batch));
  PSG_HIP(hipMalloc(&c->wnuzn, sizeof(double) * c->nktw));
  PSG_HIP(hipMalloc(&c->wnu1, sizeof(double) * c->nktw));
  PSG_HIP(hipMalloc(&c->wmask, c->nktw));
  return ns3d_tables_w(c);
}

}  // namespace psg

