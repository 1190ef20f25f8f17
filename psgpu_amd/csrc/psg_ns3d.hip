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

template <int PHASE, bool EULER> struct UpdHeun {
  cpx<double> *st;
  cpx<double> *rh;
  long ms, fs;  // member stride, field stride (same for st and rh)
  const double *nuzn, *nu1;
  const int8_t *L;
  Force F;
  double dt;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky, double kz,
                                             cpx<double> (&N)[3]) const {
    PSG_CURL
    cpx<double> *s = st + (long)m * ms + ik;
    cpx<double> *r = rh + (long)m * ms + ik;
    const double ll = (double)L[ik];
    if constexpr (PHASE == 0) {
      if constexpr (EULER) {
        r[0] = cpx<double>{ll * xiX, ll * xiY};
        r[fs] = cpx<double>{ll * etX, ll * etY};
        r[2 * fs] = cpx<double>{ll * ztX, ll * ztY};
        cpx<double> a = s[0], b = s[fs], c = s[2 * fs];
        s[0] = cpx<double>{a.x + dt * xiX, a.y + dt * xiY};
        s[fs] = cpx<double>{b.x + dt * etX, b.y + dt * etY};
        s[2 * fs] = cpx<double>{c.x + dt * ztX, c.y + dt * ztY};
      } else {
        const double NUZN = nuzn[ik];
        const double NU1 = ll * nu1[ik];
        add_forcing(F, kx, ky, kz, xiX, etY, ztX, ztY);
        r[0] = cpx<double>{xiX, xiY};
        r[fs] = cpx<double>{etX, etY};
        r[2 * fs] = cpx<double>{ztX, ztY};
        cpx<double> a = s[0], b = s[fs], c = s[2 * fs];
        s[0] = cpx<double>{NU1 * (NUZN * a.x + xiX), NU1 * (NUZN * a.y + xiY)};
        s[fs] = cpx<double>{NU1 * (NUZN * b.x + etX), NU1 * (NUZN * b.y + etY)};
        s[2 * fs] = cpx<double>{NU1 * (NUZN * c.x + ztX), NU1 * (NUZN * c.y + ztY)};
      }
    } else {
      double NU2;
      if constexpr (EULER) NU2 = 0.5 * dt;
      else {
        NU2 = 0.5 * nu1[ik];
        add_forcing(F, kx, ky, kz, xiX, etY, ztX, ztY);
      }
      cpx<double> a = s[0], b = s[fs], c = s[2 * fs];
      cpx<double> ra = r[0], rb = r[fs], rc = r[2 * fs];
      s[0] = cpx<double>{ll * (a.x + NU2 * (xiX - ra.x)), ll * (a.y + NU2 * (xiY - ra.y))};
      s[fs] = cpx<double>{ll * (b.x + NU2 * (etX - rb.x)), ll * (b.y + NU2 * (etY - rb.y))};
      s[2 * fs] = cpx<double>{ll * (c.x + NU2 * (ztX - rc.x)), ll * (c.y + NU2 * (ztY - rc.y))};
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

// ------------------------------------------------------------------ launches
template <typename P>
static int launch_zinv(Ctx *c, const cpx<double> *st, long st_ms, cpx<P> *out, long out_ms, bool bs, int umask,
                       int nm) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N>::T;
    dim3 grid((g.iktx + T - 1) / T, g.ikty, nm);
    size_t lds = (size_t)N * T * sizeof(cpx<P>);
    if (bs) {
      auto k = k_zinv<P, N, T, true>;
      PSG_HIP(lds_attr(k, lds));
      hipLaunchKernelGGL(k, grid, dim3(Col<N>::THREADS), lds, c->stream, st, st_ms, g.nkt, out, out_ms, g,
                         twp<P>(c, 2), c->mask, umask);
    } else {
      auto k = k_zinv<P, N, T, false>;
      PSG_HIP(lds_attr(k, lds));
      hipLaunchKernelGGL(k, grid, dim3(Col<N>::THREADS), lds, c->stream, st, st_ms, g.nkt, out, out_ms, g,
                         twp<P>(c, 2), c->mask, umask);
    }
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P> int launch_yinv_g(Ctx *c, const cpx<P> *in, cpx<P> *out, int nmf) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = Col<N>::T;
    dim3 grid((g.iktx + T - 1) / T, g.nz, nmf);
    size_t lds = (size_t)N * T * sizeof(cpx<P>);
    auto k = k_yinv<P, N, T>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N>::THREADS), lds, c->stream, in, out, g, twp<P>(c, 1));
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P> int launch_yfwd_g(Ctx *c, const cpx<P> *in, cpx<P> *out, int nmf) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = Col<N>::T;
    dim3 grid((g.iktx + T - 1) / T, g.nz, nmf);
    size_t lds = (size_t)N * T * sizeof(cpx<P>);
    auto k = k_yfwd<P, N, T>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N>::THREADS), lds, c->stream, in, out, g, twp<P>(c, 1));
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template int launch_yinv_g<float>(Ctx *, const cpx<float> *, cpx<float> *, int);
template int launch_yinv_g<double>(Ctx *, const cpx<double> *, cpx<double> *, int);
template int launch_yfwd_g<float>(Ctx *, const cpx<float> *, cpx<float> *, int);
template int launch_yfwd_g<double>(Ctx *, const cpx<double> *, cpx<double> *, int);

template <typename P, int FI, int FO, class PROD>
static int launch_xpass(Ctx *c, const cpx<P> *in, cpx<P> *out, long nrow_m, int nm, PROD prod) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nx, {
    constexpr int TE = Row<N>::TEAMS;
    long nrows = nrow_m * nm;
    dim3 grid((unsigned)((nrows + TE - 1) / TE));
    size_t lds = (size_t)padded_len<N>() * TE * sizeof(cpx<P>);
    auto k = k_xpass<P, N, TE, FI, FO, PROD>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Row<N>::THREADS), lds, c->stream, in, out, nrow_m, nrows, g, twp<P>(c, 0),
                       prod);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P, int FO, class UPD> static int launch_zfwd(Ctx *c, const cpx<P> *in, UPD upd, int nm) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N>::T;
    dim3 grid((g.iktx + T - 1) / T, g.ikty, nm);
    size_t lds = (size_t)N * T * sizeof(cpx<P>);
    auto k = k_zfwd<P, N, T, FO, UPD>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N>::THREADS), lds, c->stream, in, g, twp<P>(c, 2), upd);
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
static void pass_bytes(const Ctx *c, int fi, int fo, size_t b, long out[5]) {
  const Grid &g = c->g;
  long Z = (long)g.iktx * g.ikty * g.nz, Y = (long)g.iktx * g.ny * g.nz;
  out[0] = 3 * g.nkt * 16 + fi * Z * (long)b;
  out[1] = fi * (Z + Y) * (long)b;
  out[2] = (fi + fo) * Y * (long)b;
  out[3] = fo * (Y + Z) * (long)b;
  out[4] = fo * Z * (long)b;
}

// One right-hand side evaluation for all members, ending in functor upd.
template <typename P, class UPD>
static int rhs3(Ctx *c, const cpx<double> *st, long st_ms, int umask, UPD upd, long state_bytes, int nm) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->bufA, *B = (cpx<P> *)c->bufB, *Cb = (cpx<P> *)c->bufC;
  long pb[5];
  pass_bytes(c, 6, 3, sizeof(cpx<P>), pb);
  prof_begin(c, "zinv", pb[0] * nm);
  PSG_TRY(launch_zinv<P>(c, st, st_ms, A, 6 * g.zsz, true, umask, nm));
  prof_end(c);
  prof_begin(c, "yinv", pb[1] * nm);
  PSG_TRY(launch_yinv_g<P>(c, A, B, 6 * nm));
  prof_end(c);
  prof_begin(c, "xpass", pb[2] * nm);
  PSG_TRY((launch_xpass<P, 6, 3>(c, B, Cb, (long)g.ny * g.nz, nm, ProdNS<P>{})));
  prof_end(c);
  prof_begin(c, "yfwd", pb[3] * nm);
  PSG_TRY(launch_yfwd_g<P>(c, Cb, A, 3 * nm));
  prof_end(c);
  prof_begin(c, "zfwd_update", (pb[4] + state_bytes) * nm);
  PSG_TRY((launch_zfwd<P, 3>(c, A, upd, nm)));
  prof_end(c);
  return 0;
}

template <typename P> static int run_t(Ctx *c, int nsteps) {
  const Grid &g = c->g;
  const bool euler = c->cfg.flavour == PSG_EULER3D;
  Force F{c->cfg.force, c->cfg.kf, c->cfg.ampfor};
  const long ms = 3 * g.nkt, fs = g.nkt;
  const long sb_pre = 9 * g.nkt * 16, sb_cor = 9 * g.nkt * 16;  // state + RH traffic
  for (int nt = 0; nt < nsteps; nt++) {
    const int umask = nt == 0 ? 1 : 0;
    if (euler) {
      UpdHeun<0, true> up{c->state, c->rh, ms, fs, c->nuzn, c->nu1, c->mask, F, c->cfg.dt};
      PSG_TRY(rhs3<P>(c, c->state, ms, umask, up, sb_pre, c->batch));
      UpdHeun<1, true> uc{c->state, c->rh, ms, fs, c->nuzn, c->nu1, c->mask, F, c->cfg.dt};
      PSG_TRY(rhs3<P>(c, c->state, ms, 0, uc, sb_cor, c->batch));
    } else {
      UpdHeun<0, false> up{c->state, c->rh, ms, fs, c->nuzn, c->nu1, c->mask, F, c->cfg.dt};
      PSG_TRY(rhs3<P>(c, c->state, ms, umask, up, sb_pre, c->batch));
      UpdHeun<1, false> uc{c->state, c->rh, ms, fs, c->nuzn, c->nu1, c->mask, F, c->cfg.dt};
      PSG_TRY(rhs3<P>(c, c->state, ms, 0, uc, sb_cor, c->batch));
    }
    c->time += c->cfg.dt;
  }
  return 0;
}

int ns3d_run(Ctx *c, int nsteps) {
  return c->cfg.precision == PSG_FP32BULK ? run_t<float>(c, nsteps) : run_t<double>(c, nsteps);
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
  extern __shared__ char smem_raw[];
  cpx<P> *lds = reinterpret_cast<cpx<P> *>(smem_raw);
  constexpr int TPX = N / EPT;
  const int team = threadIdx.x / TPX, j = threadIdx.x % TPX;
  const long row = (long)blockIdx.x * TE + team;
  const bool valid = row < nrow;
  const int f = blockIdx.y;
  const cpx<P> *ra = in + (long)f * g.ysz + (valid ? row : 0) * g.xp;
  cpx<P> v[EPT];
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) {
    const int k = j + TPX * mm;
    cpx<P> a{0, 0};
    if (valid) {
      if (k < g.iktx) { a = ra[k]; if (k == 0) a.y = 0; }
      else if (k > N - g.iktx) { a = ra[N - k]; a.y = -a.y; }
    }
    v[mm] = a;
  }
  fft_pencil<P, N, EPT, true>(v, lds + team * N, j, 1, 0, tw);
  if (valid) {
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) out[((long)f * nrow + row) * N + j + TPX * mm] = (double)v[mm].x;
  }
}

template <typename P> static int physical_t(Ctx *c, int member, double *out) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->bufA, *B = (cpx<P> *)c->bufB;
  PSG_TRY(launch_zinv<P>(c, c->state + (long)member * 3 * g.nkt, 3 * g.nkt, A, 6 * g.zsz, true, 1, 1));
  PSG_TRY(launch_yinv_g<P>(c, A, B, 6));
  long nrow = (long)g.ny * g.nz;
  PSG_DISPATCH_N(g.nx, {
    constexpr int TE = Row<N>::TEAMS;
    dim3 grid((unsigned)((nrow + TE - 1) / TE), 6);
    size_t lds = (size_t)N * TE * sizeof(cpx<P>);
    auto k = k_xphys<P, N, TE>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Row<N>::THREADS), lds, c->stream, B, out, 6, nrow, g, twp<P>(c, 0));
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

int ns3d_physical(Ctx *c, int member, double *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? physical_t<float>(c, member, out_dev)
                                          : physical_t<double>(c, member, out_dev);
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
  PSG_HIP(hipMalloc(&c->rh, sizeof(cpx<double>) * 3 * g.nkt * c->batch));
  PSG_HIP(hipMemsetAsync(c->rh, 0, sizeof(cpx<double>) * 3 * g.nkt * c->batch, c->stream));
  return 0;
}

}  // namespace psg

namespace psg {
// 3D tangent linear: provided by the tangent-linear/adjoint translation unit
// when built; until then the Navier-Stokes 3D Newton solver uses finite differences.
int ns3d_tl_run(Ctx *, int) {
  set_error("3D tangent-linear stepping is not built");
  return PSG_EINVAL;
}
}  // namespace psg
