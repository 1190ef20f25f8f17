// 2D Navier-Stokes vorticity stepper: four RK4 stages without viscosity and
// an end-of-step Crank-Nicolson factor.
//   stage 1,2 (Step1):  Zs = dt N(Z0);  Z0 = Z + Zs/2
//   stage 3   (Step2):  Z3 = dt N(Z0);  Z0 = Z + Z3
//   final     (Step):   Z = [(1 - dt nu k^2/2) Z + (Z1 + 2 Z2 + 2 Z3 + dt N(Z0))/6] / (1 + dt nu k^2/2)
// with N = -i(kx FFT(u w) + ky FFT(v w)) and forcing NZKX -= A at (0,+-KF);
// off-mask modes of Z are zeroed at every stage.
// (2D_NavierStokes/src_RK4/src_gmres/cuda.cu:44-76, kernels.h:247-435)
// Each stage is three kernels: y-inverse from the state (fill, mask,
// Biot-Savart), the shared x-pass, and a y-forward pass ending in the stage
// update.
#include "psg_launch.hpp"

namespace psg {

// 2D y-pass column tiles: 4 columns per block from N = 512 on (the launches
// have few blocks; half the LDS per block lets them spread over more CUs)
template <int N, typename P> struct Col2 {
  static constexpr int T = N >= 512 ? 4 : Col<N, P>::T;
  static constexpr int THREADS = T * (N / EPT);
};

template <int STAGE> struct Upd2D {
  cpx<double> *z, *z0, *zs;  // state, stage input, stage arrays Z1,Z2,Z3 ([m][3][nkt])
  long nkt;
  const int8_t *L;
  double dt, v2, famp, kf;
  const double *dtm = nullptr;   // per-member steps (batched Newton flows); dt 0 leaves a member unchanged
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky,
                                             cpx<double> (&N)[2]) const {
    cpx<double> *Z = z + (long)m * nkt + ik;
    const double dt = dtm ? dtm[m] : this->dt;
    if (L[ik] == 1) {
      // NZK[ik] = FFT(v w), NZK[nkt+ik] = FFT(u w)
      double NZKX = kx * N[1].y + ky * N[0].y;
      double NZKY = -kx * N[1].x - ky * N[0].x;
      if (kx == 0.0 && fabs(ky) == kf) NZKX += famp;
      const double ZX = dt * NZKX, ZY = dt * NZKY;
      cpx<double> *Z0 = z0 + (long)m * nkt + ik;
      cpx<double> *S = zs + (long)m * 3 * nkt + ik;
      const cpx<double> zn = *Z;
      if constexpr (STAGE == 0 || STAGE == 1) {
        S[STAGE * nkt] = cpx<double>{ZX, ZY};
        *Z0 = cpx<double>{zn.x + 0.5 * ZX, zn.y + 0.5 * ZY};
      } else if constexpr (STAGE == 2) {
        S[2 * nkt] = cpx<double>{ZX, ZY};
        *Z0 = cpx<double>{zn.x + ZX, zn.y + ZY};
      } else {
        const double wk = kx * kx + ky * ky;
        const double damp1 = 1 - 0.5 * dt * v2 * wk;
        const double damp2 = 1. / (1 + 0.5 * dt * v2 * wk);
        const cpx<double> z1 = S[0], z2 = S[nkt], z3 = S[2 * nkt];
        cpx<double> r;
        r.x = damp2 * (damp1 * zn.x + (z1.x + 2 * z2.x + 2 * z3.x + ZX) / 6.);
        r.y = damp2 * (damp1 * zn.y + (z1.y + 2 * z2.y + 2 * z3.y + ZY) / 6.);
        *Z = r;
        *Z0 = r;
      }
    } else {
      *Z = cpx<double>{0.0, 0.0};
    }
  }
};

struct Upd2DStoreNL {
  cpx<double> *out;
  long nkt;
  const int8_t *L;
  double famp, kf;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky,
                                             cpx<double> (&N)[2]) const {
    double NZKX = kx * N[1].y + ky * N[0].y;
    double NZKY = -kx * N[1].x - ky * N[0].x;
    if (kx == 0.0 && fabs(ky) == kf) NZKX += famp;
    const double ll = (double)L[ik];
    out[(long)m * nkt + ik] = cpx<double>{ll * NZKX, ll * NZKY};
  }
};

// masked copy Z0 = L Z (setVelocity, kernels.h:112-147)
__global__ void k2_setz0(const cpx<double> *__restrict__ z, cpx<double> *__restrict__ z0, const int8_t *L,
                         long nkt, long total) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long ik = i % nkt;
    z0[i] = L[ik] ? z[i] : cpx<double>{0.0, 0.0};
  }
}

template <typename P> static int launch_yinv2(Ctx *c, const cpx<double> *st, cpx<P> *out, int nm) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = Col2<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, 3, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k2_yinv<P, N, T, 1, true>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col2<N, P>::THREADS), lds, c->stream, st, g.nkt, out, 3 * g.ysz, g,
                       twp<P>(c, 1), c->mask, (const cpx<double> *)nullptr);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P, class UPD> static int launch_yfwd2(Ctx *c, const cpx<P> *in, UPD upd, int nm) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = Col2<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, 1, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k2_yfwd<P, N, T, UPD>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col2<N, P>::THREADS), lds, c->stream, in, 2 * g.ysz, g, twp<P>(c, 1), upd);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P, int FI, int FO, class PROD>
static int launch_xpass2(Ctx *c, const cpx<P> *in, cpx<P> *out, long nrow_m, int nm, PROD prod) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nx, {
    if (g.iktx != xband<N>()) {
      set_error("x-pass: IKTX %d is not NX/3+1", g.iktx);
      return PSG_EINVAL;
    }
    long nrows = nrow_m * nm;
    if (nrows >= (1L << 31)) {
      set_error("x-pass: %ld rows exceed the 32-bit row index", nrows);
      return PSG_EINVAL;
    }
    if (nrows < 2048) {
      // few rows: one row per block, its field pairs transformed side by side
      constexpr int NT = xsplit_teams<FI, FO>() * Row<N>::TPX;
      static_assert(NT <= 1024, "split-row x-pass block size");
      const size_t lds = xsplit_lds<P, N, FI, FO>();
      auto k = k_xpass_split<P, N, FI, FO, PROD>;
      PSG_HIP(lds_attr(k, lds));
      hipLaunchKernelGGL(k, dim3((unsigned)nrows), dim3(NT), lds, c->stream, in, out, nrow_m, nrows, g,
                         twp<P>(c, 0), prod);
    } else {
      constexpr int TE = Row<N>::TEAMS;
      dim3 grid((unsigned)((nrows + TE - 1) / TE));
      size_t lds = (size_t)(row_len<N>() * TE + N) * sizeof(cpx<P>);
      auto k = k_xpass<P, N, TE, FI, FO, PROD>;
      PSG_HIP(lds_attr(k, lds));
      hipLaunchKernelGGL(k, grid, dim3(Row<N>::THREADS), lds, c->stream, in, out, nrow_m, nrows, g,
                         twp<P>(c, 0), prod, 0);
    }
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

#define PSG_TRY2(x)     \
  do {                  \
    int r_ = (x);       \
    if (r_) return r_;  \
  } while (0)

template <typename P, class UPD> static int rhs2(Ctx *c, const cpx<double> *z0, UPD upd, int nm) {
  const Grid &g = c->g;
  cpx<P> *B = (cpx<P> *)c->bufB, *Cb = (cpx<P> *)c->bufC;
  const long Y = (long)g.iktx * g.ny;
  const size_t b = sizeof(cpx<P>);
  prof_begin(c, "yinv2", (long)(g.nkt * 16 + 3 * Y * b) * nm);
  PSG_TRY2(launch_yinv2<P>(c, z0, B, nm));
  prof_end(c);
  prof_begin(c, "xpass2", (long)(5 * Y * b) * nm);
  PSG_TRY2((launch_xpass2<P, 3, 2>(c, B, Cb, g.ny, nm, Prod2D<P>{})));
  prof_end(c);
  prof_begin(c, "yfwd2_update", (long)(2 * Y * b + 4 * g.nkt * 16) * nm);
  PSG_TRY2((launch_yfwd2<P>(c, Cb, upd, nm)));
  prof_end(c);
  return 0;
}

template <typename P> static int run2_t(Ctx *c, int nsteps, bool first) {
  const Grid &g = c->g;
  const long total = g.nkt * c->batch;
  if (first) {
    // stage input Z0 = masked state; later steps leave Z0 = state themselves
    hipLaunchKernelGGL(k2_setz0, dim3((unsigned)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096)),
                       dim3(256), 0, c->stream, c->state, c->stage, c->mask, g.nkt, total);
    PSG_HIP(hipGetLastError());
  }
  const double famp = c->cfg.force == PSG_FORCE_NONE ? 0.0 : c->cfg.ampfor;
  for (int nt = 0; nt < nsteps; nt++) {
    Upd2D<0> u0{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, famp, c->cfg.kf, c->dtm_on ? c->dtm : nullptr};
    PSG_TRY2(rhs2<P>(c, c->stage, u0, c->batch));
    Upd2D<1> u1{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, famp, c->cfg.kf, c->dtm_on ? c->dtm : nullptr};
    PSG_TRY2(rhs2<P>(c, c->stage, u1, c->batch));
    Upd2D<2> u2{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, famp, c->cfg.kf, c->dtm_on ? c->dtm : nullptr};
    PSG_TRY2(rhs2<P>(c, c->stage, u2, c->batch));
    Upd2D<3> u3{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, famp, c->cfg.kf, c->dtm_on ? c->dtm : nullptr};
    PSG_TRY2(rhs2<P>(c, c->stage, u3, c->batch));
  }
  return 0;
}

int ns2d_run(Ctx *c, int nsteps, bool first) {
  return c->cfg.precision == PSG_FP32BULK ? run2_t<float>(c, nsteps, first) : run2_t<double>(c, nsteps, first);
}

template <typename P> static int nonlin2_t(Ctx *c, const cpx<double> *in, cpx<double> *out) {
  const double famp = c->cfg.force == PSG_FORCE_NONE ? 0.0 : c->cfg.ampfor;
  Upd2DStoreNL u{out, c->g.nkt, c->mask, famp, c->cfg.kf};
  return rhs2<P>(c, in, u, 1);
}

int ns2d_nonlin(Ctx *c, int, const cpx<double> *in_dev, cpx<double> *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? nonlin2_t<float>(c, in_dev, out_dev)
                                          : nonlin2_t<double>(c, in_dev, out_dev);
}

template <typename P, int N, int TE>
__global__ void __launch_bounds__(TE *(N / EPT))
k2_xphys(const cpx<P> *__restrict__ in, double *__restrict__ out, long nrow, Grid g, const cpx<P> *__restrict__ tw) {
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

template <typename P> static int physical2_t(Ctx *c, int member, double *out) {
  const Grid &g = c->g;
  cpx<P> *B = (cpx<P> *)c->bufB;
  // physical fields of the masked state (w, u, v)
  const long total = g.nkt;
  hipLaunchKernelGGL(k2_setz0, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream,
                     c->state + (long)member * g.nkt, c->tmp, c->mask, g.nkt, total);
  PSG_TRY2(launch_yinv2<P>(c, c->tmp, B, 1));
  long nrow = g.ny;
  PSG_DISPATCH_N(g.nx, {
    constexpr int TE = Row<N>::TEAMS;
    dim3 grid((unsigned)((nrow + TE - 1) / TE), 3);
    size_t lds = (size_t)(N * TE + N) * sizeof(cpx<P>);
    auto k = k2_xphys<P, N, TE>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Row<N>::THREADS), lds, c->stream, B, out, nrow, g, twp<P>(c, 0));
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

int ns2d_physical(Ctx *c, int member, double *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? physical2_t<float>(c, member, out_dev)
                                          : physical2_t<double>(c, member, out_dev);
}

// ------------------------------------------------------------------ tangent linear
// Base and perturbation advance together through the same RK4/CN stages:
// the perturbation stage increments use the linearised nonlinear term
// dN = -i(kx FFT(du w + u dw) + ky FFT(dv w + v dw)) (no forcing) with the
// same damping factors and masks as the base update.
template <int STAGE> struct Upd2DTL {
  Upd2D<STAGE> base;
  cpx<double> *dz, *dz0, *dzs;
  __device__ __forceinline__ void operator()(int m, long ik, int i0, double kx, double ky,
                                             cpx<double> (&N)[4]) const {
    cpx<double> nb[2] = {N[0], N[1]}, np[2] = {N[2], N[3]};
    base(m, ik, i0, kx, ky, nb);
    pert(m, ik, kx, ky, np);
  }
  // group 0: the base update from N[0..1]; group 1: the perturbation's from N[2..3]
  __device__ __forceinline__ void part(int grp, int m, long ik, int i0, double kx, double ky,
                                       cpx<double> (&N)[2]) const {
    if (grp == 0) base(m, ik, i0, kx, ky, N);
    else pert(m, ik, kx, ky, N);
  }
  // N = {FFT(dv w + v dw), FFT(du w + u dw)} / NR
  __device__ __forceinline__ void pert(int m, long ik, double kx, double ky, cpx<double> (&N)[2]) const {
    const long nkt = base.nkt;
    cpx<double> *Z = dz + (long)m * nkt + ik;
    if (base.L[ik] == 1) {
      const double NX = kx * N[1].y + ky * N[0].y;
      const double NY = -kx * N[1].x - ky * N[0].x;
      const double dt = base.dt;
      const double ZX = dt * NX, ZY = dt * NY;
      cpx<double> *Z0 = dz0 + (long)m * nkt + ik;
      cpx<double> *S = dzs + (long)m * 3 * nkt + ik;
      const cpx<double> zn = *Z;
      if constexpr (STAGE == 0 || STAGE == 1) {
        S[STAGE * nkt] = cpx<double>{ZX, ZY};
        *Z0 = cpx<double>{zn.x + 0.5 * ZX, zn.y + 0.5 * ZY};
      } else if constexpr (STAGE == 2) {
        S[2 * nkt] = cpx<double>{ZX, ZY};
        *Z0 = cpx<double>{zn.x + ZX, zn.y + ZY};
      } else {
        const double wk = kx * kx + ky * ky;
        const double damp1 = 1 - 0.5 * dt * base.v2 * wk;
        const double damp2 = 1. / (1 + 0.5 * dt * base.v2 * wk);
        const cpx<double> z1 = S[0], z2 = S[nkt], z3 = S[2 * nkt];
        cpx<double> r;
        r.x = damp2 * (damp1 * zn.x + (z1.x + 2 * z2.x + 2 * z3.x + ZX) / 6.);
        r.y = damp2 * (damp1 * zn.y + (z1.y + 2 * z2.y + 2 * z3.y + ZY) / 6.);
        *Z = r;
        *Z0 = r;
      }
    } else {
      *Z = cpx<double>{0.0, 0.0};
    }
  }
};

template <typename P, class UPD> static int rhs2_tl(Ctx *c, UPD upd, int nm) {
  const Grid &g = c->g;
  cpx<P> *B = (cpx<P> *)c->tl_bufB, *Cb = (cpx<P> *)c->tl_bufC;
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = Col2<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, 6, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k2_yinv<P, N, T, 2, true>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col2<N, P>::THREADS), lds, c->stream, c->stage, g.nkt, B, 6 * g.ysz, g,
                       twp<P>(c, 1), c->mask, c->tl_stage);
  });
  PSG_HIP(hipGetLastError());
  PSG_TRY2((launch_xpass2<P, 6, 4>(c, B, Cb, g.ny, nm, Prod2DTL<P>{})));
  PSG_DISPATCH_N(g.ny, {
    constexpr int T = Col2<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, 2, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k2_yfwd<P, N, T, UPD, 4, 2>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col2<N, P>::THREADS), lds, c->stream, Cb, 4 * g.ysz, g, twp<P>(c, 1), upd);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P> static int run2tl_t(Ctx *c, int nsteps, bool first) {
  const Grid &g = c->g;
  const long total = g.nkt * c->batch;
  const unsigned gb = (unsigned)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
  if (first) {
    hipLaunchKernelGGL(k2_setz0, dim3(gb), dim3(256), 0, c->stream, c->state, c->stage, c->mask, g.nkt, total);
    hipLaunchKernelGGL(k2_setz0, dim3(gb), dim3(256), 0, c->stream, c->tl_state, c->tl_stage, c->mask, g.nkt,
                       total);
    PSG_HIP(hipGetLastError());
  }
  const double famp = c->cfg.force == PSG_FORCE_NONE ? 0.0 : c->cfg.ampfor;
  for (int nt = 0; nt < nsteps; nt++) {
#define PSG_TL_STAGE(S)                                                                                    \
    {                                                                                                      \
      Upd2DTL<S> u{{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, famp, c->cfg.kf},     \
                   c->tl_state, c->tl_stage, c->tl_rh};                                                    \
      PSG_TRY2(rhs2_tl<P>(c, u, c->batch));                                                                 \
    }
    PSG_TL_STAGE(0) PSG_TL_STAGE(1) PSG_TL_STAGE(2) PSG_TL_STAGE(3)
#undef PSG_TL_STAGE
  }
  return 0;
}

int ns2d_tl_run(Ctx *c, int nsteps, bool first) {
  return (c->cfg.precision == PSG_FP32BULK || c->tl_fp32) ? run2tl_t<float>(c, nsteps, first)
                                                          : run2tl_t<double>(c, nsteps, first);
}

// ------------------------------------------------------------------ recorded base
// A base flow whose x-pass uses Prod2DRec stores the physical (w, u, v) of
// every stage at slot 4*step + stage, step being the device counter that the
// last stage's y-forward advances (StepEnd).  The perturbation-only flow
// (Prod2DTLRec + Upd2DPert) reads them back instead of transforming the base
// again, so one tangent-linear step costs the transforms of a plain step.
// Past the record's last slot nothing is stored and zeros are read; the host
// sizes the record so that this never happens.
__device__ __forceinline__ long rec_offset(const int *cnt, int stage, long slots, long plane) {
  const long slot = 4L * (long)(*cnt) + stage;
  return slot < slots ? slot * 3 * plane : -1;
}

template <typename P, typename RP> struct Prod2DRec {
  static constexpr bool ROW = true;
  RP *rec;
  const int *cnt;
  int stage;
  long slots;
  int nx, ny;
  struct Bound {
    RP *p;
    long plane;
    bool on;
    // in: w, u, v; out: (v w, u w) as Prod2D
    __device__ __forceinline__ void operator()(int n, const P *in, P *out) const {
      if (on) {
        p[n] = (RP)in[0];
        p[plane + n] = (RP)in[1];
        p[2 * plane + n] = (RP)in[2];
      }
      out[0] = in[2] * in[0];
      out[1] = in[1] * in[0];
    }
  };
  // every team forms the products; team 0 alone stores the row
  __device__ __forceinline__ Bound bind(unsigned, unsigned row, int team) const {
    const long plane = (long)ny * nx;
    const long o = rec_offset(cnt, stage, slots, plane);
    return Bound{o >= 0 ? rec + o + (long)row * nx : rec, plane, o >= 0 && team == 0};
  }
};

template <typename P, typename RP> struct Prod2DTLRec {
  static constexpr bool ROW = true;
  const RP *rec;
  const int *cnt;
  int stage;
  long slots;
  int nx, ny;
  struct Bound {
    const RP *p;
    long plane;
    bool on;
    // in: dw, du, dv; out: (dv w + v dw, du w + u dw) about the recorded (w, u, v)
    __device__ __forceinline__ void operator()(int n, const P *in, P *out) const {
      P w = 0, u = 0, v = 0;
      if (on) {
        w = (P)p[n];
        u = (P)p[plane + n];
        v = (P)p[2 * plane + n];
      }
      out[0] = in[2] * w + v * in[0];
      out[1] = in[1] * w + u * in[0];
    }
  };
  // only team 0's output pair is stored: the other team reads nothing
  __device__ __forceinline__ Bound bind(unsigned, unsigned row, int team) const {
    const long plane = (long)ny * nx;
    const long o = rec_offset(cnt, stage, slots, plane);
    return Bound{o >= 0 ? rec + o + (long)row * nx : rec, plane, o >= 0 && team == 0};
  }
};

// the perturbation half of the tangent-linear update
template <int STAGE> struct Upd2DPert {
  Upd2DTL<STAGE> t;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky, cpx<double> (&N)[2]) const {
    t.pert(m, ik, kx, ky, N);
  }
};

// the last stage of a recording or replaying step advances the step counter
template <class U> struct StepEnd : U {
  static constexpr bool ONCE = true;
  int *cnt;
  __device__ __forceinline__ void once() const { *cnt = *cnt + 1; }
};

__global__ void k2_zero_int(int *p) { *p = 0; }

// the split-row x-pass alone (the record is indexed by row)
template <typename P, int FI, int FO, class PROD>
static int launch_xpass2_split(Ctx *c, const cpx<P> *in, cpx<P> *out, PROD prod) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nx, {
    if (g.iktx != xband<N>()) {
      set_error("x-pass: IKTX %d is not NX/3+1", g.iktx);
      return PSG_EINVAL;
    }
    if (g.ny >= 2048) {
      set_error("recorded base: %d rows exceed the split-row x-pass", g.ny);
      return PSG_EINVAL;
    }
    constexpr int NT = xsplit_teams<FI, FO>() * Row<N>::TPX;
    const size_t lds = xsplit_lds<P, N, FI, FO>();
    auto k = k_xpass_split<P, N, FI, FO, PROD>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, dim3((unsigned)g.ny), dim3(NT), lds, c->stream, in, out, (long)g.ny, (long)g.ny, g,
                       twp<P>(c, 0), prod);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P, typename RP> static int run2rec_t(Ctx *c, int nsteps, bool first) {
  const Grid &g = c->g;
  if (c->batch != 1 || !c->rec || !c->rec_cnt) {
    set_error("recorded base: one member and an allocated record only");
    return PSG_EINVAL;
  }
  if (first) {
    hipLaunchKernelGGL(k2_setz0, dim3((unsigned)((g.nkt + 255) / 256 < 4096 ? (g.nkt + 255) / 256 : 4096)),
                       dim3(256), 0, c->stream, c->state, c->stage, c->mask, g.nkt, g.nkt);
    PSG_HIP(hipGetLastError());
  }
  const double famp = c->cfg.force == PSG_FORCE_NONE ? 0.0 : c->cfg.ampfor;
  cpx<P> *B = (cpx<P> *)c->bufB, *Cb = (cpx<P> *)c->bufC;
  for (int nt = 0; nt < nsteps; nt++) {
#define PSG_REC_STAGE(S, UPD)                                                                                \
    {                                                                                                        \
      Upd2D<S> u{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, famp, c->cfg.kf, nullptr}; \
      Prod2DRec<P, RP> pr{(RP *)c->rec, c->rec_cnt, S, c->rec_slots, g.nx, g.ny};                            \
      PSG_TRY2(launch_yinv2<P>(c, c->stage, B, 1));                                                          \
      PSG_TRY2((launch_xpass2_split<P, 3, 2>(c, B, Cb, pr)));                                               \
      PSG_TRY2(launch_yfwd2<P>(c, Cb, UPD, 1));                                                              \
    }
    PSG_REC_STAGE(0, u) PSG_REC_STAGE(1, u) PSG_REC_STAGE(2, u)
    PSG_REC_STAGE(3, (StepEnd<Upd2D<3>>{u, c->rec_cnt}))
#undef PSG_REC_STAGE
  }
  return 0;
}

template <typename P> static int run2tlrec_t(Ctx *c, int nsteps, bool first) {
  const Grid &g = c->g;
  if (c->batch != 1 || !c->rec || !c->rec_cnt || c->rec_p != (int)sizeof(P)) {
    set_error("recorded tangent linear: one member and a record of the pass precision only");
    return PSG_EINVAL;
  }
  if (first) {
    hipLaunchKernelGGL(k2_setz0, dim3((unsigned)((g.nkt + 255) / 256 < 4096 ? (g.nkt + 255) / 256 : 4096)),
                       dim3(256), 0, c->stream, c->tl_state, c->tl_stage, c->mask, g.nkt, g.nkt);
    PSG_HIP(hipGetLastError());
  }
  cpx<P> *B = (cpx<P> *)c->tl_bufB, *Cb = (cpx<P> *)c->tl_bufC;
  for (int nt = 0; nt < nsteps; nt++) {
#define PSG_TLR_STAGE(S, UPD)                                                                                  \
    {                                                                                                          \
      Upd2DPert<S> u{{{c->state, c->stage, c->rh, g.nkt, c->mask, c->cfg.dt, c->cfg.v2, 0.0, c->cfg.kf, nullptr}, \
                      c->tl_state, c->tl_stage, c->tl_rh}};                                                    \
      Prod2DTLRec<P, P> pr{(const P *)c->rec, c->rec_cnt, S, c->rec_slots, g.nx, g.ny};                       \
      PSG_TRY2(launch_yinv2<P>(c, c->tl_stage, B, 1));                                                         \
      PSG_TRY2((launch_xpass2_split<P, 3, 2>(c, B, Cb, pr)));                                                 \
      PSG_TRY2(launch_yfwd2<P>(c, Cb, UPD, 1));                                                                \
    }
    PSG_TLR_STAGE(0, u) PSG_TLR_STAGE(1, u) PSG_TLR_STAGE(2, u)
    PSG_TLR_STAGE(3, (StepEnd<Upd2DPert<3>>{u, c->rec_cnt}))
#undef PSG_TLR_STAGE
  }
  return 0;
}

int ns2d_rec_run(Ctx *c, int nsteps, bool first) {
  if (c->cfg.precision == PSG_FP32BULK) {
    if (c->rec_p != 4) { set_error("recorded base: an FP32-bulk context records FP32"); return PSG_EINVAL; }
    return run2rec_t<float, float>(c, nsteps, first);
  }
  return c->rec_p == 4 ? run2rec_t<double, float>(c, nsteps, first) : run2rec_t<double, double>(c, nsteps, first);
}

int ns2d_tlrec_run(Ctx *c, int nsteps, bool first) {
  return c->rec_p == 4 ? run2tlrec_t<float>(c, nsteps, first) : run2tlrec_t<double>(c, nsteps, first);
}

int ns2d_rec_reset(Ctx *c) {
  hipLaunchKernelGGL(k2_zero_int, dim3(1), dim3(1), 0, c->stream, c->rec_cnt);
  PSG_HIP(hipGetLastError());
  return 0;
}

int ns2d_init(Ctx *c) {
  const Grid &g = c->g;
  size_t b = c->cfg.precision == PSG_FP32BULK ? sizeof(cpx<float>) : sizeof(cpx<double>);
  c->szB = (size_t)3 * g.ysz * c->batch * b;
  c->szC = (size_t)2 * g.ysz * c->batch * b;
  PSG_HIP(hipMalloc(&c->bufB, c->szB));
  PSG_HIP(hipMalloc(&c->bufC, c->szC));
  PSG_HIP(hipMalloc(&c->rh, sizeof(cpx<double>) * 3 * g.nkt * c->batch));
  PSG_HIP(hipMalloc(&c->stage, sizeof(cpx<double>) * g.nkt * c->batch));
  PSG_HIP(hipMemsetAsync(c->rh, 0, sizeof(cpx<double>) * 3 * g.nkt * c->batch, c->stream));
  PSG_HIP(hipMemsetAsync(c->stage, 0, sizeof(cpx<double>) * g.nkt * c->batch, c->stream));
  return 0;
}

}  // namespace psg
