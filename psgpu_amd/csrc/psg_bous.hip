// 3D Boussinesq stepper: RK4 with per-stage Crank-Nicolson factors, buoyancy
// with gravity tilted by theta (radians), Leray projection after each step.
//   stage 1,2 (Step1): X_s = dt C(Z0);  Z0 = d_{1/4} (X^n + dt/2 C)
//   stage 3   (Step2): X_3 = dt C(Z0);  Z0 = d_{1/2} (X^n + dt C)
//   final     (Step) : X^{n+1} = L [(1 - dt nu k^2/2) X^n + (X_1 + dt C)/6 + (X_2 + X_3)/3] / (1 + dt nu k^2/2)
//                      Z0 = X^{n+1} (before projection); w^{n+1} <- w - k (k.w)/|k|^2
// with d_q = (1 - q dt nu k^2)/(1 + q dt nu k^2) (nu/Sc for the buoyancy),
//   C_w   = L [curl(u x w) - Ri (buoyancy torque with sin/cos theta)] + F,
//   C_rho = -L [FFT(u.grad rho) - v cos theta - w sin theta].
// (3D_Boussinesq/src/cuda.cu:65-101, kernels.h:331-621; RHS for Newton kernels.h:624-692)
// One right-hand side: pass 1 from the stage state (vorticity, Biot-Savart
// velocity and i k rho), y-inverse of 9 fields, x-pass product (-(u x w),
// u.grad rho), y-forward of 4 fields, z-forward ending in the stage update.
#include "psg_ns3d.hpp"

namespace psg {

// pass 1: outputs w(3), u(3), i kx rho, i ky rho, i kz rho; with nout = 7 the
// seventh field is rho itself (physical-space dumps, 3D_Boussinesq/src/out_low.h:79-110)
template <typename P, int N, int T>
__global__ void __launch_bounds__(T *(N / EPT))
k_zinv_bous(const cpx<double> *__restrict__ st, long st_ms, long fs, cpx<P> *__restrict__ out, long out_ms, Grid g,
            const cpx<P> *__restrict__ tw, const int8_t *__restrict__ mask, int umask, int nout) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, iky = tid.y, m = tid.z;
  const bool valid = ikx < g.iktx;
  const int ky = iky - g.kty;
  const double kx = (double)ikx * g.alpha, kyd = (double)ky;
  const cpx<double> *s0 = st + (long)m * st_ms;
  cpx<double> w[3][EPT];
  double lm[EPT];
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) {
    const int kz = valid ? wrapk(j + TP * mm, g.nz, g.ktz) : PSG_NOK;
    const FillIdx q = fill_index(g, ikx, iky, ky, kz);
    w[0][mm] = ld_fill(s0, q);
    w[1][mm] = ld_fill(s0 + fs, q);
    w[2][mm] = ld_fill(s0 + 2 * fs, q);
    lm[mm] = (umask && q.idx >= 0) ? (double)mask[q.tidx] : 1.0;
  }
  // i kx rho and i ky rho are one transform of rho scaled by constants of the
  // pencil, so nine outputs take eight transforms
#pragma unroll 1
  for (int f = 0; f < nout; f++) {
    cpx<P> v[EPT];
    if (f == 7 && nout == 9) continue;   // written with field 6
    if (f == 6) {
      // vorticity no longer needed: reuse its registers for rho
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) {
        const int kz = valid ? wrapk(j + TP * mm, g.nz, g.ktz) : PSG_NOK;
        w[0][mm] = ld_fill(s0 + 3 * fs, fill_index(g, ikx, iky, ky, kz));
      }
    }
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) {
      const int kz = wrapk(j + TP * mm, g.nz, g.ktz);
      const double kzd = kz == PSG_NOK ? 0.0 : (double)kz;
      cpx<double> r;
      if (f < 3) {
        r = w[f][mm];
      } else if (f < 6) {
        const double wk = kx * kx + kyd * kyd + kzd * kzd;
        const double wki = lm[mm] / (wk > 0.001 ? wk : 0.001);
        cpx<double> u, vv, ww;
        biot_savart(kx, kyd, kzd, w[0][mm], w[1][mm], w[2][mm], wki, u, vv, ww);
        r = f == 3 ? u : (f == 4 ? vv : ww);
      } else if (nout == 7 || f == 6) {
        r = w[0][mm];   // rho (the gradient's x and y parts are scaled after the transform)
      } else {
        r = cpx<double>{-kzd * w[0][mm].y, kzd * w[0][mm].x};   // setFFK: i kz rho
      }
      v[mm] = to<P>(r);
    }
    fft_pencil<P, N, EPT, true>(v, lds, j, T, col, tw);
    if (valid) {
      cpx<P> *o = out + (long)m * out_ms + (long)f * g.zsz;
      if (f == 6 && nout == 9) {
        // setFFK: i kx rho and i ky rho
        const P px = (P)kx, py = (P)kyd;
        cpx<P> *o7 = o + g.zsz;
#pragma unroll
        for (int mm = 0; mm < EPT; mm++) {
          const long at = zoff<P>(g, ikx, iky, j + TP * mm);
          st_pass<NT<P, PASS_ZINV>::st>(&o[at], cpx<P>{-px * v[mm].y, px * v[mm].x});
          st_pass<NT<P, PASS_ZINV>::st>(&o7[at], cpx<P>{-py * v[mm].y, py * v[mm].x});
        }
      } else {
#pragma unroll
        for (int mm = 0; mm < EPT; mm++) st_pass<NT<P, PASS_ZINV>::st>(&o[zoff<P>(g, ikx, iky, j + TP * mm)], v[mm]);
      }
    }
  }
}

// x-pass product: (U,V,W) <- -(u x w), rx <- u . grad rho  (3D_Boussinesq/src/kernels.h:298-328)
template <typename P> struct ProdBous {
  __device__ __forceinline__ void operator()(const P *in, P *out) const {
    const P tx = in[0], te = in[1], tz = in[2], tu = in[3], tv = in[4], tw = in[5];
    out[0] = -(tv * tz - tw * te);
    out[1] = -(tw * tx - tu * tz);
    out[2] = -(tu * te - tv * tx);
    out[3] = tu * in[6] + tv * in[7] + tw * in[8];
  }
};

struct BousPar {
  double dt, v2, sc, ri, sinth, costh, kf, ampfor;
};

// C (8 reals: xi, eta, zeta, rho as re/im) from the transformed products and the stage state
__device__ __forceinline__ void bous_rhs(const BousPar &p, double kx, double ky, double kz, double LC,
                                         const cpx<double> (&N)[4], cpx<double> x0, cpx<double> e0,
                                         cpx<double> z0, cpx<double> r0, double *C) {
  const double UKX = N[0].x, UKY = N[0].y, VKX = N[1].x, VKY = N[1].y, WKX = N[2].x, WKY = N[2].y;
  const double wk = kx * kx + ky * ky + kz * kz;
  const double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
  const double rXR = -kx * r0.y, rXC = kx * r0.x, rYR = -ky * r0.y, rYC = ky * r0.x, rZR = -kz * r0.y,
               rZC = kz * r0.x;
  const double ri = p.ri, s = p.sinth, co = p.costh;
  C[0] = LC * (ky * WKY - kz * VKY - ri * (rYR * s - rZR * co));
  C[1] = LC * (kz * VKX - ky * WKX - ri * (rYC * s - rZC * co));
  C[2] = LC * (kz * UKY - kx * WKY + ri * rXR * s);
  C[3] = LC * (kx * WKX - kz * UKX + ri * rXC * s);
  C[4] = LC * (kx * VKY - ky * UKY - ri * rXR * co);
  C[5] = LC * (ky * UKX - kx * VKX - ri * rXC * co);
  C[6] = -LC * (N[3].x + (kz * x0.y - kx * z0.y) * wki * co + (kx * e0.y - ky * x0.y) * wki * s);
  C[7] = -LC * (N[3].y + (kx * z0.x - kz * x0.x) * wki * co + (ky * x0.x - kx * e0.x) * wki * s);
  if (kx == 0.0 && kz == 0.0 && ky == p.kf) C[4] += p.ampfor;
}

template <int STAGE> struct UpdBous {
  cpx<double> *st, *s0, *xs;  // state [m][4][nkt], stage input [m][4][nkt], increments [m][3][4][nkt]
  long nkt;
  const int8_t *L;
  BousPar p;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky, double kz,
                                             cpx<double> (&N)[4]) const {
    cpx<double> *S = st + (long)m * 4 * nkt + ik;
    cpx<double> *Z = s0 + (long)m * 4 * nkt + ik;
    cpx<double> *X = xs + (long)m * 12 * nkt + ik;
    const double wk = kx * kx + ky * ky + kz * kz;
    const double dt = p.dt;
    const double LC = (double)L[ik];
    double C[8];
    if constexpr (STAGE < 3) {
      bous_rhs(p, kx, ky, kz, LC, N, Z[0], Z[nkt], Z[2 * nkt], Z[3 * nkt], C);
      const double q = STAGE < 2 ? 0.25 : 0.5, h = STAGE < 2 ? 0.5 : 1.0;
      const double d1 = (1. - q * dt * p.v2 * wk) / (1. + q * dt * p.v2 * wk);
      const double d2 = (1. - q * dt * p.v2 * wk / p.sc) / (1. + q * dt * p.v2 * wk / p.sc);
#pragma unroll
      for (int f = 0; f < 4; f++) {
        X[(STAGE * 4 + f) * nkt] = cpx<double>{dt * C[2 * f], dt * C[2 * f + 1]};
        const double d = f < 3 ? d1 : d2;
        const cpx<double> sn = S[f * nkt];
        Z[f * nkt] = cpx<double>{d * (sn.x + h * dt * C[2 * f]), d * (sn.y + h * dt * C[2 * f + 1])};
      }
    } else {
      bous_rhs(p, kx, ky, kz, 1.0, N, Z[0], Z[nkt], Z[2 * nkt], Z[3 * nkt], C);
      const double d1 = 1. - 0.5 * dt * p.v2 * wk, d2 = 1. - 0.5 * dt * p.v2 * wk / p.sc;
      const double den1 = 1. + 0.5 * dt * p.v2 * wk, den2 = 1. + 0.5 * dt * p.v2 * wk / p.sc;
      const double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
      cpx<double> r[4];
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const double d = f < 3 ? d1 : d2, den = f < 3 ? den1 : den2;
        const cpx<double> sn = S[f * nkt], x1 = X[f * nkt], x2 = X[(4 + f) * nkt], x3 = X[(8 + f) * nkt];
        r[f].x = LC * (d * sn.x + sixth * (x1.x + dt * C[2 * f]) + third * (x2.x + x3.x)) / den;
        r[f].y = LC * (d * sn.y + sixth * (x1.y + dt * C[2 * f + 1]) + third * (x2.y + x3.y)) / den;
        Z[f * nkt] = r[f];
      }
      const double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
      const double phiR = (kx * r[0].y + ky * r[1].y + kz * r[2].y) * wki;
      const double phiC = -(kx * r[0].x + ky * r[1].x + kz * r[2].x) * wki;
      r[0].x -= -kx * phiC; r[0].y -= kx * phiR;
      r[1].x -= -ky * phiC; r[1].y -= ky * phiR;
      r[2].x -= -kz * phiC; r[2].y -= kz * phiR;
#pragma unroll
      for (int f = 0; f < 4; f++) S[f * nkt] = r[f];
    }
  }
};

// dz/dt for Newton (3D_Boussinesq/src/kernels.h:624-692): out = -nu k^2 z + C
struct UpdBousRHS {
  const cpx<double> *st, *s0;
  cpx<double> *out;
  long nkt;
  const int8_t *L;
  BousPar p;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky, double kz,
                                             cpx<double> (&N)[4]) const {
    const cpx<double> *S = st + (long)m * 4 * nkt + ik;
    const cpx<double> *Z = s0 + (long)m * 4 * nkt + ik;
    double C[8];
    bous_rhs(p, kx, ky, kz, (double)L[ik], N, Z[0], Z[nkt], Z[2 * nkt], Z[3 * nkt], C);
    const double wk = kx * kx + ky * ky + kz * kz;
    const double d1 = -p.v2 * wk, d2 = -p.v2 * wk / p.sc;
    cpx<double> *o = out + (long)m * 4 * nkt + ik;
#pragma unroll
    for (int f = 0; f < 4; f++) {
      const double d = f < 3 ? d1 : d2;
      o[f * nkt] = cpx<double>{d * S[f * nkt].x + C[2 * f], d * S[f * nkt].y + C[2 * f + 1]};
    }
  }
};

__global__ void k_copy(const cpx<double> *__restrict__ a, cpx<double> *__restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

// ------------------------------------------------------------------ launches
#define PSG_TRYB(x)     \
  do {                  \
    int r_ = (x);       \
    if (r_) return r_;  \
  } while (0)

template <typename P>
static int launch_zinv_b(Ctx *c, const cpx<double> *st, long st_ms, int umask, int nm, int nout = 9) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->bufA;
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, g.ikty, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k_zinv_bous<P, N, T>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N, P>::THREADS), lds, c->stream, st, st_ms, g.nkt, A, 9 * g.zsz, g,
                       twp<P>(c, 2), c->mask, umask, nout);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P> int launch_xphys(Ctx *c, const cpx<P> *in, double *out, int nf);

// x-pass shapes of the Boussinesq product (9 inputs, 4 outputs), chosen per
// context by psg_autotune: the variants of the Navier-Stokes x-pass (points per
// thread, input front-loading, per-pair product re-evaluation, rows per block,
// register caps)
template <typename P> static int xpass_bous(Ctx *c, const cpx<P> *B, cpx<P> *Cb, int nm) {
  const long nr = (long)c->g.ny * c->g.nz;
  const ProdBous<P> pb{};
  switch (c->xvar) {
    case 1: return launch_xpass<P, 9, 4, ProdBous<P>, 8, true, true>(c, B, Cb, nr, nm, pb);
    case 2: return launch_xpass<P, 9, 4, ProdBous<P>, 8, false, true>(c, B, Cb, nr, nm, pb);
    case 3: return launch_xpass<P, 9, 4, ProdBous<P>, 4, true, false>(c, B, Cb, nr, nm, pb);
    case 4: return launch_xpass<P, 9, 4, ProdBous<P>, 4, false, true>(c, B, Cb, nr, nm, pb);
    case 5: return launch_xpass<P, 9, 4, ProdBous<P>, 8, false, true, 2>(c, B, Cb, nr, nm, pb);
    case 6: return launch_xpass<P, 9, 4, ProdBous<P>, 8, false, true, 4>(c, B, Cb, nr, nm, pb);
    case 7: return launch_xpass<P, 9, 4, ProdBous<P>, 4, false, true, 1>(c, B, Cb, nr, nm, pb);
    case 8:
    case 9:
      if constexpr (std::is_same<P, float>::value) {
        if (c->xvar == 8) return launch_xpass<P, 9, 4, ProdBous<P>, 16, false, true>(c, B, Cb, nr, nm, pb);
        return launch_xpass<P, 9, 4, ProdBous<P>, 16, false, true, 4>(c, B, Cb, nr, nm, pb);
      }
      return launch_xpass<P, 9, 4, ProdBous<P>, 4, false, true, 2>(c, B, Cb, nr, nm, pb);
    case 10:
      return launch_xpass<P, 9, 4, ProdBous<P>, 8, false, true, 0, sizeof(P) == 8 ? 4 : 5>(c, B, Cb, nr, nm, pb);
    case 11:
      return launch_xpass<P, 9, 4, ProdBous<P>, 4, false, true, 0, sizeof(P) == 8 ? 5 : 6>(c, B, Cb, nr, nm, pb);
    default: return launch_xpass<P, 9, 4, ProdBous<P>>(c, B, Cb, nr, nm, pb);
  }
}

template <typename P, class UPD> static int launch_zfwd_b(Ctx *c, const cpx<P> *in, UPD upd, int nm) {
  const Grid &g = c->g;
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, g.ikty, nm);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k_zfwd<P, N, T, 4, UPD>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N, P>::THREADS), lds, c->stream, in, g, twp<P>(c, 2), upd);
  });
  PSG_HIP(hipGetLastError());
  return 0;
}

template <typename P, class UPD> static int rhs_b(Ctx *c, const cpx<double> *s0, int umask, UPD upd, int nm) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->bufA, *B = (cpx<P> *)c->bufB, *Cb = (cpx<P> *)c->bufC;
  const long Z = (long)g.iktx * g.ikty * g.nz, Y = (long)g.iktx * g.ny * g.nz;
  const long b = sizeof(cpx<P>);
  prof_begin(c, "zinv_bous", (4 * g.nkt * 16 + 9 * Z * b) * nm);
  PSG_TRYB(launch_zinv_b<P>(c, s0, 4 * g.nkt, umask, nm));
  prof_end(c);
  prof_begin(c, "yinv", 9 * (Z + Y) * b * nm);
  PSG_TRYB(launch_yinv_g<P>(c, A, B, 9 * nm));
  prof_end(c);
  prof_begin(c, "xpass_bous", 13 * Y * b * nm);
  PSG_TRYB(xpass_bous<P>(c, B, Cb, nm));
  prof_end(c);
  prof_begin(c, "yfwd", 4 * (Y + Z) * b * nm);
  PSG_TRYB(launch_yfwd_g<P>(c, Cb, A, 4 * nm));
  prof_end(c);
  prof_begin(c, "zfwd_update_bous", (4 * Z * b + 16 * g.nkt * 16) * nm);
  PSG_TRYB((launch_zfwd_b<P>(c, A, upd, nm)));
  prof_end(c);
  return 0;
}

static BousPar bpar(const Ctx *c) {
  return BousPar{c->cfg.dt, c->cfg.v2, c->cfg.sc, c->cfg.ri, sin(c->cfg.theta), cos(c->cfg.theta), c->cfg.kf,
                 c->cfg.ampfor};
}

template <typename P> static int runb_t(Ctx *c, int nsteps, bool first) {
  const Grid &g = c->g;
  const long tot = 4 * g.nkt * c->batch;
  if (first) {
    // stage input = state (setVelocity); later steps leave it equal to the state themselves
    hipLaunchKernelGGL(k_copy, dim3((unsigned)std::min<long>(4096, (tot + 255) / 256)), dim3(256), 0, c->stream,
                       c->state, c->stage, tot);
    PSG_HIP(hipGetLastError());
  }
  const BousPar p = bpar(c);
  for (int nt = 0; nt < nsteps; nt++) {
    UpdBous<0> u0{c->state, c->stage, c->rh, g.nkt, c->mask, p};
    PSG_TRYB(rhs_b<P>(c, c->stage, (first && nt == 0) ? 1 : 0, u0, c->batch));
    UpdBous<1> u1{c->state, c->stage, c->rh, g.nkt, c->mask, p};
    PSG_TRYB(rhs_b<P>(c, c->stage, 0, u1, c->batch));
    UpdBous<2> u2{c->state, c->stage, c->rh, g.nkt, c->mask, p};
    PSG_TRYB(rhs_b<P>(c, c->stage, 0, u2, c->batch));
    UpdBous<3> u3{c->state, c->stage, c->rh, g.nkt, c->mask, p};
    PSG_TRYB(rhs_b<P>(c, c->stage, 0, u3, c->batch));
  }
  return 0;
}

int bous_run(Ctx *c, int nsteps, bool first) {
  return c->cfg.precision == PSG_FP32BULK ? runb_t<float>(c, nsteps, first) : runb_t<double>(c, nsteps, first);
}

template <typename P> static int rhsb_t(Ctx *c, const cpx<double> *in, cpx<double> *out) {
  const Grid &g = c->g;
  // stage copy of the input (the reference's setVelocity), then one RHS
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((4 * g.nkt + 255) / 256)), dim3(256), 0, c->stream, in, c->stage,
                     4 * g.nkt);
  PSG_HIP(hipGetLastError());
  UpdBousRHS u{in, c->stage, out, g.nkt, c->mask, bpar(c)};
  return rhs_b<P>(c, c->stage, 1, u, 1);
}

int bous_timederiv(Ctx *c, const cpx<double> *in_dev, cpx<double> *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? rhsb_t<float>(c, in_dev, out_dev) : rhsb_t<double>(c, in_dev, out_dev);
}

// physical fields of one member: w(3), u(3), rho, NR doubles each, x fastest
template <typename P> static int physb_t(Ctx *c, const cpx<double> *st, double *out) {
  const Grid &g = c->g;
  PSG_TRYB(launch_zinv_b<P>(c, st, 4 * g.nkt, 1, 1, 7));
  PSG_TRYB(launch_yinv_g<P>(c, (const cpx<P> *)c->bufA, (cpx<P> *)c->bufB, 7));
  return launch_xphys<P>(c, (const cpx<P> *)c->bufB, out, 7);
}

int bous_physical_state(Ctx *c, const cpx<double> *st, double *out_dev) {
  return c->cfg.precision == PSG_FP32BULK ? physb_t<float>(c, st, out_dev) : physb_t<double>(c, st, out_dev);
}

int bous_physical(Ctx *c, int member, double *out_dev) {
  return bous_physical_state(c, c->state + (long)member * 4 * c->g.nkt, out_dev);
}

int bous_init(Ctx *c) {
  const Grid &g = c->g;
  size_t b = c->cfg.precision == PSG_FP32BULK ? sizeof(cpx<float>) : sizeof(cpx<double>);
  c->szA = (size_t)9 * g.zsz * c->batch * b;
  c->szB = (size_t)9 * g.ysz * c->batch * b;
  c->szC = (size_t)4 * g.ysz * c->batch * b;
  PSG_HIP(hipMalloc(&c->bufA, c->szA));
  PSG_HIP(hipMalloc(&c->bufB, c->szB));
  PSG_HIP(hipMalloc(&c->bufC, c->szC));
  PSG_HIP(hipMalloc(&c->rh, sizeof(cpx<double>) * 12 * g.nkt * c->batch));
  PSG_HIP(hipMalloc(&c->stage, sizeof(cpx<double>) * 4 * g.nkt * c->batch));
  PSG_HIP(hipMemsetAsync(c->rh, 0, sizeof(cpx<double>) * 12 * g.nkt * c->batch, c->stream));
  PSG_HIP(hipMemsetAsync(c->stage, 0, sizeof(cpx<double>) * 4 * g.nkt * c->batch, c->stream));
  return 0;
}

}  // namespace psg
