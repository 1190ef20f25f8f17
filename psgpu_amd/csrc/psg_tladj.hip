// Tangent linear and adjoint of the 3D Navier-Stokes Heun + Crank-Nicolson
// step (3D_NavierStokes/src/kernels.h:309-464; no counterpart in the
// reference, which differentiates the flow map only by finite differences in
// src_gmres/main.F90:1096-1107).
//
// With C(w) = K Fwd(P(w)): P(w) = -(u x w) in physical space (u = BS w the
// Biot-Savart velocity), Fwd = (1/NR) R2C truncated to the stored modes and
// K = -i k x (the curl as written in kernels.h:326-334), one step is
//   w*  = L NU1 (NUZN w + C(w) + F)
//   w'  = L (w* + NU1/2 (C(w*) - C(w)))
// Tangent linear (base and perturbation stepped together, 12 inverse and 6
// forward fields per right-hand side):
//   dC(w; d) = K Fwd(-(du x w + u x dw))
// Adjoint in the real inner product <a,b> = sum_L Re(conj(a) b) of the active
// modes.  Inv (spectral -> physical with the Hermitian fill) and Fwd satisfy
// Inv^T = 2 NR Fwd and Fwd^T = Inv / (2 NR) on the active modes; K and BS are
// Hermitian per mode.  Hence
//   C'(w)^T l = Fwd(u x G) + BS Fwd(G x w),   G = Inv(K l)
// (9 inverse fields w, u, G; 6 forward fields), and the step transposes to
//   l*  = L (l' + C'(w*)^T (NU1/2 l'))
//   l   = L (NU1 NUZN l* + C'(w)^T (NU1 (l* - l'/2)))
// The adjoint sweep over n steps keeps a checkpoint every `ckpt` steps and
// recomputes each segment forward, recording w^n and w*, before sweeping it
// backward.
#include "psg_ns3d.hpp"
#include "../../include/psg_adjoint.h"
#include <algorithm>
#include <cstring>

namespace psg {

// -------------------------------------------------------------- products
// inputs w(0..2) u(3..5) dw(6..8) du(9..11); outputs -(u x w), -(du x w + u x dw)
template <typename P> struct ProdNSTL {
  __device__ __forceinline__ void operator()(const P *in, P *out) const {
    const P wx = in[0], wy = in[1], wz = in[2], ux = in[3], uy = in[4], uz = in[5];
    const P dwx = in[6], dwy = in[7], dwz = in[8], dux = in[9], duy = in[10], duz = in[11];
    out[0] = -(uy * wz - uz * wy);
    out[1] = -(uz * wx - ux * wz);
    out[2] = -(ux * wy - uy * wx);
    out[3] = -((duy * wz - duz * wy) + (uy * dwz - uz * dwy));
    out[4] = -((duz * wx - dux * wz) + (uz * dwx - ux * dwz));
    out[5] = -((dux * wy - duy * wx) + (ux * dwy - uy * dwx));
  }
};

// inputs w(0..2) u(3..5) G(6..8); outputs u x G, G x w
template <typename P> struct ProdNSAdj {
  __device__ __forceinline__ void operator()(const P *in, P *out) const {
    const P wx = in[0], wy = in[1], wz = in[2], ux = in[3], uy = in[4], uz = in[5];
    const P gx = in[6], gy = in[7], gz = in[8];
    out[0] = uy * gz - uz * gy;
    out[1] = uz * gx - ux * gz;
    out[2] = ux * gy - uy * gx;
    out[3] = gy * wz - gz * wy;
    out[4] = gz * wx - gx * wz;
    out[5] = gx * wy - gy * wx;
  }
};

// K v = -i k x v for one mode
__device__ __forceinline__ void curl_k(double kx, double ky, double kz, cpx<double> vx, cpx<double> vy,
                                       cpx<double> vz, cpx<double> &ox, cpx<double> &oy, cpx<double> &oz) {
  ox = cpx<double>{ky * vz.y - kz * vy.y, kz * vy.x - ky * vz.x};
  oy = cpx<double>{kz * vx.y - kx * vz.y, kx * vz.x - kz * vx.x};
  oz = cpx<double>{kx * vy.y - ky * vx.y, ky * vx.x - kx * vy.x};
}

// -------------------------------------------------------------- pass 1 for G
// out fields 0..2 = z-inverse of K (s l), s = fac * tab[ik] (tab may be null),
// with the Hermitian fill of the kx = 0 plane; one member.
template <typename P, int N, int T>
__global__ void __launch_bounds__(T *(N / EPT))
k_zinv_curl(const cpx<double> *__restrict__ st, long fs, cpx<P> *__restrict__ out, Grid g,
            const cpx<P> *__restrict__ tw, const double *__restrict__ tab, double fac) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, iky = tid.y;
  const bool valid = ikx < g.iktx;
  const int ky = iky - g.kty;
  const double kx = (double)ikx * g.alpha, kyd = (double)ky;
  cpx<double> w[3][EPT];
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) {
    const int kz = valid ? wrapk(j + TP * mm, g.nz, g.ktz) : PSG_NOK;
    const FillIdx q = fill_index(g, ikx, iky, ky, kz);
    const double s = q.idx >= 0 ? (tab ? fac * tab[q.idx] : fac) : 0.0;
    cpx<double> a = ld_fill(st, q), b = ld_fill(st + fs, q), d = ld_fill(st + 2 * fs, q);
    const double kzd = kz == PSG_NOK ? 0.0 : (double)kz;
    curl_k(kx, kyd, kzd, cpx<double>{s * a.x, s * a.y}, cpx<double>{s * b.x, s * b.y},
           cpx<double>{s * d.x, s * d.y}, w[0][mm], w[1][mm], w[2][mm]);
  }
#pragma unroll
  for (int f = 0; f < 3; f++) {
    cpx<P> v[EPT];
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) v[mm] = to<P>(w[f][mm]);
    fft_pencil<P, N, EPT, true>(v, lds, j, T, col, tw);
    if (valid) {
      cpx<P> *o = out + (long)f * g.zsz;
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) st_pass<NT<P, PASS_ZINV>::st>(&o[zoff<P>(g, ikx, iky, j + TP * mm)], v[mm]);
    }
  }
}

// -------------------------------------------------------------- updates
// Heun + CN on the base (N[0..2], forced) and the perturbation (N[3..5])
template <int PHASE> struct UpdTL {
  cpx<double> *st, *rh, *dst, *drh;
  long ms, fs;
  const double *nuzn, *nu1;
  const int8_t *L;
  Force F;
  __device__ __forceinline__ void operator()(int m, long ik, int, double kx, double ky, double kz,
                                             cpx<double> (&N)[6]) const {
    const double ll = (double)L[ik];
    cpx<double> C[2][3];
    curl_k(kx, ky, kz, N[0], N[1], N[2], C[0][0], C[0][1], C[0][2]);
    curl_k(kx, ky, kz, N[3], N[4], N[5], C[1][0], C[1][1], C[1][2]);
    add_forcing(F, kx, ky, kz, C[0][0].x, C[0][1].y, C[0][2].x, C[0][2].y);
    cpx<double> *S[2] = {st + (long)m * ms + ik, dst + (long)m * ms + ik};
    cpx<double> *R[2] = {rh + (long)m * ms + ik, drh + (long)m * ms + ik};
#pragma unroll
    for (int v = 0; v < 2; v++) {
#pragma unroll
      for (int f = 0; f < 3; f++) {
        cpx<double> a = S[v][f * fs], c = C[v][f];
        if constexpr (PHASE == 0) {
          const double NUZN = nuzn[ik], NU1 = ll * nu1[ik];
          R[v][f * fs] = c;
          S[v][f * fs] = cpx<double>{NU1 * (NUZN * a.x + c.x), NU1 * (NUZN * a.y + c.y)};
        } else {
          const double NU2 = 0.5 * nu1[ik];
          const cpx<double> r = R[v][f * fs];
          S[v][f * fs] = cpx<double>{ll * (a.x + NU2 * (c.x - r.x)), ll * (a.y + NU2 * (c.y - r.y))};
        }
      }
    }
  }
};

// adjoint step halves; A = L (N[0..2] + BS N[3..5]) = C'(w)^T applied to the zinv input
template <int PHASE> struct UpdAdj {
  cpx<double> *lam, *lstar, *nuv;  // l (in/out), l*, the next zinv input
  long fs;
  const double *nuzn, *nu1;
  const int8_t *L;
  __device__ __forceinline__ void operator()(int, long ik, int, double kx, double ky, double kz,
                                             cpx<double> (&N)[6]) const {
    const double ll = (double)L[ik];
    const double wk = kx * kx + ky * ky + kz * kz;
    const double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
    cpx<double> b[3];
    biot_savart(kx, ky, kz, N[3], N[4], N[5], wki, b[0], b[1], b[2]);
    const double NU1 = ll * nu1[ik];
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const cpx<double> A{ll * (N[f].x + b[f].x), ll * (N[f].y + b[f].y)};
      if constexpr (PHASE == 0) {
        const cpx<double> lp = lam[ik + f * fs];
        const cpx<double> ls{ll * lp.x + A.x, ll * lp.y + A.y};
        lstar[ik + f * fs] = ls;
        nuv[ik + f * fs] = cpx<double>{NU1 * (ls.x - 0.5 * lp.x), NU1 * (ls.y - 0.5 * lp.y)};
      } else {
        const cpx<double> ls = lstar[ik + f * fs];
        const double c = NU1 * nuzn[ik];
        lam[ik + f * fs] = cpx<double>{c * ls.x + A.x, c * ls.y + A.y};
      }
    }
  }
};

// -------------------------------------------------------------- drivers
template <typename P, class UPD> static int rhs_tl(Ctx *c, int umask, UPD upd, int nm) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->tl_bufA, *B = (cpx<P> *)c->tl_bufB, *Cb = (cpx<P> *)c->tl_bufC;
  const long ms = 3 * g.nkt;
  long pb[5];
  pass_bytes(c, 12, 6, sizeof(cpx<P>), pb);
  prof_begin(c, "tl_zinv", (pb[0] + 3 * g.nkt * 16) * nm);
  PSG_TRY(launch_zinv<P>(c, c->state, ms, A, 12 * g.zsz, true, umask, nm));
  PSG_TRY(launch_zinv<P>(c, c->tl_state, ms, A + 6 * g.zsz, 12 * g.zsz, true, umask, nm));
  prof_end(c);
  prof_begin(c, "tl_yinv", pb[1] * nm);
  PSG_TRY(launch_yinv_g<P>(c, A, B, 12 * nm));
  prof_end(c);
  prof_begin(c, "tl_xpass", pb[2] * nm);
  PSG_TRY((launch_xpass<P, 12, 6, ProdNSTL<P>, 4, false>(c, B, Cb, (long)g.ny * g.nz, nm, ProdNSTL<P>{})));
  prof_end(c);
  prof_begin(c, "tl_yfwd", pb[3] * nm);
  PSG_TRY(launch_yfwd_g<P>(c, Cb, A, 6 * nm));
  prof_end(c);
  prof_begin(c, "tl_zfwd_update", (pb[4] + 18 * g.nkt * 16) * nm);
  PSG_TRY((launch_zfwd<P, 6>(c, A, upd, nm)));
  prof_end(c);
  return 0;
}

template <typename P> static int tl_run_t(Ctx *c, int nsteps, bool first) {
  const Grid &g = c->g;
  Force F{c->cfg.force, c->cfg.kf, c->cfg.ampfor};
  const long ms = 3 * g.nkt, fs = g.nkt;
  for (int nt = 0; nt < nsteps; nt++) {
    UpdTL<0> up{c->state, c->rh, c->tl_state, c->tl_rh, ms, fs, c->nuzn, c->nu1, c->mask, F};
    PSG_TRY(rhs_tl<P>(c, (first && nt == 0) ? 1 : 0, up, c->batch));
    UpdTL<1> uc{c->state, c->rh, c->tl_state, c->tl_rh, ms, fs, c->nuzn, c->nu1, c->mask, F};
    PSG_TRY(rhs_tl<P>(c, 0, uc, c->batch));
  }
  return 0;
}

int ns3d_tl_run(Ctx *c, int nsteps, bool first) {
  if (c->cfg.flavour != PSG_NS3D) {
    set_error("3D tangent linear: Navier-Stokes flavour only");
    return PSG_EINVAL;
  }
  if (int r = tl_alloc(c)) return r;
  return (c->cfg.precision == PSG_FP32BULK || c->tl_fp32) ? tl_run_t<float>(c, nsteps, first)
                                                          : tl_run_t<double>(c, nsteps, first);
}

// one application of C'(w)^T to the zinv input v (scaled by fac*tab), ending in upd
template <typename P, class UPD>
static int rhs_adj(Ctx *c, const cpx<double> *w, const cpx<double> *v, const double *tab, double fac, UPD upd) {
  const Grid &g = c->g;
  cpx<P> *A = (cpx<P> *)c->tl_bufA, *B = (cpx<P> *)c->tl_bufB, *Cb = (cpx<P> *)c->tl_bufC;
  long pb[5];
  pass_bytes(c, 9, 6, sizeof(cpx<P>), pb);
  prof_begin(c, "adj_zinv", pb[0] + 3 * g.nkt * 16);
  PSG_TRY(launch_zinv<P>(c, w, 3 * g.nkt, A, 9 * g.zsz, true, 0, 1));
  PSG_DISPATCH_N(g.nz, {
    constexpr int T = Col<N, P>::T;
    dim3 grid((g.iktx + T - 1) / T, g.ikty, 1);
    size_t lds = (size_t)(N * T + N) * sizeof(cpx<P>);
    auto k = k_zinv_curl<P, N, T>;
    PSG_HIP(lds_attr(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(Col<N, P>::THREADS), lds, c->stream, v, g.nkt, A + 6 * g.zsz, g, twp<P>(c, 2),
                       tab, fac);
  });
  PSG_HIP(hipGetLastError());
  prof_end(c);
  prof_begin(c, "adj_yinv", pb[1]);
  PSG_TRY(launch_yinv_g<P>(c, A, B, 9));
  prof_end(c);
  prof_begin(c, "adj_xpass", pb[2]);
  PSG_TRY((launch_xpass<P, 9, 6, ProdNSAdj<P>, 4, false>(c, B, Cb, (long)g.ny * g.nz, 1, ProdNSAdj<P>{})));
  prof_end(c);
  prof_begin(c, "adj_yfwd", pb[3]);
  PSG_TRY(launch_yfwd_g<P>(c, Cb, A, 6));
  prof_end(c);
  prof_begin(c, "adj_zfwd_update", pb[4] + 9 * g.nkt * 16);
  PSG_TRY((launch_zfwd<P, 6>(c, A, upd, 1)));
  prof_end(c);
  return 0;
}

__global__ void k_mask_copy(const cpx<double> *__restrict__ a, cpx<double> *__restrict__ b,
                            const int8_t *__restrict__ L, long nkt, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double l = (double)L[i % nkt];
    const cpx<double> v = a[i];
    b[i] = cpx<double>{l * v.x, l * v.y};
  }
}

// masked copy of z0 into member 0, then nsteps recorded steps, keeping a
// checkpoint every K steps in ck (when not null)
static int fwd_masked(Ctx *c, const cpx<double> *z0, int nsteps, cpx<double> *ck, int K) {
  const Grid &g = c->g;
  const long n3 = 3 * g.nkt;
  const size_t sb = sizeof(cpx<double>) * n3;
  const unsigned nb = (unsigned)std::min<long>(2048, (n3 + 255) / 256);
  hipLaunchKernelGGL(k_mask_copy, dim3(nb), dim3(256), 0, c->stream, z0, c->state, c->mask, g.nkt, n3);
  PSG_HIP(hipGetLastError());
  for (int n = 0; n < nsteps; n++) {
    if (ck && n % K == 0)
      PSG_HIP(hipMemcpyAsync(ck + (long)(n / K) * n3, c->state, sb, hipMemcpyDeviceToDevice, c->stream));
    PSG_TRY(ns3d_step_rec(c, n == 0 ? 1 : 0, nullptr, nullptr));
  }
  return 0;
}

template <typename P>
static int adjoint_t(Ctx *c, const cpx<double> *z0, const cpx<double> *lamT, int nsteps, int K, cpx<double> *zT,
                     cpx<double> *lam0) {
  const Grid &g = c->g;
  const long n3 = 3 * g.nkt;
  const size_t sb = sizeof(cpx<double>) * n3;
  const int nck = (nsteps + K - 1) / K;
  cpx<double> *ck = nullptr, *traj = nullptr, *lam = nullptr, *lstar = nullptr, *nuv = nullptr, *fin = nullptr;
  int r = 0;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    for (cpx<double> *p : {ck, traj, lam, lstar, nuv, fin})
      if (p) (void)hipFree(p);
  };
#define PSG_ADJ_TRY(x)             \
  do {                             \
    if ((r = (x))) { cleanup(); return r; } \
  } while (0)
#define PSG_ADJ_HIP(x)                                                             \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      set_error("%s failed: %s", #x, hipGetErrorString(e_));                       \
      cleanup();                                                                   \
      return PSG_EHIP;                                                             \
    }                                                                              \
  } while (0)
  PSG_ADJ_HIP(hipMalloc(&ck, sb * (size_t)nck));
  PSG_ADJ_HIP(hipMalloc(&traj, sb * 2 * (size_t)K));
  PSG_ADJ_HIP(hipMalloc(&lam, sb));
  PSG_ADJ_HIP(hipMalloc(&lstar, sb));
  PSG_ADJ_HIP(hipMalloc(&nuv, sb));
  PSG_ADJ_HIP(hipMalloc(&fin, sb));
  const unsigned nb = (unsigned)std::min<long>(2048, (n3 + 255) / 256);
  // forward sweep from the masked initial state, keeping checkpoints
  PSG_ADJ_TRY(fwd_masked(c, z0, nsteps, ck, K));
  PSG_ADJ_HIP(hipMemcpyAsync(fin, c->state, sb, hipMemcpyDeviceToDevice, c->stream));
  if (zT) PSG_ADJ_HIP(hipMemcpyAsync(zT, fin, sb, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_mask_copy, dim3(nb), dim3(256), 0, c->stream, lamT, lam, c->mask, g.nkt, n3);
  PSG_ADJ_HIP(hipGetLastError());
  // backward sweep, one checkpoint segment at a time
  for (int s = nck - 1; s >= 0; s--) {
    const int n0 = s * K, n1 = std::min(nsteps, n0 + K);
    PSG_ADJ_HIP(hipMemcpyAsync(c->state, ck + (long)s * n3, sb, hipMemcpyDeviceToDevice, c->stream));
    for (int n = n0; n < n1; n++)
      PSG_ADJ_TRY(ns3d_step_rec(c, n == 0 ? 1 : 0, traj + (long)(2 * (n - n0)) * n3,
                                traj + (long)(2 * (n - n0) + 1) * n3));
    for (int n = n1 - 1; n >= n0; n--) {
      const cpx<double> *wn = traj + (long)(2 * (n - n0)) * n3, *ws = traj + (long)(2 * (n - n0) + 1) * n3;
      UpdAdj<0> a0{lam, lstar, nuv, g.nkt, c->nuzn, c->nu1, c->mask};
      PSG_ADJ_TRY(rhs_adj<P>(c, ws, lam, c->nu1, 0.5, a0));
      UpdAdj<1> a1{lam, lstar, nuv, g.nkt, c->nuzn, c->nu1, c->mask};
      PSG_ADJ_TRY(rhs_adj<P>(c, wn, nuv, nullptr, 1.0, a1));
    }
  }
  PSG_ADJ_HIP(hipMemcpyAsync(lam0, lam, sb, hipMemcpyDeviceToDevice, c->stream));
  // leave the context holding Phi(z0), as after a forward run
  PSG_ADJ_HIP(hipMemcpyAsync(c->state, fin, sb, hipMemcpyDeviceToDevice, c->stream));
#undef PSG_ADJ_TRY
#undef PSG_ADJ_HIP
  cleanup();
  return 0;
}

int ns3d_adjoint(Ctx *c, const cpx<double> *z0, const cpx<double> *lamT, int nsteps, int ckpt, cpx<double> *zT,
                 cpx<double> *lam0) {
  if (c->cfg.flavour != PSG_NS3D) {
    set_error("3D adjoint: Navier-Stokes flavour only");
    return PSG_EINVAL;
  }
  if (nsteps < 1 || ckpt < 1) {
    set_error("adjoint: nsteps and ckpt must be >= 1");
    return PSG_EINVAL;
  }
  if (int r = tl_alloc(c)) return r;
  return c->cfg.precision == PSG_FP32BULK ? adjoint_t<float>(c, z0, lamT, nsteps, ckpt, zT, lam0)
                                          : adjoint_t<double>(c, z0, lamT, nsteps, ckpt, zT, lam0);
}

}  // namespace psg

// ------------------------------------------------------------------ C-ABI
using namespace psg;

namespace {
// host staging of nf*nkt complex arrays for one member
struct Dev {
  cpx<double> *p = nullptr;
  ~Dev() { if (p) (void)hipFree(p); }
};
int dev_alloc(Dev &d, size_t bytes) {
  if (hipMalloc(&d.p, bytes) != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed", bytes);
    return PSG_ENOMEM;
  }
  return 0;
}
}  // namespace

extern "C" {

int psg_tangent_steps(psg_ctx *c, const double *z0, const double *dz0, int nsteps, double *zT, double *dzT) {
  if (!c || !z0 || !dz0 || !dzT || nsteps < 0) { set_error("psg_tangent_steps: bad argument"); return PSG_EINVAL; }
  if (c->g.dim != 3) { set_error("psg_tangent_steps: 3D Navier-Stokes only"); return PSG_EINVAL; }
  if (int r = tl_alloc(c)) return r;
  const size_t sb = sizeof(cpx<double>) * 3 * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->state, z0, sb, hipMemcpyHostToDevice, c->stream));
  PSG_HIP(hipMemcpyAsync(c->tl_state, dz0, sb, hipMemcpyHostToDevice, c->stream));
  if (int r = run_steps(c, 1, nsteps)) return r;
  if (zT) PSG_HIP(hipMemcpyAsync(zT, c->state, sb, hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipMemcpyAsync(dzT, c->tl_state, sb, hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_adjoint(psg_ctx *c, const double *z0, const double *lamT, int nsteps, int ckpt, double *zT, double *lam0) {
  if (!c || !z0 || !lamT || !lam0) { set_error("psg_adjoint: bad argument"); return PSG_EINVAL; }
  if (c->g.dim != 3) { set_error("psg_adjoint: 3D Navier-Stokes only"); return PSG_EINVAL; }
  const size_t sb = sizeof(cpx<double>) * 3 * c->g.nkt;
  Dev dz, dl, dzT, dl0;
  if (int r = dev_alloc(dz, sb)) return r;
  if (int r = dev_alloc(dl, sb)) return r;
  if (int r = dev_alloc(dzT, sb)) return r;
  if (int r = dev_alloc(dl0, sb)) return r;
  PSG_HIP(hipMemcpyAsync(dz.p, z0, sb, hipMemcpyHostToDevice, c->stream));
  PSG_HIP(hipMemcpyAsync(dl.p, lamT, sb, hipMemcpyHostToDevice, c->stream));
  if (int r = ns3d_adjoint(c, dz.p, dl.p, nsteps, ckpt, dzT.p, dl0.p)) return r;
  if (zT) PSG_HIP(hipMemcpyAsync(zT, dzT.p, sb, hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipMemcpyAsync(lam0, dl0.p, sb, hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_residual_grad(psg_ctx *c, const double *z0, int nsteps, int ckpt, double sx, double sy, double sz,
                      double *J, double *grad, double *zT) {
  if (!c || !z0 || !J || !grad) { set_error("psg_residual_grad: bad argument"); return PSG_EINVAL; }
  if (c->g.dim != 3) { set_error("psg_residual_grad: 3D Navier-Stokes only"); return PSG_EINVAL; }
  const Grid &g = c->g;
  const long nkt = g.nkt;
  std::vector<int8_t> L(nkt);
  PSG_HIP(hipMemcpy(L.data(), c->mask, nkt, hipMemcpyDeviceToHost));
  std::vector<double> fin(6 * nkt), r(6 * nkt), lam0(6 * nkt);
  {
    const size_t sb = sizeof(cpx<double>) * 3 * nkt;
    Dev dz, dzT, dl, dl0;
    if (int e = dev_alloc(dz, sb)) return e;
    if (int e = dev_alloc(dzT, sb)) return e;
    if (int e = dev_alloc(dl, sb)) return e;
    if (int e = dev_alloc(dl0, sb)) return e;
    PSG_HIP(hipMemcpyAsync(dz.p, z0, sb, hipMemcpyHostToDevice, c->stream));
    // residual r = Phi(z0) - S z0 needs Phi(z0) first: the same masked forward sweep
    if (c->cfg.flavour != PSG_NS3D) { set_error("psg_residual_grad: Navier-Stokes flavour only"); return PSG_EINVAL; }
    if (int e = fwd_masked(c, dz.p, nsteps, nullptr, 1)) return e;
    PSG_HIP(hipMemcpyAsync(fin.data(), c->state, sb, hipMemcpyDeviceToHost, c->stream));
    PSG_HIP(hipStreamSynchronize(c->stream));
    double Jv = 0.0;
    for (long f = 0; f < 3; f++)
      for (long ik = 0; ik < nkt; ik++) {
        const long ikx = ik % g.iktx, iky = (ik / g.iktx) % g.ikty, ikz = ik / ((long)g.iktx * g.ikty);
        const double ph = -(ikx * g.alpha * sx + (iky - g.kty) * sy + (ikz - g.ktz) * sz);
        const double cr = cos(ph), ci = sin(ph);
        const long i = 2 * (f * nkt + ik);
        const double ar = z0[i], ai = z0[i + 1];
        const double l = L[ik] ? 1.0 : 0.0;
        const double rr = l * (fin[i] - (cr * ar - ci * ai)), ri = l * (fin[i + 1] - (cr * ai + ci * ar));
        r[i] = 2.0 * rr;
        r[i + 1] = 2.0 * ri;
        Jv += rr * rr + ri * ri;
      }
    *J = Jv;
    PSG_HIP(hipMemcpyAsync(dl.p, r.data(), sb, hipMemcpyHostToDevice, c->stream));
    if (int e = ns3d_adjoint(c, dz.p, dl.p, nsteps, ckpt, dzT.p, dl0.p)) return e;
    PSG_HIP(hipMemcpyAsync(lam0.data(), dl0.p, sb, hipMemcpyDeviceToHost, c->stream));
    PSG_HIP(hipStreamSynchronize(c->stream));
  }
  // grad = Phi'^T (2r) - S^{-1} (2r)
  for (long f = 0; f < 3; f++)
    for (long ik = 0; ik < nkt; ik++) {
      const long ikx = ik % g.iktx, iky = (ik / g.iktx) % g.ikty, ikz = ik / ((long)g.iktx * g.ikty);
      const double ph = ikx * g.alpha * sx + (iky - g.kty) * sy + (ikz - g.ktz) * sz;
      const double cr = cos(ph), ci = sin(ph);
      const long i = 2 * (f * nkt + ik);
      const double l = L[ik] ? 1.0 : 0.0;
      grad[i] = l * (lam0[i] - (cr * r[i] - ci * r[i + 1]));
      grad[i + 1] = l * (lam0[i + 1] - (cr * r[i + 1] + ci * r[i]));
    }
  if (zT) std::memcpy(zT, fin.data(), sizeof(double) * 6 * nkt);
  return 0;
}

}  // extern "C"
