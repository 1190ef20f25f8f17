// Run statistics on the GPU (include/psg_stats.h).
//
// Reference: 3D_NavierStokes/src/functions.h:128-209 and
// 3D_Boussinesq/src/functions.h:275-411 (computeStats), 3D_Boussinesq/src/
// kernels.h:96-135 (avgVelocity, subtractReal), out_low.h:5-76 (adaptive dt
// and RiG), finalise.h:17-190 (spectra and mean fields).
//
// Every reduction has a fixed order: per-thread compensated sums over a
// grid-stride range, a block tree in LDS, and one ordered pass over the block
// partials, so that repeated runs give identical sums.
#include <math.h>
#include <string.h>
#include <vector>
#include "psg_internal.hpp"
#include "../../include/psg_stats.h"

using namespace psg;

namespace {
constexpr int SB = 1024;   // blocks of the mode reductions
constexpr int PB = 16;     // blocks per plane of the RiG sums
constexpr int NVMAX = 16;

struct StPar {
  double ampfor, kf, ri, sinth, costh;
};

struct KSum {
  double s = 0.0, c = 0.0;
  __device__ void add(double x) {
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ double val() const { return s + c; }
};

// NV values per thread reduced over the block into part[blockIdx.x * NVMAX + v]
template <int NV> __device__ void block_reduce_store(KSum (&acc)[NV], double *part) {
  __shared__ double red[NV][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int v = 0; v < NV; v++) red[v][t] = acc[v].val();
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int v = 0; v < NV; v++) red[v][t] += red[v][t + w];
    }
    __syncthreads();
  }
  if (t < NV) part[(long)blockIdx.x * NVMAX + t] = red[t][0];
}

__device__ inline void kxyz(const Grid &g, long ik, double &kx, double &ky, double &kz) {
  const long ikx = ik % g.iktx, r = ik / g.iktx;
  const long iky = r % g.ikty, ikz = r / g.ikty;
  kx = (double)ikx * g.alpha;
  ky = (double)(iky - g.kty);
  kz = (double)(ikz - g.ktz);
}

__device__ inline cpx<double> avg_update(cpx<double> *t, cpx<double> z, double cnt) {
  cpx<double> a = *t;
  a.x = (a.x * cnt + z.x) / (cnt + 1.0);
  a.y = (a.y * cnt + z.y) / (cnt + 1.0);
  *t = a;
  return a;
}

// computeStats sums of one state (3D_NavierStokes/src/functions.h:141-180;
// 3D_Boussinesq/src/functions.h:290-360), running averages updated in place
template <bool BOUS>
__global__ void __launch_bounds__(256) k_stats(const cpx<double> *__restrict__ st, long nkt, Grid g,
                                               const int8_t *__restrict__ mask, cpx<double> *__restrict__ tavg,
                                               double cnt, StPar p, double *__restrict__ part) {
  constexpr int NV = BOUS ? (int)PSG_STB_N : (int)PSG_STNS_N;
  KSum acc[NV];
  for (long ik = (long)blockIdx.x * 256 + threadIdx.x; ik < nkt; ik += (long)gridDim.x * 256) {
    if (mask[ik] != 1) continue;
    double kx, ky, kz;
    kxyz(g, ik, kx, ky, kz);
    double wk = kx * kx + ky * ky + kz * kz;
    wk = wk > 0.001 ? wk : 0.001;
    const cpx<double> X = st[ik], E = st[nkt + ik], Z = st[2 * nkt + ik];
    const double vzx = X.x * X.x + X.y * X.y, vzy = E.x * E.x + E.y * E.y, vzz = Z.x * Z.x + Z.y * Z.y;
    const double vz = vzx + vzy + vzz;
    acc[0].add(vz);
    acc[1].add(vz / wk);
    const cpx<double> Xt = avg_update(tavg + ik, X, cnt), Et = avg_update(tavg + nkt + ik, E, cnt),
                      Zt = avg_update(tavg + 2 * nkt + ik, Z, cnt);
    const double vzt = (Xt.x * Xt.x + Xt.y * Xt.y) + (Et.x * Et.x + Et.y * Et.y) + (Zt.x * Zt.x + Zt.y * Zt.y);
    acc[2].add(vzt / wk);
    acc[3].add(vzt);
    if constexpr (!BOUS) {
      if (kx == 0.0 && ky == p.kf && kz == p.kf) acc[4].add(p.ampfor * X.x);
      if (kx == 0.0 && ky == p.kf && kz == -p.kf) acc[4].add(-p.ampfor * X.x);
      acc[5].add(vzx);
      acc[6].add(vzy);
      acc[7].add(vzz);
      const double u1 = (ky * Z.y - kz * E.y) / wk, u2 = (kz * E.x - ky * Z.x) / wk;
      const double v1 = (kz * X.y - kx * Z.y) / wk, v2 = (kx * Z.x - kz * X.x) / wk;
      const double w1 = (kx * E.y - ky * X.y) / wk, w2 = (ky * X.x - kx * E.x) / wk;
      acc[8].add(u1 * u1 + u2 * u2);
      acc[9].add(v1 * v1 + v2 * v2);
      acc[10].add(w1 * w1 + w2 * w2);
    } else {
      if (kx == 0.0 && ky == p.kf && kz == 0.0) acc[4].add(-p.ampfor * Z.x / p.kf);
      const cpx<double> R = st[3 * nkt + ik];
      const double vr = R.x * R.x + R.y * R.y;
      acc[5].add(vr);
      acc[6].add(vr * wk);
      acc[7].add(kx * kx * vr);
      acc[8].add(ky * ky * vr);
      const cpx<double> Rt = avg_update(tavg + 3 * nkt + ik, R, cnt);
      const double vrt = Rt.x * Rt.x + Rt.y * Rt.y;
      acc[9].add(vrt);
      acc[10].add(vrt * wk);
      const double vpR = (kz * X.y - kx * Z.y) / wk, vpC = (kx * Z.x - kz * X.x) / wk;
      const double wpR = (kx * E.y - ky * X.y) / wk, wpC = (ky * X.x - kx * E.x) / wk;
      const double ubR = vpR * p.sinth + wpR * p.costh, ubC = vpC * p.sinth + wpC * p.costh;
      const double phR = atan2(R.y, R.x), phW = atan2(ubC, ubR);
      acc[11].add(p.ri * sqrt(ubR * ubR + ubC * ubC) * sqrt(R.x * R.x + R.y * R.y) * cos(phR - phW));
    }
  }
  block_reduce_store<NV>(acc, part);
}

__global__ void k_final(const double *__restrict__ part, int nblk, int nv, double *__restrict__ vals) {
  const int v = threadIdx.x;
  if (v >= nv) return;
  KSum s;
  for (int b = 0; b < nblk; b++) s.add(part[(long)b * NVMAX + v]);
  vals[v] = s.val();
}

// avgVelocity (kernels.h:96-107) on the physical u, v, w of the current state
__global__ void k_rms(const double *__restrict__ phys, long nr, double *__restrict__ rms, double cnt) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nr; i += (long)gridDim.x * 256) {
    const double u = phys[3 * nr + i], v = phys[4 * nr + i], w = phys[5 * nr + i];
    rms[i] = (rms[i] * cnt + u * u) / (cnt + 1.0);
    rms[nr + i] = (rms[nr + i] * cnt + v * v) / (cnt + 1.0);
    rms[2 * nr + i] = (rms[2 * nr + i] * cnt + w * w) / (cnt + 1.0);
  }
}

__global__ void __launch_bounds__(256) k_umax(const double *__restrict__ u, long nr, double *__restrict__ part) {
  __shared__ double red[256];
  double m = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nr; i += (long)gridDim.x * 256) m = fmax(m, fabs(u[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_umax_final(const double *__restrict__ part, int n, double *__restrict__ out) {
  if (threadIdx.x) return;
  double m = 0.0;
  for (int b = 0; b < n; b++) m = fmax(m, part[b]);
  *out = m;
}

// RiG plane sums (out_low.h:18-76).  zdir: planes in z with periodic neighbours
// (theta >= pi/4); else planes in y with the reference's neighbour indices
// (iy = ny-1 takes iy+1 = 1, iy = 0 takes iy+1 = ny-1 and iy-1 = -1 through the
// flat index, which here wraps modulo the grid)
__global__ void __launch_bounds__(256) k_rig(const double *__restrict__ phys, int nx, int ny, int nz, int zdir,
                                             double *__restrict__ part) {
  const long nr = (long)nx * ny * nz;
  const double *U = phys + 3 * nr, *V = phys + 4 * nr, *RO = phys + 6 * nr;
  const double twopi = 4.0 * asin(1.0);
  const int pl = blockIdx.x, b = blockIdx.y;
  KSum acc[2];
  if (zdir) {
    const int iz = pl, izp = iz == nz - 1 ? 0 : iz + 1, izm = iz == 0 ? nz - 1 : iz - 1;
    const long npt = (long)nx * ny;
    for (long q = (long)b * 256 + threadIdx.x; q < npt; q += (long)PB * 256) {
      const long ikp = (long)izp * ny * nx + q, ikm = (long)izm * ny * nx + q;
      acc[0].add(0.5 * nz * (RO[ikp] - RO[ikm]) / twopi + 1.0);
      const double sv = 0.5 * nz * (V[ikp] - V[ikm]) / twopi, su = 0.5 * nz * (U[ikp] - U[ikm]) / twopi;
      acc[1].add(sv * sv + su * su);
    }
  } else {
    const int iy = pl;
    int iyp = iy + 1;
    if (iy == ny - 1) iyp = 1;
    const int iym = iy - 1;
    if (iy == 0) iyp = ny - 1;
    const long npt = (long)nx * nz;
    for (long q = (long)b * 256 + threadIdx.x; q < npt; q += (long)PB * 256) {
      const long ix = q % nx, iz = q / nx;
      const long ikp = (iz * ny + iyp) * nx + ix;
      long ikm = (iz * ny + iym) * nx + ix;
      if (ikm < 0) ikm += nr;
      acc[0].add(0.5 * ny * (RO[ikp] - RO[ikm]) / twopi + 1.0);
      const double su = 0.5 * ny * (U[ikp] - U[ikm]) / twopi;
      acc[1].add(su * su);
    }
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = acc[0].val();
  red[1][threadIdx.x] = acc[1].val();
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long)pl * PB + b) * 2] = red[0][0];
    part[((long)pl * PB + b) * 2 + 1] = red[1][0];
  }
}

__global__ void k_rig_final(const double *__restrict__ part, int npl, double *__restrict__ out) {
  const int pl = blockIdx.x * 64 + threadIdx.x;
  if (pl >= npl) return;
  KSum a, s;
  for (int b = 0; b < PB; b++) {
    a.add(part[((long)pl * PB + b) * 2]);
    s.add(part[((long)pl * PB + b) * 2 + 1]);
  }
  out[2 * pl] = a.val();
  out[2 * pl + 1] = s.val();
}

// spectra of the averaged vorticity |<w>|^2 / wk: sums along the vertical index
// per horizontal column, and over the columns per vertical index (vdir 2: kz
// vertical, columns (ikx, iky); vdir 1: ky vertical, columns (ikx, ikz))
__device__ inline double spec_term(const cpx<double> *t, long nkt, const Grid &g, long ik) {
  double kx, ky, kz;
  kxyz(g, ik, kx, ky, kz);
  double wk = kx * kx + ky * ky + kz * kz;
  wk = wk > 0.001 ? wk : 0.001;
  const cpx<double> X = t[ik], E = t[nkt + ik], Z = t[2 * nkt + ik];
  return ((X.x * X.x + X.y * X.y) + (E.x * E.x + E.y * E.y) + (Z.x * Z.x + Z.y * Z.y)) / wk;
}

__global__ void k_spec_cols(const cpx<double> *__restrict__ t, long nkt, Grid g, int vdir, double *__restrict__ cols) {
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  const long ncol = (long)g.iktx * (vdir == 2 ? g.ikty : g.iktz);
  if (c >= ncol) return;
  const long ikx = c % g.iktx, o = c / g.iktx;
  KSum s;
  const int nv = vdir == 2 ? g.iktz : g.ikty;
  for (int v = 0; v < nv; v++) {
    const long ik = vdir == 2 ? ((long)v * g.ikty + o) * g.iktx + ikx : (o * g.ikty + v) * g.iktx + ikx;
    s.add(spec_term(t, nkt, g, ik));
  }
  cols[c] = s.val();
}

__global__ void __launch_bounds__(256) k_spec_rows(const cpx<double> *__restrict__ t, long nkt, Grid g, int vdir,
                                                   double *__restrict__ rows) {
  const int v = blockIdx.x;
  const long ncol = (long)g.iktx * (vdir == 2 ? g.ikty : g.iktz);
  KSum s;
  for (long c = threadIdx.x; c < ncol; c += 256) {
    const long ikx = c % g.iktx, o = c / g.iktx;
    const long ik = vdir == 2 ? ((long)v * g.ikty + o) * g.iktx + ikx : (o * g.ikty + v) * g.iktx + ikx;
    s.add(spec_term(t, nkt, g, ik));
  }
  __shared__ double red[256];
  red[threadIdx.x] = s.val();
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rows[v] = red[0];
}

// subtractReal (kernels.h:125-134): <u^2> - <u>^2 into out
__global__ void k_uvar(const double *__restrict__ rms_u, const double *__restrict__ u, long nr, double *__restrict__ out) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nr; i += (long)gridDim.x * 256) out[i] = rms_u[i] - u[i] * u[i];
}

#define ST_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);    \
      return PSG_EHIP;                                                                        \
    }                                                                                         \
  } while (0)

}  // namespace

struct psg_stats {
  Ctx *c = nullptr;
  int member = 0, nf = 0;
  bool bous = false;
  long nkt = 0, nr = 0;
  cpx<double> *tavg = nullptr;
  double *part = nullptr;    // SB * NVMAX
  double *vals = nullptr;    // NVMAX + 1
  double *rms = nullptr;     // 3 nr, lazily
  double *phys = nullptr;    // 7 nr, lazily
  double *pl = nullptr;      // plane partials and results
  int npl = 0;
};

static int st_alloc_phys(psg_stats *s) {
  if (!s->phys) {
    ST_HIP(hipMalloc(&s->phys, sizeof(double) * 7 * s->nr));
  }
  if (!s->rms) {
    ST_HIP(hipMalloc(&s->rms, sizeof(double) * 3 * s->nr));
    ST_HIP(hipMemsetAsync(s->rms, 0, sizeof(double) * 3 * s->nr, s->c->stream));
  }
  return 0;
}

static bool zplanes(const Ctx *c) { return !(c->cfg.theta < 4.0 * asin(1.0) / 8.0); }

extern "C" {

int psg_set_ampfor(psg_ctx *ctx, double ampfor) {
  if (!ctx) { set_error("psg_set_ampfor: null context"); return PSG_EINVAL; }
  ST_HIP(hipStreamSynchronize(ctx->stream));
  ctx->cfg.ampfor = ampfor;
  ctx->graph_gen++;   // captured steps hold the amplitude by value
  return 0;
}

int psg_stats_create(psg_ctx *ctx, int member, psg_stats **out) {
  if (!ctx || !out) { set_error("psg_stats_create: null argument"); return PSG_EINVAL; }
  *out = nullptr;
  Ctx *c = ctx;
  if (c->g.dim != 3 || (c->cfg.flavour != PSG_NS3D && c->cfg.flavour != PSG_EULER3D && c->cfg.flavour != PSG_BOUS3D)) {
    set_error("psg_stats_create: 3D flavours only");
    return PSG_EINVAL;
  }
  if (member < 0 || member >= c->batch) { set_error("psg_stats_create: member %d out of range", member); return PSG_EINVAL; }
  psg_stats *s = new psg_stats();
  s->c = c;
  s->member = member;
  s->nf = c->nf;
  s->bous = c->cfg.flavour == PSG_BOUS3D;
  s->nkt = c->g.nkt;
  s->nr = (long)c->g.nx * c->g.ny * c->g.nz;
  s->npl = zplanes(c) ? c->g.nz : c->g.ny;
  const size_t tb = sizeof(cpx<double>) * s->nf * s->nkt;
  if (hipMalloc(&s->tavg, tb) != hipSuccess || hipMalloc(&s->part, sizeof(double) * SB * NVMAX) != hipSuccess ||
      hipMalloc(&s->vals, sizeof(double) * (NVMAX + 1)) != hipSuccess ||
      hipMalloc(&s->pl, sizeof(double) * ((size_t)s->npl * PB * 2 + 2 * s->npl)) != hipSuccess ||
      hipMemsetAsync(s->tavg, 0, tb, c->stream) != hipSuccess) {
    set_error("psg_stats_create: device allocation");
    psg_stats_destroy(s);
    return PSG_ENOMEM;
  }
  *out = s;
  return 0;
}

int psg_stats_update(psg_stats *s, long count, double *vals) {
  if (!s || !vals) { set_error("psg_stats_update: null argument"); return PSG_EINVAL; }
  Ctx *c = s->c;
  const cpx<double> *st = c->state + (long)s->member * s->nf * s->nkt;
  const double th = c->cfg.theta;
  StPar p{c->cfg.ampfor, c->cfg.kf, c->cfg.ri, sin(th), cos(th)};
  const int nv = s->bous ? (int)PSG_STB_N : (int)PSG_STNS_N;
  if (s->bous)
    hipLaunchKernelGGL(k_stats<true>, dim3(SB), dim3(256), 0, c->stream, st, s->nkt, c->g, c->mask, s->tavg,
                       (double)count, p, s->part);
  else
    hipLaunchKernelGGL(k_stats<false>, dim3(SB), dim3(256), 0, c->stream, st, s->nkt, c->g, c->mask, s->tavg,
                       (double)count, p, s->part);
  ST_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, c->stream, s->part, SB, nv, s->vals);
  ST_HIP(hipGetLastError());
  ST_HIP(hipMemcpyAsync(vals, s->vals, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
  ST_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// physical fields of a state: (xi, eta, zeta, u, v, w[, rho]) as the flavour has them
static int phys_state(psg_stats *s, const cpx<double> *st) {
  return s->bous ? bous_physical_state(s->c, st, s->phys) : ns3d_physical_state(s->c, st, s->phys);
}

int psg_stats_physical(psg_stats *s, int what, long count, double *umax, double *navg, double *savg, int *nplanes) {
  if (!s) { set_error("psg_stats_physical: null handle"); return PSG_EINVAL; }
  if (!s->bous && (what & PSG_STP_RIG)) { set_error("psg_stats_physical: RiG needs a Boussinesq context"); return PSG_EINVAL; }
  Ctx *c = s->c;
  if (int r = st_alloc_phys(s)) return r;
  if (int r = phys_state(s, c->state + (long)s->member * s->nf * s->nkt)) return r;
  const int nb = 2048;
  if (what & PSG_STP_RMS) {
    hipLaunchKernelGGL(k_rms, dim3(nb), dim3(256), 0, c->stream, s->phys, s->nr, s->rms, (double)count);
    ST_HIP(hipGetLastError());
  }
  if (what & PSG_STP_UMAX) {
    if (!umax) { set_error("psg_stats_physical: umax is NULL"); return PSG_EINVAL; }
    hipLaunchKernelGGL(k_umax, dim3(SB), dim3(256), 0, c->stream, s->phys + 3 * s->nr, s->nr, s->part);
    hipLaunchKernelGGL(k_umax_final, dim3(1), dim3(64), 0, c->stream, s->part, SB, s->vals + NVMAX);
    ST_HIP(hipGetLastError());
    ST_HIP(hipMemcpyAsync(umax, s->vals + NVMAX, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (what & PSG_STP_RIG) {
    if (!navg || !savg) { set_error("psg_stats_physical: navg/savg are NULL"); return PSG_EINVAL; }
    const int zd = zplanes(c) ? 1 : 0;
    hipLaunchKernelGGL(k_rig, dim3(s->npl, PB), dim3(256), 0, c->stream, s->phys, c->g.nx, c->g.ny, c->g.nz, zd,
                       s->pl);
    double *res = s->pl + (size_t)s->npl * PB * 2;
    hipLaunchKernelGGL(k_rig_final, dim3((s->npl + 63) / 64), dim3(64), 0, c->stream, s->pl, s->npl, res);
    ST_HIP(hipGetLastError());
    std::vector<double> h(2 * s->npl);
    ST_HIP(hipMemcpyAsync(h.data(), res, sizeof(double) * 2 * s->npl, hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < s->npl; i++) {
      navg[i] = h[2 * i];
      savg[i] = h[2 * i + 1];
    }
    if (nplanes) *nplanes = s->npl;
  }
  ST_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_stats_spectra(psg_stats *s, double *spch, double *spcv, int cap, int *size) {
  if (!s || !spch || !spcv || !size) { set_error("psg_stats_spectra: null argument"); return PSG_EINVAL; }
  Ctx *c = s->c;
  const Grid &g = c->g;
  if (cap < g.nx + 1) { set_error("psg_stats_spectra: cap %d < nx + 1", cap); return PSG_EINVAL; }
  const int vdir = zplanes(c) ? 2 : 1;
  const long ncol = (long)g.iktx * (vdir == 2 ? g.ikty : g.iktz);
  const int nv = vdir == 2 ? g.iktz : g.ikty;
  double *d = nullptr;
  ST_HIP(hipMalloc(&d, sizeof(double) * (ncol + nv)));
  hipLaunchKernelGGL(k_spec_cols, dim3((ncol + 255) / 256), dim3(256), 0, c->stream, s->tavg, s->nkt, g, vdir, d);
  hipLaunchKernelGGL(k_spec_rows, dim3(nv), dim3(256), 0, c->stream, s->tavg, s->nkt, g, vdir, d + ncol);
  std::vector<double> h(ncol + nv);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * (ncol + nv), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) { set_error("psg_stats_spectra: %s", hipGetErrorString(e)); return PSG_EHIP; }
  // binning of finalise.h:32-55: horizontal |k_h|, vertical max(k_v, 1e-3), rounded
  for (int i = 0; i < cap; i++) spch[i] = spcv[i] = 0.0;
  int sz = 0;
  for (long col = 0; col < ncol; col++) {
    const long ikx = col % g.iktx, o = col / g.iktx;
    const double kx = ikx * g.alpha, kh2 = vdir == 2 ? (double)(o - g.kty) : (double)(o - g.ktz);
    double wkh = sqrt(kx * kx + kh2 * kh2);
    wkh = wkh > 0.001 ? wkh : 0.001;
    const int jh = (int)(wkh + 0.5);
    if (jh > g.nx) continue;
    spch[jh] += h[col];
    sz = sz > jh ? sz : jh;
  }
  for (int v = 0; v < nv; v++) {
    const double kv = vdir == 2 ? (double)(v - g.ktz) : (double)(v - g.kty);
    const double wkv = kv > 0.001 ? kv : 0.001;
    const int jv = (int)(wkv + 0.5);
    if (jv > g.nx) continue;
    spcv[jv] += h[ncol + v];
    sz = sz > jv ? sz : jv;
  }
  *size = sz;
  return 0;
}

int psg_stats_mean_fields(psg_stats *s, double *out) {
  if (!s || !out) { set_error("psg_stats_mean_fields: null argument"); return PSG_EINVAL; }
  Ctx *c = s->c;
  if (int r = st_alloc_phys(s)) return r;
  if (int r = phys_state(s, s->tavg)) return r;
  const long nr = s->nr;
  hipLaunchKernelGGL(k_uvar, dim3(2048), dim3(256), 0, c->stream, s->rms, s->phys + 3 * nr, nr, s->phys);
  ST_HIP(hipGetLastError());
  // Boussinesq: rho, u, v, w, <v^2>, <w^2>, <u^2>-<u>^2 (3D_Boussinesq/src/finalise.h:62-160);
  // Navier-Stokes: u, v, w, <v^2>, <u^2>-<u>^2 (3D_NavierStokes/src/finalise.h:25-117)
  const double *srcb[7] = {s->phys + 6 * nr, s->phys + 3 * nr, s->phys + 4 * nr, s->phys + 5 * nr,
                           s->rms + nr, s->rms + 2 * nr, s->phys};
  const double *srcn[5] = {s->phys + 3 * nr, s->phys + 4 * nr, s->phys + 5 * nr, s->rms + nr, s->phys};
  const int nfo = s->bous ? 7 : 5;
  for (int f = 0; f < nfo; f++)
    ST_HIP(hipMemcpyAsync(out + f * nr, s->bous ? srcb[f] : srcn[f], sizeof(double) * nr, hipMemcpyDeviceToHost,
                          c->stream));
  ST_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_stats_tavg(psg_stats *s, double *host) {
  if (!s || !host) { set_error("psg_stats_tavg: null argument"); return PSG_EINVAL; }
  ST_HIP(hipStreamSynchronize(s->c->stream));
  ST_HIP(hipMemcpy(host, s->tavg, sizeof(cpx<double>) * s->nf * s->nkt, hipMemcpyDeviceToHost));
  return 0;
}

int psg_stats_destroy(psg_stats *s) {
  if (!s) return 0;
  if (s->c) (void)hipStreamSynchronize(s->c->stream);
  void *p[] = {s->tavg, s->part, s->vals, s->rms, s->phys, s->pl};
  for (void *q : p)
    if (q) (void)hipFree(q);
  delete s;
  return 0;
}

}  // extern "C"
