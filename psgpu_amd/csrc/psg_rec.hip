// Recurrence search over a DNS history (include/psg_rec.h).
//
// Reference: 3D_NavierStokes/src/functions.h:111-125 (computeNorm), :213-389
// (computeZdash, recurrence_check), cuda.cu:114-139 (the NOUT block), init.h:
// 182-213 (subset and shifts); 3D_Boussinesq/src/functions.h:190-208,
// :415-619, out_high.h:8-33, init.h:275-310; 2D_NavierStokes/src_RK4/
// functions.h:59-67, :105-309.
//
// Device data (HBM): the subset ring sub[store][field][i] (complex), the full
// states ring (optional), the phase table e^{-i k.s} per (shift, i), the
// shifted current subset dash[s][field][i] (density pre-multiplied by ri).
// A check is three kernels on the context's stream:
//   k_rec_dash   one block per shift: dash and |dash|^2 (fixed-order sums)
//   k_rec_cross  split-K FP64 product C[s][j] = sum_k Re(dash[s][k] conj sub[j][k])
//                over (32 shifts x 32 stores) LDS tiles, partial sums per K slice
//   k_rec_min    per store: residuals in the reference's shift order and the
//                first minimum (strict <)
// and the per-store minima come back to the host, which runs the reference's
// turning-point and output-sequence logic on them.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <vector>
#include "psg_internal.hpp"
#include "../../include/psg_rec.h"

using namespace psg;

struct psg_rec {
  Ctx *c = nullptr;
  psg_rec_opts o;
  int nf = 0, nnorm = 0, ns = 0;
  long nkt = 0, K = 0;                 // K = nf * nnorm complex per subset vector
  double v2 = 0.0, freq2 = 0.0;
  std::vector<int> idx;                // subset mode indices (iNorm)
  std::vector<double> shx, shy, shz;   // shift grids
  std::vector<double> normz;           // normZStore
  // device
  int *d_idx = nullptr;
  double *d_nw = nullptr;              // per-(field, i) weight of the history norm
  double *d_fw = nullptr;              // per-field weight of the comparison (ri for density)
  cpx<double> *d_phase = nullptr;      // [ns][nnorm] (cos, sin)
  cpx<double> *d_sub = nullptr;        // [store_size][K]
  cpx<double> *d_full = nullptr;       // [store_size][nf*nkt] or null
  cpx<double> *d_dash = nullptr;       // [ns][K]
  double *d_nzd = nullptr;             // [ns]
  double *d_norm = nullptr;            // [store_size + 1]: history norms, then the last push's
  double *d_part = nullptr;            // [ks_max][ns][store_size]
  double *d_res = nullptr;             // [store_size][ns]
  int *d_slots = nullptr;              // [store_size]
  double *d_rmin = nullptr;            // [store_size]
  int *d_smin = nullptr;               // [store_size]
  int ks_max = 0;
  // reference bookkeeping
  int nStore = 0, nCopy = 0, iStart = 0, RecOUTflag = 0, minSTOREouter = 0, izkout = 0;
  double minRESIDUALouter = 0, minTIMEouter = 0, RecPERIOD = 0, minTIMEinner = 0;
  double minSHIFTXouter = 0, minSHIFTYouter = 0, minSHIFTZouter = 0;
  double minSHIFTXinner = 0, minSHIFTYinner = 0, minSHIFTZinner = 0;
  // last check
  std::vector<int> slots, counts;
  std::vector<double> periods;
  int last_nj = 0;
};

namespace {

constexpr int RT = 32;    // shifts x stores per cross tile
constexpr int RK = 16;    // complex K per LDS stage
constexpr int KS_MAX = 16;

__device__ inline double block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

// subset of member state -> sub[slot]; history norm (computeNorm) -> *norm
__global__ void __launch_bounds__(256) k_rec_gather(const cpx<double> *__restrict__ st, long nkt, int nf,
                                                    const int *__restrict__ idx, int nnorm,
                                                    const double *__restrict__ nw, int quirk_field, double ri,
                                                    cpx<double> *__restrict__ sub, double *__restrict__ norm) {
  __shared__ double red[256];
  double acc = 0.0;
  const int K = nf * nnorm;
  for (int e = threadIdx.x; e < K; e += 256) {
    const int f = e / nnorm, i = e - f * nnorm;
    const cpx<double> z = st[(long)f * nkt + idx[i]];
    sub[e] = z;
    // the Boussinesq history norm weights only the density's real part by ri
    // (3D_Boussinesq/src/functions.h:204: ri*x*x + y*y)
    acc += f == quirk_field ? ri * z.x * z.x + z.y * z.y : nw[e] * (z.x * z.x + z.y * z.y);
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) *norm = s;
}

// dash[s] = fw_f * Z e^{-i theta_s} (computeZdash) and nzd[s] = sum fw_f |Z e^{-i theta_s}|^2
__global__ void __launch_bounds__(256) k_rec_dash(const cpx<double> *__restrict__ z, int nf, int nnorm,
                                                  const cpx<double> *__restrict__ phase,
                                                  const double *__restrict__ fw, cpx<double> *__restrict__ dash,
                                                  double *__restrict__ nzd) {
  __shared__ double red[256];
  const int s = blockIdx.x;
  const int K = nf * nnorm;
  double acc = 0.0;
  for (int e = threadIdx.x; e < K; e += 256) {
    const int f = e / nnorm, i = e - f * nnorm;
    const cpx<double> p = phase[(long)s * nnorm + i], v = z[e];
    const double dx = p.x * v.x + p.y * v.y, dy = p.x * v.y - p.y * v.x;
    const double w = fw[f];
    dash[(long)s * K + e] = cpx<double>{w * dx, w * dy};
    acc += w * (dx * dx + dy * dy);
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) nzd[s] = t;
}

// part[z][s][j] = sum over K slice z of Re(dash[s][k] conj sub[slot_j][k])
__global__ void __launch_bounds__(256) k_rec_cross(const cpx<double> *__restrict__ dash, int ns,
                                                   const cpx<double> *__restrict__ sub,
                                                   const int *__restrict__ slots, int nj, long K, long kslice,
                                                   double *__restrict__ part) {
  __shared__ cpx<double> A[RK][RT + 1];
  __shared__ cpx<double> B[RK][RT + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int s0 = blockIdx.x * RT, j0 = blockIdx.y * RT;
  const long k0 = (long)blockIdx.z * kslice;
  const long k1 = k0 + kslice < K ? k0 + kslice : K;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (long kb = k0; kb < k1; kb += RK) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int e = threadIdx.x + 256 * r;
      const int kk = e % RK, q = e / RK;
      const long k = kb + kk;
      const int s = s0 + q, j = j0 + q;
      A[kk][q] = (k < k1 && s < ns) ? dash[(long)s * K + k] : cpx<double>{0.0, 0.0};
      B[kk][q] = (k < k1 && j < nj) ? sub[(long)slots[j] * K + k] : cpx<double>{0.0, 0.0};
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < RK; kk++) {
      const cpx<double> a0 = A[kk][ty], a1 = A[kk][ty + 16], b0 = B[kk][tx], b1 = B[kk][tx + 16];
      acc[0][0] += a0.x * b0.x + a0.y * b0.y;
      acc[0][1] += a0.x * b1.x + a0.y * b1.y;
      acc[1][0] += a1.x * b0.x + a1.y * b0.y;
      acc[1][1] += a1.x * b1.x + a1.y * b1.y;
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int s = s0 + ty + 16 * a, j = j0 + tx + 16 * b;
      if (s < ns && j < nj) part[((long)blockIdx.z * ns + s) * nj + j] = acc[a][b];
    }
}

// residuals of store j over the shifts in the reference's loop order
// (y outer, then z, then x; is = x + NX (y + NY z)); first strict minimum
__global__ void __launch_bounds__(64) k_rec_min(const double *__restrict__ part, int nks, int ns, int nj,
                                                const int *__restrict__ slots, const double *__restrict__ nzd,
                                                const double *__restrict__ norms_dev, double norm_copy,
                                                int rel_both, int nxs, int nys, int nzs,
                                                double *__restrict__ res, double *__restrict__ rmin,
                                                int *__restrict__ smin) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= nj) return;
  const double nj_norm = norms_dev[slots[j]];
  const double den = rel_both ? nj_norm + norm_copy : norm_copy;
  double best = 10000.0;
  int bs = -1;
  for (int y = 0; y < nys; y++)
    for (int zz = 0; zz < nzs; zz++)
      for (int x = 0; x < nxs; x++) {
        const int s = x + nxs * (y + nys * zz);
        double c = 0.0;
        for (int q = 0; q < nks; q++) c += part[((long)q * ns + s) * nj + j];
        const double fn = -2.0 * c + nzd[s] + nj_norm;
        const double r = sqrt(fn / den);
        res[(long)j * ns + s] = r;
        if (r < best) { best = r; bs = s; }
      }
  rmin[j] = best;
  smin[j] = bs;
}

void rec_free(psg_rec *r) {
  void *p[] = {r->d_idx, r->d_nw, r->d_fw, r->d_phase, r->d_sub, r->d_full, r->d_dash, r->d_nzd,
               r->d_norm, r->d_part, r->d_res, r->d_slots, r->d_rmin, r->d_smin};
  for (void *q : p)
    if (q) (void)hipFree(q);
}

#define REC_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return PSG_EHIP;                                                              \
    }                                                                               \
  } while (0)

// push the member's current state into slot `slot`: subset, norm, full copy
int rec_push(psg_rec *r, int slot) {
  Ctx *c = r->c;
  const cpx<double> *st = c->state + (long)r->o.member * r->nf * r->nkt;
  const int quirk = r->o.rule == PSG_REC_BOUS ? 3 : -1;
  hipLaunchKernelGGL(k_rec_gather, dim3(1), dim3(256), 0, c->stream, st, r->nkt, r->nf, r->d_idx, r->nnorm,
                     r->d_nw, quirk, r->o.ri, r->d_sub + (long)slot * r->K, r->d_norm + r->o.store_size);
  REC_HIP(hipGetLastError());
  if (r->d_full)
    REC_HIP(hipMemcpyAsync(r->d_full + (long)slot * r->nf * r->nkt, st, sizeof(cpx<double>) * r->nf * r->nkt,
                           hipMemcpyDeviceToDevice, c->stream));
  double nz = 0.0;
  REC_HIP(hipMemcpyAsync(&nz, r->d_norm + r->o.store_size, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  REC_HIP(hipStreamSynchronize(c->stream));
  r->normz[slot] = nz;
  return 0;
}

// the stores the check compares, in the reference's order (iCount = nStore-1 .. 1)
void rec_plan(psg_rec *r, double delt) {
  r->slots.clear();
  r->counts.clear();
  r->periods.clear();
  const int S = r->o.store_size;
  for (int iCount = r->nStore - 1; iCount >= 1; iCount--) {
    const double PERIODinner = ((r->nStore - iCount) * r->o.nout) * delt;
    int iStore = iCount + r->iStart;
    iStore -= (iStore / S) * S;
    if (r->o.rule == PSG_REC_BOUS && PERIODinner < 0.5) continue;
    if (PERIODinner > 20.0 && fabs(PERIODinner - r->freq2 * floor(PERIODinner / r->freq2 + 0.5)) > 0.0001) continue;
    r->slots.push_back(iStore);
    r->counts.push_back(iCount);
    r->periods.push_back(PERIODinner);
  }
}

}  // namespace

extern "C" {

int psg_rec_defaults(psg_ctx *ctx, psg_rec_opts *o) {
  if (!ctx || !o) { set_error("psg_rec_defaults: null argument"); return PSG_EINVAL; }
  memset(o, 0, sizeof *o);
  const int fl = ctx->cfg.flavour;
  o->store_size = 500;
  o->threshold = 1.0;
  o->tstart = 0.0;
  o->keep_states = 1;
  const int nout = (int)(0.1 / ctx->cfg.dt);   // NOUT = INT(outfreq1/DELT), outfreq1 = 0.1
  o->nout = nout > 0 ? nout : 1;
  if (fl == PSG_NS2D) {
    o->rule = PSG_REC_NS2D; o->norm_width = 42; o->nshift_x = 30; o->nshift_y = 4; o->nshift_z = 1;
  } else if (fl == PSG_BOUS3D) {
    o->rule = PSG_REC_BOUS; o->norm_width = 8; o->nshift_x = 64; o->nshift_y = 1; o->nshift_z = 1;
    o->ri = ctx->cfg.ri;
  } else {
    o->rule = PSG_REC_NS; o->norm_width = 8; o->nshift_x = 30; o->nshift_y = 2; o->nshift_z = 2;
  }
  return 0;
}

int psg_rec_create(psg_ctx *ctx, const psg_rec_opts *o, psg_rec **out) {
  if (!ctx || !o || !out) { set_error("psg_rec_create: null argument"); return PSG_EINVAL; }
  *out = nullptr;
  Ctx *c = ctx;
  const Grid &g = c->g;
  const int W = o->norm_width;
  if (o->member < 0 || o->member >= c->batch) { set_error("psg_rec_create: member %d out of range", o->member); return PSG_EINVAL; }
  if (o->store_size < 2 || o->store_size > 100000) { set_error("psg_rec_create: store_size %d", o->store_size); return PSG_EINVAL; }
  if (o->nout < 1) { set_error("psg_rec_create: nout must be >= 1"); return PSG_EINVAL; }
  if (o->nshift_x < 1 || o->nshift_y < 1 || o->nshift_z < 1 || (long)o->nshift_x * o->nshift_y * o->nshift_z > 4096) {
    set_error("psg_rec_create: shift counts %d x %d x %d (1 .. 4096 shifts)", o->nshift_x, o->nshift_y, o->nshift_z);
    return PSG_EINVAL;
  }
  const bool two = g.dim == 2;
  if (two != (o->rule == PSG_REC_NS2D) || (o->rule == PSG_REC_BOUS) != (c->cfg.flavour == PSG_BOUS3D) ||
      o->rule < 0 || o->rule > 2) {
    set_error("psg_rec_create: rule %d does not fit the context's flavour", o->rule);
    return PSG_EINVAL;
  }
  if (two && o->nshift_z != 1) { set_error("psg_rec_create: 2D takes nshift_z = 1"); return PSG_EINVAL; }
  // subset bounds: ikx < W; 3D: ikz, iky = KTZ - W .. KTZ + W - 1 (init.h:184-190 offsets
  // both by KTZ); 2D: iky = KTY - W .. KTY + W - 1
  const int off = two ? g.kty : g.ktz;
  if (W < 1 || W > g.iktx || off - W < 0 || off + W - 1 >= g.ikty || (!two && g.ktz + W - 1 >= g.iktz)) {
    set_error("psg_rec_create: norm_width %d does not fit the grid (IKTX %d, KTY %d, KTZ %d)", W, g.iktx, g.kty, g.ktz);
    return PSG_EINVAL;
  }
  psg_rec *r = new psg_rec();
  r->c = c;
  r->o = *o;
  r->nf = c->nf;
  r->nkt = g.nkt;
  r->v2 = c->cfg.v2;
  r->freq2 = 2.0 * o->nout * c->cfg.dt;   // init.h: freq2 = 2*nOut*delt, fixed at set-up
  r->nnorm = two ? 2 * W * W : 4 * W * W * W;
  r->ns = o->nshift_x * o->nshift_y * o->nshift_z;
  r->K = (long)r->nf * r->nnorm;
  r->idx.resize(r->nnorm);
  std::vector<double> kx(r->nnorm), ky(r->nnorm), kz(r->nnorm, 0.0);
  if (two) {
    for (int iy = 0; iy < 2 * W; iy++)
      for (int ix = 0; ix < W; ix++) {
        const int i = iy * W + ix, iky = g.kty - W + iy;
        r->idx[i] = iky * g.iktx + ix;
        kx[i] = ix * g.alpha;
        ky[i] = iky - g.kty;
      }
  } else {
    for (int iz = 0; iz < 2 * W; iz++)
      for (int iy = 0; iy < 2 * W; iy++)
        for (int ix = 0; ix < W; ix++) {
          const int i = (iz * 2 * W + iy) * W + ix, ikz = g.ktz - W + iz, iky = g.ktz - W + iy;
          r->idx[i] = (ikz * g.ikty + iky) * g.iktx + ix;
          kx[i] = ix * g.alpha;
          ky[i] = iky - g.kty;
          kz[i] = ikz - g.ktz;
        }
  }
  const double twopi = 4.0 * asin(1.0);
  r->shx.resize(o->nshift_x);
  r->shy.resize(o->nshift_y);
  r->shz.resize(o->nshift_z);
  for (int i = 0; i < o->nshift_y; i++) r->shy[i] = twopi * i / (o->nshift_y);
  for (int i = 0; i < o->nshift_x; i++) r->shx[i] = twopi * i / (o->nshift_x * g.alpha);
  for (int i = 0; i < o->nshift_z; i++) r->shz[i] = twopi * i / (o->nshift_z);
  if (two) r->shz[0] = 0.0;
  // phase table (cos, sin) of theta = kx sx + ky sy + kz sz, on the host's libm
  std::vector<cpx<double>> ph((size_t)r->ns * r->nnorm);
  for (int z = 0; z < o->nshift_z; z++)
    for (int y = 0; y < o->nshift_y; y++)
      for (int x = 0; x < o->nshift_x; x++) {
        const int s = x + o->nshift_x * (y + o->nshift_y * z);
        for (int i = 0; i < r->nnorm; i++) {
          const double th = kx[i] * r->shx[x] + ky[i] * r->shy[y] + kz[i] * r->shz[z];
          ph[(size_t)s * r->nnorm + i] = cpx<double>{cos(th), sin(th)};
        }
      }
  // history norm weights: the NS search masks by the active set (functions.h:118),
  // the Boussinesq and 2D ones do not
  std::vector<int8_t> hm(g.nkt);
  if (hipMemcpy(hm.data(), c->mask, g.nkt, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("psg_rec_create: mask copy");
    delete r;
    return PSG_EHIP;
  }
  std::vector<double> nw(r->K, 1.0), fw(r->nf, 1.0);
  if (o->rule == PSG_REC_NS)
    for (int f = 0; f < r->nf; f++)
      for (int i = 0; i < r->nnorm; i++) nw[(size_t)f * r->nnorm + i] = hm[r->idx[i]] == 1 ? 1.0 : 0.0;
  if (o->rule == PSG_REC_BOUS) fw[3] = o->ri;
  const int S = o->store_size;
  r->ks_max = KS_MAX;
  r->normz.assign(S, 0.0);
  auto fail = [&](const char *what) {
    set_error("psg_rec_create: %s", what);
    rec_free(r);
    delete r;
    return PSG_ENOMEM;
  };
  if (hipMalloc(&r->d_idx, sizeof(int) * r->nnorm) != hipSuccess ||
      hipMalloc(&r->d_nw, sizeof(double) * r->K) != hipSuccess ||
      hipMalloc(&r->d_fw, sizeof(double) * r->nf) != hipSuccess ||
      hipMalloc(&r->d_phase, sizeof(cpx<double>) * ph.size()) != hipSuccess ||
      hipMalloc(&r->d_sub, sizeof(cpx<double>) * S * r->K) != hipSuccess ||
      hipMalloc(&r->d_dash, sizeof(cpx<double>) * r->ns * r->K) != hipSuccess ||
      hipMalloc(&r->d_nzd, sizeof(double) * r->ns) != hipSuccess ||
      hipMalloc(&r->d_norm, sizeof(double) * (S + 1)) != hipSuccess ||
      hipMalloc(&r->d_part, sizeof(double) * KS_MAX * r->ns * S) != hipSuccess ||
      hipMalloc(&r->d_res, sizeof(double) * r->ns * S) != hipSuccess ||
      hipMalloc(&r->d_slots, sizeof(int) * S) != hipSuccess ||
      hipMalloc(&r->d_rmin, sizeof(double) * S) != hipSuccess ||
      hipMalloc(&r->d_smin, sizeof(int) * S) != hipSuccess)
    return fail("device allocation");
  if (o->keep_states && hipMalloc(&r->d_full, sizeof(cpx<double>) * S * r->nf * g.nkt) != hipSuccess)
    return fail("device allocation of the state history (store_size full states; keep_states = 0 keeps subsets only)");
  if (hipMemcpy(r->d_idx, r->idx.data(), sizeof(int) * r->nnorm, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(r->d_nw, nw.data(), sizeof(double) * r->K, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(r->d_fw, fw.data(), sizeof(double) * r->nf, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(r->d_phase, ph.data(), sizeof(cpx<double>) * ph.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail("table upload");
  // init.h:165-212: iStart = nCopy = 0, the current state is entry 0
  if (int e = rec_push(r, 0)) {
    rec_free(r);
    delete r;
    return e;
  }
  *out = r;
  return 0;
}

int psg_rec_update(psg_rec *r, long nt, double delt, psg_rec_guess *guess, double *guess_state,
                   psg_rec_check *chk) {
  if (!r) { set_error("psg_rec_update: null handle"); return PSG_EINVAL; }
  if (!(delt > 0.0)) { set_error("psg_rec_update: dt must be > 0"); return PSG_EINVAL; }
  Ctx *c = r->c;
  const psg_rec_opts &o = r->o;
  const int S = o.store_size;
  int produced = 0;
  if (chk) {
    memset(chk, 0, sizeof *chk);
    chk->residual = 10000.0;
  }
  r->last_nj = 0;
  if (r->nStore >= 1) {
    // ---- recurrence_check(NT, nCopy)
    const double tme = o.tstart + (nt - o.nout) * delt;
    rec_plan(r, delt);
    const int nj = (int)r->slots.size();
    std::vector<double> rmin(nj);
    std::vector<int> smin(nj);
    if (nj > 0) {
      hipLaunchKernelGGL(k_rec_dash, dim3(r->ns), dim3(256), 0, c->stream, r->d_sub + (long)r->nCopy * r->K, r->nf,
                         r->nnorm, r->d_phase, r->d_fw, r->d_dash, r->d_nzd);
      REC_HIP(hipGetLastError());
      REC_HIP(hipMemcpyAsync(r->d_slots, r->slots.data(), sizeof(int) * nj, hipMemcpyHostToDevice, c->stream));
      REC_HIP(hipMemcpyAsync(r->d_norm, r->normz.data(), sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
      const int ts = (r->ns + RT - 1) / RT, tj = (nj + RT - 1) / RT;
      const long nchunk = (r->K + RK - 1) / RK;
      long ks = (1024 + ts * tj - 1) / (ts * tj);
      if (ks > KS_MAX) ks = KS_MAX;
      if (ks > nchunk) ks = nchunk;
      if (ks < 1) ks = 1;
      const long kslice = ((nchunk + ks - 1) / ks) * RK;
      ks = (r->K + kslice - 1) / kslice;
      hipLaunchKernelGGL(k_rec_cross, dim3(ts, tj, (unsigned)ks), dim3(256), 0, c->stream, r->d_dash, r->ns,
                         r->d_sub, r->d_slots, nj, r->K, kslice, r->d_part);
      REC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_rec_min, dim3((nj + 63) / 64), dim3(64), 0, c->stream, r->d_part, (int)ks, r->ns, nj,
                         r->d_slots, r->d_nzd, r->d_norm, r->normz[r->nCopy], o.rule == PSG_REC_BOUS ? 1 : 0,
                         o.nshift_x, o.nshift_y, o.nshift_z, r->d_res, r->d_rmin, r->d_smin);
      REC_HIP(hipGetLastError());
      REC_HIP(hipMemcpyAsync(rmin.data(), r->d_rmin, sizeof(double) * nj, hipMemcpyDeviceToHost, c->stream));
      REC_HIP(hipMemcpyAsync(smin.data(), r->d_smin, sizeof(int) * nj, hipMemcpyDeviceToHost, c->stream));
      REC_HIP(hipStreamSynchronize(c->stream));
    }
    r->last_nj = nj;
    // the sequential part of recurrence_check (functions.h:283-340 in 3D)
    double minSx = 0.0, minSy = 0.0, minSz = 0.0, resPrev = 0.0, res = 0.0;
    double minRESIDUALinner = 10000.0, minPERIOD = 10000.0;
    int minPflag = 1, minSTOREinner = 0, minCOUNTinner = 0;
    for (int j = 0; j < nj; j++) {
      const double PERIODinner = r->periods[j];
      resPrev = res;
      res = rmin[j];
      if (smin[j] >= 0) {
        const int s = smin[j], x = s % o.nshift_x, yz = s / o.nshift_x;
        minSx = r->shx[x];
        minSy = r->shy[yz % o.nshift_y];
        minSz = r->shz[yz / o.nshift_y];
      }
      if (res > resPrev && minPflag == 1) {
        minPERIOD = PERIODinner;
      } else {
        minPflag = 0;
      }
      if (res < minRESIDUALinner && PERIODinner > minPERIOD) {
        minRESIDUALinner = res;
        minSTOREinner = r->slots[j];
        minCOUNTinner = r->nStore - r->counts[j];
        r->minSHIFTXinner = minSx;
        r->minSHIFTYinner = minSy;
        r->minSHIFTZinner = minSz;
      }
    }
    if (chk) {
      chk->checked = 1;
      chk->compared = nj;
      chk->time = tme;
      chk->residual = minRESIDUALinner;
      if (minRESIDUALinner < 10000.0) {
        chk->store = minSTOREinner;
        chk->count = minCOUNTinner;
        chk->sx = r->minSHIFTXinner;
        chk->sy = r->minSHIFTYinner;
        chk->sz = r->minSHIFTZinner;
      }
    }
    const double thr = o.threshold;
    if (minRESIDUALinner < thr && r->RecOUTflag == 0) {
      r->minRESIDUALouter = thr;
      if (o.rule == PSG_REC_NS2D) r->minTIMEinner = tme + minCOUNTinner * o.nout * delt;
      r->RecOUTflag = 1;
    }
    if (minRESIDUALinner < r->minRESIDUALouter && r->RecOUTflag == 1) {
      r->minRESIDUALouter = minRESIDUALinner;
      r->minSTOREouter = minSTOREinner;
      r->minTIMEouter = tme + minCOUNTinner * o.nout * delt;
      r->minSHIFTXouter = r->minSHIFTXinner;
      r->minSHIFTYouter = r->minSHIFTYinner;
      r->minSHIFTZouter = r->minSHIFTZinner;
      r->RecPERIOD = ((minCOUNTinner) * o.nout) * (delt);
    }
    if (minRESIDUALinner > thr && r->RecOUTflag == 1) {
      r->RecOUTflag = 0;
      double timeUNDER = 0.0;
      if (o.rule == PSG_REC_NS2D) {
        r->minTIMEinner = tme + minCOUNTinner * o.nout * delt - r->minTIMEinner;
        timeUNDER = sqrt(r->minTIMEinner * r->minTIMEinner / (r->RecPERIOD * r->RecPERIOD));
      }
      r->izkout++;
      if (guess) {
        guess->index = r->izkout;
        guess->store = r->minSTOREouter;
        guess->time = r->minTIMEouter;
        guess->v2 = r->v2;
        guess->period = r->RecPERIOD;
        guess->sx = r->minSHIFTXouter;
        guess->sy = r->minSHIFTYouter;
        guess->sz = o.rule == PSG_REC_NS2D ? 0.0 : r->minSHIFTZouter;
        guess->residual = r->minRESIDUALouter;
        guess->time_under = timeUNDER;
      }
      if (guess_state) {
        if (!r->d_full) { set_error("psg_rec_update: guess state needs keep_states"); return PSG_EINVAL; }
        REC_HIP(hipMemcpy(guess_state, r->d_full + (long)r->minSTOREouter * r->nf * r->nkt,
                          sizeof(cpx<double>) * r->nf * r->nkt, hipMemcpyDeviceToHost));
      }
      produced = 1;
    }
  }
  // ---- ring bookkeeping (cuda.cu:120-130) and the push of the current state
  r->nStore++;
  if (r->nStore >= S) {
    r->nStore = S;
    r->nCopy++;
    r->nCopy -= (r->nCopy / S) * S;
    r->iStart = r->nCopy;
  } else {
    r->nCopy = r->nStore;
  }
  if (int e = rec_push(r, r->nCopy)) return e;
  return produced;
}

int psg_rec_last(psg_rec *r, double *res, int *slots, int *nj, int cap) {
  if (!r || !nj) { set_error("psg_rec_last: null argument"); return PSG_EINVAL; }
  *nj = r->last_nj;
  if (r->last_nj > cap) { set_error("psg_rec_last: %d stores exceed cap %d", r->last_nj, cap); return PSG_EINVAL; }
  if (slots && r->last_nj) memcpy(slots, r->slots.data(), sizeof(int) * r->last_nj);
  if (res && r->last_nj)
    REC_HIP(hipMemcpy(res, r->d_res, sizeof(double) * r->last_nj * r->ns, hipMemcpyDeviceToHost));
  return 0;
}

int psg_rec_state(psg_rec *r, int store, double *host) {
  if (!r || !host) { set_error("psg_rec_state: null argument"); return PSG_EINVAL; }
  if (!r->d_full) { set_error("psg_rec_state: the history keeps no full states (keep_states = 0)"); return PSG_EINVAL; }
  if (store < 0 || store >= r->o.store_size) { set_error("psg_rec_state: store %d out of range", store); return PSG_EINVAL; }
  REC_HIP(hipStreamSynchronize(r->c->stream));
  REC_HIP(hipMemcpy(host, r->d_full + (long)store * r->nf * r->nkt, sizeof(cpx<double>) * r->nf * r->nkt,
                    hipMemcpyDeviceToHost));
  return 0;
}

int psg_rec_destroy(psg_rec *r) {
  if (!r) return 0;
  (void)hipStreamSynchronize(r->c->stream);
  rec_free(r);
  delete r;
  return 0;
}

}  // extern "C"
