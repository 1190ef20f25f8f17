// Context management, state I/O, diagnostics and the namespaced C-ABI.
//
// A context owns every device allocation for the life of a run (state, CN
// tables, mask, twiddles, pass buffers), so repeated calls pay no allocation
// or planning cost; the reference allocates, plans and frees on every
// timestep_cuda_ call and inside every transform wrapper
// (3D_NavierStokes/src/init.h:38-101, functions.h:46-47,73-74,93,107).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <mutex>
#include "psg_launch.hpp"
#include "../../include/psg_host.h"

namespace psg {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

static hipEvent_t get_event(Ctx *c) {
  if (c->ev_used >= (int)c->evs.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->evs.push_back(e);
  }
  return c->evs[c->ev_used++];
}

void prof_begin(Ctx *c, const char *name, long bytes) {
  if (!c->prof) return;
  int idx = -1;
  for (size_t i = 0; i < c->profs.size(); i++)
    if (c->profs[i].name == name) idx = (int)i;
  if (idx < 0) {
    ProfEntry e;
    e.name = name;
    c->profs.push_back(e);
    idx = (int)c->profs.size() - 1;
  }
  hipEvent_t e0 = get_event(c);
  if (!e0) return;
  hipEventRecord(e0, c->stream);
  c->prof_cur_name = idx;
  c->prof_cur_bytes = bytes;
  c->prof_cur_e0 = c->ev_used - 1;
}

void prof_end(Ctx *c) {
  if (!c->prof) return;
  if (c->prof_cur_name < 0) return;
  hipEvent_t e1 = get_event(c);
  if (!e1) return;
  hipEventRecord(e1, c->stream);
  c->prof_pend.push_back(ProfPending{c->prof_cur_name, c->prof_cur_bytes, c->prof_cur_e0, c->ev_used - 1});
  c->prof_cur_name = -1;
}

static void prof_collect(Ctx *c) {
  if (c->prof_pend.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto &p : c->prof_pend) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->evs[p.e0], c->evs[p.e1]) == hipSuccess) {
      ProfEntry &e = c->profs[p.name];
      e.ms += ms;
      e.bytes += p.bytes;
      e.count += 1;
    }
  }
  c->prof_pend.clear();
  c->ev_used = 0;
}

// ------------------------------------------------------------------ diag
__global__ void k_diag(const cpx<double> *__restrict__ st, long fs, int nf, const int8_t *__restrict__ L, Grid g,
                       int nbins, double *__restrict__ part) {
  PSG_DYN_SMEM(sh_raw);  // [2*blockDim] + nbins doubles
  double *sh = reinterpret_cast<double *>(sh_raw);
  double *bins = sh + 2 * blockDim.x;
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) bins[i] = 0.0;
  __syncthreads();
  // per-thread sums compensated (Neumaier), so that the total over up to 2e7
  // modes carries tree-reduction rounding only
  double e = 0.0, z = 0.0, ec = 0.0, zc = 0.0;
  auto kadd = [](double &s, double &c, double x) {
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  };
  for (long ik = blockIdx.x * (long)blockDim.x + threadIdx.x; ik < g.nkt; ik += (long)gridDim.x * blockDim.x) {
    if (L[ik] != 1) continue;
    const int ikx = (int)(ik % g.iktx);
    const long t = ik / g.iktx;
    const int iky = (int)(t % g.ikty);
    const int ikz = (int)(t / g.ikty);
    const double kx = ikx * g.alpha, ky = iky - g.kty, kz = g.dim == 3 ? (double)(ikz - g.ktz) : 0.0;
    const double kk = kx * kx + ky * ky + kz * kz;
    const double wk = kk > 0.001 ? kk : 0.001;
    double vz = 0.0;
    for (int f = 0; f < nf; f++) {
      cpx<double> a = st[f * fs + ik];
      vz += a.x * a.x + a.y * a.y;
    }
    kadd(z, zc, vz);
    kadd(e, ec, vz / wk);
    int j = (int)(sqrt(kk) + 0.5);
    if (j < nbins) atomicAdd(&bins[j], vz / wk);
  }
  sh[threadIdx.x] = e + ec;
  sh[blockDim.x + threadIdx.x] = z + zc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[threadIdx.x] += sh[threadIdx.x + s];
      sh[blockDim.x + threadIdx.x] += sh[blockDim.x + threadIdx.x + s];
    }
    __syncthreads();
  }
  double *o = part + (long)blockIdx.x * (2 + nbins);
  if (threadIdx.x == 0) { o[0] = sh[0]; o[1] = sh[blockDim.x]; }
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) o[2 + i] = bins[i];
}

// blocks of k_diag and its launch: partial rows of 2 + KTY + 2 doubles into part
static int diag_blocks(const Ctx *c) { return (int)std::min<long>(256, (c->g.nkt + 255) / 256); }
static int launch_diag(Ctx *c, int member, double *part) {
  const Grid &g = c->g;
  const int nbins = g.kty + 2;
  const int nfv = c->cfg.flavour == PSG_BOUS3D ? 3 : (c->g.dim == 3 ? 3 : 1);
  size_t sh = (2 * 256 + nbins) * sizeof(double);
  hipLaunchKernelGGL(k_diag, dim3(diag_blocks(c)), dim3(256), sh, c->stream,
                     c->state + (long)member * c->nf * g.nkt, g.nkt, nfv, c->mask, g, nbins, part);
  PSG_HIP(hipGetLastError());
  return 0;
}

int diag_member(Ctx *c, int member, double *e, double *z, double *spec) {
  const Grid &g = c->g;
  const int nbins = g.kty + 2;
  const int blocks = diag_blocks(c);
  const size_t need = (size_t)blocks * (2 + nbins);
  if (c->dred_n < need) {
    if (c->dred) hipFree(c->dred);
    PSG_HIP(hipMalloc(&c->dred, need * sizeof(double)));
    c->dred_n = need;
  }
  if (int r = launch_diag(c, member, c->dred)) return r;
  std::vector<double> h(need);
  PSG_HIP(hipMemcpyAsync(h.data(), c->dred, need * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipStreamSynchronize(c->stream));
  double E = 0, Z = 0, Ec = 0, Zc = 0;
  auto kadd = [](double &s, double &c, double x) {
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  };
  std::vector<double> sp(nbins, 0.0), spc(nbins, 0.0);
  for (int b = 0; b < blocks; b++) {
    kadd(E, Ec, h[(size_t)b * (2 + nbins)]);
    kadd(Z, Zc, h[(size_t)b * (2 + nbins) + 1]);
    for (int i = 0; i < nbins; i++) kadd(sp[i], spc[i], h[(size_t)b * (2 + nbins) + 2 + i]);
  }
  E += Ec;
  Z += Zc;
  for (int i = 0; i < nbins; i++) sp[i] += spc[i];
  if (g.dim == 2) Z *= 2.0;
  if (e) *e = E;
  if (z) *z = Z;
  if (spec) memcpy(spec, sp.data(), sizeof(double) * nbins);
  return 0;
}

// ------------------------------------------------------------------ diag without allocation
// psg_diag_noalloc keeps its partial sums in this fixed device array instead of
// the context's lazily allocated reduction buffer: one row of 2 + bins
// partials per k_diag block, then one row of totals.  Each device holds its
// own copy at its own address; host code reaches it only through the address
// hipGetSymbolAddress returns for that device (na_scratch).  Every row a call
// reads is written earlier in the same call, so nothing needs clearing; calls
// that overlap on one device share the array and race.
constexpr int NA_BLOCKS = 256;
constexpr int NA_MAX_BINS = 344;   // KTY + 2 of the largest supported grid (1024 points) is 343
constexpr int NA_ROW = 2 + NA_MAX_BINS;
static __device__ __align__(256) double scratch[(NA_BLOCKS + 1) * NA_ROW];
constexpr int NA_MAX_DEV = 64;
static std::mutex g_na_mu;
static double *g_na_ptr[NA_MAX_DEV];

// the calling thread's current device must be dev
static int na_scratch(int dev, double **p) {
  if (dev < 0 || dev >= NA_MAX_DEV) { set_error("psg_diag_noalloc: device %d out of range", dev); return PSG_EINVAL; }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("psg_diag_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
      return PSG_EHIP;
    }
    g_na_ptr[dev] = (double *)ptr;
  }
  *p = g_na_ptr[dev];
  return 0;
}

// totals of the partial rows, one thread per column, in the block order and
// compensated arithmetic of diag_member's host sum (so both give the same bits)
__global__ void k_diag_total(double *__restrict__ part, int blocks, int ncol, double zscale) {
  for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
    double s = 0.0, c = 0.0;
    for (int b = 0; b < blocks; b++) {
      const double x = part[(long)b * ncol + col];
      const double t = s + x;
      c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
      s = t;
    }
    s += c;
    part[(long)blocks * ncol + col] = col == 1 ? s * zscale : s;
  }
}

static int upload_tables(Ctx *c) {
  psgh_grid hg;
  const Grid &g = c->g;
  if (g.dim == 3) psgh_grid3(&hg, g.nx, g.ny, g.nz, g.alpha);
  else psgh_grid2(&hg, g.nx, g.ny, g.alpha);
  std::vector<int> L(g.nkt);
  psgh_tables(&hg, nullptr, nullptr, nullptr, L.data());
  std::vector<int8_t> L8(g.nkt);
  for (long i = 0; i < g.nkt; i++) L8[i] = (int8_t)L[i];
  PSG_HIP(hipMemcpy(c->mask, L8.data(), g.nkt, hipMemcpyHostToDevice));
  std::vector<double> nu(g.nkt), nuzn(g.nkt), nu1(g.nkt);
  psgh_nutab(&hg, c->cfg.v2 > 0.0 ? c->cfg.v2 : 1e-300, c->cfg.dt, nu.data(), nuzn.data(), nu1.data());
  if (c->cfg.v2 <= 0.0) std::fill(nu.begin(), nu.end(), 0.0);
  PSG_HIP(hipMemcpy(c->nu, nu.data(), sizeof(double) * g.nkt, hipMemcpyHostToDevice));
  if (g.dim == 3) {
    PSG_HIP(hipMemcpy(c->nuzn, nuzn.data(), sizeof(double) * g.nkt, hipMemcpyHostToDevice));
    PSG_HIP(hipMemcpy(c->nu1, nu1.data(), sizeof(double) * g.nkt, hipMemcpyHostToDevice));
  }
  return ns3d_tables_w(c);
}

static int upload_twiddles(Ctx *c) {
  const int lens[3] = {c->g.nx, c->g.ny, c->g.dim == 3 ? c->g.nz : 16};
  for (int a = 0; a < 3; a++) {
    const int n = lens[a];
    std::vector<double> w(2 * n);
    psgh_twiddles(n, w.data());
    std::vector<float> wf(2 * n);
    for (int i = 0; i < 2 * n; i++) wf[i] = (float)w[i];
    PSG_HIP(hipMalloc(&c->tw64[a], sizeof(cpx<double>) * n));
    PSG_HIP(hipMalloc(&c->tw32[a], sizeof(cpx<float>) * n));
    PSG_HIP(hipMemcpy(c->tw64[a], w.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
    PSG_HIP(hipMemcpy(c->tw32[a], wf.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
  }
  return 0;
}

static void free_ctx(Ctx *c) {
  if (!c) return;
  // no launch of a graph may be in flight when its executable is destroyed
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->graph_exec)
    if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
  void *ptrs[] = {c->state, c->rh, c->stage, c->tmp, c->st32, c->stw, c->wnuzn, c->wnu1, c->wmask, c->nuzn, c->nu1, c->nu, c->mask, c->tl_state, c->tl_stage,
                  c->tl_rh, c->tl_bufA, c->tl_bufB, c->tl_bufC, c->bufA, c->bufB, c->bufC,
                  c->tw64[0], c->tw64[1], c->tw64[2], c->tw32[0], c->tw32[1], c->tw32[2], c->dred, c->io, c->dtm,
                  c->rec, c->rec_cnt};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (auto e : c->evs) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
}

// out = L (-NU z + out): the viscous part of TIMEDERIV (3D_NavierStokes/src_gmres/timederiv.F90:31-42)
__global__ void k_add_visc(const cpx<double> *__restrict__ z, cpx<double> *__restrict__ out,
                           const double *__restrict__ nu, const int8_t *__restrict__ L, long nkt, int nf) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nkt * nf; i += (long)gridDim.x * blockDim.x) {
    const long ik = i % nkt;
    const cpx<double> a = z[i], b = out[i];
    out[i] = L[ik] ? cpx<double>{-nu[ik] * a.x + b.x, -nu[ik] * a.y + b.y} : cpx<double>{0.0, 0.0};
  }
}

int timederiv_dev(Ctx *c, const cpx<double> *in, cpx<double> *out) {
  if (c->cfg.flavour == PSG_BOUS3D) return bous_timederiv(c, in, out);
  int r = c->g.dim == 2 ? ns2d_nonlin(c, 0, in, out) : ns3d_nonlin(c, 0, in, out);
  if (r) return r;
  const long tot = c->nf * c->g.nkt;
  hipLaunchKernelGGL(k_add_visc, dim3((unsigned)std::min<long>(4096, (tot + 255) / 256)), dim3(256), 0, c->stream,
                     in, out, c->nu, c->mask, c->g.nkt, c->nf);
  PSG_HIP(hipGetLastError());
  return 0;
}

static int run_eager(Ctx *c, int kind, int n, bool first) {
  if (n <= 0) return 0;
  if (kind == 1) return c->g.dim == 2 ? ns2d_tl_run(c, n, first) : ns3d_tl_run(c, n, first);
  if (kind == 2) return ns2d_rec_run(c, n, first);
  if (kind == 3) return ns2d_tlrec_run(c, n, first);
  if (c->g.dim == 2) return ns2d_run(c, n, first);
  return c->cfg.flavour == PSG_BOUS3D ? bous_run(c, n, first) : ns3d_run(c, n, first);
}

// Capture graph_steps steps (all launches, no memsets or host work) into an
// executable graph for `kind`; a failed capture switches the context to
// eager launches for good.
static int build_graph(Ctx *c, int kind) {
  if (c->graph_exec[kind]) {
    PSG_HIP(hipStreamSynchronize(c->stream));   // the old executable may still be running
    (void)hipGraphExecDestroy(c->graph_exec[kind]);
    c->graph_exec[kind] = nullptr;
  }
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
  const int r = run_eager(c, kind, c->graph_steps, false);
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (r || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return -1;
  }
  const hipError_t ei = hipGraphInstantiate(&c->graph_exec[kind], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    c->graph_exec[kind] = nullptr;
    (void)hipGetLastError();
    return -1;
  }
  c->graph_built[kind] = c->graph_gen;
  return 0;
}

static int run_steps_in(Ctx *c, int kind, int nsteps);

// the working copy of the 3D Navier-Stokes/Euler state (layout LayW; FP32 in
// FP32-bulk) is converted once around all of a call's steps, outside the
// captured graphs
int run_steps(Ctx *c, int kind, int nsteps) {
  if (nsteps <= 0) return 0;
  const bool work = kind == 0 && (c->st32 || c->stw);
  if (work)
    if (int r0 = ns3d_work_begin(c)) return r0;
  const int r = run_steps_in(c, kind, nsteps);
  if (!work) return r;
  const int r2 = ns3d_work_end(c);
  return r ? r : r2;
}

static int run_steps_in(Ctx *c, int kind, int nsteps) {
  const int G = c->graph_steps;
  if (G <= 0 || c->prof || c->graph_failed || nsteps < G + 1) return run_eager(c, kind, nsteps, true);
  // the first step of a call is eager: it may mask the velocity / copy the stage input
  if (int r = run_eager(c, kind, 1, true)) return r;
  int left = nsteps - 1;
  if (c->graph_built[kind] != c->graph_gen || !c->graph_exec[kind]) {
    if (build_graph(c, kind)) {
      c->graph_failed = true;
      return run_eager(c, kind, left, false);
    }
  }
  for (; left >= G; left -= G) PSG_HIP(hipGraphLaunch(c->graph_exec[kind], c->stream));
  return run_eager(c, kind, left, false);
}

void tl_free(Ctx *c) {
  for (void **p : {(void **)&c->tl_state, (void **)&c->tl_stage, (void **)&c->tl_rh, &c->tl_bufA, &c->tl_bufB,
                   &c->tl_bufC}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

// All-or-nothing: on any failure every tangent-linear buffer is released and
// the pointers are null again, so a later call retries instead of running
// kernels on a partial set.
int tl_alloc(Ctx *c) {
  if (c->tl_state && c->tl_stage && c->tl_rh && c->tl_bufB && c->tl_bufC && (c->g.dim == 2 || c->tl_bufA))
    return 0;
  tl_free(c);
  const Grid &g = c->g;
  const size_t sb = sizeof(cpx<double>) * c->nf * g.nkt * c->batch;
  const size_t b = c->cfg.precision == PSG_FP32BULK ? sizeof(cpx<float>) : sizeof(cpx<double>);
  bool ok = hipMalloc(&c->tl_state, sb) == hipSuccess && hipMalloc(&c->tl_stage, sb) == hipSuccess &&
            hipMalloc(&c->tl_rh, 3 * sb) == hipSuccess;
  if (ok && g.dim == 2)
    ok = hipMalloc(&c->tl_bufB, 6 * g.ysz * c->batch * b) == hipSuccess &&
         hipMalloc(&c->tl_bufC, 4 * g.ysz * c->batch * b) == hipSuccess;
  else if (ok)
    ok = hipMalloc(&c->tl_bufA, 12 * g.zsz * c->batch * b) == hipSuccess &&
         hipMalloc(&c->tl_bufB, 12 * g.ysz * c->batch * b) == hipSuccess &&
         hipMalloc(&c->tl_bufC, 6 * g.ysz * c->batch * b) == hipSuccess;
  ok = ok && hipMemsetAsync(c->tl_state, 0, sb, c->stream) == hipSuccess &&
       hipMemsetAsync(c->tl_stage, 0, sb, c->stream) == hipSuccess &&
       hipMemsetAsync(c->tl_rh, 0, 3 * sb, c->stream) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    tl_free(c);
    set_error("tangent-linear buffers: device allocation failed");
    return PSG_ENOMEM;
  }
  return 0;
}

// z1/dz1 = Phi_T(z0) and its derivative applied to dz0, with the Newton time policy
int tl_flow_map(Ctx *c, const cpx<double> *z0, const cpx<double> *dz0, cpx<double> *z1, cpx<double> *dz1, double T,
                double dt0, int policy, bool) {
  if (int r = tl_alloc(c)) return r;
  if (c->g.dim == 3 && c->cfg.flavour != PSG_NS3D) { set_error("tangent linear: 2D and 3D Navier-Stokes only"); return PSG_EINVAL; }
  const size_t bytes = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->state, z0, bytes, hipMemcpyDeviceToDevice, c->stream));
  PSG_HIP(hipMemcpyAsync(c->tl_state, dz0, bytes, hipMemcpyDeviceToDevice, c->stream));
  auto run = [&](int n) { return run_steps(c, 1, n); };
  int nstop = (int)(T / dt0), r = 0;
  if (policy == 0) {
    if (nstop < 1) nstop = 1;
    const double dt = T / (double)nstop;
    if (c->cfg.dt != dt && (r = psg_set_dt(static_cast<psg_ctx *>(c), dt))) return r;
    if ((r = run(nstop))) return r;
  } else {
    if (nstop > 0) {
      if (c->cfg.dt != dt0 && (r = psg_set_dt(static_cast<psg_ctx *>(c), dt0))) return r;
      if ((r = run(nstop))) return r;
    }
    const double last = last_step(T, dt0, nstop);
    if (last != 0.0) {
      if ((r = psg_set_dt(static_cast<psg_ctx *>(c), last))) return r;
      if ((r = run(1))) return r;
    }
  }
  if (z1) PSG_HIP(hipMemcpyAsync(z1, c->state, bytes, hipMemcpyDeviceToDevice, c->stream));
  PSG_HIP(hipMemcpyAsync(dz1, c->tl_state, bytes, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// ------------------------------------------------------------------ recorded base
// steps of a flow over T under the Newton time policy
static long flow_steps(double T, double dt0, int policy) {
  long n = (long)(T / dt0);
  if (policy == 0) return n < 1 ? 1 : n;
  return n + (last_step(T, dt0, n) != 0.0 ? 1 : 0);
}

bool rec_usable(Ctx *c, double T, double dt0, int policy, int rp) {
  const Grid &g = c->g;
  if (g.dim != 2 || c->batch != 1 || g.ny >= 2048 || c->dtm_on || (rp != 4 && rp != 8)) return false;
  if (c->cfg.precision == PSG_FP32BULK && rp != 4) return false;
  if (!(T > 0.0) || !(dt0 > 0.0)) return false;
  const long slots = 4 * flow_steps(T, dt0, policy);
  if (c->rec && c->rec_p == rp && c->rec_slots >= slots) return true;
  if (!c->rec_cnt && hipMalloc(&c->rec_cnt, sizeof(int)) != hipSuccess) {
    c->rec_cnt = nullptr;
    (void)hipGetLastError();
    return false;
  }
  // room for a quarter more steps, so that small period updates keep the buffer
  const long want = slots + slots / 4;
  const size_t per = (size_t)3 * g.ny * g.nx * rp;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
  if (c->rec) (void)hipFree(c->rec);
  c->rec = nullptr;
  c->rec_slots = 0;
  c->graph_gen++;   // captured launches hold the record's address
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) { (void)hipGetLastError(); return false; }
  // leave a tenth of the device (and at least 2 GiB) to everything else
  const size_t reserve = std::max<size_t>(tot / 10, (size_t)2 << 30);
  if (fr < reserve || (size_t)want * per > fr - reserve) return false;
  if (hipMalloc(&c->rec, (size_t)want * per) != hipSuccess) {
    c->rec = nullptr;
    (void)hipGetLastError();
    return false;
  }
  c->rec_slots = want;
  c->rec_p = rp;
  return true;
}

// c->state = in, the time-policy steps of `kind`, out = c->state (tl: on tl_state)
static int policy_run(Ctx *c, int kind, double T, double dt0, int policy) {
  int nstop = (int)(T / dt0), r = 0;
  if (policy == 0) {
    if (nstop < 1) nstop = 1;
    const double dt = T / (double)nstop;
    if (c->cfg.dt != dt && (r = psg_set_dt(static_cast<psg_ctx *>(c), dt))) return r;
    return run_steps(c, kind, nstop);
  }
  if (nstop > 0) {
    if (c->cfg.dt != dt0 && (r = psg_set_dt(static_cast<psg_ctx *>(c), dt0))) return r;
    if ((r = run_steps(c, kind, nstop))) return r;
  }
  const double last = last_step(T, dt0, nstop);
  if (last != 0.0) {
    if ((r = psg_set_dt(static_cast<psg_ctx *>(c), last))) return r;
    if ((r = run_steps(c, kind, 1))) return r;
  }
  return 0;
}

int ns2d_rec_reset(Ctx *c);

int rec_flow(Ctx *c, const cpx<double> *in, cpx<double> *out, double T, double dt0, int policy) {
  if (!c->rec || (long)4 * flow_steps(T, dt0, policy) > c->rec_slots) {
    set_error("recorded base: the record holds fewer steps than the flow");
    return PSG_EINVAL;
  }
  const size_t bytes = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->state, in, bytes, hipMemcpyDeviceToDevice, c->stream));
  if (int r = ns2d_rec_reset(c)) return r;
  if (int r = policy_run(c, 2, T, dt0, policy)) return r;
  PSG_HIP(hipMemcpyAsync(out, c->state, bytes, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

int tl_replay(Ctx *c, const cpx<double> *dz0, cpx<double> *dz1, double T, double dt0, int policy) {
  if (int r = tl_alloc(c)) return r;
  if (!c->rec || (long)4 * flow_steps(T, dt0, policy) > c->rec_slots) {
    set_error("recorded tangent linear: the record holds fewer steps than the flow");
    return PSG_EINVAL;
  }
  const size_t bytes = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->tl_state, dz0, bytes, hipMemcpyDeviceToDevice, c->stream));
  if (int r = ns2d_rec_reset(c)) return r;
  if (int r = policy_run(c, 3, T, dt0, policy)) return r;
  PSG_HIP(hipMemcpyAsync(dz1, c->tl_state, bytes, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// per-member steps of a 2D context (host array of batch doubles, or NULL to go back
// to cfg.dt); used by the batched Newton flows.  Synchronous: the array is read by
// every later step, captured graphs included.
int set_member_dt(Ctx *c, const double *host) {
  if (c->g.dim != 2) { set_error("per-member steps: 2D contexts only"); return PSG_EINVAL; }
  PSG_HIP(hipStreamSynchronize(c->stream));
  if (!host) {
    if (c->dtm_on) c->graph_gen++;
    c->dtm_on = false;
    return 0;
  }
  if (!c->dtm) PSG_HIP(hipMalloc(&c->dtm, sizeof(double) * c->batch));
  PSG_HIP(hipMemcpy(c->dtm, host, sizeof(double) * c->batch, hipMemcpyHostToDevice));
  if (!c->dtm_on) c->graph_gen++;
  c->dtm_on = true;
  return 0;
}
}  // namespace psg

using namespace psg;

extern "C" {

const char *psg_last_error(void) { return psg::g_err; }

int psg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int psg_sizes(const psg_config *cfg, long *iktx, long *ikty, long *iktz, long *nkt, long *nactive) {
  if (!cfg) return PSG_EINVAL;
  psgh_grid hg;
  int r = cfg->flavour == PSG_NS2D ? psgh_grid2(&hg, cfg->nx, cfg->ny, cfg->alpha)
                                   : psgh_grid3(&hg, cfg->nx, cfg->ny, cfg->nz, cfg->alpha);
  if (r) { set_error("grid too small"); return PSG_EINVAL; }
  if (iktx) *iktx = hg.iktx;
  if (ikty) *ikty = hg.ikty;
  if (iktz) *iktz = hg.iktz;
  if (nkt) *nkt = hg.nkt;
  if (nactive) *nactive = psgh_count_active(&hg);
  return 0;
}

int psg_create(const psg_config *cfg, psg_ctx **out) {
  if (!cfg || !out) { set_error("null argument"); return PSG_EINVAL; }
  *out = nullptr;
  const bool is2d = cfg->flavour == PSG_NS2D;
  if (cfg->flavour != PSG_NS2D && cfg->flavour != PSG_NS3D && cfg->flavour != PSG_EULER3D &&
      cfg->flavour != PSG_BOUS3D) {
    set_error("flavour %d not available in this build", cfg->flavour);
    return PSG_EINVAL;
  }
  if (!supported_len(cfg->nx) || !supported_len(cfg->ny) || (!is2d && !supported_len(cfg->nz))) {
    set_error("grid %dx%dx%d: each extent must be a power of two in [16,1024]", cfg->nx, cfg->ny, cfg->nz);
    return PSG_EINVAL;
  }
  if (cfg->batch < 1 || (cfg->precision != PSG_FP64 && cfg->precision != PSG_FP32BULK) || !(cfg->dt > 0.0) ||
      !(cfg->alpha > 0.0)) {
    set_error("invalid batch/precision/dt/alpha");
    return PSG_EINVAL;
  }
  if (psg_device_count() < 1) { set_error("no HIP device visible"); return PSG_ENODEV; }
  psg_ctx *c = new psg_ctx();
  c->cfg = *cfg;
  c->batch = cfg->batch;
  c->nf = is2d ? 1 : (cfg->flavour == PSG_BOUS3D ? 4 : 3);
  psgh_grid hg;
  if (is2d) psgh_grid2(&hg, cfg->nx, cfg->ny, cfg->alpha);
  else psgh_grid3(&hg, cfg->nx, cfg->ny, cfg->nz, cfg->alpha);
  Grid &g = c->g;
  g.dim = hg.dim; g.nx = hg.nx; g.ny = hg.ny; g.nz = hg.nz;
  g.ktx = hg.ktx; g.kty = hg.kty; g.ktz = hg.ktz;
  g.iktx = hg.iktx; g.ikty = hg.ikty; g.iktz = hg.iktz;
  g.alpha = hg.alpha; g.nkt = hg.nkt;
  g.xp = (g.iktx + 15) / 16 * 16;
  g.zsz = is2d ? 0 : (long)g.nz * g.ikty * g.xp;
  g.ysz = (long)g.nz * g.ny * g.xp;
  g.inr = 1.0 / (double)hg.nr;
  g.kmax2 = ((double)g.kty - 0.5) * ((double)g.kty - 0.5);
  hipGetDevice(&c->device);
  int rc = 0;
  do {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { rc = PSG_EHIP; set_error("stream"); break; }
    const size_t ssz = sizeof(cpx<double>) * c->nf * g.nkt * c->batch;
    if (hipMalloc(&c->state, ssz) != hipSuccess || hipMalloc(&c->tmp, sizeof(cpx<double>) * c->nf * g.nkt) != hipSuccess ||
        hipMalloc(&c->mask, g.nkt) != hipSuccess) { rc = PSG_ENOMEM; set_error("state allocation failed"); break; }
    hipMemset(c->state, 0, ssz);
    if (hipMalloc(&c->nu, sizeof(double) * g.nkt) != hipSuccess) { rc = PSG_ENOMEM; set_error("tables"); break; }
    if (!is2d) {
      if (hipMalloc(&c->nuzn, sizeof(double) * g.nkt) != hipSuccess ||
          hipMalloc(&c->nu1, sizeof(double) * g.nkt) != hipSuccess) { rc = PSG_ENOMEM; set_error("tables"); break; }
    }
    if ((rc = upload_tables(c))) break;
    if ((rc = upload_twiddles(c))) break;
    rc = is2d ? ns2d_init(c) : (cfg->flavour == PSG_BOUS3D ? bous_init(c) : ns3d_init(c));
    if (rc) break;
    if (hipStreamSynchronize(c->stream) != hipSuccess) { rc = PSG_EHIP; set_error("init sync"); break; }
  } while (0);
  if (rc) { free_ctx(c); delete c; return rc; }
  *out = c;
  return 0;
}

int psg_destroy(psg_ctx *c) {
  if (!c) return 0;
  if (c->stream) hipStreamSynchronize(c->stream);
  free_ctx(c);
  delete c;
  return 0;
}

static int check_member(psg_ctx *c, int m) {
  if (!c) { set_error("null context"); return PSG_EINVAL; }
  if (m < 0 || m >= c->batch) { set_error("member %d out of range [0,%d)", m, c->batch); return PSG_EINVAL; }
  return 0;
}

int psg_set_state(psg_ctx *c, int m, const double *host) {
  if (int r = check_member(c, m)) return r;
  const size_t n = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->state + (long)m * c->nf * c->g.nkt, host, n, hipMemcpyHostToDevice, c->stream));
  PSG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_get_state(psg_ctx *c, int m, double *host) {
  if (int r = check_member(c, m)) return r;
  const size_t n = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(host, c->state + (long)m * c->nf * c->g.nkt, n, hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_state_ptr(psg_ctx *c, int m, void **dev) {
  if (int r = check_member(c, m)) return r;
  *dev = c->state + (long)m * c->nf * c->g.nkt;
  return 0;
}

int psg_set_state_device(psg_ctx *c, int m, const void *dev) {
  if (int r = check_member(c, m)) return r;
  const size_t n = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->state + (long)m * c->nf * c->g.nkt, dev, n, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

int psg_get_state_device(psg_ctx *c, int m, void *dev) {
  if (int r = check_member(c, m)) return r;
  const size_t n = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(dev, c->state + (long)m * c->nf * c->g.nkt, n, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

int psg_run(psg_ctx *c, int nsteps) {
  if (!c || nsteps < 0) { set_error("invalid run arguments"); return PSG_EINVAL; }
  if (nsteps == 0) return 0;
  int r = run_steps(c, 0, nsteps);
  if (!r) c->time += nsteps * c->cfg.dt;
  if (c->prof) prof_collect(c);
  return r;
}

int psg_set_dt(psg_ctx *c, double dt) {
  if (!c || !(dt > 0.0)) { set_error("invalid dt"); return PSG_EINVAL; }
  PSG_HIP(hipStreamSynchronize(c->stream));
  c->cfg.dt = dt;
  c->graph_gen++;   // captured launches hold dt by value
  return upload_tables(c);
}

int psg_set_v2(psg_ctx *c, double v2) {
  if (!c || !(v2 > 0.0)) { set_error("invalid v2"); return PSG_EINVAL; }
  PSG_HIP(hipStreamSynchronize(c->stream));
  c->cfg.v2 = v2;
  c->graph_gen++;   // the 2D update holds v2 by value
  return upload_tables(c);
}

int psg_set_tables(psg_ctx *c, const double *nuzn, const double *nu1, const int *L) {
  if (!c) { set_error("null context"); return PSG_EINVAL; }
  const long n = c->g.nkt;
  if ((nuzn || nu1) && c->g.dim != 3) { set_error("CN tables are 3D-only"); return PSG_EINVAL; }
  if (nuzn) PSG_HIP(hipMemcpyAsync(c->nuzn, nuzn, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (nu1) PSG_HIP(hipMemcpyAsync(c->nu1, nu1, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (L) {
    std::vector<int8_t> l8(n);
    for (long i = 0; i < n; i++) l8[i] = (int8_t)(L[i] != 0);
    PSG_HIP(hipMemcpyAsync(c->mask, l8.data(), n, hipMemcpyHostToDevice, c->stream));
    PSG_HIP(hipStreamSynchronize(c->stream));
  }
  return ns3d_tables_w(c);
}

int psg_sync(psg_ctx *c) {
  if (!c) return PSG_EINVAL;
  PSG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int psg_diag(psg_ctx *c, int m, double *e, double *z, double *spec) {
  if (int r = check_member(c, m)) return r;
  return diag_member(c, m, e, z, spec);
}

int psg_diag_noalloc(psg_ctx *c, int m, double *e, double *z, double *spec) {
  if (int r = check_member(c, m)) return r;
  const int nbins = c->g.kty + 2, ncol = 2 + nbins, blocks = diag_blocks(c);
  if (nbins > NA_MAX_BINS || blocks > NA_BLOCKS) {
    set_error("psg_diag_noalloc: %d spectrum bins exceed the fixed scratch of %d", nbins, NA_MAX_BINS);
    return PSG_EINVAL;
  }
  int cur = -1;
  PSG_HIP(hipGetDevice(&cur));
  if (cur != c->device) {
    set_error("psg_diag_noalloc: the context's device %d is not current (%d)", c->device, cur);
    return PSG_EINVAL;
  }
  double *part = nullptr;
  if (int r = na_scratch(cur, &part)) return r;
  if (int r = launch_diag(c, m, part)) return r;
  hipLaunchKernelGGL(k_diag_total, dim3(1), dim3(256), 0, c->stream, part, blocks, ncol, c->g.dim == 2 ? 2.0 : 1.0);
  PSG_HIP(hipGetLastError());
  const double *tot = part + (long)blocks * ncol;
  if (e) PSG_HIP(hipMemcpyAsync(e, tot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (z) PSG_HIP(hipMemcpyAsync(z, tot + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (spec) PSG_HIP(hipMemcpyAsync(spec, tot + 2, sizeof(double) * nbins, hipMemcpyDeviceToHost, c->stream));
  PSG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// Device buffer for the host-array entry points (psg_nonlin, psg_timederiv,
// psg_physical), grown on demand and kept with the context, so repeated
// calls (a legacy timederiv_cuda_ loop, DNS dumps) allocate nothing.
static int io_buffer(Ctx *c, size_t bytes, void **p) {
  if (c->io_n < bytes) {
    PSG_HIP(hipStreamSynchronize(c->stream));
    if (c->io) (void)hipFree(c->io);
    c->io = nullptr;
    c->io_n = 0;
    PSG_HIP(hipMalloc(&c->io, bytes));
    c->io_n = bytes;
  }
  *p = c->io;
  return 0;
}

int psg_nonlin(psg_ctx *c, int m, const double *host_in, double *host_out) {
  if (int r = check_member(c, m)) return r;
  if (c->cfg.flavour == PSG_BOUS3D) { set_error("psg_nonlin: use psg_timederiv for the Boussinesq flavour"); return PSG_EINVAL; }
  const size_t n = sizeof(cpx<double>) * c->nf * c->g.nkt;
  cpx<double> *dout = nullptr;
  if (int r = io_buffer(c, n, (void **)&dout)) return r;
  PSG_HIP(hipMemcpyAsync(c->tmp, host_in, n, hipMemcpyHostToDevice, c->stream));
  int r = c->g.dim == 2 ? ns2d_nonlin(c, 0, c->tmp, dout) : ns3d_nonlin(c, 0, c->tmp, dout);
  if (!r) {
    if (hipMemcpyAsync(host_out, dout, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) { set_error("nonlin copy"); r = PSG_EHIP; }
  }
  hipStreamSynchronize(c->stream);
  return r;
}

int psg_timederiv(psg_ctx *c, int m, const double *host_in, double *host_out) {
  if (int r = check_member(c, m)) return r;
  const size_t n = sizeof(cpx<double>) * c->nf * c->g.nkt;
  cpx<double> *dout = nullptr;
  if (int r = io_buffer(c, n, (void **)&dout)) return r;
  PSG_HIP(hipMemcpyAsync(c->tmp, host_in, n, hipMemcpyHostToDevice, c->stream));
  int r = timederiv_dev(c, c->tmp, dout);
  if (!r) {
    if (hipMemcpyAsync(host_out, dout, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) { set_error("timederiv copy"); r = PSG_EHIP; }
  }
  hipStreamSynchronize(c->stream);
  return r;
}

int psg_physical(psg_ctx *c, int m, double *host_out) {
  if (int r = check_member(c, m)) return r;
  const bool bous = c->cfg.flavour == PSG_BOUS3D;
  const long nr = (long)c->g.nx * c->g.ny * c->g.nz;
  const int nfo = c->g.dim == 2 ? 3 : (bous ? 7 : 6);
  double *d = nullptr;
  if (int r = io_buffer(c, sizeof(double) * nr * nfo, (void **)&d)) return r;
  int r = c->g.dim == 2 ? ns2d_physical(c, m, d) : (bous ? bous_physical(c, m, d) : ns3d_physical(c, m, d));
  if (!r) {
    if (hipMemcpyAsync(host_out, d, sizeof(double) * nr * nfo, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) { set_error("physical copy"); r = PSG_EHIP; }
  }
  hipStreamSynchronize(c->stream);
  return r;
}

void *psg_stream(psg_ctx *c) { return c ? (void *)c->stream : nullptr; }

int psg_set_variant(psg_ctx *c, int pass, int variant) {
  if (!c) return PSG_EINVAL;
  const int nv = pass == PSG_PASS_XPASS ? PSG_XPASS_VARIANTS
                 : pass == PSG_PASS_ZFWD ? PSG_ZFWD_VARIANTS
                 : pass == PSG_PASS_SLAB ? PSG_SLAB_VARIANTS : 0;
  if (variant < 0 || variant >= nv) {
    set_error("psg_set_variant: unknown pass %d or variant %d", pass, variant);
    return PSG_EINVAL;
  }
  (pass == PSG_PASS_XPASS ? c->xvar : pass == PSG_PASS_ZFWD ? c->zvar : c->slvar) = variant;
  c->graph_gen++;
  return 0;
}

int psg_set_graph(psg_ctx *c, int steps) {
  if (!c || steps < 0) { set_error("psg_set_graph: bad argument"); return PSG_EINVAL; }
  PSG_HIP(hipStreamSynchronize(c->stream));
  c->graph_steps = steps;
  c->graph_gen++;
  return 0;
}

int psg_get_variant(psg_ctx *c, int pass) {
  if (!c || (pass != PSG_PASS_XPASS && pass != PSG_PASS_ZFWD && pass != PSG_PASS_SLAB)) return PSG_EINVAL;
  return pass == PSG_PASS_XPASS ? c->xvar : pass == PSG_PASS_ZFWD ? c->zvar : c->slvar;
}

int psg_autotune(psg_ctx *c, int steps, char *report, int report_len) {
  if (!c) return PSG_EINVAL;
  std::string rep;
  if (c->g.dim != 3) {
    rep = "no shape variants for this flavour";
  } else {
    if (steps < 1) steps = 1;
    const size_t sb = sizeof(cpx<double>) * c->nf * c->g.nkt * c->batch;
    void *save = nullptr;
    PSG_HIP(hipMalloc(&save, sb));
    PSG_HIP(hipMemcpyAsync(save, c->state, sb, hipMemcpyDeviceToDevice, c->stream));
    const double t0 = c->time;
    const bool prof = c->prof;
    c->prof = false;
    hipEvent_t e0, e1;
    PSG_HIP(hipEventCreate(&e0));
    PSG_HIP(hipEventCreate(&e1));
    int rc = 0;
    // passes are tuned one after the other, each with the others' choices fixed
    // (the Boussinesq flavour has x-pass shapes only)
    struct { const char *name; int *var; int n; } passes[] = {{"xpass", &c->xvar, PSG_XPASS_VARIANTS},
                                                                {"zfwd", &c->zvar, PSG_ZFWD_VARIANTS},
                                                                {"slab", &c->slvar, PSG_SLAB_VARIANTS}};
    const int npass = c->cfg.flavour == PSG_BOUS3D ? 1 : 3;
    for (int ip = 0; ip < npass; ip++) {
      auto &ps = passes[ip];
      double best = 1e300;
      int bestv = *ps.var;
      for (int v = 0; v < ps.n && rc == 0; v++) {
        *ps.var = v;
        c->graph_gen++;
        if ((rc = run_steps(c, 0, 1))) break;
        hipEventRecord(e0, c->stream);
        if ((rc = run_steps(c, 0, steps))) break;
        hipEventRecord(e1, c->stream);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / steps;
        char line[96];
        snprintf(line, sizeof line, "%s variant %d: %.4f ms/step\n", ps.name, v, per);
        rep += line;
        if (per < best) { best = per; bestv = v; }
        hipMemcpyAsync(c->state, save, sb, hipMemcpyDeviceToDevice, c->stream);
      }
      *ps.var = bestv;
      char line[64];
      snprintf(line, sizeof line, "chosen %s variant %d\n", ps.name, bestv);
      rep += line;
    }
    hipStreamSynchronize(c->stream);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(save);
    c->time = t0;
    c->prof = prof;
    c->graph_gen++;
    if (rc) return rc;
  }
  if (report && report_len > 0) {
    strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return 0;
}

int psg_profile_enable(psg_ctx *c, int on) {
  if (!c) return PSG_EINVAL;
  hipStreamSynchronize(c->stream);
  c->prof = on != 0;
  // pending brackets index profs and evs: drop them together with the table
  c->prof_pend.clear();
  c->prof_cur_name = -1;
  c->profs.clear();
  c->ev_used = 0;
  return 0;
}

int psg_profile_get(psg_ctx *c, int idx, const char **name, double *avg_ms, long *bytes, int *count) {
  if (!c) return PSG_EINVAL;
  prof_collect(c);
  if (idx < 0 || idx >= (int)c->profs.size()) return PSG_EINVAL;
  const ProfEntry &e = c->profs[idx];
  if (name) *name = e.name.c_str();
  if (avg_ms) *avg_ms = e.count ? e.ms / e.count : 0.0;
  if (bytes) *bytes = e.count ? e.bytes / e.count : 0;
  if (count) *count = e.count;
  return 0;
}

}  // extern "C"
