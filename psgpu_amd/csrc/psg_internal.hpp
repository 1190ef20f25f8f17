// Internal structures shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/psg.h"
#include "psg_fft.hpp"

namespace psg {

// Grid description passed by value to every kernel.
struct Grid {
  int dim;
  int nx, ny, nz;
  int ktx, kty, ktz;
  int iktx, ikty, iktz;
  int xp;          // padded kx pitch of the pass buffers (multiple of 8)
  double alpha;
  long nkt;        // stored modes per field
  long zsz;        // per-field size of z-transformed buffers: nz*ikty*xp (2D: 0)
  long ysz;        // per-field size of y-transformed buffers: nz*ny*xp
  double inr;      // 1/NR
  double kmax2;    // (KTY-0.5)^2, only used for documentation/tests
};

enum { PREC64 = 0, PREC32 = 1 };
constexpr int PSG_XPASS_VARIANTS = 12;
constexpr int PSG_ZFWD_VARIANTS = 2;
constexpr int PSG_SLAB_VARIANTS = 6;

struct ProfEntry {
  std::string name;
  double ms = 0.0;
  long bytes = 0;
  int count = 0;
};

// one launch bracketed by events, not yet read back
struct ProfPending {
  int name;
  long bytes;
  int e0, e1;
};

struct Ctx {
  psg_config cfg;
  Grid g;
  int nf;          // state fields per member
  int batch;
  hipStream_t stream = nullptr;
  int device = 0;
  // state and scheme arrays (FP64 complex), member-major [batch][nf][nkt]
  cpx<double> *state = nullptr;
  cpx<double> *rh = nullptr;        // Heun: RH store; RK4: stage arrays
  cpx<double> *stage = nullptr;     // RK4: Z0 stage input / generic scratch
  cpx<double> *tmp = nullptr;       // scratch state-sized buffer
  // 3D Navier-Stokes/Euler: working copy of the state during a run, in the
  // z-run layout LayW (FP64, or FP32 in FP32-bulk), and the CN tables and
  // mask in the same order (nktw modes per field, IKTX rounded up to 8)
  cpx<double> *stw = nullptr;
  cpx<float> *st32 = nullptr;
  bool work_live = false;           // the working copy holds the current state (inside run_steps)
  long nktw = 0;
  double *wnuzn = nullptr, *wnu1 = nullptr;
  int8_t *wmask = nullptr;
  double *nuzn = nullptr, *nu1 = nullptr;  // CN tables (host-built, reference formula)
  double *nu = nullptr;                    // NU = 10^(log10 v2 + 2 log10|k|), for dz/dt
  int8_t *mask = nullptr;           // active mask L
  // pass buffers (precision P)
  void *bufA = nullptr, *bufB = nullptr, *bufC = nullptr;
  size_t szA = 0, szB = 0, szC = 0;
  // twiddle tables: [0]=x, [1]=y, [2]=z lengths; FP64 and FP32 copies
  cpx<double> *tw64[3] = {nullptr, nullptr, nullptr};
  cpx<float> *tw32[3] = {nullptr, nullptr, nullptr};
  double time = 0.0;
  // 2D: per-member time steps (device, batch doubles) used instead of cfg.dt while dtm_on
  double *dtm = nullptr;
  bool dtm_on = false;
  // tangent-linear companion state (allocated on first use): perturbation,
  // its stage input and stage increments, and pass buffers sized for the
  // combined base + perturbation transforms
  cpx<double> *tl_state = nullptr, *tl_stage = nullptr, *tl_rh = nullptr;
  bool tl_fp32 = false;             // tangent-linear passes in FP32-bulk in an FP64 context
  void *tl_bufA = nullptr, *tl_bufB = nullptr, *tl_bufC = nullptr;
  // 2D recorded base trajectory: the physical (w, u, v) of every RK stage of
  // one base flow, [slot = 4 step + stage][field][y][x] in rec_p bytes per
  // value (8, or 4 for FP32 tangent-linear passes); rec_cnt is the device
  // step counter the passes index it by.  A tangent-linear flow that replays
  // it transforms only the perturbation.
  void *rec = nullptr;
  long rec_slots = 0;
  int rec_p = 0;
  int *rec_cnt = nullptr;
  // captured step graphs: [0] plain stepper, [1] tangent linear, [2] plain
  // stepper recording the base, [3] tangent linear replaying it; rebuilt when
  // graph_gen changes (dt, variants, graph length, record buffer)
  int graph_steps = 16;            // steps per captured graph, 0 = eager launches only
  unsigned graph_gen = 1;
  hipGraphExec_t graph_exec[4] = {nullptr, nullptr, nullptr, nullptr};
  unsigned graph_built[4] = {0, 0, 0, 0};
  bool graph_failed = false;       // a capture failed once: eager from then on
  // kernel shape variants (psg_autotune / psg_set_variant)
  int xvar = 0;    // Navier-Stokes x-pass: 0..PSG_XPASS_VARIANTS-1
  int zvar = 0;    // Navier-Stokes z-forward/update: 0..PSG_ZFWD_VARIANTS-1
  int slvar = 0;   // Navier-Stokes passes 2-4 by z slabs: 0 whole grid, v >= 1: 128 >> (v-1) planes
  // diagnostics scratch
  double *dred = nullptr;
  size_t dred_n = 0;
  // device side of the host-array entry points (nonlin, timederiv, physical)
  void *io = nullptr;
  size_t io_n = 0;
  // profiling
  bool prof = false;
  std::vector<ProfEntry> profs;
  std::vector<hipEvent_t> evs;
  int ev_used = 0;
  // launches recorded since the last read-back, and the open bracket
  std::vector<ProfPending> prof_pend;
  int prof_cur_name = -1;
  long prof_cur_bytes = 0;
  int prof_cur_e0 = -1;
};

// The shortened last step of the 3D time policy, T - INT(T/dt0) dt0
// (3D_NavierStokes/src_gmres/main.F90:771-788).  When T/dt0 rounds up to an
// integer it comes out a hair below zero; the reference then takes a step of
// that sub-ulp negative length, here it is skipped.
inline double last_step(double T, double dt0, long nstop) {
  const double l = T - (double)nstop * dt0;
  return l > 0.0 ? l : 0.0;
}

// z planes per slab of passes 2-4 (the whole grid when slabs are off)
inline int slab_planes(const Ctx *c) {
  const int s = c->slvar <= 0 ? c->g.nz : (128 >> (c->slvar - 1));
  return s < c->g.nz ? s : c->g.nz;
}

// error helpers (defined in psg_api.hip)
void set_error(const char *fmt, ...);
#define PSG_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::psg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                    \
      return PSG_EHIP;                                                         \
    }                                                                          \
  } while (0)

// profiling marks around launches (no-ops unless ctx->prof)
void prof_begin(Ctx *c, const char *name, long bytes);
void prof_end(Ctx *c);

// per-flavour drivers
int ns3d_init(Ctx *c);
int ns3d_run(Ctx *c, int nsteps, bool first);
int ns3d_work_begin(Ctx *c);
int ns3d_work_end(Ctx *c);
int ns3d_tables_w(Ctx *c);
int ns3d_nonlin(Ctx *c, int member, const cpx<double> *in_dev, cpx<double> *out_dev);
int ns3d_physical(Ctx *c, int member, double *out_dev);
// physical fields (xi, eta, zeta, u, v, w) of a 3D Navier-Stokes state held anywhere in device memory
int ns3d_physical_state(Ctx *c, const cpx<double> *st, double *out_dev);
int set_member_dt(Ctx *c, const double *host);
int ns2d_init(Ctx *c);
int ns2d_run(Ctx *c, int nsteps, bool first);
int ns2d_nonlin(Ctx *c, int member, const cpx<double> *in_dev, cpx<double> *out_dev);
int ns2d_physical(Ctx *c, int member, double *out_dev);
int diag_member(Ctx *c, int member, double *e, double *z, double *spec);
int bous_init(Ctx *c);
int bous_run(Ctx *c, int nsteps, bool first);
int bous_timederiv(Ctx *c, const cpx<double> *in_dev, cpx<double> *out_dev);
int bous_physical(Ctx *c, int member, double *out_dev);
// physical fields (xi, eta, zeta, u, v, w, rho) of a Boussinesq state held anywhere in device memory
int bous_physical_state(Ctx *c, const cpx<double> *st, double *out_dev);
int timederiv_dev(Ctx *c, const cpx<double> *in, cpx<double> *out);
int ns2d_tl_run(Ctx *c, int nsteps, bool first);
int ns3d_tl_run(Ctx *c, int nsteps, bool first);
int ns2d_rec_run(Ctx *c, int nsteps, bool first);
int ns2d_tlrec_run(Ctx *c, int nsteps, bool first);
// nsteps of the plain (kind 0), tangent-linear (kind 1), base-recording (kind 2)
// or recorded tangent-linear (kind 3) stepper: an eager first step, then
// replays of a captured graph of graph_steps steps
int run_steps(Ctx *c, int kind, int nsteps);
// recorded-base flows (2D, one member): rec_flow = Phi_T(in) recording its
// stages (rp: 8 or 4 bytes per recorded value); tl_replay = dPhi_T dz0 about
// the recorded base, with the same T, dt0 and policy.  rec_usable says
// whether a record of T/dt0 steps fits (allocating it if need be).
bool rec_usable(Ctx *c, double T, double dt0, int policy, int rp);
int rec_flow(Ctx *c, const cpx<double> *in, cpx<double> *out, double T, double dt0, int policy);
int tl_replay(Ctx *c, const cpx<double> *dz0, cpx<double> *dz1, double T, double dt0, int policy);
int ns3d_step_rec(Ctx *c, int umask, cpx<double> *wn, cpx<double> *wstar);
int ns3d_adjoint(Ctx *c, const cpx<double> *z0, const cpx<double> *lamT, int nsteps, int ckpt, cpx<double> *zT,
                 cpx<double> *lam0);
int tl_alloc(Ctx *c);

}  // namespace psg

struct psg_ctx : public psg::Ctx {};
