/*
 * psgpu-mi355x C-ABI: persistent-context pseudospectral steppers on gfx950.
 *
 * Every entry point takes plain pointers and sizes; no torch or HIP types
 * appear in the signatures.  State arrays use the reference storage order:
 * per field NKT complex doubles (re,im interleaved), index
 *   ik = (ikz*IKTY + iky)*IKTX + ikx      (3D; 2D drops ikz)
 * with kx = ikx*alpha, ky = iky-KTY, kz = ikz-KTZ
 * (3D_NavierStokes/src/init.F90:25-36, 2D_NavierStokes/src_RK4/main.F90:146-166).
 * Fields are consecutive: 3D (xii, eta, zet), Boussinesq (xii, eta, zet, rho),
 * 2D (z).  Batched contexts hold `batch` independent members.
 *
 * The reference has no persistent context: each timestep_cuda_ call
 * allocates, plans, copies, steps and frees (3D_NavierStokes/src/cuda.cu:18-151,
 * init.h:38-101).  psg_create/psg_run replace that call sequence; the legacy
 * symbols in psg_legacy.h map the reference argument lists onto them.
 *
 * Errors: every function returns 0 on success or a negative PSG_E* code;
 * psg_last_error() returns a message for the calling thread's last failure.
 *
 * Threads: a context owns its HIP stream, buffers, graphs and profiling
 * events, and must be driven by one host thread at a time (calls on one
 * context from several threads at once are not serialised and race).
 * Different contexts may be driven from different threads; the state shared
 * between contexts (the LDS attribute cache, the per-device scratch of the
 * allocation-free diagnostics) is guarded by mutexes.
 */
#ifndef PSG_H
#define PSG_H
#ifdef __cplusplus
extern "C" {
#endif

#define PSG_OK 0
#define PSG_EINVAL -1   /* bad argument / unsupported size */
#define PSG_EHIP -2     /* HIP runtime failure */
#define PSG_ENOMEM -3
#define PSG_ENODEV -4   /* no GPU */

enum psg_flavour {
  PSG_NS2D = 1,     /* 2D vorticity, RK4 + end-of-step CN  (2D_NavierStokes/src_RK4) */
  PSG_NS3D = 2,     /* 3D vorticity, Heun + CN             (3D_NavierStokes/src, src_gmres) */
  PSG_EULER3D = 3,  /* 3D inviscid, Heun                   (3D_Euler/src) */
  PSG_BOUS3D = 4    /* 3D Boussinesq, RK4 per-stage CN     (3D_Boussinesq/src) */
};
enum psg_precision {
  PSG_FP64 = 0,     /* plain form: FFTs and products in FP64 */
  PSG_FP32BULK = 1  /* FFTs and products in FP32, state/update/diagnostics FP64 */
};
enum psg_forcing {
  PSG_FORCE_NONE = 0,
  PSG_FORCE_DNS = 1,    /* xi.re +-A at (0,KF,+-KF)  3D_NavierStokes/src/kernels.h:342-343 */
  PSG_FORCE_GMRES = 2,  /* eta.im, zet.im at (0,KF,+-KF)  src_gmres/kernels.h:234-238 */
  PSG_FORCE_KOLMO = 3   /* f=(sin KF y,0,0): zet.re(0,KF,0) += A  3D_Boussinesq/src/kernels.h:382-384;
                           2D: NZKX += A at (0,+-KF), A=-2 is the reference's "NZKX -= 2"
                           (2D_NavierStokes/src_RK4/kernels.h:264-266) */
};

typedef struct {
  int flavour;        /* psg_flavour */
  int precision;      /* psg_precision */
  int batch;          /* members stepped together (>=1) */
  int nx, ny, nz;     /* physical grid (nz ignored in 2D); powers of two 16..1024 */
  double alpha;       /* kx = ikx*alpha */
  double v2;          /* viscosity 1/Re (ignored for Euler) */
  double dt;          /* time step */
  int force;          /* psg_forcing */
  double kf;          /* forcing wavenumber */
  double ampfor;      /* forcing amplitude A */
  double ri, sc, theta; /* Boussinesq: Richardson number, Schmidt number, gravity tilt in radians
                           (passed straight to sin/cos as 3D_Boussinesq/src/init.h:66-67 does) */
} psg_config;

typedef struct psg_ctx psg_ctx;

/* Sizes implied by a config (IKTX, IKTY, IKTZ, NKT, active-mode count). */
int psg_sizes(const psg_config *cfg, long *iktx, long *ikty, long *iktz, long *nkt, long *nactive);

int psg_create(const psg_config *cfg, psg_ctx **out);
int psg_destroy(psg_ctx *ctx);
/* copy member state host<->device: nf*nkt complex (re,im) doubles */
int psg_set_state(psg_ctx *ctx, int member, const double *host);
int psg_get_state(psg_ctx *ctx, int member, double *host);
/* device pointer to a member's state (nf*nkt complex doubles), for zero-copy callers */
int psg_state_ptr(psg_ctx *ctx, int member, void **dev);
/* device-to-device state copies (ptr may be any device allocation) */
int psg_set_state_device(psg_ctx *ctx, int member, const void *dev);
int psg_get_state_device(psg_ctx *ctx, int member, void *dev);
/* advance every member by nsteps steps of dt (asynchronous on the context stream) */
int psg_run(psg_ctx *ctx, int nsteps);
/* change dt (rebuilds the CN tables, as src_gmres/main.F90:759-768 does per period) */
int psg_set_dt(psg_ctx *ctx, double dt);
/* change the viscosity v2 = 1/Re (rebuilds the CN tables, as the reference's
 * arc-length continuation does at every change of Re,
 * 2D_NavierStokes/src_RK4/src_gmres/main.F90:588-599) */
int psg_set_v2(psg_ctx *ctx, double v2);
/* replace the CN tables (NUZN, NU1; nkt doubles each, 3D only) and the mask L
 * (nkt ints) with caller-built ones, as the reference receives them from its
 * Fortran driver (3D_NavierStokes/src/cuda.cu:24-25,39).  Any pointer may be NULL. */
int psg_set_tables(psg_ctx *ctx, const double *nuzn, const double *nu1, const int *L);
int psg_sync(psg_ctx *ctx);
/* FP64 diagnostics of one member: energy = sum_L |w|^2/|k|^2, enstrophy = sum_L |w|^2
 * (2D: 2*sum_L|z|^2 as out.F90), spectrum[j] = sum over shell j=INT(|k|+0.5) of |w|^2/|k|^2
 * (spec may be NULL, else length KTY+2).  3D_NavierStokes/src/functions.h:128-161 */
int psg_diag(psg_ctx *ctx, int member, double *energy, double *enstrophy, double *spec);
/* psg_diag without any device allocation: the per-block partial sums live in a
 * fixed device array of the library (at most 257 x 346 doubles, 711 KB, held
 * by each device the library runs on until the process ends) instead of the
 * context's lazily allocated reduction buffer.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * between captured graph launches, tight serving loops, harnesses that forbid
 * hipMalloc).  The context's device must be current.  Calls that overlap on
 * one GPU share the array and race.  Results equal psg_diag's.  Grids whose
 * spectrum needs more than 344 bins are rejected with PSG_EINVAL. */
int psg_diag_noalloc(psg_ctx *ctx, int member, double *energy, double *enstrophy, double *spec);
/* spectral nonlinear term dz/dt = N(z)+F (no viscous term), masked, for one member,
 * written into out (nf*nkt complex); cf. TIMEDERIV src_gmres/timederiv.F90:2-44 */
int psg_nonlin(psg_ctx *ctx, int member, const double *host_in, double *host_out);
/* dz/dt = -NU z + N(z) + F, masked, for one member (host in/out, nf*nkt complex), with
 * NU = 10^(log10 v2 + 2 log10|k|): TIMEDERIV (3D_NavierStokes/src_gmres/timederiv.F90:2-44,
 * 2D_NavierStokes/src_RK4/src_gmres/timederiv.F90:2-32) and the Boussinesq GPU RHS kernel
 * (3D_Boussinesq/src/kernels.h:624-692, timederiv_cuda_ src_gmres/cuda.cu:319-533). */
int psg_timederiv(psg_ctx *ctx, int member, const double *host_in, double *host_out);
/* physical-space fields of a member: 6 (3D: w then u), 7 (Boussinesq: w, u, rho) or
 * 3 (2D: w,u,v) arrays of NR doubles, x fastest, as KR_FFT_ALL produces
 * (3D_NavierStokes/src/functions.h:26-75; 3D_Boussinesq/src/out_low.h:79-110) */
int psg_physical(psg_ctx *ctx, int member, double *host_out);
/* stream used by the context (hipStream_t as void*) */
void *psg_stream(psg_ctx *ctx);
const char *psg_last_error(void);
/* number of visible devices (0 when none) */
int psg_device_count(void);

/* ---------- kernel shapes and autotuning ----------
 * Shape variants of a pass (launch geometry, points per thread, register
 * staging); results are identical up to FP rounding of the transform order.
 * x-pass of the 3D Navier-Stokes and Euler right-hand side, variants 0..11
 * (rows per block: as many as fit in 256 threads unless stated):
 *   0  8 points/thread, inputs front-loaded, product kept for all outputs
 *   1  8 points/thread, inputs front-loaded, product re-evaluated per output pair
 *   2  8 points/thread, inputs loaded per pair, product per output pair
 *   3  4 points/thread, inputs front-loaded, product kept
 *   4  4 points/thread, inputs loaded per pair, product per output pair
 *   5  as 1 with 2 rows per block
 *   6  as 1 with 4 rows per block
 *   7  4 points/thread, front-loaded, product per pair, 1 row per block
 *   8  16 points/thread (radix-16), inputs per pair, product per pair (FP32-bulk;
 *      in the FP64 form 8 and 9 run variant 1)
 *   9  as 8 with 4 rows per block
 *  10  as 1 with registers capped for 4 waves/SIMD (FP64) or 5 (FP32-bulk)
 *  11  as 1 with registers capped for 5 waves/SIMD (FP64) or 6 (FP32-bulk)
 * x-pass of the Boussinesq right-hand side (9 inputs, 4 outputs), variants 0..11:
 *   0-4 as above; 5, 6 as 2 with 2 and 4 rows per block; 7 as 4 with 1 row per
 *   block; 8, 9 as above in FP32-bulk (in FP64: 4 points/thread, inputs per pair,
 *   2 rows per block); 10 as 2 with registers capped for 4 waves/SIMD (FP64) or 5;
 *   11 as 4 with registers capped for 5 waves/SIMD (FP64) or 6
 *   (the Boussinesq flavour has no z-forward or slab variants)
 * z-forward + update pass of the 3D Navier-Stokes and Euler step (PSG_PASS_ZFWD),
 * variants 0..1:
 *   0  the three fields' inputs front-loaded
 *   1  each field loaded just before its transform (fewer registers)
 * y-inverse, x-pass and y-forward of the 3D Navier-Stokes and Euler step as
 * slabs of z planes (PSG_PASS_SLAB), variants 0..5:
 *   0  each pass over the whole grid in one launch
 *   v  the three passes over 128 >> (v-1) planes at a time (128, 64, 32, 16, 8),
 *      so that a slab's intermediate buffers are read back from the device's
 *      last-level cache instead of HBM
 * psg_autotune times `steps` steps of every variant of each pass in turn on the
 * context's own state (restored afterwards), keeps the fastest of each and
 * writes a text report. */
#define PSG_PASS_XPASS 0
#define PSG_PASS_ZFWD 1
#define PSG_PASS_SLAB 2
/* Step graphs: psg_run launches its first step directly, then replays a
 * captured HIP graph of `steps` steps (default 16) as often as fits, then
 * launches the remainder directly; results are bitwise those of direct
 * launches.  steps = 0 disables capture.  Graphs are rebuilt after psg_set_dt
 * or a variant change, and bypassed while profiling is on. */
int psg_set_graph(psg_ctx *ctx, int steps);
int psg_set_variant(psg_ctx *ctx, int pass, int variant);
int psg_get_variant(psg_ctx *ctx, int pass);
int psg_autotune(psg_ctx *ctx, int steps, char *report, int report_len);

/* ---------- timing hooks for measurement ----------
 * Kernel names of one step in launch order and their average duration (ms)
 * measured with HIP events on the context stream over the last psg_run call
 * when profiling is enabled. */
int psg_profile_enable(psg_ctx *ctx, int on);
int psg_profile_get(psg_ctx *ctx, int idx, const char **name, double *avg_ms, long *bytes, int *count);

#ifdef __cplusplus
}
#endif
#endif
