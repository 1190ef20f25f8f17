/*
 * FP64 CPU oracle for the psgpu pseudospectral steppers.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; it is loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in
 * psgpu_amd/ links, loads or calls it.
 *
 * It restates the reference scheme step by step in plain C99 + OpenMP:
 * the conj-fill / zero-pad / C2R / product / R2C / gather pipeline of
 * KR_FFT_ALL + multReal + RK_FFT, followed by the per-flavour update
 * kernels, with a project-written FP64 FFT (radix-2, FP64 twiddles from
 * sincos).  Each function cites the reference file:line it follows
 * (paths relative to the reference repository root).
 *
 * Parity pins: or_fft_* are pinned against numpy.fft golden vectors
 * (tests/golden/fft_*.npz) and closed forms; the tables/masks against the
 * reference-held constants (nkt=310675, SzV=22428, SzV=288872, Zk.out
 * header 1,86,171,171); ran1/ranno against the Park-Miller check value;
 * the steppers against closed forms (single-mode CN decay, Kolmogorov
 * laminar fixed point, inviscid energy conservation, time-step order).
 * Bit-for-bit agreement with the reference binary (cuFFT/FFTW round-off)
 * is "parity unpinned": the reference cannot be built here.
 */
#ifndef PSG_ORACLE_H
#define PSG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc;

/* Neumaier (Kahan-Babuska) compensated sum.  The diagnostics add up to
 * ~2e7 terms per field; a plain running sum loses up to n*eps of relative
 * accuracy there (1e-13 at 256^3), which would swamp a 1e-12 comparison. */
typedef struct { double s, c; } or_ksum;
static inline void or_kadd(or_ksum *k, double x) {
  const double t = k->s + x;
  const double as = k->s < 0 ? -k->s : k->s, ax = x < 0 ? -x : x;
  if (as >= ax) k->c += (k->s - t) + x;
  else k->c += (x - t) + k->s;
  k->s = t;
}
static inline double or_kval(const or_ksum *k) { return k->s + k->c; }

/* ---------------- FFT (or_fft.c) ---------------- */
/* In-place complex FFT of length n (power of two), unnormalised.
 * sign = -1: forward  X_k = sum x_j e^{-2 pi i jk/n}
 * sign = +1: inverse  x_j = sum X_k e^{+2 pi i jk/n}                  */
void or_fft_c2c(orc *x, int n, int sign);
/* Strided batched version: nb transforms, element stride es, batch stride bs */
void or_fft_many(orc *x, int n, int sign, long es, long nb, long bs);
/* Real <-> half-complex of length n (n even):
 * r2c: out[0..n/2] = sum r_j e^{-2pi i jk/n}          (cuFFT D2Z / numpy rfft)
 * c2r: r_j = sum_{k=0}^{n-1} X_k e^{+2pi i jk/n} with X_{n-k}=conj(X_k),
 *      Im X_0 and Im X_{n/2} ignored                 (cuFFT Z2D / n*numpy irfft) */
void or_fft_r2c(const double *r, orc *out, int n);
void or_fft_c2r(const orc *in, double *r, int n);
/* full 3D on the padded NX2 x NY x NZ half-complex layout
 * (index ((iz*ny)+iy)*nx2+ix), as cufftPlan3d(nz,ny,nx,Z2D/D2Z) */
void or_fft3_c2r(orc *F, double *r, int nx, int ny, int nz);   /* F is destroyed */
void or_fft3_r2c(const double *r, orc *F, int nx, int ny, int nz);
void or_fft2_c2r(orc *F, double *r, int nx, int ny);           /* F is destroyed */
void or_fft2_r2c(const double *r, orc *F, int nx, int ny);

/* ---------------- grid, tables, IC (or_spec.c) ---------------- */
typedef struct {
  int dim;                    /* 2 or 3 */
  int nx, ny, nz;             /* physical grid (nz=1 in 2D) */
  double alpha;               /* kx = ikx*alpha */
  int ktx, kty, ktz;
  int iktx, ikty, iktz;       /* iktz=1 in 2D */
  int nx2;                    /* nx/2+1 */
  long nkt, nr, n2;           /* truncated, physical, padded counts */
} or_grid;

int  or_grid3(or_grid *g, int nx, int ny, int nz, double alpha);
int  or_grid2(or_grid *g, int nx, int ny, double alpha);
/* k-vectors and active mask L (0/1); kz may be NULL in 2D */
void or_tables(const or_grid *g, double *kx, double *ky, double *kz, int *L);
long or_count_active(const or_grid *g);
/* NU=10**(log10(v2)+2log10|k|), NUZN=(1-NU dt/2)/dt, NU1=dt/(1+NU dt/2); 0 off-mask */
void or_nutab(const or_grid *g, double v2, double dt, double *nu, double *nuzn, double *nu1);
/* padded<->truncated index maps (-1 = zero pad) exactly as main.F90 */
void or_ikf_ikn(const or_grid *g, int *ikf, int *ikn);

/* Numerical-Recipes ran1 + reference ranno (re-seed on every draw) */
void   or_rng_seed(int iseed);           /* = RANNO(ISEED)  */
double or_ranno0(void);                  /* = RANNO(0)      */
int    or_pm_next(int idum);             /* one Park-Miller step (Schrage) */
/* reference random ICs (enstrophy-normalised random-phase shells) */
void or_ic_ns3d_ref(const or_grid *g, int iseed, double ampv, orc *xii, orc *eta, orc *zet);
void or_ic_ns2d_ref(const or_grid *g, int iseed, double ampv, orc *z);
void or_ic_euler_ref(const or_grid *g, int iseed, double ampv, orc *xii, orc *eta, orc *zet);
void or_ic_bous_ref(const or_grid *g, int iseed, double ampv, orc *xii, orc *eta, orc *zet, orc *rho);
/* benchmark ICs: SplitMix64 Gaussian phases, documented spectra */
void or_ic_solenoidal(const or_grid *g, uint64_t seed, double energy, double amp_scale,
                      orc *xii, orc *eta, orc *zet, orc *rho /* may be NULL */);
void or_ic_ns2d_cfg(const or_grid *g, uint64_t seed, double ampv, orc *z);
uint64_t or_splitmix64(uint64_t *s);
double   or_gauss(uint64_t key, uint64_t idx);

/* ---------------- 3D steppers (or_ns3d.c) ---------------- */
enum { OR_FORCE_NONE = 0, OR_FORCE_DNS = 1, OR_FORCE_GMRES = 2, OR_FORCE_KOLMO = 3 };
typedef struct or3_ctx or3_ctx;
/* flavour: 0 = Navier-Stokes Heun+CN, 1 = Euler Heun */
or3_ctx *or3_create(int nx, int ny, int nz, double alpha, int flavour,
                    double v2, double dt, int force_mode, double kf, double ampfor);
void or3_destroy(or3_ctx *c);
const or_grid *or3_grid(const or3_ctx *c);
/* advance state (3 x nkt complex, in/out) by nsteps */
void or3_run(or3_ctx *c, orc *xii, orc *eta, orc *zet, int nsteps);
/* one evaluation of the spectral right-hand side curl(u x w)+F (masked), as
 * TIMEDERIV in src_gmres/timederiv.F90 without the viscous term */
void or3_nonlin(or3_ctx *c, const orc *xii, const orc *eta, const orc *zet,
                orc *nx_, orc *ny_, orc *nz_);
/* physical velocity/vorticity (nr reals each), as KR_FFT_ALL output */
void or3_physical(or3_ctx *c, const orc *xii, const orc *eta, const orc *zet, double *out6);
/* diagnostics: E = sum_L |w|^2/|k|^2, Z = sum_L |w|^2 (spec[j], j=INT(|k|+0.5)) */
void or3_diag(const or3_ctx *c, const orc *xii, const orc *eta, const orc *zet,
              double *energy, double *enstrophy, double *spec /* len kty+2, may be NULL */);

/* ---------------- 2D stepper (or_ns2d.c) ---------------- */
typedef struct or2_ctx or2_ctx;
or2_ctx *or2_create(int nx, int ny, double alpha, double v2, double dt, int kf, double famp);
void or2_destroy(or2_ctx *c);
const or_grid *or2_grid(const or2_ctx *c);
void or2_run(or2_ctx *c, orc *z, int nsteps);
void or2_nonlin(or2_ctx *c, const orc *z, orc *nz_);
void or2_diag(const or2_ctx *c, const orc *z, double *energy, double *enstrophy, double *spec);

/* ---------------- Boussinesq (or_bous.c) ---------------- */
typedef struct orb_ctx orb_ctx;
orb_ctx *orb_create(int nx, int ny, int nz, double alpha, double v2, double sc, double ri,
                    double theta, double dt, double kf, double ampfor);
void orb_destroy(orb_ctx *c);
const or_grid *orb_grid(const orb_ctx *c);
void orb_run(orb_ctx *c, orc *xii, orc *eta, orc *zet, orc *rho, int nsteps);
void orb_rhs(orb_ctx *c, const orc *xii, const orc *eta, const orc *zet, const orc *rho,
             orc *dxii, orc *deta, orc *dzet, orc *drho);

/* ---------------- recurrence search (or_rec.c) ----------------
 * rule 0: 3D Navier-Stokes/Euler, 1: Boussinesq, 2: 2D.  L: active mask per
 * stored mode (NKT bytes).  state0/state: nf*NKT complex.  See or_rec.c. */
typedef struct or_rec or_rec;
or_rec *or_rec_new(int rule, int nf, int iktx, int ikty, int iktz, int kty, int ktz, double alpha, const int8_t *L,
                   int store_size, int norm_width, int nsx, int nsy, int nsz, int nout, double dt0, double tstart,
                   double thr, double ri, double v2, const double *state0);
void or_rec_free(or_rec *r);
int or_rec_update(or_rec *r, long NT, double delt, const double *state, double *guess, double *guess_state,
                  double *chk);
int or_rec_last(const or_rec *r, double *res, int *slots, int cap);

/* ---------------- run statistics (or_stats.c) ---------------- */
void or_stats_ns3d(long nkt, const double *kx, const double *ky, const double *kz, const int *LL, const orc *X,
                   const orc *E, const orc *Z, orc *Xt, orc *Et, orc *Zt, double cnt, double ampfor, double kfy,
                   double *out);
void or_stats_bous(long nkt, const double *kx, const double *ky, const double *kz, const int *LL, const orc *X,
                   const orc *E, const orc *Z, const orc *R, orc *Xt, orc *Et, orc *Zt, orc *Rt, double cnt,
                   double ampfor, double kfy, double ri, double theta, double *out);
void or_rig(int nx, int ny, int nz, double theta, const double *U, const double *V, const double *W,
            const double *RO, double *navg, double *savg);
int or_spectra(long nkt, int nx, const double *kx, const double *ky, const double *kz, const orc *Xt, const orc *Et,
               const orc *Zt, double theta, double *spch, double *spcv);

/* ---------------- threading ---------------- */
void or_set_threads(int n);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
