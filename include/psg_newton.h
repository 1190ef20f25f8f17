/*
 * Newton-GMRES-hookstep search for (relative) periodic orbits
 *   S_s Phi_T(z) = z,   S_s z = exp(-i (kx sx + ky sy + kz sz)) z
 * with unknowns (z, T, sx[, sz]) and the period/shift phase conditions
 * <dz/dt, dz> = 0, <dz/ds, dz> = 0, in the real inner product over the
 * active half-space modes, <a,b> = sum_L Re(conj(a) b).
 * Restates the reference solver (2D_NavierStokes/src_RK4/src_gmres/main.F90:
 * residual 714-760, GMRES/Arnoldi 853-1255, hookstep 1264-1623, update and
 * statuses 1632-1845; 3D_NavierStokes/src_gmres/main.F90:740-1880;
 * 3D_Boussinesq/src/src_gmres/main.F90 two shifts) on the persistent GPU
 * context: Krylov vectors, dot products and axpys are FP64 GPU kernels,
 * Phi_T is psg_run, the (k+1) x k Hessenberg SVD and the trust-region
 * search are host FP64.
 * Documented deviations: the Krylov dimension is capped (max_gmres) and at
 * the cap the hookstep is taken in the space built so far instead of
 * failing; Jv may come from the tangent-linear stepper (jv = 1) instead of
 * the finite difference (jv = 0, the reference's epsilon = 1e-7), in FP64
 * or FP32-bulk passes (jv = 2: the FP64 residual then acts as iterative
 * refinement of the FP32 Jacobian).
 */
#ifndef PSG_NEWTON_H
#define PSG_NEWTON_H
#include "psg.h"
#ifdef __cplusplus
extern "C" {
#endif

enum {
  PSG_NEWTON_CONVERGED = 0,
  PSG_NEWTON_GMRES_FAILED = 1,   /* "GMRES FAILED"   main.F90:1245-1253 */
  PSG_NEWTON_HSTEP_FAILED = 2,   /* "H-STEP FAILED"  main.F90:1268-1276 */
  PSG_NEWTON_TSTEP_FAILED = 3,   /* "T-STEP FAILED"  main.F90:1636-1665 */
  PSG_NEWTON_STALLED = 4,        /* "NEWTON STALLED" main.F90:1752-1765 */
  PSG_NEWTON_PERIOD_SMALL = 5,   /* "PERIOD TOO SMALL" main.F90:1807-1813 */
  PSG_NEWTON_FAILED = 6          /* "NEWTON FAILED"  main.F90:1815-1820 */
};

enum { PSG_TIME_UNIFORM = 0,    /* dt = T / INT(T/dt0)           (2D, main.F90:696-699) */
       PSG_TIME_LASTSTEP = 1 }; /* INT(T/dt0) steps of dt0 + one of T - n dt0 (3D, main.F90:771-788) */

typedef struct {
  int max_newton;     /* NEWTONMAX (2D 100, 3D 75) */
  int max_gmres;      /* Krylov cap (reference GMRESMAX 2000/500) */
  int max_hook;       /* IHOOKMAX 50 */
  double netol;       /* NETOL (2D 1e-12, 3D 1e-10) */
  double gmtol;       /* GMTOL 1e-3 (Boussinesq 1e-7) */
  double eps;         /* finite-difference epsilon 1e-7 */
  double delta0;      /* initial trust region GMRESdelta 0.1 */
  int nshift;         /* 1: sx unknown; 2: sx and sz unknown (Boussinesq) */
  int fixed_period;   /* FixedP: T not an unknown */
  int jv;             /* 0 finite difference, 1 tangent linear, 2 tangent linear with FP32-bulk
                         passes (the residual flow keeps the context's precision).  With a
                         tangent-linear Jv the hookstep's predicted residual F + J dx is formed
                         from the Arnoldi relation J Q_k = Q_{k+1} H_k instead of one more Jv */
  int time_policy;    /* PSG_TIME_* */
  double dt0;         /* TRUEDELT */
  double sy;          /* fixed y shift SNEWTY */
  int stall_count;    /* NERcount limit (2D 3, 3D 5) */
  int verbose;
  int tl_record;      /* tangent-linear Jv on a one-member 2D context: 1 (default) records the
                         physical fields of every stage of the base flow Phi_T(z) once per Newton
                         step, and each Jv then integrates the perturbation alone against the
                         record; 0 integrates base and perturbation together at every Jv.  The
                         record takes 4*steps*3*NX*NY values (8 bytes, 4 with jv = 2) and is
                         skipped when the device lacks the room */
} psg_newton_opts;

typedef struct {
  double T, sx, sz;
  double ner;         /* ||F|| / ||Phi_T(z)|| (Boussinesq: ||F|| / ||(Phi_T(z), z)||) */
  double ner2;        /* low-wavenumber residual NER2 (|k|^2 <= KYmax2) */
  int newton_iters, gmres_iters, hook_iters, status;
  double seconds;
  long phi_evals;     /* stepper integrations over a period (incl. Jv) */
} psg_newton_result;

/* defaults of the 2D reference solver (main.F90:131-136) */
void psg_newton_defaults(int flavour, psg_newton_opts *o);
/* z (nf*nkt complex, host) in: guess, out: solution; T, sx, sz in/out */
int psg_newton_solve(psg_ctx *ctx, double *z, double *T, double *sx, double *sz, const psg_newton_opts *o,
                     psg_newton_result *res);
/* Pseudo-arclength continuation in v2 = 1/Re (the reference's arclcFLAG2
 * Newton branches, 2D_NavierStokes/src_RK4/src_gmres/main.F90:805-812,
 * 1049-1062, 1123-1155, 1350-1366, 1441-1530, 1678, 1720-1726): v2 is a
 * further unknown and <dX/dr, X - X_r> = ds a further row, with
 * X = (z, T, sx, v2) and the inner product of the Newton vectors.  The
 * caller (the continuation driver) supplies the tangent dX/dr and the last
 * point X_r of the branch; Jv in v2 is the finite difference of the
 * reference (jv = 0) or, with the tangent-linear Jv, a central difference
 * S dPhi/dv2 taken once per Newton step.  Acceptance compares
 * |b+| = sqrt(|F+|^2 + Narc+^2) with sqrt(|F|^2 + Narc^2) (main.F90:1539). */
typedef struct {
  const double *dzdr;  /* host, nf*nkt complex: dZr/dr */
  double dTdr, dsdr, dv2dr;
  const double *zr;    /* host, nf*nkt complex: the branch point Zr */
  double Tr, sr, v2r;
  double ds;           /* arclength step Ndelta_r */
} psg_alc;
int psg_newton_solve_alc(psg_ctx *ctx, double *z, double *T, double *sx, double *v2, const psg_alc *alc,
                         const psg_newton_opts *o, psg_newton_result *res);
/* Many guesses on one batched 2D context.  The reference solves its guesses one
 * after the other (2D_NavierStokes/src_RK4/src_gmres/main.F90:663-1845, the loop
 * over IGUESS); here the context's `batch` members are lanes: guess g
 * (0 <= g < nguess) runs psg_newton_solve's iteration on a free lane, and each
 * time every active solver waits for a flow map (residual, finite-difference Jv,
 * hookstep candidate) one batched run of the context serves them all, member m
 * stepping with its own dt = T_m / INT(T_m/dt0) and standing still once done;
 * a finished guess frees its lane for the next.  z: nguess * nf*nkt complex
 * (host, in/out); T, sx (and sz, may be NULL): nguess each, in/out; res: nguess
 * results.  Finite-difference Jv (jv = 0), uniform steps, no continuation.
 * Returns 0, or the first error of any guess. */
int psg_newton_batch(psg_ctx *ctx, int nguess, double *z, double *T, double *sx, double *sz,
                     const psg_newton_opts *o, psg_newton_result *res);
/* Phi_T with the solver's time policy: z (host, in/out) advanced by T */
int psg_flow_map(psg_ctx *ctx, double *z, double T, double dt0, int time_policy);

/* dz1 = dPhi_T(z) dz0 from the tangent-linear stepper (host arrays, nf*nkt complex);
 * z1 (may be NULL) receives Phi_T(z).  2D and 3D Navier-Stokes. */
int psg_tangent(psg_ctx *ctx, const double *z, const double *dz0, double T, double dt0, int time_policy, double *z1,
                double *dz1);
/* dz1[i] = dPhi_T(z) dz0[i] for ndir directions about one base (host arrays,
 * ndir * nf*nkt complex), z1 (may be NULL) = Phi_T(z).  On a one-member 2D
 * context the base flow records the physical fields of its stages once
 * (fp32 != 0: as FP32, and the directions then run FP32 passes) and every
 * direction integrates the perturbation alone against the record; elsewhere,
 * or when the device lacks room for the record, each direction runs the
 * combined base + perturbation stepper of psg_tangent.  No counterpart in the
 * reference (whose Jv are finite differences, main.F90:1084-1100). */
int psg_tangent_many(psg_ctx *ctx, const double *z, int ndir, const double *dz0, double T, double dt0,
                     int time_policy, int fp32, double *z1, double *dz1);

/* ---------- Floquet (stability) analysis of a periodic orbit ----------
 * Leading eigenvalues mu of the monodromy operator A q = S_s dPhi_T(z) q
 * (the Jacobian action of the Newton residual without its -q term), by
 * Arnoldi iteration on the GPU with FP64 masked real inner products.
 * Replaces the reference's ARPACK dnaupd reverse-communication loop over the
 * GMRES Jv code (2D_NavierStokes/src_RK4/src_gmres/main.F90:381-400 setup,
 * 859-1037 iteration and output, 1084-1100 the Jv it feeds; defaults nev = 10,
 * ncv = 200, TOL = 1e-7, which = 'LR', main.F90:63-67,388).  The reference
 * restarts implicitly (ARPACK's IRAM, exact shifts); here a cycle of ncv
 * Arnoldi steps is restarted explicitly from the wanted Ritz vectors, and Jv
 * may come from the tangent-linear stepper.  Ritz values come from the
 * Hessenberg matrix by Francis double-shift QR (psg_hess_eig), Ritz vectors
 * by inverse iteration; a pair converges when |h_{m+1,m}| |y_m| <= tol |mu|.
 * Floquet exponents are log|mu|/T and atan2(mu_im, mu_re)/T
 * (main.F90:929-930). */
enum { PSG_FLOQUET_LR = 0, PSG_FLOQUET_LM = 1 };
typedef struct {
  int nev;            /* wanted multipliers */
  int ncv;            /* Arnoldi steps per cycle (nev + 2 .. 400) */
  int max_restarts;   /* cycles after the first */
  double tol;         /* relative Ritz residual tolerance */
  int which;          /* PSG_FLOQUET_LR: largest real part; PSG_FLOQUET_LM: largest magnitude */
  int jv;             /* 0 finite difference (eps), 1 tangent linear, 2 tangent linear in FP32-bulk
                         passes; on a one-member 2D context the tangent linear runs against a record
                         of the orbit's flow taken once (as psg_tangent_many) when the device has room */
  double eps;         /* finite-difference epsilon, 1e-7 */
  int time_policy;    /* PSG_TIME_* */
  double dt0;
  double sy;          /* fixed y shift */
  unsigned long long seed;  /* start vector */
  int verbose;
} psg_floquet_opts;

typedef struct {
  int nconv;          /* wanted pairs that met tol */
  int matvecs;        /* Jacobian actions (flows over T) */
  int cycles;
  double seconds;
} psg_floquet_result;

void psg_floquet_defaults(psg_floquet_opts *o);
/* z (host, nf*nkt complex): the orbit; T, sx, sz its period and shifts.
 * mu_re, mu_im, err: nev entries, ordered by `which`, err = |h_{m+1,m}| |y_m|
 * (absolute Ritz residual).  vecs (may be NULL): nev host vectors of nf*nkt
 * complex, the Ritz vector of a real mu, and for a complex pair (mu, conj mu)
 * its real part then its imaginary part (the dneupd convention). */
int psg_floquet(psg_ctx *ctx, const double *z, double T, double sx, double sz, const psg_floquet_opts *o,
                double *mu_re, double *mu_im, double *err, double *vecs, psg_floquet_result *res);

/* Dense helpers exposed for tests: one-sided Jacobi SVD of a column-major
 * m x n (m >= n) matrix A: U (m x n), s (n, descending), V (n x n); and the
 * hookstep trust-region search on (s, p): x_i = p_i s_i/(mu + s_i^2) with
 * Delta/sqrt2 < |x| < sqrt2 Delta (main.F90:1283-1333). Returns mu. */
int psg_svd(int m, int n, const double *A, double *U, double *s, double *V);
double psg_hook_mu(int k, const double *s, const double *p, double delta, double *x);
/* eigenvalues (wr + i wi) of the n x n upper Hessenberg matrix H (column-major,
 * entries below the subdiagonal ignored) by Francis double-shift QR; complex
 * pairs are adjacent, positive imaginary part first.  0, or -1 if the
 * iteration does not converge. */
int psg_hess_eig(int n, const double *H, double *wr, double *wi);

#ifdef __cplusplus
}
#endif
#endif
