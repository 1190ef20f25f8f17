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
                         passes (the residual flow keeps the context's precision) */
  int time_policy;    /* PSG_TIME_* */
  double dt0;         /* TRUEDELT */
  double sy;          /* fixed y shift SNEWTY */
  int stall_count;    /* NERcount limit (2D 3, 3D 5) */
  int verbose;
} psg_newton_opts;

typedef struct {
  double T, sx, sz;
  double ner;         /* ||F|| / ||Phi_T(z)|| */
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
/* Phi_T with the solver's time policy: z (host, in/out) advanced by T */
int psg_flow_map(psg_ctx *ctx, double *z, double T, double dt0, int time_policy);

/* dz1 = dPhi_T(z) dz0 from the tangent-linear stepper (host arrays, nf*nkt complex);
 * z1 (may be NULL) receives Phi_T(z).  2D and 3D Navier-Stokes. */
int psg_tangent(psg_ctx *ctx, const double *z, const double *dz0, double T, double dt0, int time_policy, double *z1,
                double *dz1);

/* Dense helpers exposed for tests: one-sided Jacobi SVD of a column-major
 * m x n (m >= n) matrix A: U (m x n), s (n, descending), V (n x n); and the
 * hookstep trust-region search on (s, p): x_i = p_i s_i/(mu + s_i^2) with
 * Delta/sqrt2 < |x| < sqrt2 Delta (main.F90:1283-1333). Returns mu. */
int psg_svd(int m, int n, const double *A, double *U, double *s, double *V);
double psg_hook_mu(int k, const double *s, const double *p, double delta, double *x);

#ifdef __cplusplus
}
#endif
#endif
