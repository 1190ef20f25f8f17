/*
 * Recurrence search for Newton guesses over a DNS history, on the GPU.
 *
 * Restates the reference's near-recurrence check that its DNS drivers run
 * every NOUT steps (3D_NavierStokes/src/cuda.cu:114-139 with
 * functions.h:111-125 computeNorm, :213-389 computeZdash/recurrence_check and
 * init.h:182-213 set-up; 3D_Boussinesq/src/out_high.h:8-33 with
 * functions.h:190-208, :415-619 and init.h:275-310; the 2D variant
 * 2D_NavierStokes/src_RK4/functions.h:59-67, :105-309).  A ring of the last
 * `store_size` states is kept; at each check the current state is phase
 * shifted by every (sx, sy, sz) of a uniform grid over the periodic box and
 * compared, on the mode subset 0 <= ikx < W, -W <= ky, kz < W, with every
 * stored state:
 *   res(j, s) = sqrt(( -2 <Z_s, Z_j> + |Z_s|^2 + |Z_j|^2 ) / |Z|^2)
 * (Boussinesq: rho weighted by ri, divided by |Z_j|^2 + |Z|^2), Z_s = Z e^{-i k.s}.
 * The per-store minimum over shifts, the first turning point of the residual
 * as the period grows, and the output sequences that start when the best
 * residual drops below the threshold and end when it rises above it follow
 * the reference statement by statement; a sequence's best match becomes a
 * guess (the UPO_ZK.out record).
 *
 * MI355X layout: the ring (subset vectors, and the full states when
 * keep_states is set) lives in HBM; the shift sweep is one split-K FP64
 * product of the (shifts x subset) and (stores x subset) matrices staged
 * through LDS, reduced per store in the reference's shift order.  The
 * reference copies every history state to the host and sweeps on one core.
 */
#ifndef PSG_REC_H
#define PSG_REC_H
#include "psg.h"
#ifdef __cplusplus
extern "C" {
#endif

enum {
  PSG_REC_NS = 0,    /* 3D Navier-Stokes / Euler (masked history norms) */
  PSG_REC_BOUS = 1,  /* 3D Boussinesq: ri-weighted density, periods < 0.5 skipped,
                        residual relative to |Z_j|^2 + |Z|^2 */
  PSG_REC_NS2D = 2   /* 2D: unmasked norms, time under threshold in the guess */
};

typedef struct psg_rec psg_rec;

typedef struct {
  int rule;          /* PSG_REC_* */
  int member;        /* context member whose state is searched */
  int store_size;    /* history length storeSize (500) */
  int norm_width;    /* normWidth W: 3D 8, 2D 42 (W <= KTY, KTZ, IKTX) */
  int nshift_x, nshift_y, nshift_z;  /* NS 30,2,2; Boussinesq 64,1,1; 2D 30,4,1 */
  int nout;          /* NOUT: steps between history entries */
  double tstart;     /* TSTART */
  double threshold;  /* ResidualThreshold */
  double ri;         /* Boussinesq weight of the density field (Ri) */
  int keep_states;   /* 1: full states kept on the device for the guess records */
} psg_rec_opts;

/* one completed guess (the UPO_ZK.out / UPOinfo_ts.dat record) */
typedef struct {
  int index;         /* izkout: 1, 2, ... */
  int store;         /* history slot of the guess state */
  double time;       /* minTIMEouter */
  double v2;
  double period;     /* RecPERIOD */
  double sx, sy, sz;
  double residual;   /* minRESIDUALouter */
  double time_under; /* 2D: time under threshold over the period (timeUNDER) */
} psg_rec_guess;

/* the best match of the last check (minRESIDUALinner and companions) */
typedef struct {
  int checked;       /* 1 when this update ran a check */
  int compared;      /* stores compared */
  int store, count;  /* slot and lag (in history entries) of the best match; 0 if none */
  double residual;   /* 10000 if none */
  double sx, sy, sz;
  double time;       /* tme = tstart + (nt - nout) dt */
} psg_rec_check;

/* defaults of the reference for the context's flavour (storeSize 500,
 * normWidth and shift counts of its kernels.h, NOUT 10, threshold 0.1) */
int psg_rec_defaults(psg_ctx *ctx, psg_rec_opts *o);
/* allocates the ring and makes the context's current state entry 0 */
int psg_rec_create(psg_ctx *ctx, const psg_rec_opts *o, psg_rec **out);
/* The NOUT block at loop index nt with step dt: a check of the newest entry
 * against the history (once there is one), then the context's current state
 * is pushed.  Returns 1 when an output sequence ended with a guess (filled into
 * *guess; its state, nf*nkt complex, into guess_state when not NULL and the
 * full states are kept), 0 otherwise, or a negative PSG_E* code.  chk (may be
 * NULL) receives the check's best match. */
int psg_rec_update(psg_rec *r, long nt, double dt, psg_rec_guess *guess, double *guess_state,
                   psg_rec_check *chk);
/* residual table of the last check: res[j * nshift + s] over the compared
 * stores j (slots into `slots`) and shifts s = x + NX (y + NY z); sizes in *nj */
int psg_rec_last(psg_rec *r, double *res, int *slots, int *nj, int cap);
/* full state of history slot `store` (needs keep_states) */
int psg_rec_state(psg_rec *r, int store, double *host);
int psg_rec_destroy(psg_rec *r);

#ifdef __cplusplus
}
#endif
#endif
