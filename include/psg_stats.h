/*
 * Run statistics of a DNS on the GPU: the reference's computeStats sums, the
 * running time averages of the state, the physical-space velocity variances,
 * gradient-Richardson plane averages, max|u| for the adaptive step, and the
 * horizontal/vertical spectra and mean fields of the run's end.
 *
 * Reference: 3D_NavierStokes/src/functions.h:128-209 (computeStats);
 * 3D_Boussinesq/src/functions.h:275-411 (computeStats with density and
 * buoyancy flux, the Dtarget throttle), kernels.h:96-135 (avgVelocity,
 * subtractReal), out_low.h:5-76 (adaptive dt, RiG), finalise.h:17-230
 * (spectra, avgs/means, pdfs).  The reference copies every state to the host
 * and loops over it on one core; here each quantity is a device reduction in a
 * fixed order (per-thread compensated sums, block trees, an ordered final
 * pass), the time averages stay in HBM, and only the sums come back.
 * The caller keeps the reference's scalar bookkeeping (averages, extrema,
 * normalisations, file layouts), as the legacy Boussinesq and 3D DNS
 * entry points do.
 */
#ifndef PSG_STATS_H
#define PSG_STATS_H
#include "psg.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct psg_stats psg_stats;

/* raw sums of psg_stats_update (before the reference's normalisations).
 * 3D Navier-Stokes (masked modes, wk = max(|k|^2, 1e-3)): */
enum {
  PSG_STNS_DISS = 0,    /* sum |w|^2                               */
  PSG_STNS_ENERGY = 1,  /* sum |w|^2 / wk                          */
  PSG_STNS_ETURB = 2,   /* sum |<w>|^2 / wk (after the average update) */
  PSG_STNS_DTURB = 3,   /* sum |<w>|^2                             */
  PSG_STNS_INPUT = 4,   /* ampfor (Re xi(0,kf,kf) - Re xi(0,kf,-kf)) */
  PSG_STNS_XIA = 5, PSG_STNS_ETA = 6, PSG_STNS_ZTA = 7,   /* sum |xi|^2, |eta|^2, |zeta|^2 */
  PSG_STNS_UA = 8, PSG_STNS_VA = 9, PSG_STNS_WA = 10,      /* sum |u|^2, |v|^2, |w|^2 (spectral) */
  PSG_STNS_N = 11
};
/* Boussinesq: */
enum {
  PSG_STB_DISS = 0, PSG_STB_ENERGY = 1, PSG_STB_ETURB = 2, PSG_STB_DTURB = 3,
  PSG_STB_INPUT = 4,    /* -ampfor Re zeta(0,kf,0) / kf            */
  PSG_STB_PE = 5,       /* sum |rho|^2                             */
  PSG_STB_PED = 6,      /* sum |rho|^2 wk                          */
  PSG_STB_PED1 = 7,     /* sum kx^2 |rho|^2                        */
  PSG_STB_PED2 = 8,     /* sum ky^2 |rho|^2                        */
  PSG_STB_PETURB = 9,   /* sum |<rho>|^2                           */
  PSG_STB_PEDTURB = 10, /* sum |<rho>|^2 wk                        */
  PSG_STB_BFLUX = 11,   /* sum ri |u_b| |rho| cos(arg rho - arg u_b), u_b = v sin(theta) + w cos(theta) */
  PSG_STB_N = 12
};

/* time averages start at zero; member: the context member whose state is used */
int psg_stats_create(psg_ctx *ctx, int member, psg_stats **out);
/* computeStats: sums over the active modes of the current state, and the
 * running averages <z> <- (<z> count + z)/(count + 1) updated on the device.
 * vals: PSG_STNS_N or PSG_STB_N doubles. */
int psg_stats_update(psg_stats *s, long count, double *vals);
/* At a Tout crossing (Boussinesq, out_low.h:5-76), or every NOUT steps
 * (3D Navier-Stokes avgVelocity, 3D_NavierStokes/src/cuda.cu:70 and
 * kernels.h:70-80): physical fields of the current state, then, as `what` asks,
 *   PSG_STP_RMS   <u^2>, <v^2>, <w^2> <- (. count + u^2)/(count + 1) per grid point
 *   PSG_STP_UMAX  *umax = max |u| over the grid (the adaptive step)
 *   PSG_STP_RIG   navg/savg (nz entries, or ny when theta < pi/4) of
 *                 sum (N^2) and sum (shear^2) per plane, *nplanes set
 *                 (Boussinesq only) */
enum { PSG_STP_RMS = 1, PSG_STP_UMAX = 2, PSG_STP_RIG = 4 };
int psg_stats_physical(psg_stats *s, int what, long count, double *umax, double *navg, double *savg,
                       int *nplanes);
/* finalise.h spectra from the time-averaged vorticity: spch[j], spcv[j] for
 * j <= size (horizontal |k_h| and vertical k_v bins, k_v = ky when theta < pi/4,
 * else kz; negative k_v in bin 0 as max(k_v, 1e-3) puts them); cap >= nx + 1 */
int psg_stats_spectra(psg_stats *s, double *spch, double *spcv, int cap, int *size);
/* finalise.h mean fields, physical fields of nx*ny*nz doubles, x fastest, in
 * the avgs.dat order.  Boussinesq: 7 fields, rho, u, v, w of the time-averaged
 * state, <v^2>, <w^2>, and <u^2> - <u>^2 (3D_Boussinesq/src/finalise.h);
 * 3D Navier-Stokes: 5 fields, u, v, w of the time-averaged state, <v^2> and
 * <u^2> - <u>^2 (3D_NavierStokes/src/finalise.h:25-117) */
int psg_stats_mean_fields(psg_stats *s, double *out);
/* the time-averaged state (nf*nkt complex) */
int psg_stats_tavg(psg_stats *s, double *host);
int psg_stats_destroy(psg_stats *s);
/* change the forcing amplitude of a context (the Dtarget throttle) */
int psg_set_ampfor(psg_ctx *ctx, double ampfor);

#ifdef __cplusplus
}
#endif
#endif
