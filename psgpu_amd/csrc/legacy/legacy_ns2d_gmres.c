/* Drop-in timestep_cuda_ for the 2D Newton-GMRES driver
 * (2D_NavierStokes/src_RK4/src_gmres/cuda.cu:16-83, called from
 * src_gmres/main.F90:708,1076,1387,1483).  NSTOP RK4+CN steps of DELT with
 * viscosity V2 and the forcing NZKX -= 2 at (0,+-4); *Time is reset to 0 and
 * advanced by DELT per step; ZR, TSTART and TNEWT are not used, as in the
 * reference. */
#include "legacy_common.h"

void timestep_cuda_ns2d_gmres(psg_double2 *Z, double *ZR, double *kx, double *ky, double *Time, double *TSTART,
                              double *TNEWT, double *DELT, double *alpha, double *V2, int *ikF, int *ikN,
                              int *IKTX, int *IKTY, int *KTY, int *NX, int *NY, int *NSTOP, int *LC) {
  (void)ZR; (void)kx; (void)ky; (void)TSTART; (void)TNEWT; (void)ikF; (void)ikN; (void)KTY;
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_NS2D; c.precision = lg_precision(); c.batch = 1;
  c.nx = *NX; c.ny = *NY; c.nz = 1; c.alpha = *alpha; c.v2 = *V2; c.dt = *DELT;
  c.force = PSG_FORCE_KOLMO; c.kf = 4.0; c.ampfor = -2.0;
  psg_ctx *ctx = lg_context(&c);
  if (!ctx) return;
  const long nkt = (long)*IKTX * *IKTY;
  if (psg_set_tables(ctx, NULL, NULL, LC) || psg_set_state(ctx, 0, (const double *)Z) || psg_run(ctx, *NSTOP) ||
      psg_get_state(ctx, 0, (double *)Z)) {
    lg_fail("timestep");
  }
  (void)nkt;
  *Time = 0.0;
  for (int i = 0; i < *NSTOP; i++) *Time += *DELT;
}

void timestep_cuda_(psg_double2 *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, int *, int *, int *, int *, int *, int *, int *, int *, int *)
    __attribute__((alias("timestep_cuda_ns2d_gmres")));
