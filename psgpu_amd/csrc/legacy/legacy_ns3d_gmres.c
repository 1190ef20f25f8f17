/* Drop-in timestep_cuda_ for the 3D Navier-Stokes Newton-GMRES driver
 * (3D_NavierStokes/src_gmres/cuda.cu:10-176; called from src_gmres/main.F90
 * with a fixed dt and again for the shortened last step).  NSTOP Heun+CN
 * steps with the eta/zeta forcing at (0,KF,+-KF) (src_gmres/kernels.h:234-238),
 * no file output.  The grid is compiled into the reference
 * (src_gmres/kernels.h:1-12); here it comes from PSG_NX/PSG_NY/PSG_NZ/PSG_ALPHA
 * (default 128^3, alpha 1). */
#include "legacy_common.h"

static int env_int(const char *k, int d) { const char *v = getenv(k); return v ? atoi(v) : d; }
static double env_dbl(const char *k, double d) { const char *v = getenv(k); return v ? atof(v) : d; }

void timestep_cuda_ns3d_gmres(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *nuzn, double *nu1,
                              double *kx, double *ky, double *kz, double *Time, double *TSTART, double *AMPFOR,
                              double *DELT, int *KFY, double *V2, int *ikF, int *ikN, int *LC, int *NSTOP) {
  (void)kx; (void)ky; (void)kz; (void)ikF; (void)ikN;
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_NS3D; c.precision = lg_precision(); c.batch = 1;
  c.nz = env_int("PSG_NZ", 128); c.ny = env_int("PSG_NY", c.nz);
  c.alpha = env_dbl("PSG_ALPHA", 1.0); c.nx = env_int("PSG_NX", (int)(c.nz / c.alpha));
  c.v2 = *V2; c.dt = *DELT; c.force = PSG_FORCE_GMRES; c.kf = (double)*KFY; c.ampfor = *AMPFOR;
  psg_ctx *ctx = lg_context(&c);
  if (!ctx) return;
  long nkt = 0;
  psg_sizes(&c, NULL, NULL, NULL, &nkt, NULL);
  double *buf = malloc(sizeof(double) * 6 * nkt);
  if (!buf) { lg_fail("host allocation"); return; }
  psg_double2 *f[3] = {xii, eta, zet};
  lg_pack(buf, nkt, 3, f);
  if (psg_set_tables(ctx, nuzn, nu1, LC) || psg_set_state(ctx, 0, buf) || psg_run(ctx, *NSTOP) ||
      psg_get_state(ctx, 0, buf)) {
    lg_fail("timestep");
  } else {
    lg_unpack(buf, nkt, 3, f);
  }
  *Time = *TSTART + *NSTOP * *DELT;
  free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, int *, double *, int *, int *, int *, int *)
    __attribute__((alias("timestep_cuda_ns3d_gmres")));
