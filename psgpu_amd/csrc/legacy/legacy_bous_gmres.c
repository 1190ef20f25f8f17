/* Drop-in timestep_cuda_ and timederiv_cuda_ for the 3D Boussinesq
 * Newton-GMRES driver (3D_Boussinesq/src/src_gmres/cuda.cu:12-314 and
 * :319-533, called from 3D_Boussinesq/src/src_gmres/main.F90).
 *
 * timestep_cuda_: dt = DELT; RK4 steps while Tstep - tme > dt with tme from
 * TSTART, *Time advanced by dt per step; then, when tme != Tstep, one more
 * step of Tstep - tme that does not advance *Time (cuda.cu:188-262).  No file
 * output.
 * timederiv_cuda_: dz/dt = -NU z + N(z) + F of the 4-field state, written
 * over xii, eta, zet, rho (the RHS kernel, 3D_Boussinesq/src/kernels.h:624-692).
 * The reference's timederiv_cuda_ reads the buoyancy angle from the sin/cos
 * globals left by an earlier timestep_cuda_ call; here they come from *Theta
 * of the same call. */
#include "legacy_common.h"

static psg_ctx *bous_context(double *KFY, double *AMPFOR, double *V2, double *Ri, double *Sc, double *Theta,
                             double *alpha, double dt, int *NX, int *NY, int *NZ) {
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_BOUS3D; c.precision = lg_precision(); c.batch = 1;
  c.nx = *NX; c.ny = *NY; c.nz = *NZ; c.alpha = *alpha; c.v2 = *V2; c.dt = dt;
  c.force = PSG_FORCE_KOLMO; c.kf = *KFY; c.ampfor = *AMPFOR;
  c.ri = *Ri; c.sc = *Sc; c.theta = *Theta;
  return lg_context(&c);
}

void timestep_cuda_bous_gmres(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, psg_double2 *rho, double *kx,
                              double *ky, double *kz, double *Time, double *TSTART, double *Tstep, double *AMPFOR,
                              double *DELT, double *ResThresh, double *KFY, double *V2, double *Ri, double *Sc,
                              double *Theta, double *alpha, double *Dtarget, int *ikF, int *ikN, int *LC,
                              int *IKTX, int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ,
                              int *NOUT, int *NOUTV, int *statsFLAG, int *RecFLAG, int *adaptFLAG, int *RANK) {
  (void)kx; (void)ky; (void)kz; (void)ResThresh; (void)Dtarget; (void)ikF; (void)ikN; (void)IKTX; (void)IKTY;
  (void)IKTZ; (void)KTZ; (void)NOUT; (void)NOUTV; (void)statsFLAG; (void)RecFLAG; (void)adaptFLAG; (void)RANK;
  const double dt = *DELT;
  psg_ctx *ctx = bous_context(KFY, AMPFOR, V2, Ri, Sc, Theta, alpha, dt, NX, NY, NZ);
  if (!ctx) return;
  const long nkt = *NKT;
  double *buf = malloc(sizeof(double) * 8 * nkt);
  if (!buf) { lg_fail("host allocation"); return; }
  psg_double2 *f[4] = {xii, eta, zet, rho};
  lg_pack(buf, nkt, 4, f);
  if (psg_set_tables(ctx, NULL, NULL, LC) || psg_set_state(ctx, 0, buf)) { lg_fail("upload"); free(buf); return; }
  double tme = *TSTART;
  int n = 0;
  while (*Tstep - tme > dt) { tme += dt; n++; }
  int bad = n > 0 && psg_run(ctx, n);
  for (int i = 0; i < n; i++) *Time += dt;
  if (!bad && tme != *Tstep) {
    /* the last, shortened step; the context's dt is restored for the next call */
    bad = lg_set_dt(ctx, *Tstep - tme) || psg_run(ctx, 1) || lg_set_dt(ctx, dt);
  }
  if (bad) lg_fail("run");
  if (psg_get_state(ctx, 0, buf) == 0) lg_unpack(buf, nkt, 4, f);
  else lg_fail("download");
  free(buf);
}

void timederiv_cuda_(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, psg_double2 *rho, double *kx,
                     double *ky, double *kz, double *Time, double *TSTART, double *Tstep, double *AMPFOR,
                     double *DELT, double *ResThresh, double *KFY, double *V2, double *Ri, double *Sc,
                     double *Theta, double *alpha, double *Dtarget, int *ikF, int *ikN, int *LC, int *IKTX,
                     int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NOUT, int *NOUTV,
                     int *statsFLAG, int *RecFLAG, int *adaptFLAG, int *RANK) {
  (void)kx; (void)ky; (void)kz; (void)Time; (void)TSTART; (void)Tstep; (void)ResThresh; (void)Dtarget;
  (void)ikF; (void)ikN; (void)IKTX; (void)IKTY; (void)IKTZ; (void)KTZ; (void)NOUT; (void)NOUTV; (void)statsFLAG;
  (void)RecFLAG; (void)adaptFLAG; (void)RANK;
  psg_ctx *ctx = bous_context(KFY, AMPFOR, V2, Ri, Sc, Theta, alpha, *DELT, NX, NY, NZ);
  if (!ctx) return;
  const long nkt = *NKT;
  double *buf = malloc(sizeof(double) * 16 * nkt);
  if (!buf) { lg_fail("host allocation"); return; }
  psg_double2 *f[4] = {xii, eta, zet, rho};
  lg_pack(buf, nkt, 4, f);
  if (psg_set_tables(ctx, NULL, NULL, LC) || psg_timederiv(ctx, 0, buf, buf + 8 * nkt)) lg_fail("timederiv");
  else lg_unpack(buf + 8 * nkt, nkt, 4, f);
  free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *,
                    double *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, int *, int *, int *, int *, int *, int *, int *, int *,
                    int *, int *, int *, int *, int *, int *, int *, int *, int *)
    __attribute__((alias("timestep_cuda_bous_gmres")));
