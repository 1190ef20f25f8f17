/* Drop-in for the 23-argument timestep_cuda_ the 2D DNS driver calls
 * (2D_NavierStokes/src_RK4/main.F90:219-221).  The reference repository does
 * not contain this function (its makefile lists a cuda.cu that is absent,
 * 2D_NavierStokes/src_RK4/makefile:11-15); the behaviour is defined from the
 * call site: NIT RK4+CN steps of DELT with viscosity v2 and the forcing
 * NZKX -= 2 at (0,+-4) on ZO (in place); with STFLAG=1, stats.dat receives
 * "t E D I" every NOUT steps with the definitions of
 * 2D_NavierStokes/src_RK4/src_gmres/out.F90:2-56 (E = sum |z|^2/k^2,
 * D = 2 v2 sum |z|^2, I = -Im(i KF z(0,KF))/KF^2 ... as OUT).  RCFLAG (recurrence
 * search) and SHIFTXin are accepted and ignored. */
#include "legacy_common.h"

void timestep_cuda_ns2d_dns(psg_double2 *ZO, double *ZR, double *KX, double *KY, double *TSTEP, double *TSTART,
                            double *DELT, double *alpha, double *v2, double *SHIFTXin, double *ResThresh,
                            int *IKF, int *IKN, int *IKTX, int *IKTY, int *KTY, int *NX, int *NY, int *NIT,
                            int *NOUT, int *L, int *RCFLAG, int *STFLAG) {
  (void)ZR; (void)KX; (void)KY; (void)TSTEP; (void)SHIFTXin; (void)ResThresh; (void)IKF; (void)IKN; (void)RCFLAG;
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_NS2D; c.precision = lg_precision(); c.batch = 1;
  c.nx = *NX; c.ny = *NY; c.nz = 1; c.alpha = *alpha; c.v2 = *v2; c.dt = *DELT;
  c.force = PSG_FORCE_KOLMO; c.kf = 4.0; c.ampfor = -2.0;
  psg_ctx *ctx = lg_context(&c);
  if (!ctx) return;
  const long iktx = *IKTX, ikty = *IKTY, kty = *KTY, nkt = iktx * ikty;
  if (psg_set_tables(ctx, NULL, NULL, L) || psg_set_state(ctx, 0, (const double *)ZO)) { lg_fail("upload"); return; }
  FILE *fs = (*STFLAG == 1) ? fopen("stats.dat", "w") : NULL;
  const int nout = *NOUT > 0 ? *NOUT : 1, nit = *NIT;
  int nt = 0;
  while (nt < nit) {
    int e = fs ? ((nt / nout) + 1) * nout : nit;
    if (e > nit) e = nit;
    if (psg_run(ctx, e - nt)) { lg_fail("run"); break; }
    nt = e;
    if (fs && nt % nout == 0) {
      double en = 0, ens = 0;
      if (psg_diag(ctx, 0, &en, &ens, NULL) == 0 && psg_get_state(ctx, 0, (double *)ZO) == 0) {
        /* UK(0,KF) = i KF z(0,KF)/KF^2, I = -Im(UK) (out.F90:43-45) */
        const long ik = (4 + kty) * iktx;
        const double I = -(4.0 * ZO[ik].x) / 16.0 * (L[ik] ? 1.0 : 0.0);
        fprintf(fs, " %12.5E    %12.5E    %12.5E    %12.5E\n", *TSTART + nt * *DELT, en, ens * *v2, I);
        fflush(fs);
      }
    }
  }
  if (psg_get_state(ctx, 0, (double *)ZO)) lg_fail("download");
  (void)nkt;
  if (fs) fclose(fs);
}

void timestep_cuda_(psg_double2 *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, double *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
                    int *) __attribute__((alias("timestep_cuda_ns2d_dns")));
