/* Drop-in timestep_cuda_ for the 3D Euler driver (3D_Euler/src/cuda.cu:17-123,
 * called from 3D_Euler/src/main.F90:182).  NSTOP inviscid Heun steps; every
 * NOUTV steps (from step 0, when NSTOP > 1) the physical vorticity goes to
 * xir/etr/ztr and {xii,eta,zet}NNN.raw in the working directory
 * (3D_Euler/src/cuda.cu:58-76); with statsFLAG != 0 stats.dat gets
 * "t energy enstrophy" every NOUT steps. */
#include "legacy_common.h"

void timestep_cuda_euler(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *xir, double *etr,
                         double *ztr, double *kx, double *ky, double *kz, double *Time, double *TSTART,
                         double *DELT, double *alpha, int *ikF, int *ikN, int *LC, int *IKTX, int *IKTY,
                         int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NSTOP, int *NOUT,
                         int *NOUTV, int *statsFLAG) {
  (void)kx; (void)ky; (void)kz; (void)ikF; (void)ikN; (void)IKTX; (void)IKTY; (void)IKTZ; (void)KTZ;
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_EULER3D; c.precision = lg_precision(); c.batch = 1;
  c.nx = *NX; c.ny = *NY; c.nz = *NZ; c.alpha = *alpha; c.v2 = 0.0; c.dt = *DELT; c.force = PSG_FORCE_NONE;
  psg_ctx *ctx = lg_context(&c);
  if (!ctx) return;
  const long nkt = *NKT, nr = (long)*NX * *NY * *NZ;
  double *buf = malloc(sizeof(double) * 6 * nkt);
  if (!buf) { lg_fail("host allocation"); return; }
  psg_double2 *f[3] = {xii, eta, zet};
  lg_pack(buf, nkt, 3, f);
  if (psg_set_tables(ctx, NULL, NULL, LC) || psg_set_state(ctx, 0, buf)) { lg_fail("upload"); free(buf); return; }
  FILE *fs = *statsFLAG ? fopen("stats.dat", "w") : NULL;
  const int nout = *NOUT > 0 ? *NOUT : 1, nvort = *NOUTV > 0 ? *NOUTV : 1, nstop = *NSTOP;
  double *phys = NULL, tme = 0.0;
  int count = 0, nt = 0;
  while (nt < nstop) {
    if (nt % nvort == 0 && nstop > 1) {
      if (!phys) phys = malloc(sizeof(double) * 6 * nr);
      if (!phys) { lg_fail("host allocation"); break; }
      if (psg_physical(ctx, 0, phys)) { lg_fail("physical"); break; }
      memcpy(xir, phys, sizeof(double) * nr);
      memcpy(etr, phys + nr, sizeof(double) * nr);
      memcpy(ztr, phys + 2 * nr, sizeof(double) * nr);
      count++;
      const char *nm[3] = {"xii", "eta", "zet"};
      for (int i = 0; i < 3; i++) {
        char p[64];
        snprintf(p, sizeof p, "%s%03d.raw", nm[i], count);
        FILE *r = fopen(p, "wb");
        if (r) { fwrite(phys + i * nr, sizeof(double), (size_t)nr, r); fclose(r); }
      }
    }
    int e = nt + 1;
    while (e < nstop && !((e % nvort == 0 && nstop > 1) || (fs && (e - 1) % nout == 0 && e - 1 != 0))) e++;
    if (psg_run(ctx, e - nt)) { lg_fail("run"); break; }
    tme += (e - nt) * *DELT;
    nt = e;
    if (fs && (nt - 1) % nout == 0 && nt - 1 != 0) {
      double en = 0, ens = 0;
      if (psg_diag(ctx, 0, &en, &ens, NULL) == 0) { fprintf(fs, "%e %e %e \n", tme, en, ens); fflush(fs); }
    }
  }
  if (psg_get_state(ctx, 0, buf) == 0) lg_unpack(buf, nkt, 3, f);
  else lg_fail("download");
  *Time = *TSTART + nstop * *DELT;
  if (fs) fclose(fs);
  free(phys); free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, double *, int *, int *, int *, int *, int *, int *,
                    int *, int *, int *, int *, int *, int *, int *, int *, int *)
    __attribute__((alias("timestep_cuda_euler")));
