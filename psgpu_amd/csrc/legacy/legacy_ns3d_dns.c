/* Drop-in timestep_cuda_ for the 3D Navier-Stokes DNS driver
 * (3D_NavierStokes/src/cuda.cu:18-151, called from src/main.F90:203-204).
 * NSTOP Heun+CN steps with the xi forcing at (0,KF,+-KF); every NOUTV steps
 * (from step 0, when NSTOP > 1) the physical vorticity is copied to xir, etr,
 * ztr and written to JOB<RANK>/{xii,eta,zet}NNN.raw; with statsFLAG != 0,
 * JOB<RANK>/stats.dat and engy.dat get one line every NOUT steps
 * (computeStats, 3D_NavierStokes/src/functions.h:128-209; statsFLAG = 2 from
 * tme > 500), from device reductions and device time averages (psg_stats).  With RecFLAG = 1
 * every NOUT block runs the recurrence search on the GPU (psg_rec: check of the
 * newest history entry, then the push of the current state, cuda.cu:114-139)
 * and writes its guesses to JOB<RANK>/UPO_ZK.out and UPOinfo_ts.dat; as in the
 * reference, a check resets tme (the time column of stats.dat) to
 * TSTART + (NT - NOUT) DELT.  The time-averaged means/pdfs files are not produced. */
#include "legacy_common.h"
#include "../../../include/psg_stats.h"

/* computeStats (3D_NavierStokes/src/functions.h:128-209) from the device sums of
 * psg_stats: the normalisations and the two output lines */
static int stats_line(psg_stats *st, long tcount, double v2, double kfy, double tme, FILE *fs, FILE *fe) {
  double v[PSG_STNS_N];
  if (psg_stats_update(st, tcount, v)) return -1;
  const double input = v[PSG_STNS_INPUT] * v2 * kfy * kfy;
  const double diss = v[PSG_STNS_DISS] * v2 * v2 * 8.0 * kfy * kfy;
  const double dturb = v[PSG_STNS_DTURB] * v2 * 16.0 * kfy * kfy;
  fprintf(fs, "%e %e %e %e %e %e \n", tme, v[PSG_STNS_ENERGY], v[PSG_STNS_ETURB], diss, dturb, input);
  fflush(fs);
  fprintf(fe, "%e %e %e %e %e %e %e \n", tme, v[PSG_STNS_UA], v[PSG_STNS_VA], v[PSG_STNS_WA], v[PSG_STNS_XIA],
          v[PSG_STNS_ETA], v[PSG_STNS_ZTA]);
  fflush(fe);
  return 0;
}

void timestep_cuda_ns3d_dns(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *xir, double *etr,
                            double *ztr, double *nuzn, double *nu1, double *kx, double *ky, double *kz,
                            double *Time, double *TSTART, double *AMPFOR, double *DELT, double *ResThresh,
                            double *KFY, double *V2, double *alpha, int *ikF, int *ikN, int *LC, int *IKTX,
                            int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NSTOP,
                            int *NOUT, int *NOUTV, int *statsFLAG, int *RecFLAG, int *RANK) {
  (void)kx; (void)ky; (void)kz; (void)ikF; (void)ikN; (void)IKTX; (void)IKTY; (void)IKTZ; (void)KTZ;
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_NS3D; c.precision = lg_precision(); c.batch = 1;
  c.nx = *NX; c.ny = *NY; c.nz = *NZ; c.alpha = *alpha; c.v2 = *V2; c.dt = *DELT;
  c.force = PSG_FORCE_DNS; c.kf = *KFY; c.ampfor = *AMPFOR;
  psg_ctx *ctx = lg_context(&c);
  if (!ctx) return;
  const long nkt = *NKT, nr = (long)*NX * *NY * *NZ;
  double *buf = malloc(sizeof(double) * 6 * nkt);
  if (!buf) { lg_fail("host allocation"); return; }
  psg_double2 *f[3] = {xii, eta, zet};
  lg_pack(buf, nkt, 3, f);
  if (psg_set_tables(ctx, nuzn, nu1, LC) || psg_set_state(ctx, 0, buf)) { lg_fail("upload"); free(buf); return; }
  FILE *fs = NULL, *fe = NULL;
  psg_stats *st = NULL;
  long tcount = 0;
  if (*statsFLAG != 0) {
    char p[64];
    snprintf(p, sizeof p, "JOB%1d/stats.dat", *RANK);
    fs = fopen(p, "w");
    snprintf(p, sizeof p, "JOB%1d/engy.dat", *RANK);
    fe = fopen(p, "w");
    if (!fs || !fe || psg_stats_create(ctx, 0, &st)) {
      if (!fs || !fe) fprintf(stderr, "psgpu legacy: cannot open JOB%d/stats.dat\n", *RANK);
      else lg_fail("statistics");
      if (fs) fclose(fs);
      if (fe) fclose(fe);
      fs = fe = NULL;
    }
  }
  const int nout = *NOUT > 0 ? *NOUT : 1, nvort = *NOUTV > 0 ? *NOUTV : 1, nstop = *NSTOP;
  lg_upo upo;
  memset(&upo, 0, sizeof upo);
  if (*RecFLAG == 1 && lg_upo_open(&upo, ctx, *RANK, 3, nkt, nout, *TSTART, *ResThresh)) {
    lg_fail("recurrence search");
    lg_upo_close(&upo);
  }
  const int hf = fs || upo.rec;   /* high-frequency blocks every NOUT steps */
  double *phys = NULL;
  int count = 0;
  double tme = *TSTART;
  int nt = 0;
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
        snprintf(p, sizeof p, "JOB%1d/%s%03d.raw", *RANK, nm[i], count);
        FILE *r = fopen(p, "wb");
        if (r) { fwrite(phys + i * nr, sizeof(double), (size_t)nr, r); fclose(r); }
      }
    }
    /* run up to the next step with an output event: a dump at the top of
     * iteration e, or statistics after iteration e-1 */
    int e = nt + 1;
    while (e < nstop && !((e % nvort == 0 && nstop > 1) || (hf && (e - 1) % nout == 0 && e - 1 != 0))) e++;
    if (psg_run(ctx, e - nt)) { lg_fail("run"); break; }
    tme += (e - nt) * *DELT;
    nt = e;
    if (hf && (nt - 1) % nout == 0 && nt - 1 != 0) {
      if (upo.rec && lg_upo_update(&upo, nt - 1, *DELT, &tme)) { lg_fail("recurrence search"); break; }
      if (fs && (*statsFLAG == 1 || (*statsFLAG == 2 && tme > 500.0))) {
        if (stats_line(st, tcount, *V2, *KFY, tme, fs, fe)) { lg_fail("statistics"); break; }
        tcount++;
      }
    }
  }
  if (psg_get_state(ctx, 0, buf) == 0) lg_unpack(buf, nkt, 3, f);
  else lg_fail("download");
  *Time = *TSTART + nstop * *DELT;
  if (fs) fclose(fs);
  if (fe) fclose(fe);
  if (st) psg_stats_destroy(st);
  lg_upo_close(&upo);
  free(phys); free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, double *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
                    int *, int *, int *, int *, int *, int *) __attribute__((alias("timestep_cuda_ns3d_dns")));
