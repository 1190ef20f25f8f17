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
 * TSTART + (NT - NOUT) DELT.  With statsFLAG != 0 the velocity variances
 * are accumulated every NOUT steps (avgVelocity, cuda.cu:70), and with
 * statsFLAG = 1 the end of the run (finalise.h:16-167) takes a last
 * computeStats and writes avgs.dat, means.dat and pdfs.dat and the AVERAGE
 * STATS lines.  Deviations: Diss/Energy/Input entries that computeStats never
 * fills read as zero (the reference's are uninitialised), and a value at the
 * maximum lands in a histogram slot past the last printed bin instead of past
 * the array. */
#include "legacy_common.h"
#include "../../../include/psg_stats.h"

/* the scalar bookkeeping of computeStats and finalise (the reference's
 * file-scope doubles, which start at zero) */
typedef struct {
  double dissMax, dissMin, enerMax, enerMin, inputMax, inputMin, dissAvg, enerAvg, inputAvg;
  double *Diss, *Energy, *Input;
  int nStats;
  FILE *avgs, *means, *pdfs;
} ns_book;

/* computeStats (3D_NavierStokes/src/functions.h:128-209) from the device sums of
 * psg_stats: the normalisations, the bookkeeping and the two output lines; NT is
 * the reference's step index at the call */
static int stats_line(psg_stats *st, long tcount, double v2, double kfy, double tme, FILE *fs, FILE *fe, ns_book *b,
                      int NT, int nout) {
  double v[PSG_STNS_N];
  if (psg_stats_update(st, tcount, v)) return -1;
  const double input = v[PSG_STNS_INPUT] * v2 * kfy * kfy;
  const double diss = v[PSG_STNS_DISS] * v2 * v2 * 8.0 * kfy * kfy;
  const double dturb = v[PSG_STNS_DTURB] * v2 * 16.0 * kfy * kfy;
  const double energy = v[PSG_STNS_ENERGY];
  if (b) {
    b->dissMax = fmax(diss, b->dissMax);
    b->dissMin = fmin(diss, b->dissMin);
    b->enerMax = fmax(energy, b->enerMax);
    b->enerMin = fmin(energy, b->enerMin);
    b->inputMax = fmax(input, b->inputMax);
    b->inputMin = fmin(input, b->inputMin);
    b->dissAvg = (b->dissAvg * tcount + diss) / (tcount + 1.0);
    b->enerAvg = (b->enerAvg * tcount + energy) / (tcount + 1.0);
    b->inputAvg = (b->inputAvg * tcount + input) / (tcount + 1.0);
    const int istat = NT / nout - 1;
    if (istat >= 0 && istat < b->nStats) {
      b->Diss[istat] = diss;
      b->Energy[istat] = energy;
      b->Input[istat] = input;
    }
  }
  fprintf(fs, "%e %e %e %e %e %e \n", tme, energy, v[PSG_STNS_ETURB], diss, dturb, input);
  fflush(fs);
  fprintf(fe, "%e %e %e %e %e %e %e \n", tme, v[PSG_STNS_UA], v[PSG_STNS_VA], v[PSG_STNS_WA], v[PSG_STNS_XIA],
          v[PSG_STNS_ETA], v[PSG_STNS_ZTA]);
  fflush(fe);
  return 0;
}

/* finalise.h:25-167 after the last computeStats: the five mean fields of the
 * time-averaged state (avgs.dat), their x-means (means.dat), the AVERAGE
 * STATS lines and the p.d.f.s of energy, dissipation and input (pdfs.dat) */
static int ns_finalise(psg_stats *st, ns_book *b, double Time, int nx, int ny, int nz) {
  const long nr = (long)nx * ny * nz;
  double *f = malloc(sizeof(double) * 5 * nr);
  if (!f || psg_stats_mean_fields(st, f)) { free(f); return -1; }
  for (int k = 0; k < 5; k++) {
    fwrite(&Time, sizeof(double), 1, b->avgs);
    fwrite(f + k * nr, sizeof(double), (size_t)nr, b->avgs);
  }
  fflush(b->avgs);
  const double twopi = 4.0 * asin(1.0);
  /* columns: y, z, <u>, <v>, x-means of <u^2>-<u>^2 and <v^2> */
  const int src[4] = {0, 1, 4, 3};
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++) {
      double m[4];
      for (int q = 0; q < 4; q++) {
        const double *row = f + src[q] * nr + ((long)ny * iz + iy) * nx;
        double a = 0.0;
        for (int ix = 0; ix < nx; ix++) a += row[ix];
        m[q] = a / (double)nx;
      }
      fprintf(b->means, "%.15e %.15e %.15e %.15e %.15e %.15e \n", iy * twopi / (double)ny, iz * twopi / (double)nz,
              m[0], m[1], m[2], m[3]);
    }
  fflush(b->means);
  free(f);
  printf("AVERAGE STATS *****\n");
  printf("energy = %e,  eturb = %e \n", b->enerAvg, 0.0);
  printf("dissipation = %e,  dturb = %e \n", b->dissAvg, 0.0);
  printf("input E = %e \n", b->inputAvg);
  printf("max/min dissipation %e, %e \n ", b->dissMax, b->dissMin);
  printf("max/min energy %e, %e \n", b->enerMax, b->enerMin);
  printf("max/min input %e, %e \n", b->inputMax, b->inputMin);
  if (b->dissMax - b->dissMin > 0.001) {
    enum { nBin = 100 };
    const double correct = 1.0000001;
    int I_E[nBin + 2], I_D[nBin + 2], I_I[nBin + 2];
    const double dD = (b->dissMax * correct - b->dissMin) / (double)nBin;
    const double dI = (b->inputMax * correct - b->inputMin) / (double)nBin;
    const double dE = (b->enerMax * correct - b->enerMin) / (double)nBin;
    memset(I_E, 0, sizeof I_E); memset(I_D, 0, sizeof I_D); memset(I_I, 0, sizeof I_I);
    for (int i = 0; i < b->nStats - 1; i++) {
      int j = (int)((b->Energy[i] - b->enerMin) / dE) + 1;
      if (j >= 0 && j <= nBin + 1) I_E[j]++;
      j = (int)((b->Diss[i] - b->dissMin) / dD) + 1;
      if (j > nBin) printf("error in j!");
      if (j >= 0 && j <= nBin + 1) I_D[j]++;
      j = (int)((b->Input[i] - b->inputMin) / dI) + 1;
      if (j > nBin) printf("error in j!");
      if (j >= 0 && j <= nBin + 1) I_I[j]++;
    }
    for (int ib = 0; ib < nBin; ib++)
      fprintf(b->pdfs, "%.15e %.15e %.15e %.15e %.15e %.15e \n", b->enerMin + (ib - 0.5) * dE,
              I_E[ib] / ((b->nStats - 1) * dE), b->dissMin + (ib - 0.5) * dD, I_D[ib] / ((b->nStats - 1) * dD),
              b->inputMin + (ib - 0.5) * dI, I_I[ib] / ((b->nStats - 1) * dI));
    fflush(b->pdfs);
  }
  printf("pdfs and averages done \n");
  fflush(stdout);
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
  const int nout = *NOUT > 0 ? *NOUT : 1, nvort = *NOUTV > 0 ? *NOUTV : 1, nstop = *NSTOP;
  ns_book book;
  memset(&book, 0, sizeof book);
  if (*statsFLAG != 0) {
    char p[64];
    snprintf(p, sizeof p, "JOB%1d/stats.dat", *RANK);
    fs = fopen(p, "w");
    snprintf(p, sizeof p, "JOB%1d/engy.dat", *RANK);
    fe = fopen(p, "w");
    snprintf(p, sizeof p, "JOB%1d/avgs.dat", *RANK);
    book.avgs = fopen(p, "w");
    snprintf(p, sizeof p, "JOB%1d/means.dat", *RANK);
    book.means = fopen(p, "w");
    snprintf(p, sizeof p, "JOB%1d/pdfs.dat", *RANK);
    book.pdfs = fopen(p, "w");
    book.nStats = nstop / nout + 1;
    book.Diss = calloc((size_t)book.nStats, sizeof(double));
    book.Energy = calloc((size_t)book.nStats, sizeof(double));
    book.Input = calloc((size_t)book.nStats, sizeof(double));
    if (!fs || !fe || !book.avgs || !book.means || !book.pdfs || !book.Diss || !book.Energy || !book.Input ||
        psg_stats_create(ctx, 0, &st)) {
      if (!fs || !fe || !book.avgs || !book.means || !book.pdfs)
        fprintf(stderr, "psgpu legacy: cannot open the JOB%d statistics files\n", *RANK);
      else lg_fail("statistics");
      if (fs) fclose(fs);
      if (fe) fclose(fe);
      fs = fe = NULL;
    }
  }
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
    /* avgVelocity at the top of every NOUT-th step, with the computeStats count so far */
    if (fs && nt % nout == 0 && psg_stats_physical(st, PSG_STP_RMS, tcount, NULL, NULL, NULL, NULL)) {
      lg_fail("statistics");
      break;
    }
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
    while (e < nstop &&
           !((e % nvort == 0 && nstop > 1) || (fs && e % nout == 0) || (hf && (e - 1) % nout == 0 && e - 1 != 0)))
      e++;
    if (psg_run(ctx, e - nt)) { lg_fail("run"); break; }
    tme += (e - nt) * *DELT;
    nt = e;
    if (hf && (nt - 1) % nout == 0 && nt - 1 != 0) {
      if (upo.rec && lg_upo_update(&upo, nt - 1, *DELT, &tme)) { lg_fail("recurrence search"); break; }
      if (fs && (*statsFLAG == 1 || (*statsFLAG == 2 && tme > 500.0))) {
        if (stats_line(st, tcount, *V2, *KFY, tme, fs, fe, &book, nt - 1, nout)) { lg_fail("statistics"); break; }
        tcount++;
      }
    }
  }
  *Time = *TSTART + nstop * *DELT;
  if (fs && *statsFLAG == 1 && nstop > 1) {
    /* finalise.h:16-21: NT++ past the loop's end, then a last computeStats */
    if (stats_line(st, tcount, *V2, *KFY, tme, fs, fe, &book, nstop + 1, nout) ||
        ns_finalise(st, &book, *Time, *NX, *NY, *NZ))
      lg_fail("statistics");
    tcount++;
  }
  if (psg_get_state(ctx, 0, buf) == 0) lg_unpack(buf, nkt, 3, f);
  else lg_fail("download");
  if (fs) fclose(fs);
  if (fe) fclose(fe);
  if (book.avgs) fclose(book.avgs);
  if (book.means) fclose(book.means);
  if (book.pdfs) fclose(book.pdfs);
  free(book.Diss); free(book.Energy); free(book.Input);
  if (st) psg_stats_destroy(st);
  lg_upo_close(&upo);
  free(phys); free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, double *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
                    int *, int *, int *, int *, int *, int *) __attribute__((alias("timestep_cuda_ns3d_dns")));
