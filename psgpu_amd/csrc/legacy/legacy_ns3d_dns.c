/* Drop-in timestep_cuda_ for the 3D Navier-Stokes DNS driver
 * (3D_NavierStokes/src/cuda.cu:18-151, called from src/main.F90:203-204).
 * NSTOP Heun+CN steps with the xi forcing at (0,KF,+-KF); every NOUTV steps
 * (from step 0, when NSTOP > 1) the physical vorticity is copied to xir, etr,
 * ztr and written to JOB<RANK>/{xii,eta,zet}NNN.raw; with statsFLAG != 0,
 * JOB<RANK>/stats.dat and engy.dat get one line every NOUT steps
 * (computeStats, 3D_NavierStokes/src/functions.h:128-209).  With RecFLAG = 1
 * every NOUT block runs the recurrence search on the GPU (psg_rec: check of the
 * newest history entry, then the push of the current state, cuda.cu:114-139)
 * and writes its guesses to JOB<RANK>/UPO_ZK.out and UPOinfo_ts.dat; as in the
 * reference, a check resets tme (the time column of stats.dat) to
 * TSTART + (NT - NOUT) DELT.  The time-averaged means/pdfs files are not produced. */
#include "legacy_common.h"

static void stats_line(const double *st, long nkt, const double *kx, const double *ky, const double *kz,
                       const int *LL, double *tavg, int *tcount, double v2, double kfy, double ampfor, double tme,
                       FILE *fs, FILE *fe) {
  const double *X = st, *E = st + 2 * nkt, *Z = st + 4 * nkt;
  double energy = 0, diss = 0, input = 0, eturb = 0, dturb = 0, XIa = 0, ETa = 0, ZTa = 0, Ua = 0, Va = 0, Wa = 0;
  double *XT = tavg, *ET = tavg + 2 * nkt, *ZT = tavg + 4 * nkt;
  const double tc = (double)*tcount;
  for (long i = 0; i < nkt; i++) {
    if (LL[i] != 1) continue;
    double wk = kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i];
    wk = wk > 0.001 ? wk : 0.001;
    const double xr = X[2 * i], xi = X[2 * i + 1], er = E[2 * i], ei = E[2 * i + 1], zr = Z[2 * i], zi = Z[2 * i + 1];
    const double vzx = xr * xr + xi * xi, vzy = er * er + ei * ei, vzz = zr * zr + zi * zi;
    const double vz = vzx + vzy + vzz;
    diss += vz;
    energy += vz / wk;
    XIa += vzx; ETa += vzy; ZTa += vzz;
    Ua += pow((ky[i] * zi - kz[i] * ei) / wk, 2) + pow((kz[i] * er - ky[i] * zr) / wk, 2);
    Va += pow((kz[i] * xi - kx[i] * zi) / wk, 2) + pow((kx[i] * zr - kz[i] * xr) / wk, 2);
    Wa += pow((kx[i] * ei - ky[i] * xi) / wk, 2) + pow((ky[i] * xr - kx[i] * er) / wk, 2);
    XT[2 * i] = (XT[2 * i] * tc + xr) / (tc + 1.0); XT[2 * i + 1] = (XT[2 * i + 1] * tc + xi) / (tc + 1.0);
    ET[2 * i] = (ET[2 * i] * tc + er) / (tc + 1.0); ET[2 * i + 1] = (ET[2 * i + 1] * tc + ei) / (tc + 1.0);
    ZT[2 * i] = (ZT[2 * i] * tc + zr) / (tc + 1.0); ZT[2 * i + 1] = (ZT[2 * i + 1] * tc + zi) / (tc + 1.0);
    const double v = XT[2 * i] * XT[2 * i] + XT[2 * i + 1] * XT[2 * i + 1] + ET[2 * i] * ET[2 * i] +
                     ET[2 * i + 1] * ET[2 * i + 1] + ZT[2 * i] * ZT[2 * i] + ZT[2 * i + 1] * ZT[2 * i + 1];
    eturb += v / wk;
    dturb += v;
    if (kx[i] == 0.0 && ky[i] == kfy && kz[i] == kfy) input += ampfor * xr;
    if (kx[i] == 0.0 && ky[i] == kfy && kz[i] == -kfy) input -= ampfor * xr;
  }
  input *= v2 * kfy * kfy;
  diss *= v2 * v2 * 8.0 * kfy * kfy;
  dturb *= v2 * 16.0 * kfy * kfy;
  (*tcount)++;
  fprintf(fs, "%e %e %e %e %e %e \n", tme, energy, eturb, diss, dturb, input);
  fflush(fs);
  fprintf(fe, "%e %e %e %e %e %e %e \n", tme, Ua, Va, Wa, XIa, ETa, ZTa);
  fflush(fe);
}

void timestep_cuda_ns3d_dns(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *xir, double *etr,
                            double *ztr, double *nuzn, double *nu1, double *kx, double *ky, double *kz,
                            double *Time, double *TSTART, double *AMPFOR, double *DELT, double *ResThresh,
                            double *KFY, double *V2, double *alpha, int *ikF, int *ikN, int *LC, int *IKTX,
                            int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NSTOP,
                            int *NOUT, int *NOUTV, int *statsFLAG, int *RecFLAG, int *RANK) {
  (void)ikF; (void)ikN; (void)IKTX; (void)IKTY; (void)IKTZ; (void)KTZ;
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
  double *tavg = NULL;
  int tcount = 0;
  if (*statsFLAG != 0) {
    char p[64];
    snprintf(p, sizeof p, "JOB%1d/stats.dat", *RANK);
    fs = fopen(p, "w");
    snprintf(p, sizeof p, "JOB%1d/engy.dat", *RANK);
    fe = fopen(p, "w");
    tavg = calloc(6 * nkt, sizeof(double));
    if (!fs || !fe) { fprintf(stderr, "psgpu legacy: cannot open JOB%d/stats.dat\n", *RANK); }
    if (!tavg) {
      lg_fail("host allocation");
      if (fs) fclose(fs);
      if (fe) fclose(fe);
      free(buf);
      return;
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
      if (fs) {
        if (psg_get_state(ctx, 0, buf)) { lg_fail("download"); break; }
        stats_line(buf, nkt, kx, ky, kz, LC, tavg, &tcount, *V2, *KFY, *AMPFOR, tme, fs, fe);
      }
    }
  }
  if (psg_get_state(ctx, 0, buf) == 0) lg_unpack(buf, nkt, 3, f);
  else lg_fail("download");
  *Time = *TSTART + nstop * *DELT;
  if (fs) fclose(fs);
  if (fe) fclose(fe);
  lg_upo_close(&upo);
  free(tavg); free(phys); free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, double *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
                    int *, int *, int *, int *, int *, int *) __attribute__((alias("timestep_cuda_ns3d_dns")));
