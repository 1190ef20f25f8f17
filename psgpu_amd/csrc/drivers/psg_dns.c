/* DNS driver: the caller side of the reference's time-stepping runs, restated
 * in C over the legacy drop-in entry points (include/psg_legacy.h).
 *
 *   ns3d   3D_NavierStokes/src/main.F90:37-220: JOB<rank>/params.txt
 *          (parameters.F90:129-156, params_update 31-35), JOB<rank>/Zk.out header,
 *          loop IREST = ISTART..ISTOP of INIT + IKF/IKN + TIMESTEP_CUDA.
 *   euler  3D_Euler/src/main.F90:37-185: params.txt and Zk.out in the run
 *          directory, no viscosity or forcing.
 *   bous   3D_Boussinesq/src/main.F90:43-230: JOB<rank>/params.txt
 *          (parameters.F90:118-140, params_update: AMPFOR = -KF/2, Theta in
 *          degrees converted to radians), stepping while t <= TSTEP.
 *   ns2d   2D_NavierStokes/src_RK4/main.F90:1-305: params.txt
 *          (parameters.F90:75-101), one run of INT(TSTEP/DELT) RK4 steps, or
 *          of one period PERIODin of a restart record when UPOflag = 1; the
 *          final state, and with UPOflag = 1 its shifted copy and the
 *          residual ||F||/||Z||, appended to Zk.out.
 *
 * INIT (3D_NavierStokes/src/init.F90:25-230 and the flavours' counterparts):
 * wavenumbers and the mask from psgh_tables, the CN tables from psgh_nutab,
 * and with IREST = 0 the flavour's random initial condition; with IREST > 0
 * record IREST of `infile` (infile type D, the Zk.out layout, same grid).
 * The time stepping, the JOB<rank>/ statistics and raw dumps are the legacy
 * entry point's.  The grid, which the reference compiles in
 * (global_params.F90), is given on the command line; rank 0.
 *
 * usage: psg_dns ns3d|euler|bous NX [NY NZ ALPHA] [--dir D]
 *        psg_dns ns2d NX [NY ALPHA] [--dir D]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "psg.h"
#include "psg_host.h"
#include "psg_legacy.h"

static int die(const char *what) {
  fprintf(stderr, "psg_dns: %s\n", what);
  return 1;
}

static int run_ns2d(int nx, int ny, double alpha) {
  char msg[256];
  static psgh_kv kv;
  if (psgh_params_read("params.txt", PSGH_PARAMS_NS2D_DNS, &kv, msg, sizeof msg)) return die(msg);
  const int ISEED = (int)psgh_kv_int(&kv, "ISEED", 1), IREST = (int)psgh_kv_int(&kv, "IREST", 0);
  const int UPOflag = (int)psgh_kv_int(&kv, "UPOflag", 0);
  int STFLAG = (int)psgh_kv_int(&kv, "STFLAG", 0), RCFLAG = 0;
  const char *infile = psgh_kv_str(&kv, "infile"), *itype = psgh_kv_str(&kv, "infile type");
  double v2 = 1.0 / psgh_kv_float(&kv, "Re", 1.0), DELT = psgh_kv_float(&kv, "DELT", 0.01);
  double TSTEP = psgh_kv_float(&kv, "TSTEP", 1.0), ResThresh = psgh_kv_float(&kv, "ResidualThreshold", 1.0);
  const double AMPV = 1.0;   /* params_update */
  int NOUT = (int)(psgh_kv_float(&kv, "outfreq1", DELT) / DELT);
  if (NOUT == 0) NOUT = 1;

  psgh_grid g;
  if (psgh_grid2(&g, nx, ny, alpha)) return die("bad grid");
  const long nkt = g.nkt;
  double *kx = malloc(sizeof(double) * nkt), *ky = malloc(sizeof(double) * nkt);
  double *z = calloc(2 * nkt, sizeof(double)), *z0 = calloc(2 * nkt, sizeof(double));
  double *zr = malloc(sizeof(double) * (long)nx * ny);
  int *L = malloc(sizeof(int) * nkt), *ikf = malloc(sizeof(int) * (long)g.nx2 * ny), *ikn = malloc(sizeof(int) * nkt);
  if (!kx || !ky || !z || !z0 || !zr || !L || !ikf || !ikn) return die("out of memory");
  int IKTX = g.iktx, IKTY = g.ikty, KTY = g.kty, NX = nx, NY = ny;

  psgh_rng_seed(ISEED);   /* TIME = RANNO(ISEED) */
  FILE *fzk = fopen("Zk.out", "wb");
  if (!fzk) return die("cannot create Zk.out");
  psgh_zk_header(fzk, 1, IKTX, IKTY, 0);
  double TSTART = TSTEP, PERIODin = 0, SHIFTXin = 0, SHIFTYin = 0, NERin = 0;
  psgh_tables(&g, kx, ky, NULL, L);
  if (IREST == 0) {
    psgh_ic_ns2d_ref(&g, ISEED, AMPV, z);
  } else {
    /* 2D_NavierStokes/src_RK4/init.F90:41-106: record IREST of the infile type's layout */
    psgh_restart rr;
    memset(&rr, 0, sizeof rr);
    if (psgh_restart_read(infile ? infile : "", PSGH_REC_NS2D, itype, IREST, IKTX, IKTY, 0, 1, nkt, z, &rr, msg,
                          sizeof msg))
      return die(msg);
    printf(" Just read time= %.15g\n", rr.time);
    PERIODin = rr.period;
    SHIFTXin = rr.shiftx;
    SHIFTYin = rr.shifty;
    NERin = rr.ner;
    if (rr.has_v2) v2 = rr.v2;
    for (long i = 0; i < nkt; i++)
      if (L[i] != 1) z[2 * i] = z[2 * i + 1] = 0.0;
  }
  if (UPOflag == 1) TSTEP = PERIODin;
  const double TIME = TSTART;
  for (long i = 0; i < nkt; i++)
    if (L[i]) { z0[2 * i] = z[2 * i]; z0[2 * i + 1] = z[2 * i + 1]; }
  int NIT = (int)(TSTEP / DELT);
  psgh_ikf_ikn(&g, ikf, ikn);
  if (UPOflag == 1) {
    if (PERIODin <= 0.0) return die("UPOflag = 1 needs a restart record with a period");
    NIT = (int)(PERIODin / DELT);
    DELT = PERIODin / (double)NIT;
    TSTEP = PERIODin;
  }
  const clock_t c0 = clock();
  timestep_cuda_ns2d_dns((psg_double2 *)z, zr, kx, ky, &TSTEP, &TSTART, &DELT, &alpha, &v2, &SHIFTXin, &ResThresh,
                         ikf, ikn, &IKTX, &IKTY, &KTY, &NX, &NY, &NIT, &NOUT, L, &RCFLAG, &STFLAG);
  printf(" CPU time required for main loop = %7.3f s\n", (double)(clock() - c0) / CLOCKS_PER_SEC);
  const double sc[6] = {TIME, v2, TSTEP, SHIFTXin, 0.0, 0.0};
  psgh_zk_record(fzk, sc, 6, z, nkt);
  if (UPOflag == 1) {
    /* shift the final state back and compare with the initial one (main.F90:243-290) */
    double *zs = calloc(2 * nkt, sizeof(double)), nz2 = 0, nf2 = 0;
    if (!zs) return die("out of memory");
    for (long i = 0; i < nkt; i++) {
      if (!L[i]) continue;
      const double ph = -(kx[i] * SHIFTXin + ky[i] * SHIFTYin), c = cos(ph), s = sin(ph);
      const double a = z[2 * i], b = z[2 * i + 1];
      zs[2 * i] = c * a - s * b;
      zs[2 * i + 1] = s * a + c * b;
      nz2 += a * a + b * b;
      const double fr = zs[2 * i] - z0[2 * i], fi = zs[2 * i + 1] - z0[2 * i + 1];
      nf2 += fr * fr + fi * fi;
    }
    const double NER = sqrt(nf2) / sqrt(nz2);
    printf(" Newton convergence check: ||F|| / ||Z|| = %.15g\n", NER);
    const double sc2[6] = {TIME, v2, 0.0, 0.0, 0.0, 0.0};
    psgh_zk_record(fzk, sc2, 6, zs, nkt);
    if (NER != NERin) printf(" WARNING: calculated ||F||/||Z|| does not match GMRES\n");
    free(zs);
  }
  fclose(fzk);
  printf(" Finished\n");
  free(kx); free(ky); free(z); free(z0); free(zr); free(L); free(ikf); free(ikn);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) return die("usage: psg_dns ns3d|euler|bous NX [NY NZ ALPHA] [--dir D] | ns2d NX [NY ALPHA] [--dir D]");
  const char *fl = argv[1];
  const int ns3d = !strcmp(fl, "ns3d"), euler = !strcmp(fl, "euler"), bous = !strcmp(fl, "bous");
  const int ns2d = !strcmp(fl, "ns2d");
  if (!ns3d && !euler && !bous && !ns2d) return die("flavour must be ns3d, euler, bous or ns2d");
  int nx = atoi(argv[2]), ny = nx, nz = nx, ai = 3;
  double alpha = 1.0;
  if (ns2d && argc > 4 && argv[3][0] != '-') { ny = atoi(argv[3]); alpha = atof(argv[4]); ai = 5; }
  if (!ns2d && argc > 5 && argv[3][0] != '-') { ny = atoi(argv[3]); nz = atoi(argv[4]); alpha = atof(argv[5]); ai = 6; }
  for (int i = ai; i < argc; i++) {
    if (!strcmp(argv[i], "--dir") && i + 1 < argc) {
      if (chdir(argv[++i])) return die("cannot enter the run directory");
    } else {
      return die("unknown option");
    }
  }
  if (ns2d) return run_ns2d(nx, ny, alpha);
  int RANK = 0;
  const char *simdir = euler ? "." : "JOB0";
  char path[1024], msg[256];
  snprintf(path, sizeof path, "%s/params.txt", simdir);
  static psgh_kv kv;
  const int layout = ns3d ? PSGH_PARAMS_NS3D_DNS : (euler ? PSGH_PARAMS_EULER : PSGH_PARAMS_BOUS_DNS);
  if (psgh_params_read(path, layout, &kv, msg, sizeof msg)) return die(msg);
  const int ISEED = (int)psgh_kv_int(&kv, "ISEED", 1);
  const long ISTART = psgh_kv_int(&kv, "IREST", 0), ISTOP = psgh_kv_int(&kv, "ISTOP", ISTART);
  const char *infile = psgh_kv_str(&kv, "infile"), *itype = psgh_kv_str(&kv, "infile type");
  double DELT = psgh_kv_float(&kv, "DELT", 0.01), TSTEP = psgh_kv_float(&kv, "TSTEP", 1.0);
  double AMPV = psgh_kv_float(&kv, "AMPV", 1.0), ResThresh = psgh_kv_float(&kv, "ResidualThreshold", 1.0);
  const int KF = (int)psgh_kv_int(&kv, "KF", 4);
  double v2 = euler ? 0.0 : 1.0 / psgh_kv_float(&kv, "Re", 1.0);
  double AMPFOR = ns3d ? -0.25 * KF * KF : (bous ? -0.5 * KF : 0.0);
  double Ri = psgh_kv_float(&kv, "Ri", 0.0), Sc = psgh_kv_float(&kv, "Sc", 1.0);
  double Theta = psgh_kv_float(&kv, "Theta", 0.0) * (4.0 * asin(1.0)) / 360.0;
  double Dtarget = psgh_kv_float(&kv, "Dtarget", 0.0);
  int NSTOP = (int)(TSTEP / DELT), NOUT = (int)(psgh_kv_float(&kv, "outfreq1", DELT) / DELT);
  int NOUTV = (int)(psgh_kv_float(&kv, "outfreq2", DELT) / DELT);
  if (NOUT == 0) NOUT = 1;
  if (NOUTV == 0) NOUTV = 1;
  int STATSFLAG = (int)psgh_kv_int(&kv, "STATSFLAG", 0), RCFLAG = (int)psgh_kv_int(&kv, "RCFLAG", 0);
  int ADAPTFLAG = (int)psgh_kv_int(&kv, "ADAPTflag", 0);
  const int UPOflag = (int)psgh_kv_int(&kv, "UPOflag", 0);
  double FK = KF;

  psgh_grid g;
  if (psgh_grid3(&g, nx, ny, nz, alpha)) return die("bad grid");
  const int nf = bous ? 4 : 3;
  const long nkt = g.nkt, nr = (long)nx * ny * nz;
  double *kx = malloc(sizeof(double) * nkt), *ky = malloc(sizeof(double) * nkt), *kz = malloc(sizeof(double) * nkt);
  double *nu = malloc(sizeof(double) * nkt), *nuzn = malloc(sizeof(double) * nkt), *nu1 = malloc(sizeof(double) * nkt);
  int *L = malloc(sizeof(int) * nkt);
  int *ikf = malloc(sizeof(int) * (long)g.nx2 * ny * nz), *ikn = malloc(sizeof(int) * nkt);
  double *state = malloc(sizeof(double) * 2 * nf * nkt);
  double *phys = malloc(sizeof(double) * 3 * nr);
  if (!kx || !ky || !kz || !nu || !nuzn || !nu1 || !L || !ikf || !ikn || !state || !phys) return die("out of memory");
  int IKTX = g.iktx, IKTY = g.ikty, IKTZ = g.iktz, KTZ = g.ktz, NKT = (int)nkt, NX = nx, NY = ny, NZ = nz;

  psgh_rng_seed(ISEED);   /* TIME = RANNO(ISEED) */
  snprintf(path, sizeof path, "%s/Zk.out", simdir);
  FILE *fzk = fopen(path, "wb");
  if (!fzk) return die("cannot create Zk.out");
  psgh_zk_header(fzk, 1, IKTX, IKTY, IKTZ);
  fclose(fzk);

  for (long irest = ISTART; irest <= ISTOP; irest++) {
    memset(state, 0, sizeof(double) * 2 * nf * nkt);
    psgh_tables(&g, kx, ky, kz, L);
    if (ns3d) psgh_nutab(&g, v2, DELT, nu, nuzn, nu1);
    double *xii = state, *eta = state + 2 * nkt, *zet = state + 4 * nkt, *rho = state + 6 * nkt;
    if (irest == 0) {
      if (ns3d) psgh_ic_ns3d_ref(&g, ISEED, AMPV, xii, eta, zet);
      else if (euler) psgh_ic_euler_ref(&g, ISEED, AMPV, xii, eta, zet);
      else psgh_ic_bous_ref(&g, ISEED, AMPV, xii, eta, zet, rho);
    } else {
      /* record IREST of the infile type's layout: 3D_NavierStokes/src/init.F90:66-147,
       * 3D_Euler/src/init.F90:66-118, 3D_Boussinesq/src/init.F90:46-182; the record's
       * v2 (and for the Boussinesq F-H types Ri and alpha) replace the params.txt values,
       * and with UPOflag = 1 its period is the run length (main.F90:68-76) */
      psgh_restart rr;
      memset(&rr, 0, sizeof rr);
      snprintf(path, sizeof path, "%s/%s", simdir, infile ? infile : "");
      const int fam = ns3d ? PSGH_REC_NS3D : (euler ? PSGH_REC_EULER : PSGH_REC_BOUS);
      if (psgh_restart_read(path, fam, itype, irest, IKTX, IKTY, IKTZ, nf, nkt, state, &rr, msg, sizeof msg))
        return die(msg);
      printf(" Just read time= %.15g\n", rr.time);
      if (rr.has_v2 && !euler) v2 = rr.v2;
      if (rr.has_ri) Ri = rr.ri;
      if (rr.has_alpha && rr.alpha != alpha) {
        alpha = rr.alpha;
        if (psgh_grid3(&g, nx, ny, nz, alpha)) return die("bad grid");
        psgh_tables(&g, kx, ky, kz, L);
      }
      if (ns3d) psgh_nutab(&g, v2, DELT, nu, nuzn, nu1);
      if (UPOflag == 1) {
        TSTEP = rr.period;
        NSTOP = (int)(TSTEP / DELT);
      }
    }
    psgh_ikf_ikn(&g, ikf, ikn);
    double TIME = 0.0, TSTART = 0.0;
    const clock_t c0 = clock();
    if (ns3d)
      timestep_cuda_ns3d_dns((psg_double2 *)xii, (psg_double2 *)eta, (psg_double2 *)zet, phys, phys + nr,
                             phys + 2 * nr, nuzn, nu1, kx, ky, kz, &TIME, &TSTART, &AMPFOR, &DELT, &ResThresh, &FK,
                             &v2, &alpha, ikf, ikn, L, &IKTX, &IKTY, &IKTZ, &KTZ, &NKT, &NX, &NY, &NZ, &NSTOP,
                             &NOUT, &NOUTV, &STATSFLAG, &RCFLAG, &RANK);
    else if (euler)
      timestep_cuda_euler((psg_double2 *)xii, (psg_double2 *)eta, (psg_double2 *)zet, phys, phys + nr,
                          phys + 2 * nr, kx, ky, kz, &TIME, &TSTART, &DELT, &alpha, ikf, ikn, L, &IKTX, &IKTY,
                          &IKTZ, &KTZ, &NKT, &NX, &NY, &NZ, &NSTOP, &NOUT, &NOUTV, &STATSFLAG);
    else
      timestep_cuda_bous_dns((psg_double2 *)xii, (psg_double2 *)eta, (psg_double2 *)zet, (psg_double2 *)rho, kx,
                             ky, kz, &TIME, &TSTART, &TSTEP, &AMPFOR, &DELT, &ResThresh, &FK, &v2, &Ri, &Sc, &Theta,
                             &alpha, &Dtarget, ikf, ikn, L, &IKTX, &IKTY, &IKTZ, &KTZ, &NKT, &NX, &NY, &NZ, &NOUT,
                             &NOUTV, &STATSFLAG, &RCFLAG, &ADAPTFLAG, &RANK);
    printf(" CPU time required for main loop = %7.3f s\n", (double)(clock() - c0) / CLOCKS_PER_SEC);
    printf(" IREST %ld: t = %.6f\n", irest, TIME);
    /* the final state, for restarts and checks (the reference keeps it in memory only) */
    snprintf(path, sizeof path, "%s/final_%03ld.out", simdir, irest);
    FILE *ff = fopen(path, "wb");
    if (ff) {
      psgh_zk_header(ff, 1, IKTX, IKTY, IKTZ);
      const double sc[7] = {TIME, v2, 0.0, 0.0, 0.0, 0.0, 0.0};
      psgh_zk_record(ff, sc, 7, state, (long)nf * nkt);
      fclose(ff);
    }
  }
  printf(" Finished\n");
  free(kx); free(ky); free(kz); free(nu); free(nuzn); free(nu1); free(L); free(ikf); free(ikn);
  free(state); free(phys);
  return 0;
}
