/* Drop-in timestep_cuda_ for the 3D Boussinesq DNS driver
 * (3D_Boussinesq/src/cuda.cu:21-117, called from 3D_Boussinesq/src/main.F90).
 *
 * Kept from the reference:
 *  - time step: nStop = Tstep/DELT, dt = Tstep/nStop unless nStop*DELT == Tstep
 *    (init.h:24-29); steps run while tme <= Tstep with tme from TSTART
 *    (cuda.cu:68-99), *Time advanced by dt per step;
 *  - output times: the step that first reaches a multiple of Tout = NOUT*dt
 *    (tcheck < 0, out_low.h:5-8) and of Vout = NOUTV*dt (out_low.h:79-80);
 *  - adaptive dt (adaptFLAG = 1, at Tout crossings with statsFLAG != 0):
 *    dt = max(C/|Umax|, dt0) with C = 2 pi/NX and Umax = max |u_x|
 *    (init.h:31-34, out_low.h:10-15);
 *  - full-field dumps JOB<rank>/{U,V,W,rho}NNN.raw (NR doubles, x fastest) at
 *    Vout crossings when statsFLAG != 0 and nStop > 1 (out_low.h:79-110);
 *  - with RecFLAG = 1, the recurrence search after every step that started at
 *    a Tout crossing (out_high.h:8-33 with the tcheck of out_low.h:5-6), on the
 *    GPU (psg_rec, Boussinesq rule), guesses to JOB<rank>/UPO_ZK.out and
 *    UPOinfo_ts.dat.  The reference's check also overwrites the loop's time
 *    (functions.h:501 sets the global tme back by NOUT steps, so the same Tout
 *    crossing recurs and the loop does not end); here the loop keeps its time.
 * Not built (reported once on stderr when requested): the forcing throttle
 * (Dtarget != 0), the RiG, buoyancy-flux, statistics, means, pdfs and
 * spectra outputs of out_low.h/out_high.h/finalise.h. */
#include "legacy_common.h"
#include <sys/stat.h>

static const double lg_small = 1e-10; /* 3D_Boussinesq/src/kernels.h:31 */

static int lg_crossing(double tme, double dt, double period) {
  if (!(period > 0.0)) return 0;
  const double crit = period * (double)(int)((tme + lg_small) / period);
  return (tme - dt - crit + lg_small) * (tme + lg_small - crit) < 0.0;
}

void timestep_cuda_bous_dns(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, psg_double2 *rho, double *kx,
                            double *ky, double *kz, double *Time, double *TSTART, double *Tstep, double *AMPFOR,
                            double *DELT, double *ResThresh, double *KFY, double *V2, double *Ri, double *Sc,
                            double *Theta, double *alpha, double *Dtarget, int *ikF, int *ikN, int *LC, int *IKTX,
                            int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NOUT,
                            int *NOUTV, int *statsFLAG, int *RecFLAG, int *adaptFLAG, int *RANK) {
  (void)kx; (void)ky; (void)kz; (void)ikF; (void)ikN; (void)IKTX; (void)IKTY; (void)IKTZ;
  (void)KTZ;
  const double nStop = *Tstep / *DELT;
  double dt = (nStop * *DELT != *Tstep) ? *Tstep / nStop : *DELT;
  const double dt0 = dt, courant = 2.0 * M_PI / *NX;
  const double tout = *NOUT * dt, vout = *NOUTV * dt;
  if (*Dtarget != 0.0 || *statsFLAG != 0) {
    static int warned = 0;
    if (!warned)
      fprintf(stderr, "psgpu legacy Boussinesq: the forcing throttle and the statistics files are not "
                      "produced; stepping, adaptive dt, field dumps and recurrence checks are\n");
    warned = 1;
  }
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = PSG_BOUS3D; c.precision = lg_precision(); c.batch = 1;
  c.nx = *NX; c.ny = *NY; c.nz = *NZ; c.alpha = *alpha; c.v2 = *V2; c.dt = dt;
  c.force = PSG_FORCE_KOLMO; c.kf = *KFY; c.ampfor = *AMPFOR;
  c.ri = *Ri; c.sc = *Sc; c.theta = *Theta;
  psg_ctx *ctx = lg_context(&c);
  if (!ctx) return;
  const long nkt = *NKT, nr = (long)*NX * *NY * *NZ;
  double *buf = malloc(sizeof(double) * 8 * nkt), *phys = NULL;
  if (!buf) { lg_fail("host allocation"); return; }
  psg_double2 *f[4] = {xii, eta, zet, rho};
  lg_pack(buf, nkt, 4, f);
  if (psg_set_tables(ctx, NULL, NULL, LC) || psg_set_state(ctx, 0, buf)) { lg_fail("upload"); free(buf); return; }
  lg_upo upo;
  memset(&upo, 0, sizeof upo);
  if (*RecFLAG == 1) {
    char dir[32];
    snprintf(dir, sizeof dir, "JOB%1d", *RANK);
    mkdir(dir, 0755);
    if (lg_upo_open(&upo, ctx, *RANK, 4, nkt, *NOUT > 0 ? *NOUT : 1, *TSTART, *ResThresh)) {
      lg_fail("recurrence search");
      lg_upo_close(&upo);
    }
  }
  double tme = *TSTART;
  long timestep = 0;
  int count = 0;
  while (tme <= *Tstep) {
    /* the recurrence block follows a step whose start crossed a Tout multiple */
    const int hi = upo.rec && lg_crossing(tme, dt, tout);
    const int at_tout = *statsFLAG != 0 && *adaptFLAG == 1 && lg_crossing(tme, dt, tout);
    int have_phys = 0;
    if (at_tout) {
      if (!phys) phys = malloc(sizeof(double) * 7 * nr);
      if (!phys) { lg_fail("host allocation"); break; }
      if (psg_physical(ctx, 0, phys)) { lg_fail("physical"); break; }
      have_phys = 1;
      double umax = 0.0;
      for (long i = 0; i < nr; i++) umax = fmax(umax, fabs(phys[3 * nr + i]));
      const double nd = fmax(courant / umax, dt0);
      if (nd != dt) {
        if (lg_set_dt(ctx, nd)) { lg_fail("set_dt"); break; }
        dt = nd;
      }
    }
    /* the dump test uses the updated dt, as out_low.h:79-80 follows :10-15 */
    const int at_vout = *statsFLAG != 0 && nStop > 1 && lg_crossing(tme, dt, vout);
    if (at_vout && !have_phys) {
      if (!phys) phys = malloc(sizeof(double) * 7 * nr);
      if (!phys) { lg_fail("host allocation"); break; }
      if (psg_physical(ctx, 0, phys)) { lg_fail("physical"); break; }
    }
    if (at_vout) {
      count++;
      char dir[32];
      snprintf(dir, sizeof dir, "JOB%1d", *RANK);
      mkdir(dir, 0755);
      const char *nm[4] = {"U", "V", "W", "rho"};
      const long off[4] = {3 * nr, 4 * nr, 5 * nr, 6 * nr};
      for (int i = 0; i < 4; i++) {
        char p[64];
        snprintf(p, sizeof p, "JOB%1d/%s%03d.raw", *RANK, nm[i], count);
        FILE *r = fopen(p, "wb");
        if (r) { fwrite(phys + off[i], sizeof(double), (size_t)nr, r); fclose(r); }
      }
    }
    /* steps up to the next step that has work before or after it */
    int n = 1;
    double t = tme + dt;
    while (!hi && t <= *Tstep && n < 1 << 20 &&
           !(*statsFLAG != 0 && ((*adaptFLAG == 1 && lg_crossing(t, dt, tout)) ||
                                 (nStop > 1 && lg_crossing(t, dt, vout)))) &&
           !(upo.rec && lg_crossing(t, dt, tout))) {
      t += dt;
      n++;
    }
    if (psg_run(ctx, n)) { lg_fail("run"); break; }
    for (int i = 0; i < n; i++) { *Time += dt; tme += dt; }
    timestep += n;
    double tcheck_time;
    if (hi && lg_upo_update(&upo, timestep, dt, &tcheck_time)) { lg_fail("recurrence search"); break; }
  }
  lg_upo_close(&upo);
  if (psg_get_state(ctx, 0, buf) == 0) lg_unpack(buf, nkt, 4, f);
  else lg_fail("download");
  free(phys); free(buf);
}

void timestep_cuda_(psg_double2 *, psg_double2 *, psg_double2 *, psg_double2 *, double *, double *, double *,
                    double *, double *, double *, double *, double *, double *, double *, double *, double *,
                    double *, double *, double *, double *, int *, int *, int *, int *, int *, int *, int *, int *,
                    int *, int *, int *, int *, int *, int *, int *, int *, int *)
    __attribute__((alias("timestep_cuda_bous_dns")));
