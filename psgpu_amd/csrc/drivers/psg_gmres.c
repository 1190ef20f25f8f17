/* Newton-GMRES-hookstep driver: the caller side of the reference's periodic-
 * orbit search, restated in C over the C-ABI (include/psg_newton.h).
 *
 * 2D: 2D_NavierStokes/src_RK4/src_gmres/main.F90 (guess loop 160-186, output
 * files 146-153 and 224-225, per-guess records 645-646 / 1744-1745 /
 * 1834-1835, UPOinfo lines and statuses 1245-1820, formats 1501-1502);
 * parameters params_gmres.txt and paramsEXTRA.txt (parameters.F90:140-203).
 * 3D: 3D_NavierStokes/src_gmres/main.F90 (same structure, NETOL 1e-10, the
 * GMRES forcing, formats 1915-1916); parameters params.txt and
 * paramsEXTRA.txt (src_gmres/parameters.F90:140-200).
 *
 * For each guess IGUESS = IREST..IGUESSEND the state, period and shifts are
 * read from record IGUESS of `infile` in the layout its "infile type" names
 * (2D types A-L, 2D_NavierStokes/src_RK4/src_gmres/init.F90:66-106; 3D types
 * A-D, 3D_NavierStokes/src_gmres/init.F90:77-130; psgh_restart_read).  As in
 * the reference (2D main.F90:366-376, 3D main.F90:439-450): the 2D y shift
 * enters with its sign flipped, a fixed period rescales the x shift, a
 * record's v2 replaces the params.txt Re (types that carry it), and a period
 * within 1e-10 of a whole number of steps is snapped to it (main.F90:678-683).
 * The Newton solve runs on the GPU, and the driver writes
 *   UPOinfo.dat  one line per guess: guess, start time, Re, period, shifts and
 *                residual of the guess, status, UPO number, Re, period,
 *                shifts and residual reached, Newton/GMRES/hook counts;
 *   UPOs.out     header + one record per converged orbit;
 *   Zk.out       header + the guess and the final state of every guess.
 * The grid is compiled into the reference (global_params.F90); here it is given
 * on the command line.  With ArnoldiFLAG each guess, taken to be a converged
 * orbit, is analysed for stability instead of solved (floquet_guess,
 * eigenvalues.dat).  With arclcFLAG the first guess starts a pseudo-arclength
 * continuation in Re of IGUESSEND steps (trace_branch).  A type L record with
 * RsymFLAG set is rejected (the reflection symmetry is not built).
 *
 * Boussinesq: 3D_Boussinesq/src/src_gmres/main.F90 (guess loop 222-1860,
 * NETOL 1e-10, GMTOL 1e-7, trust region 0.05, the x and z shifts as unknowns;
 * Theta in degrees, params_update parameters.F90:30-47); records (0, v1, Ri,
 * alpha, T, sx, sy, sz, NER, xi, eta, zeta, rho) in UPOs.out after a header
 * and in Zk.out, which holds the last iterate only (main.F90:1739-1744,
 * 1834-1838); UPOinfo.dat lines in format 1502 (main.F90:1876-1877), whose
 * "Re" columns carry the ParFLAG parameter (v1, Ri or alpha, main.F90:262-264).
 * Its ArnoldiFLAG analysis takes the 3D settings and eigenvalues.dat layout
 * (main.F90:69-70,430,968-983, the same as 3D Navier-Stokes) with
 * finite-difference Jv.  Continuation is not built for this flavour.
 *
 * usage: psg_gmres ns2d|ns3d|bous NX [NY NZ ALPHA] [--dir D] [--jv fd|tl|tl32] [--maxnewton K] [--noshift]
 * (--jv tl: tangent-linear Jv; tl32: the same in FP32-bulk passes under the
 * FP64 residual; --noshift: the x shift is not an unknown, for equilibria and
 * other shift-free solutions; --list: print each guess as read and adjusted,
 * with no GPU work)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "psg.h"
#include "psg_host.h"
#include "psg_newton.h"

/* between " : " and " : " these give the reference's status fields, e.g.
 * ' :   "CONVERGED"   : ' and ' : "GMRES FAILED"  : ' (main.F90:1245-1810) */
static const char *status_text[] = {"  \"CONVERGED\"  ", "\"GMRES FAILED\" ", "\"H-STEP FAILED\"",
                                    "\"T-STEP FAILED\"", "\"NEWTON STALLED\"", "\"PERIOD TOO SMALL\"",
                                    "\"NEWTON FAILED\""};

/* Fortran Ew.d: [-]0.ddddddE+xx right-justified in w characters (out: 64 bytes) */
static void fort_e(char *out, int w, int d, double v) {
  char buf[64];
  if (v == 0.0) {
    snprintf(buf, sizeof buf, "0.%0*dE+00", d, 0);
  } else {
    int e = (int)floor(log10(fabs(v))) + 1;
    double m = fabs(v) / pow(10.0, e);
    if (m >= 1.0 - 0.5 * pow(10.0, -d)) { m /= 10.0; e += 1; }
    snprintf(buf, sizeof buf, "%s0.%0*.0f%c%c%02d", v < 0 ? "-" : "", d, m * pow(10.0, d), 'E', e < 0 ? '-' : '+',
             abs(e));
  }
  snprintf(out, 64, "%*s", w, buf);
}

/* Floquet analysis of guess ig (ArnoldiFLAG): the leading multipliers mu of
 * S dPhi_T and, appended to eigenvalues.dat, the exponents log|mu|/T,
 * atan2(mu_im, mu_re)/T of the converged ones (Ritz estimate < 100 TOL) with
 * positive real part, in the flavour's layout:
 *   2D  "UPO  n", then E13.5,E13.5 pairs, then 1/v2 and the largest exponent
 *       (2D_NavierStokes/src_RK4/src_gmres/main.F90:926-952; nev 10, ncv 200,
 *       TOL 1e-7, which 'LR', main.F90:63-67,388)
 *   3D  I3,1x,F13.5,F13.5 rows "i lr li", a blank row, then IREST, their sum
 *       and T (3D_NavierStokes/src_gmres/main.F90:983-1005; nev 30, ncv 120,
 *       TOL 1e-6, which 'LM', main.F90:64-68,461) */
static int floquet_guess(psg_ctx *ctx, int is2d, long ig, const double *z, double T, double sx, double sz, double sy,
                         double delt, int jv, double v2, const char *dir) {
  psg_floquet_opts fo;
  psg_floquet_defaults(&fo);
  fo.nev = is2d ? 10 : 30;
  fo.ncv = is2d ? 200 : 120;
  fo.tol = is2d ? 1e-7 : 1e-6;
  fo.which = is2d ? PSG_FLOQUET_LR : PSG_FLOQUET_LM;
  fo.jv = jv;
  fo.dt0 = delt;
  fo.time_policy = is2d ? PSG_TIME_UNIFORM : PSG_TIME_LASTSTEP;
  fo.sy = sy;
  double mr[30], mi[30], er[30];
  psg_floquet_result fr;
  if (psg_floquet(ctx, z, T, sx, sz, &fo, mr, mi, er, NULL, &fr)) {
    fprintf(stderr, "psg_gmres: %s\n", psg_last_error());
    return 1;
  }
  char path[1024];
  snprintf(path, sizeof path, "%s/eigenvalues.dat", dir);
  FILE *f = fopen(path, "a");
  if (!f) { fprintf(stderr, "psg_gmres: cannot open %s\n", path); return 1; }
  if (is2d) fprintf(f, " UPO %12ld\n", ig);
  double maxeig = 0.0, sum = 0.0;
  int cnt = 0;
  for (int i = 0; i < fo.nev; i++) {
    if (!(er[i] >= 0.0 && er[i] < 100.0 * fo.tol)) continue;
    const double lr = log(sqrt(mr[i] * mr[i] + mi[i] * mi[i])) / T, li = atan2(mi[i], mr[i]) / T;
    if (!(lr > 0.0)) continue;
    if (is2d) {
      char a[64], b[64];
      fort_e(a, 13, 5, lr);
      fort_e(b, 13, 5, li);
      fprintf(f, "%s%s\n", a, b);
    } else {
      fprintf(f, "%3d %13.5f%13.5f\n", ++cnt, lr, li);
    }
    if (lr > maxeig) maxeig = lr;
    sum += lr;
  }
  if (is2d) {
    char a[64], b[64];
    fort_e(a, 13, 5, 1.0 / v2);
    fort_e(b, 13, 5, maxeig);
    fprintf(f, "%s%s\n", a, b);
  } else {
    fprintf(f, "     \n%3ld %13.5f%13.5f\n", ig, sum, T);
  }
  fclose(f);
  printf("guess %ld: Floquet %d/%d converged in %d Jv (%d cycles, %.1f s); leading mu = %.10g%+.10gi\n", ig,
         fr.nconv, fo.nev, fr.matvecs, fr.cycles, fr.seconds, mr[0], mi[0]);
  fflush(stdout);
  return 0;
}

/* output files and the guess a result came from */
typedef struct {
  FILE *finfo, *fupo, *fzk;
  int is2d, nsc, nupo;
  long ncpx;
  double guess_t, guess_v2, guess_T, guess_sx, guess_sy, guess_sz, guess_ner;
} outputs;

/* the result of one Newton solve: Zk.out record, UPOinfo.dat line, and a
 * UPOs.out record when it converged (main.F90:1744-1745, 1774-1778, 1834-1835) */
static void report(outputs *w, long ig, const psg_newton_result *r, const double *z, double v2, double T, double sx,
                   double sy, double sz) {
  double out[7] = {0.0, v2, T, sx, sy, 0.0, 0.0};
  if (w->is2d) out[5] = r->ner;
  else { out[5] = sz; out[6] = r->ner; }
  psgh_zk_record(w->fzk, out, w->nsc, z, w->ncpx);
  char e1[64], e2[64], e3[64];
  fort_e(e1, 16, 6, w->guess_t);
  fort_e(e2, 14, 5, w->guess_ner);
  fort_e(e3, 14, 5, r->ner);
  if (w->is2d)
    fprintf(w->finfo, "%5ld    %s%10.5f%12.5f%12.5f%12.5f%s : %s : %4d%12.5f%10.5f%10.5f%10.5f%s%6d%6d%6d\n", ig, e1,
            1.0 / w->guess_v2, w->guess_T, w->guess_sx, w->guess_sy, e2, status_text[r->status], w->nupo + 1,
            1.0 / v2, T, sx, sy, e3, r->newton_iters, r->gmres_iters, r->hook_iters);
  else
    fprintf(w->finfo, "%5ld    %s%10.5f%12.5f%12.5f%12.5f%12.5f%s : %s : %4d%12.5f%10.5f%10.5f%10.5f%10.5f%s%6d%6d%6d\n",
            ig, e1, 1.0 / w->guess_v2, w->guess_T, w->guess_sx, w->guess_sy, w->guess_sz, e2, status_text[r->status],
            w->nupo + 1, 1.0 / v2, T, sx, sy, sz, e3, r->newton_iters, r->gmres_iters, r->hook_iters);
  fflush(w->finfo);
  if (r->status == PSG_NEWTON_CONVERGED) {
    psgh_zk_record(w->fupo, out, w->nsc, z, w->ncpx);
    fflush(w->fupo);
    w->nupo++;
  }
  if (w->is2d)   /* NER2: the low-wavenumber residual the 2D solver prints (main.F90:737-751) */
    printf("guess %ld: %s NER=%.3e NER2=%.3e Re=%.8g T=%.8f sx=%.6f (%d Newton, %d GMRES, %.1f s)\n", ig,
           status_text[r->status], r->ner, r->ner2, 1.0 / v2, T, sx, r->newton_iters, r->gmres_iters, r->seconds);
  else
    printf("guess %ld: %s NER=%.3e Re=%.8g T=%.8f sx=%.6f (%d Newton, %d GMRES, %.1f s)\n", ig,
           status_text[r->status], r->ner, 1.0 / v2, T, sx, r->newton_iters, r->gmres_iters, r->seconds);
  fflush(stdout);
}

/* the Boussinesq outputs of one Newton solve (3D_Boussinesq/src/src_gmres/main.F90:
 * 1739-1744 Zk.out, rewritten in place; 1754-1810 UPOinfo.dat, format 1502;
 * 1827-1838 UPOs.out); par: the ParFLAG parameter printed in the "Re" columns */
static void report_bous(outputs *w, long ig, const psg_newton_result *r, const double *z, double v1, double ri,
                        double alpha, double par, double T, double sx, double sy, double sz) {
  const double out[9] = {0.0, v1, ri, alpha, T, sx, sy, sz, r->ner};
  rewind(w->fzk);
  psgh_zk_record(w->fzk, out, 9, z, w->ncpx);
  fflush(w->fzk);
  char e1[64], e2[64], e3[64], e4[64], e5[64];
  fort_e(e1, 16, 6, w->guess_t);
  fort_e(e2, 16, 6, w->guess_v2);
  fort_e(e3, 14, 5, w->guess_ner);
  fort_e(e4, 16, 6, par);
  fort_e(e5, 14, 5, r->ner);
  fprintf(w->finfo, "%5ld    %s%s%12.5f%12.5f%12.5f%12.5f%s : %s : %4d%s%10.5f%10.5f%10.5f%10.5f%s%6d%6d%6d\n", ig, e1,
          e2, w->guess_T, w->guess_sx, w->guess_sy, w->guess_sz, e3, status_text[r->status], w->nupo + 1, e4, T, sx,
          sy, sz, e5, r->newton_iters, r->gmres_iters, r->hook_iters);
  fflush(w->finfo);
  if (r->status == PSG_NEWTON_CONVERGED) {
    psgh_zk_record(w->fupo, out, 9, z, w->ncpx);
    fflush(w->fupo);
    w->nupo++;
  }
  printf("guess %ld: %s NER=%.3e T=%.8f sx=%.6f sz=%.6f (%d Newton, %d GMRES, %.1f s)\n", ig, status_text[r->status],
         r->ner, T, sx, sz, r->newton_iters, r->gmres_iters, r->seconds);
  fflush(stdout);
}

/* masked real inner product of two states (the Newton vectors' product) */
static double sdot(const double *a, const double *b, const int *L, long nkt, int nf) {
  double s = 0.0;
  for (int f = 0; f < nf; f++)
    for (long i = 0; i < nkt; i++)
      if (L[i]) s += a[2 * (f * nkt + i)] * b[2 * (f * nkt + i)] + a[2 * (f * nkt + i) + 1] * b[2 * (f * nkt + i) + 1];
  return s;
}

/* Pseudo-arclength continuation of a converged orbit in v2 = 1/Re, as the
 * reference's arclcFLAG branch of its guess loop
 * (2D_NavierStokes/src_RK4/src_gmres/main.F90:166-183, 408-604, 1776-1803):
 * the first step changes Re by delta_v2 at fixed v2 (halving it after a
 * failure); each later step predicts X = (z, T, sx, v2) along the secant (first
 * order, or second order through the last three points), with step Ndelta_r
 * doubled after the first success, then grown by 1.1 up to its maximum after
 * successes and halved after failures, and corrects with the arclength row.
 * Every step writes its Zk.out record and UPOinfo.dat line, and each point
 * of the branch a UPOs.out record.  Steps ig .. ig + nsteps. */
static int trace_branch(psg_ctx *ctx, outputs *w, psg_newton_opts *o, const psgh_kv *kx, const int *L, long ig0,
                        long nsteps, const double *z0, double T0, double s0, double v20) {
  const int nf = w->is2d ? 1 : 3;
  const long n2 = 2 * w->ncpx, nkt = w->ncpx / nf;
  int FirstOrder = (int)psgh_kv_int(kx, "1stOrder", 0), FullSecond = (int)psgh_kv_int(kx, "Full 2ndOrder", 0);
  int drN_reset = (int)psgh_kv_int(kx, "drN_reset", 0), FixedP = o->fixed_period;
  const double FixedPsize = psgh_kv_float(kx, "Fixed Period Size", T0);
  double delta_v2 = psgh_kv_float(kx, "delta_v2", 0.0);
  double *buf = malloc(sizeof(double) * n2 * 6);
  if (!buf) { fprintf(stderr, "psg_gmres: out of memory\n"); return 1; }
  double *Zr0 = buf, *Zr1 = buf + n2, *Zr2 = buf + 2 * n2, *dZ = buf + 3 * n2, *ZO = buf + 4 * n2, *tmp = buf + 5 * n2;
  double Tr0 = T0, Tr1 = T0, Tr2 = T0, Sr0 = s0, Sr1 = s0, Sr2 = s0, v2r0 = v20, v2r1 = v20, v2r2 = v20;
  double delta_r1 = 0.0, delta_r2 = 0.0, Ndelta_r = 0.0, Ndelta_rmax = 0.0;
  int CONVERGED = 1, FirstRun = 1, SecondRun = 0, rc = 0;
  double v2 = v20, T = T0, S = s0;
  memcpy(Zr0, z0, sizeof(double) * n2);
  memcpy(Zr1, z0, sizeof(double) * n2);
  memcpy(Zr2, z0, sizeof(double) * n2);
  for (long ig = ig0; ig <= ig0 + nsteps && rc == 0; ig++) {
    psg_alc a;
    memset(&a, 0, sizeof a);
    int use_alc = 0;
    if (FirstRun) {   /* a fixed change in Re from the guess (main.F90:412-428) */
      if (CONVERGED) v2 = v2 - delta_v2;
      else { v2 = v2 + delta_v2; delta_v2 /= 2.0; v2 = v2 - delta_v2; }
      memcpy(ZO, z0, sizeof(double) * n2);
      memcpy(Zr2, z0, sizeof(double) * n2);
      T = T0; S = s0;
      if (!FixedP) Tr2 = T0;
      Sr2 = s0;
      v2r2 = v2 + delta_v2;
      CONVERGED = 0;
    } else {
      if (Tr2 < FixedPsize) {   /* an equilibrium or travelling wave: fix the period (main.F90:437-450) */
        printf("continuation: period below %g, fixing it (equilibrium or travelling wave)\n", FixedPsize);
        FixedP = 1; o->fixed_period = 1;
        Sr2 = FixedPsize * Sr2 / Tr2; Tr2 = FixedPsize;
        Sr1 = FixedPsize * Sr1 / Tr1; Tr1 = FixedPsize;
        Sr0 = FixedPsize * Sr0 / Tr0; Tr0 = FixedPsize;
      }
      for (long i = 0; i < n2; i++) tmp[i] = Zr2[i] - Zr1[i];
      delta_r2 = sdot(tmp, tmp, L, nkt, nf) + (Sr2 - Sr1) * (Sr2 - Sr1) + (v2r2 - v2r1) * (v2r2 - v2r1);
      if (!FixedP) delta_r2 += (Tr2 - Tr1) * (Tr2 - Tr1);
      delta_r2 = sqrt(delta_r2);
      for (long i = 0; i < n2; i++) dZ[i] = tmp[i] / delta_r2;
      double dT = FixedP ? 0.0 : (Tr2 - Tr1) / delta_r2, dS = (Sr2 - Sr1) / delta_r2, dV = (v2r2 - v2r1) / delta_r2;
      if (SecondRun) {   /* main.F90:467-471 */
        Ndelta_r = CONVERGED ? 2.0 * delta_r2 : Ndelta_r / 2.0;
        Ndelta_rmax = Ndelta_r;
      } else {
        if (CONVERGED) {
          Ndelta_r = delta_r2 < Ndelta_rmax ? delta_r2 : Ndelta_rmax;
          if (1.1 * Ndelta_r < Ndelta_rmax) Ndelta_r *= 1.1;
        } else {
          Ndelta_r /= 2.0;
        }
      }
      CONVERGED = 0;
      if (SecondRun || FirstOrder) {   /* first-order prediction (main.F90:488-505) */
        for (long i = 0; i < n2; i++) ZO[i] = Zr2[i] + dZ[i] * Ndelta_r;
        T = FixedP ? FixedPsize : Tr2 + dT * Ndelta_r;
        S = Sr2 + dS * Ndelta_r;
        v2 = v2r2 + dV * Ndelta_r;
      } else {   /* second order through r0, r1, r2 (main.F90:509-560) */
        const double q = delta_r2 / (delta_r2 + delta_r1), d1 = delta_r2 * (1.0 - q),
                     d2 = 0.5 * delta_r2 * delta_r2 * (1.0 - 1.0 / q), h = Ndelta_r;
#define O2(x2, x1, x0, o1, o2) \
  do { o1 = ((1.0 - q * q) * (x2) - (x1) + q * q * (x0)) / d1; o2 = ((q - 1.0) * (x2) + (x1) - q * (x0)) / d2; } while (0)
        for (long i = 0; i < n2; i++) {
          double g1, g2;
          O2(Zr2[i], Zr1[i], Zr0[i], g1, g2);
          ZO[i] = Zr2[i] + g1 * h + 0.5 * h * h * g2;
          if (FullSecond) dZ[i] = g1;
        }
        double g1, g2;
        if (FixedP) T = FixedPsize;
        else {
          O2(Tr2, Tr1, Tr0, g1, g2);
          T = Tr2 + g1 * h + 0.5 * h * h * g2;
          if (FullSecond) dT = g1;
        }
        O2(Sr2, Sr1, Sr0, g1, g2);
        S = Sr2 + g1 * h + 0.5 * h * h * g2;
        if (FullSecond) dS = g1;
        O2(v2r2, v2r1, v2r0, g1, g2);
        v2 = v2r2 + g1 * h + 0.5 * h * h * g2;
        if (FullSecond) dV = g1;
#undef O2
      }
      if (drN_reset) {   /* main.F90:567-573 */
        for (long i = 0; i < n2; i++) tmp[i] = ZO[i] - Zr2[i];
        double d = sdot(tmp, tmp, L, nkt, nf) + (S - Sr2) * (S - Sr2) + (v2 - v2r2) * (v2 - v2r2);
        if (!FixedP) d += (T - Tr2) * (T - Tr2);
        Ndelta_r = sqrt(d);
      }
      a.dzdr = dZ; a.dTdr = dT; a.dsdr = dS; a.dv2dr = dV;
      a.zr = Zr2; a.Tr = Tr2; a.sr = Sr2; a.v2r = v2r2; a.ds = Ndelta_r;
      use_alc = 1;
    }
    /* the solve, from the predicted (or Re-shifted) point */
    w->guess_t = 0.0; w->guess_v2 = v2; w->guess_T = T; w->guess_sx = S; w->guess_sy = o->sy; w->guess_sz = 0.0;
    w->guess_ner = 0.0;
    const double in_sc[7] = {0.0, v2, T, S, o->sy, 0.0, 0.0};
    psgh_zk_record(w->fzk, in_sc, w->nsc, ZO, w->ncpx);
    psg_newton_result r;
    double sz = 0.0;
    int e;
    if (use_alc) {
      e = psg_newton_solve_alc(ctx, ZO, &T, &S, &v2, &a, o, &r);
    } else {
      if ((e = psg_set_v2(ctx, v2)) == 0) e = psg_newton_solve(ctx, ZO, &T, &S, &sz, o, &r);
    }
    if (e) { fprintf(stderr, "psg_gmres: %s\n", psg_last_error()); rc = 1; break; }
    printf("continuation step %ld: %s at Re=%.10g (arclength step %.4g)\n", ig - ig0, status_text[r.status], 1.0 / v2,
           use_alc ? Ndelta_r : 0.0);
    report(w, ig, &r, ZO, v2, T, S, o->sy, sz);
    if (r.status == PSG_NEWTON_CONVERGED) {   /* shift the branch history (main.F90:1779-1802) */
      CONVERGED = 1;
      if (SecondRun) SecondRun = 0;
      if (FirstRun) { FirstRun = 0; SecondRun = 1; }
      memcpy(Zr0, Zr1, sizeof(double) * n2); Tr0 = Tr1; Sr0 = Sr1; v2r0 = v2r1;
      memcpy(Zr1, Zr2, sizeof(double) * n2); Tr1 = Tr2; Sr1 = Sr2; v2r1 = v2r2;
      delta_r1 = delta_r2;
      memcpy(Zr2, ZO, sizeof(double) * n2); Tr2 = T; Sr2 = S; v2r2 = v2;
      delta_r2 = Ndelta_r;
    }
  }
  free(buf);
  return rc;
}

static int die(const char *what) {
  fprintf(stderr, "psg_gmres: %s\n", what);
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 3)
    return die("usage: psg_gmres ns2d|ns3d|bous NX [NY NZ ALPHA] [--dir D] [--jv fd|tl|tl32] [--maxnewton K] "
               "[--noshift] [--list]");
  const int is2d = strcmp(argv[1], "ns2d") == 0, bous = strcmp(argv[1], "bous") == 0;
  if (!is2d && !bous && strcmp(argv[1], "ns3d") != 0) return die("flavour must be ns2d, ns3d or bous");
  int nx = atoi(argv[2]), ny = nx, nz = nx, ai = 3, jv = 0, maxnewton = -1, noshift = 0, list = 0;
  double alpha = 1.0;
  const char *dir = ".";
  if (argc > 5 && argv[3][0] != '-') { ny = atoi(argv[3]); nz = atoi(argv[4]); alpha = atof(argv[5]); ai = 6; }
  for (int i = ai; i < argc; i++) {
    if (!strcmp(argv[i], "--dir") && i + 1 < argc) dir = argv[++i];
    else if (!strcmp(argv[i], "--jv") && i + 1 < argc) {
      const char *f = argv[++i];
      jv = !strcmp(f, "tl") ? 1 : (!strcmp(f, "tl32") ? 2 : 0);
      if (jv == 0 && strcmp(f, "fd")) return die("--jv must be fd, tl or tl32");
    }
    else if (!strcmp(argv[i], "--maxnewton") && i + 1 < argc) maxnewton = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--noshift")) noshift = 1;
    else if (!strcmp(argv[i], "--list")) list = 1;
    else return die("unknown option");
  }
  char path[1024], msg[256];
  snprintf(path, sizeof path, "%s/%s", dir, is2d ? "params_gmres.txt" : "params.txt");
  static psgh_kv kv;
  if (psgh_params_read(path, is2d ? PSGH_PARAMS_NS2D_GMRES : (bous ? PSGH_PARAMS_BOUS_GMRES : PSGH_PARAMS_NS3D_GMRES),
                       &kv, msg, sizeof msg))
    return die(msg);
  static psgh_kv kx;   /* paramsEXTRA.txt (paramsEXTRA_read_txt) */
  snprintf(path, sizeof path, "%s/paramsEXTRA.txt", dir);
  if (psgh_params_read(path, is2d ? PSGH_EXTRA_NS2D : (bous ? PSGH_EXTRA_BOUS : PSGH_EXTRA_NS3D), &kx, msg,
                       sizeof msg))
    return die(msg);
  const double Re = psgh_kv_float(&kv, "Re", 40.0), delt = psgh_kv_float(&kv, "DELT", 0.005);
  const long irest = psgh_kv_int(&kv, "IREST", 1);
  const long iguessend = psgh_kv_int(&kx, "IGUESSEND", irest);
  const char *infile = psgh_kv_str(&kv, "infile"), *itype = psgh_kv_str(&kv, "infile type");
  const int family = is2d ? PSGH_REC_NS2D : (bous ? PSGH_REC_BOUS : PSGH_REC_NS3D);
  if (!infile || !psgh_restart_known(family, itype)) return die("infile type unknown for this flavour");
  const int arclc = (int)psgh_kv_int(&kx, "arclcFLAG", 0);
  const int arnoldi = (int)psgh_kv_int(&kx, "ArnoldiFLAG", 0);
  if (arclc && arnoldi) return die("arclcFLAG and ArnoldiFLAG are exclusive");
  if (bous && arclc) return die("continuation is not built for the Boussinesq flavour");
  if (bous && jv) return die("the Boussinesq flavour takes finite-difference Jv only");
  const int parflag = bous ? (int)psgh_kv_int(&kx, "ParFLAG", 1) : 1;
  if (irest < 1) return die("IREST must name the first guess record (>= 1)");

  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = is2d ? PSG_NS2D : (bous ? PSG_BOUS3D : PSG_NS3D);
  c.precision = PSG_FP64;
  c.batch = 1;
  c.nx = nx; c.ny = ny; c.nz = is2d ? 1 : nz; c.alpha = alpha;
  c.v2 = 1.0 / Re;
  c.dt = delt;
  if (is2d) {
    c.force = PSG_FORCE_KOLMO; c.kf = 4.0; c.ampfor = -2.0;   /* NZKX -= 2 at (0,+-4), kernels.h:264-266 */
  } else if (bous) {
    /* params_update (3D_Boussinesq/src/src_gmres/parameters.F90:30-47): Theta in degrees */
    c.force = PSG_FORCE_KOLMO; c.kf = (double)psgh_kv_int(&kv, "KF", 1);
    c.ampfor = -0.5 * c.kf;
    c.ri = psgh_kv_float(&kv, "Ri", 1.0);
    c.sc = psgh_kv_float(&kv, "Sc", 1.0);
    c.theta = psgh_kv_float(&kv, "Theta", 90.0) * 4.0 * asin(1.0) / 360.0;
  } else {
    c.force = PSG_FORCE_GMRES; c.kf = (double)psgh_kv_int(&kv, "KF", 1);
    c.ampfor = -0.25 * c.kf;                                  /* src_gmres/parameters.F90 params_update */
  }
  long iktx, ikty, iktz, nkt, nact;
  if (psg_sizes(&c, &iktx, &ikty, &iktz, &nkt, &nact)) return die(psg_last_error());
  psg_ctx *ctx = NULL;
  psg_newton_opts o;
  psg_newton_defaults(c.flavour, &o);
  o.jv = jv;
  o.dt0 = delt;
  o.fixed_period = (int)psgh_kv_int(&kx, "Fixed Period", 0);
  if (maxnewton > 0) o.max_newton = maxnewton;
  if (noshift) o.nshift = 0;
  const int nsc = is2d ? 6 : 7;      /* scalars per output record (Boussinesq: 9, report_bous) */
  const int nf = is2d ? 1 : (bous ? 4 : 3);
  const long ncpx = nf * nkt;

  snprintf(path, sizeof path, "%s/%s", dir, infile);
  psgh_restart_file *fin = psgh_restart_open(path, family, itype, (int)iktx, (int)ikty, (int)iktz, nf, nkt,
                                             msg, sizeof msg);
  if (!fin) return die(msg);
  double *z = malloc(sizeof(double) * 2 * ncpx);
  if (!z) return die("out of memory");
  snprintf(path, sizeof path, "%s/UPOinfo.dat", dir);
  FILE *finfo = fopen(path, "w");
  snprintf(path, sizeof path, "%s/UPOs.out", dir);
  FILE *fupo = fopen(path, "wb");
  snprintf(path, sizeof path, "%s/Zk.out", dir);
  FILE *fzk = fopen(path, "wb");
  if (!finfo || !fupo || !fzk) return die("cannot create output files");
  if (bous)   /* format 1501, 3D_Boussinesq/src/src_gmres/main.F90:169-172 */
    fprintf(finfo, "Guess No : Start time : guess Re :guess Period :guess Shiftx: guess Shifty: guess Shiftz: guess "
                   "Res  :      Status     : UPO No :   Re   : Period : Shift x : Shift y : Shift z : Residual     : "
                   "#NEWTON : #GMRES : #HOOK \n");
  else
    fprintf(finfo, "Guess No : Start time : guess Re :guess Period :guess Shiftx:guess Shifty:%s guess Res  :      "
                   "Status     : UPO No :   Re   : Period : Shift x : Shift y :%s Residual     : #NEWTON : #GMRES : "
                   "#HOOK \n", is2d ? "" : "guess Shiftz:", is2d ? "" : " Shift z :");
  psgh_zk_header(fupo, 1, (int)iktx, (int)ikty, is2d ? 0 : (int)iktz);
  if (!bous) psgh_zk_header(fzk, 1, (int)iktx, (int)ikty, is2d ? 0 : (int)iktz);
  outputs out;
  memset(&out, 0, sizeof out);
  out.finfo = finfo; out.fupo = fupo; out.fzk = fzk; out.is2d = is2d; out.nsc = nsc; out.ncpx = ncpx;
  int *Lmask = NULL;   /* the active mask, for the continuation's branch distances */
  if (arclc) {
    psgh_grid hg;
    if (is2d) psgh_grid2(&hg, nx, ny, alpha);
    else psgh_grid3(&hg, nx, ny, nz, alpha);
    Lmask = malloc(sizeof(int) * nkt);
    if (!Lmask) return die("out of memory");
    psgh_tables(&hg, NULL, NULL, NULL, Lmask);
  }

  int nupo = 0, rc = 0;
  long at = 0;   /* records consumed */
  psgh_restart rr;
  memset(&rr, 0, sizeof rr);
  for (long ig = irest; ig <= iguessend && rc == 0; ig++) {
    while (at < ig) {
      rr.has_v2 = 0;
      if (psgh_restart_next(fin, z, &rr, msg, sizeof msg)) {
        fprintf(stderr, "psg_gmres: %s\n", msg);
        rc = 1;
        break;
      }
      at++;
    }
    if (rc) break;
    if (rr.rsym != 0.0) {
      fprintf(stderr, "psg_gmres: record %ld asks for the reflection symmetry (RsymFLAG), which is not built\n", ig);
      rc = 1;
      break;
    }
    /* the guess's Re: the record's v2 where its type carries one (v2in = v2, main.F90:688-695);
     * Boussinesq records may also carry Ri and alpha (types F-H) */
    const double v2in = rr.has_v2 ? rr.v2 : 1.0 / Re;
    const double riin = (bous && rr.has_ri) ? rr.ri : c.ri, alin = (bous && rr.has_alpha) ? rr.alpha : c.alpha;
    if (!list && (!ctx || c.v2 != v2in || c.ri != riin || c.alpha != alin)) {
      if (ctx) psg_destroy(ctx);
      c.v2 = v2in;
      c.ri = riin;
      c.alpha = alin;
      if (psg_create(&c, &ctx)) return die(psg_last_error());
    }
    const double tin = rr.time, Tin = rr.period, sxin = rr.shiftx, syin = rr.shifty, szin = rr.shiftz;
    const double nerin = rr.ner;
    double T = Tin, sx = sxin, sz = szin;
    o.sy = is2d ? -syin : syin;            /* SNEWTY = -SNEWTYin in 2D, main.F90:370 */
    if (o.fixed_period) {                  /* main.F90:373-376 */
      const double fp = psgh_kv_float(&kx, "Fixed Period Size", Tin);
      sx = Tin != 0.0 ? fp * sx / Tin : sx;
      T = fp;
    }
    /* a guess period from the DNS is made exactly a whole number of steps (main.F90:678-683) */
    if (fabs(T - nearbyint(T / delt) * delt) < 1e-10) T = nearbyint(T / delt) * delt;
    if (list) {   /* the guess as the solver would start from it; no GPU work */
      printf("guess %ld: t=%.17g v2=%.17g T=%.17g sx=%.17g sy=%.17g sz=%.17g NER=%.17g\n", ig, tin, v2in, T, sx, o.sy,
             sz, nerin);
      continue;
    }
    if (arnoldi) {   /* stability of the guess (a converged orbit) instead of a Newton solve */
      if (floquet_guess(ctx, is2d, ig, z, T, sx, sz, o.sy, delt, jv, c.v2, dir)) { rc = 1; break; }
      continue;
    }
    if (arclc) {   /* trace the branch from this guess instead of solving it once */
      out.guess_v2 = v2in; out.guess_T = Tin; out.guess_sx = sxin; out.guess_sy = syin; out.guess_sz = szin;
      out.guess_t = tin; out.guess_ner = nerin;
      if (trace_branch(ctx, &out, &o, &kx, Lmask, ig, iguessend, z, T, sx, v2in)) rc = 1;
      break;
    }
    /* the parameter the Boussinesq solver prints as "Re" (ParFLAG, main.F90:262-264) */
    const double par = parflag == 2 ? c.ri : (parflag == 3 ? c.alpha : c.v2);
    if (!bous) {
      const double in_sc[7] = {tin, v2in, Tin, sxin, o.sy, is2d ? nerin : szin, nerin};
      psgh_zk_record(fzk, in_sc, nsc, z, ncpx);
    }
    psg_newton_result r;
    if (psg_newton_solve(ctx, z, &T, &sx, &sz, &o, &r)) {
      fprintf(stderr, "psg_gmres: %s\n", psg_last_error());
      rc = 1;
      break;
    }
    out.guess_v2 = bous ? par : v2in; out.guess_T = Tin; out.guess_sx = sxin; out.guess_sy = syin;
    out.guess_sz = szin; out.guess_t = tin; out.guess_ner = nerin;
    if (bous) report_bous(&out, ig, &r, z, c.v2, c.ri, c.alpha, par, T, sx, o.sy, sz);
    else report(&out, ig, &r, z, c.v2, T, sx, o.sy, sz);
  }
  nupo = out.nupo;
  free(Lmask);
  psgh_restart_close(fin);
  fclose(finfo);
  fclose(fupo);
  fclose(fzk);
  free(z);
  if (ctx) psg_destroy(ctx);
  printf("%d orbit(s) written to UPOs.out\n", nupo);
  return rc;
}
