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
 * on the command line.  Arc-length continuation and Arnoldi/Floquet analysis
 * are not part of this path and are rejected, as is a type L record with
 * RsymFLAG set (the reflection symmetry is not built).
 *
 * usage: psg_gmres ns2d|ns3d NX [NY NZ ALPHA] [--dir D] [--jv fd|tl|tl32] [--maxnewton K] [--noshift]
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

static const char *status_text[] = {"   \"CONVERGED\"   ", "\"GMRES FAILED\"  ", "\"H-STEP FAILED\" ",
                                    "\"T-STEP FAILED\" ", "\"NEWTON STALLED\" ", "\"PERIOD TOO SMALL\" ",
                                    "\"NEWTON FAILED\" "};

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

static int die(const char *what) {
  fprintf(stderr, "psg_gmres: %s\n", what);
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 3)
    return die("usage: psg_gmres ns2d|ns3d NX [NY NZ ALPHA] [--dir D] [--jv fd|tl|tl32] [--maxnewton K] [--noshift] "
               "[--list]");
  const int is2d = strcmp(argv[1], "ns2d") == 0;
  if (!is2d && strcmp(argv[1], "ns3d") != 0) return die("flavour must be ns2d or ns3d");
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
  if (psgh_params_read(path, is2d ? PSGH_PARAMS_NS2D_GMRES : PSGH_PARAMS_NS3D_GMRES, &kv, msg, sizeof msg))
    return die(msg);
  static psgh_kv kx;   /* paramsEXTRA.txt (paramsEXTRA_read_txt) */
  snprintf(path, sizeof path, "%s/paramsEXTRA.txt", dir);
  if (psgh_params_read(path, is2d ? PSGH_EXTRA_NS2D : PSGH_EXTRA_NS3D, &kx, msg, sizeof msg)) return die(msg);
  const double Re = psgh_kv_float(&kv, "Re", 40.0), delt = psgh_kv_float(&kv, "DELT", 0.005);
  const long irest = psgh_kv_int(&kv, "IREST", 1);
  const long iguessend = psgh_kv_int(&kx, "IGUESSEND", irest);
  const char *infile = psgh_kv_str(&kv, "infile"), *itype = psgh_kv_str(&kv, "infile type");
  const int family = is2d ? PSGH_REC_NS2D : PSGH_REC_NS3D;
  if (!infile || !psgh_restart_known(family, itype)) return die("infile type unknown for this flavour");
  if (psgh_kv_int(&kx, "arclcFLAG", 0) || psgh_kv_int(&kx, "ArnoldiFLAG", 0))
    return die("arc-length continuation and Arnoldi analysis are not part of this path");
  if (irest < 1) return die("IREST must name the first guess record (>= 1)");

  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = is2d ? PSG_NS2D : PSG_NS3D;
  c.precision = PSG_FP64;
  c.batch = 1;
  c.nx = nx; c.ny = ny; c.nz = is2d ? 1 : nz; c.alpha = alpha;
  c.v2 = 1.0 / Re;
  c.dt = delt;
  if (is2d) {
    c.force = PSG_FORCE_KOLMO; c.kf = 4.0; c.ampfor = -2.0;   /* NZKX -= 2 at (0,+-4), kernels.h:264-266 */
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
  const int nsc = is2d ? 6 : 7;      /* scalars per output record */
  const long ncpx = (is2d ? 1 : 3) * nkt;

  snprintf(path, sizeof path, "%s/%s", dir, infile);
  psgh_restart_file *fin = psgh_restart_open(path, family, itype, (int)iktx, (int)ikty, (int)iktz, is2d ? 1 : 3, nkt,
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
  fprintf(finfo, "Guess No : Start time : guess Re :guess Period :guess Shiftx:guess Shifty:%s guess Res  :      "
                 "Status     : UPO No :   Re   : Period : Shift x : Shift y :%s Residual     : #NEWTON : #GMRES : "
                 "#HOOK \n", is2d ? "" : "guess Shiftz:", is2d ? "" : " Shift z :");
  psgh_zk_header(fupo, 1, (int)iktx, (int)ikty, is2d ? 0 : (int)iktz);
  psgh_zk_header(fzk, 1, (int)iktx, (int)ikty, is2d ? 0 : (int)iktz);

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
    /* the guess's Re: the record's v2 where its type carries one (v2in = v2, main.F90:688-695) */
    const double v2in = rr.has_v2 ? rr.v2 : 1.0 / Re;
    if (!list && (!ctx || c.v2 != v2in)) {
      if (ctx) psg_destroy(ctx);
      c.v2 = v2in;
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
    const double in_sc[7] = {tin, v2in, Tin, sxin, o.sy, is2d ? nerin : szin, nerin};
    psgh_zk_record(fzk, in_sc, nsc, z, ncpx);
    psg_newton_result r;
    if (psg_newton_solve(ctx, z, &T, &sx, &sz, &o, &r)) {
      fprintf(stderr, "psg_gmres: %s\n", psg_last_error());
      rc = 1;
      break;
    }
    double out[7] = {0.0, c.v2, T, sx, o.sy, 0.0, 0.0};
    if (is2d) out[5] = r.ner;
    else { out[5] = sz; out[6] = r.ner; }
    psgh_zk_record(fzk, out, nsc, z, ncpx);
    char e1[64], e2[64], e3[64];
    fort_e(e1, 16, 6, tin);
    fort_e(e2, 14, 5, nerin);
    fort_e(e3, 14, 5, r.ner);
    if (is2d)
      fprintf(finfo, "%5ld    %s%10.5f%12.5f%12.5f%12.5f%s : %s : %4d%12.5f%10.5f%10.5f%10.5f%s%6d%6d%6d\n", ig, e1,
              1.0 / v2in, Tin, sxin, syin, e2, status_text[r.status], nupo + 1, 1.0 / c.v2, T, sx, o.sy, e3,
              r.newton_iters, r.gmres_iters, r.hook_iters);
    else
      fprintf(finfo, "%5ld    %s%10.5f%12.5f%12.5f%12.5f%12.5f%s : %s : %4d%12.5f%10.5f%10.5f%10.5f%10.5f%s%6d%6d%6d\n",
              ig, e1, 1.0 / v2in, Tin, sxin, syin, szin, e2, status_text[r.status], nupo + 1, 1.0 / c.v2, T, sx, o.sy,
              sz, e3, r.newton_iters, r.gmres_iters, r.hook_iters);
    fflush(finfo);
    if (r.status == PSG_NEWTON_CONVERGED) {
      psgh_zk_record(fupo, out, nsc, z, ncpx);
      fflush(fupo);
      nupo++;
    }
    printf("guess %ld: %s NER=%.3e T=%.8f sx=%.6f (%d Newton, %d GMRES, %.1f s)\n", ig, status_text[r.status], r.ner,
           T, sx, r.newton_iters, r.gmres_iters, r.seconds);
    fflush(stdout);
  }
  psgh_restart_close(fin);
  fclose(finfo);
  fclose(fupo);
  fclose(fzk);
  free(z);
  if (ctx) psg_destroy(ctx);
  printf("%d orbit(s) written to UPOs.out\n", nupo);
  return rc;
}
