/* Shared helpers of the legacy drop-in shims (one translation unit per
 * flavour library includes this once). */
#ifndef PSG_LEGACY_COMMON_H
#define PSG_LEGACY_COMMON_H
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/psg.h"
#include "../../../include/psg_legacy.h"
#include "../../../include/psg_rec.h"

/* One cached context per library: reused while flavour, sizes, viscosity and
 * forcing are unchanged; dt changes go through psg_set_dt. */
static psg_ctx *lg_ctx = NULL;
static psg_config lg_cfg;

static void lg_fail(const char *where) {
  fprintf(stderr, "psgpu legacy %s: %s\n", where, psg_last_error());
  fflush(stderr);
}

static psg_ctx *lg_context(const psg_config *c) {
  if (lg_ctx && lg_cfg.flavour == c->flavour && lg_cfg.nx == c->nx && lg_cfg.ny == c->ny && lg_cfg.nz == c->nz &&
      lg_cfg.alpha == c->alpha && lg_cfg.v2 == c->v2 && lg_cfg.force == c->force && lg_cfg.kf == c->kf &&
      lg_cfg.ampfor == c->ampfor && lg_cfg.precision == c->precision && lg_cfg.batch == c->batch &&
      lg_cfg.ri == c->ri && lg_cfg.sc == c->sc && lg_cfg.theta == c->theta) {
    if (lg_cfg.dt != c->dt) {
      if (psg_set_dt(lg_ctx, c->dt)) { lg_fail("set_dt"); return NULL; }
      lg_cfg.dt = c->dt;
    }
    return lg_ctx;
  }
  if (lg_ctx) psg_destroy(lg_ctx);
  lg_ctx = NULL;
  if (psg_create(c, &lg_ctx)) { lg_fail("create"); lg_ctx = NULL; return NULL; }
  lg_cfg = *c;
  return lg_ctx;
}

/* dt changes inside a call go through here so the cache key stays true */
static int lg_set_dt(psg_ctx *ctx, double dt) {
  if (psg_set_dt(ctx, dt)) return -1;
  lg_cfg.dt = dt;
  return 0;
}

static int lg_precision(void) {
  const char *p = getenv("PSG_PRECISION");
  return (p && (strcmp(p, "fp32") == 0 || strcmp(p, "fp32bulk") == 0)) ? PSG_FP32BULK : PSG_FP64;
}

/* gather separate field arrays into one [nf][nkt] buffer and back */
static void lg_pack(double *buf, long nkt, int nf, psg_double2 **f) {
  for (int i = 0; i < nf; i++) memcpy(buf + 2 * nkt * i, f[i], sizeof(double) * 2 * nkt);
}
static void lg_unpack(const double *buf, long nkt, int nf, psg_double2 **f) {
  for (int i = 0; i < nf; i++) memcpy(f[i], buf + 2 * nkt * i, sizeof(double) * 2 * nkt);
}

/* Recurrence-search outputs of the 3D DNS callers (3D_NavierStokes/src/init.h:
 * 192-201 opens them, functions.h:364-382 writes a record): JOB<rank>/UPO_ZK.out,
 * a C stream of records (time, v2, period, sx, sy, sz, residual, then the fields),
 * and JOB<rank>/UPOinfo_ts.dat, one text line per record after a header. */
typedef struct {
  psg_rec *rec;
  FILE *zk, *info;
  double *state;
  int nf;
  long nkt;
} lg_upo;

static int lg_upo_open(lg_upo *u, psg_ctx *ctx, int rank, int nf, long nkt, int nout, double tstart, double thr) {
  memset(u, 0, sizeof *u);
  psg_rec_opts o;
  if (psg_rec_defaults(ctx, &o)) return -1;
  o.nout = nout;
  o.tstart = tstart;
  o.threshold = thr;
  if (psg_rec_create(ctx, &o, &u->rec)) return -1;
  char p[64];
  snprintf(p, sizeof p, "JOB%1d/UPO_ZK.out", rank);
  u->zk = fopen(p, "wb");
  snprintf(p, sizeof p, "JOB%1d/UPOinfo_ts.dat", rank);
  u->info = fopen(p, "w");
  u->nf = nf;
  u->nkt = nkt;
  u->state = malloc(sizeof(double) * 2 * nf * nkt);
  if (!u->zk || !u->info || !u->state) {
    fprintf(stderr, "psgpu legacy: cannot open JOB%d/UPO_ZK.out / UPOinfo_ts.dat\n", rank);
    return -1;
  }
  fprintf(u->info, "Guess No : Start time : guess Period : guess Shift x: guess Shift y: guess Residual \n");
  return 0;
}

/* the NOUT block's check-and-push; *tme takes the check's time as the reference's
 * recurrence_check sets it (functions.h:280) */
static int lg_upo_update(lg_upo *u, long nt, double dt, double *tme) {
  psg_rec_guess g;
  psg_rec_check c;
  const int r = psg_rec_update(u->rec, nt, dt, &g, u->state, &c);
  if (r < 0) return r;
  if (c.checked) *tme = c.time;
  if (r == 1) {
    const double h[7] = {g.time, g.v2, g.period, g.sx, g.sy, g.sz, g.residual};
    fwrite(h, sizeof(double), 7, u->zk);
    fwrite(u->state, sizeof(double), (size_t)2 * u->nf * u->nkt, u->zk);
    printf("Wrote to UPO_Zk.out number %d at time %e \n", g.index, *tme);
    fprintf(u->info, "%d %e %e %e %e %e %e \n", g.index, g.time, g.period, g.sx, g.sy, g.sz, g.residual);
  }
  fflush(u->info);
  fflush(u->zk);
  return 0;
}

static void lg_upo_close(lg_upo *u) {
  if (u->rec) psg_rec_destroy(u->rec);
  if (u->zk) fclose(u->zk);
  if (u->info) fclose(u->info);
  free(u->state);
  memset(u, 0, sizeof *u);
}

void set_gpu_(int *go) {
  (void)go;
  int n = psg_device_count();
  if (n == 0) {
    printf("psgpu error: no HIP devices\n");
    exit(-1);
  }
  printf("\n Using HIP device 0 of %d (gfx950)\n\n", n);
}
#endif
