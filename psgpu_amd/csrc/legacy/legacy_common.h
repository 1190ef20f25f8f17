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
