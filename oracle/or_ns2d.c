/*
 * Oracle 2D Navier-Stokes vorticity stepper: RK4 stages without viscosity
 * plus an end-of-step Crank-Nicolson factor (test infrastructure only; see
 * oracle.h).  Restates
 *   loop        2D_NavierStokes/src_RK4/src_gmres/cuda.cu:44-76
 *   KR_FFT_ALL  2D_NavierStokes/src_RK4/functions.h:33-57 (in-place conj, setFFN, Z2D)
 *   conj        2D_NavierStokes/src_RK4/kernels.h:82-96
 *   setVelocity 2D_NavierStokes/src_RK4/kernels.h:112-147
 *   multReal    2D_NavierStokes/src_RK4/kernels.h:187-201
 *   RK_FFT      2D_NavierStokes/src_RK4/functions.h:137-149 + normFF1 kernels.h:151-168
 *   Step1/Step2/Step  2D_NavierStokes/src_RK4/kernels.h:247-435
 *   diagnostics 2D_NavierStokes/src_RK4/src_gmres/out.F90:2-56, spec.F90:2-57
 * The forcing "NZKX -= 2 at kx=0,|ky|=4" (kernels.h:264-266) is "NZKX += famp" with the
 * wavenumber and amplitude as parameters (kf=4, famp=-2 reproduce it).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or2_ctx {
  or_grid g;
  double v2, dt, famp;
  int kf;
  double *kx, *ky;
  int *L, *ikf, *ikn;
  orc *z0, *z1, *z2, *z3, *uk, *vk, *nzk, *FF;
  double *zr, *ur, *vr, *nzr;
};

or2_ctx *or2_create(int nx, int ny, double alpha, double v2, double dt, int kf, double famp) {
  or2_ctx *c = calloc(1, sizeof(or2_ctx));
  or_grid2(&c->g, nx, ny, alpha);
  c->v2 = v2; c->dt = dt; c->kf = kf; c->famp = famp;
  long nkt = c->g.nkt;
  c->kx = malloc(sizeof(double) * nkt); c->ky = malloc(sizeof(double) * nkt);
  c->L = malloc(sizeof(int) * nkt);
  c->ikf = malloc(sizeof(int) * c->g.n2); c->ikn = malloc(sizeof(int) * nkt);
  or_tables(&c->g, c->kx, c->ky, NULL, c->L);
  or_ikf_ikn(&c->g, c->ikf, c->ikn);
  c->z0 = calloc(nkt, sizeof(orc)); c->z1 = calloc(nkt, sizeof(orc));
  c->z2 = calloc(nkt, sizeof(orc)); c->z3 = calloc(nkt, sizeof(orc));
  c->uk = calloc(nkt, sizeof(orc)); c->vk = calloc(nkt, sizeof(orc));
  c->nzk = calloc(2 * nkt, sizeof(orc));
  c->FF = calloc(2 * c->g.n2, sizeof(orc));
  c->zr = calloc(c->g.nr, sizeof(double)); c->ur = calloc(c->g.nr, sizeof(double));
  c->vr = calloc(c->g.nr, sizeof(double)); c->nzr = calloc(2 * c->g.nr, sizeof(double));
  return c;
}

void or2_destroy(or2_ctx *c) {
  if (!c) return;
  free(c->kx); free(c->ky); free(c->L); free(c->ikf); free(c->ikn);
  free(c->z0); free(c->z1); free(c->z2); free(c->z3); free(c->uk); free(c->vk);
  free(c->nzk); free(c->FF); free(c->zr); free(c->ur); free(c->vr); free(c->nzr);
  free(c);
}

const or_grid *or2_grid(const or2_ctx *c) { return &c->g; }

static void conj2(const or2_ctx *c, orc *ZK) {
  for (long ik = 0; ik < c->g.nkt; ik++) {
    if (c->kx[ik] == 0.0 && c->ky[ik] < 0.0) {
      long iky = ik / c->g.iktx;
      long ikk = (c->g.ikty - iky - 1) * c->g.iktx;
      ZK[ik].re = ZK[ikk].re; ZK[ik].im = -ZK[ikk].im;
    }
  }
}

static void kr2_one(or2_ctx *c, orc *V, double *out) {
  conj2(c, V);
  for (long ip = 0; ip < c->g.n2; ip++) {
    int ikk = c->ikf[ip];
    if (ikk < 0) { c->FF[ip].re = 0.0; c->FF[ip].im = 0.0; } else c->FF[ip] = V[ikk];
  }
  or_fft2_c2r(c->FF, out, c->g.nx, c->g.ny);
}

static void rhs2(or2_ctx *c) {
  kr2_one(c, c->z0, c->zr);
  kr2_one(c, c->uk, c->ur);
  kr2_one(c, c->vk, c->vr);
  long nr = c->g.nr;
  for (long i = 0; i < nr; i++) { c->nzr[i] = c->vr[i] * c->zr[i]; c->nzr[nr + i] = c->ur[i] * c->zr[i]; }
  or_fft2_r2c(c->nzr, c->FF, c->g.nx, c->g.ny);
  or_fft2_r2c(c->nzr + nr, c->FF + c->g.n2, c->g.nx, c->g.ny);
  double in = 1.0 / (double)nr;
  for (long ik = 0; ik < c->g.nkt; ik++) {
    orc a = c->FF[c->ikn[ik]], b = c->FF[c->ikn[ik] + c->g.n2];
    c->nzk[ik].re = a.re * in; c->nzk[ik].im = a.im * in;
    c->nzk[c->g.nkt + ik].re = b.re * in; c->nzk[c->g.nkt + ik].im = b.im * in;
  }
}

static inline void vel2(or2_ctx *c, long ik, double kkx, double kky, double wki) {
  c->uk[ik].re = -kky * c->z0[ik].im * wki;
  c->uk[ik].im = kky * c->z0[ik].re * wki;
  c->vk[ik].re = kkx * c->z0[ik].im * wki;
  c->vk[ik].im = -kkx * c->z0[ik].re * wki;
}

/* stage = 0: Step1 with half increment, 1: Step2 with full increment, 2: Step (final) */
static void step2(or2_ctx *c, orc *Z, orc *Zs, int stage) {
  long nkt = c->g.nkt;
  for (long ik = 0; ik < nkt; ik++) {
    if (c->L[ik] == 1) {
      double kky = c->ky[ik], kkx = c->kx[ik];
      double wk = kkx * kkx + kky * kky;
      double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
      double NZKX = kkx * c->nzk[ik + nkt].im + kky * c->nzk[ik].im;
      double NZKY = -kkx * c->nzk[ik + nkt].re - kky * c->nzk[ik].re;
      if (kkx == 0.0 && fabs(kky) == (double)c->kf) NZKX += c->famp;
      double ZX = c->dt * NZKX, ZY = c->dt * NZKY;
      if (stage < 2) {
        Zs[ik].re = ZX; Zs[ik].im = ZY;
        double f = stage == 0 ? 0.5 : 1.0;
        if (stage == 0) { c->z0[ik].re = Z[ik].re + 0.5 * ZX; c->z0[ik].im = Z[ik].im + 0.5 * ZY; }
        else { c->z0[ik].re = Z[ik].re + ZX; c->z0[ik].im = Z[ik].im + ZY; }
        (void)f;
      } else {
        double damp1 = 1 - 0.5 * c->dt * c->v2 * wk;
        double damp2 = 1. / (1 + 0.5 * c->dt * c->v2 * wk);
        Z[ik].re = damp2 * (damp1 * Z[ik].re + (c->z1[ik].re + 2 * c->z2[ik].re + 2 * c->z3[ik].re + ZX) / 6.);
        Z[ik].im = damp2 * (damp1 * Z[ik].im + (c->z1[ik].im + 2 * c->z2[ik].im + 2 * c->z3[ik].im + ZY) / 6.);
        c->z0[ik] = Z[ik];
      }
      vel2(c, ik, kkx, kky, wki);
    } else {
      c->uk[ik].re = c->uk[ik].im = 0.0;
      c->vk[ik].re = c->vk[ik].im = 0.0;
      Z[ik].re = Z[ik].im = 0.0;
    }
    c->nzk[ik].re = c->nzk[ik].im = 0.0;
    c->nzk[ik + nkt].re = c->nzk[ik + nkt].im = 0.0;
  }
}

static void set_velocity2(or2_ctx *c, const orc *Z) {
  for (long ik = 0; ik < c->g.nkt; ik++) {
    if (c->L[ik] == 1) {
      double kkx = c->kx[ik], kky = c->ky[ik];
      double wk = kkx * kkx + kky * kky;
      double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
      c->z0[ik] = Z[ik];
      vel2(c, ik, kkx, kky, wki);
    } else {
      c->uk[ik].re = c->uk[ik].im = 0.0; c->vk[ik].re = c->vk[ik].im = 0.0;
      c->z0[ik].re = c->z0[ik].im = 0.0;
    }
  }
}

void or2_run(or2_ctx *c, orc *z, int nsteps) {
  set_velocity2(c, z);
  for (int nt = 0; nt < nsteps; nt++) {
    rhs2(c); step2(c, z, c->z1, 0);
    rhs2(c); step2(c, z, c->z2, 0);
    rhs2(c); step2(c, z, c->z3, 1);
    rhs2(c); step2(c, z, NULL, 2);
  }
}

/* masked N(z) = -i(kx FFT(u w) + ky FFT(v w)) + forcing, the 2D convol
 * (2D_NavierStokes/src_RK4/src_gmres/convol.F90:2-64) */
void or2_nonlin(or2_ctx *c, const orc *z, orc *nz_) {
  set_velocity2(c, z);
  rhs2(c);
  long nkt = c->g.nkt;
  for (long ik = 0; ik < nkt; ik++) {
    double kkx = c->kx[ik], kky = c->ky[ik];
    double NZKX = kkx * c->nzk[ik + nkt].im + kky * c->nzk[ik].im;
    double NZKY = -kkx * c->nzk[ik + nkt].re - kky * c->nzk[ik].re;
    if (kkx == 0.0 && fabs(kky) == (double)c->kf) NZKX += c->famp;
    nz_[ik].re = c->L[ik] ? NZKX : 0.0;
    nz_[ik].im = c->L[ik] ? NZKY : 0.0;
  }
}

void or2_diag(const or2_ctx *c, const orc *z, double *energy, double *enstrophy, double *spec) {
  /* out.F90's sums in compensated arithmetic */
  or_ksum e = {0, 0}, v = {0, 0};
  const int nb = c->g.kty + 2;
  or_ksum *bins = (or_ksum *)calloc((size_t)nb, sizeof(or_ksum));
  for (long ik = 0; ik < c->g.nkt; ik++) {
    if (!c->L[ik]) continue;
    double wk = c->kx[ik] * c->kx[ik] + c->ky[ik] * c->ky[ik];
    double vz = z[ik].re * z[ik].re + z[ik].im * z[ik].im;
    or_kadd(&e, vz / wk);
    or_kadd(&v, vz);
    int j = (int)(sqrt(wk) + 0.5);
    if (j >= 0 && j < nb) or_kadd(&bins[j], vz / wk);
  }
  if (spec)
    for (int j = 0; j < nb; j++) spec[j] = or_kval(&bins[j]);
  free(bins);
  *energy = or_kval(&e);
  *enstrophy = 2.0 * or_kval(&v);
}
