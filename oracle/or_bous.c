/*
 * Oracle 3D Boussinesq stepper: RK4 with per-stage Crank-Nicolson factors,
 * buoyancy with tilted gravity and a Leray projection of the vorticity after
 * each step (test infrastructure only; see oracle.h).  Restates
 *   loop        3D_Boussinesq/src/cuda.cu:65-101 (fixed step count here; the
 *               while(tme<=Tstep) count is the caller's)
 *   KR_FFT_ALL  3D_Boussinesq/src/functions.h:60-125 (w, u, and i k rho via setFFK)
 *   setFFK      3D_Boussinesq/src/kernels.h:75-93
 *   multReal    3D_Boussinesq/src/kernels.h:298-328 (U,V,W <- -(u x w), rx <- u.grad rho)
 *   RK_FFT      3D_Boussinesq/src/functions.h:129-165
 *   setVelocity 3D_Boussinesq/src/kernels.h:194-240
 *   Step1/Step2/Step  3D_Boussinesq/src/kernels.h:331-621
 *   RHS (dz/dt) 3D_Boussinesq/src/kernels.h:624-692
 * theta is in radians, as the reference passes Theta straight to sin/cos
 * (3D_Boussinesq/src/init.h:66-67).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orb_ctx {
  or_grid g;
  double v2, sc, ri, sinth, costh, dt, kf, ampfor;
  double *kx, *ky, *kz;
  int *L, *ikf, *ikn;
  orc *s0[4], *st[3][4], *uk[4], *tmp, *FF;   /* stage input, stage increments, N (UK,VK,WK,RK) */
  double *phys[9];
};

orb_ctx *orb_create(int nx, int ny, int nz, double alpha, double v2, double sc, double ri, double theta,
                    double dt, double kf, double ampfor) {
  orb_ctx *c = calloc(1, sizeof(orb_ctx));
  or_grid3(&c->g, nx, ny, nz, alpha);
  c->v2 = v2; c->sc = sc; c->ri = ri; c->sinth = sin(theta); c->costh = cos(theta);
  c->dt = dt; c->kf = kf; c->ampfor = ampfor;
  long n = c->g.nkt;
  c->kx = malloc(sizeof(double) * n); c->ky = malloc(sizeof(double) * n); c->kz = malloc(sizeof(double) * n);
  c->L = malloc(sizeof(int) * n);
  c->ikf = malloc(sizeof(int) * c->g.n2); c->ikn = malloc(sizeof(int) * n);
  or_tables(&c->g, c->kx, c->ky, c->kz, c->L);
  or_ikf_ikn(&c->g, c->ikf, c->ikn);
  for (int f = 0; f < 4; f++) {
    c->s0[f] = calloc(n, sizeof(orc));
    c->uk[f] = calloc(n, sizeof(orc));
    for (int s = 0; s < 3; s++) c->st[s][f] = calloc(n, sizeof(orc));
  }
  c->tmp = calloc(n, sizeof(orc));
  c->FF = calloc(c->g.n2, sizeof(orc));
  for (int i = 0; i < 9; i++) c->phys[i] = calloc(c->g.nr, sizeof(double));
  return c;
}

void orb_destroy(orb_ctx *c) {
  if (!c) return;
  free(c->kx); free(c->ky); free(c->kz); free(c->L); free(c->ikf); free(c->ikn);
  for (int f = 0; f < 4; f++) {
    free(c->s0[f]); free(c->uk[f]);
    for (int s = 0; s < 3; s++) free(c->st[s][f]);
  }
  free(c->tmp); free(c->FF);
  for (int i = 0; i < 9; i++) free(c->phys[i]);
  free(c);
}

const or_grid *orb_grid(const orb_ctx *c) { return &c->g; }

/* conj of 3D_Boussinesq/src/kernels.h:139-189 (same as the Navier-Stokes one) */
static void conj_fill(const orb_ctx *c, const orc *ZK, orc *ZZ) {
  const or_grid *g = &c->g;
  for (long ik = 0; ik < g->nkt; ik++) {
    ZZ[ik] = ZK[ik];
    if (c->kx[ik] != 0.0) continue;
    long ikt = ik / g->iktx;
    int ikz = (int)(ikt / g->ikty), iky = (int)(ikt - (long)ikz * g->ikty);
    if (iky == g->kty) {
      if (ikz == g->ktz) { ZZ[ik].re = 0.0; ZZ[ik].im = 0.0; }
      else if (ikz < g->ktz) {
        long ikk = ((long)(g->iktz - 1 - ikz) * g->ikty + iky) * g->iktx;
        ZZ[ik].re = ZK[ikk].re; ZZ[ik].im = -ZK[ikk].im;
      }
    } else if (iky < g->kty) {
      long ikk = ((long)(g->iktz - 1 - ikz) * g->ikty + (g->ikty - 1 - iky)) * g->iktx;
      ZZ[ik].re = ZK[ikk].re; ZZ[ik].im = -ZK[ikk].im;
    }
  }
}

/* setFFN (kk == NULL) or setFFK (multiply by i kk) then Z2D */
static void kr_one(orb_ctx *c, const orc *V, const double *kk, double *out) {
  for (long ip = 0; ip < c->g.n2; ip++) {
    int ikk = c->ikf[ip];
    if (ikk < 0) { c->FF[ip].re = 0.0; c->FF[ip].im = 0.0; }
    else if (!kk) c->FF[ip] = V[ikk];
    else { c->FF[ip].re = -kk[ikk] * V[ikk].im; c->FF[ip].im = kk[ikk] * V[ikk].re; }
  }
  or_fft3_c2r(c->FF, out, c->g.nx, c->g.ny, c->g.nz);
}

static void rhs_fft(orb_ctx *c) {
  for (int f = 0; f < 3; f++) { conj_fill(c, c->s0[f], c->tmp); kr_one(c, c->tmp, NULL, c->phys[f]); }
  for (int f = 0; f < 3; f++) { conj_fill(c, c->uk[f], c->tmp); kr_one(c, c->tmp, NULL, c->phys[3 + f]); }
  conj_fill(c, c->s0[3], c->tmp);
  kr_one(c, c->tmp, c->kx, c->phys[6]);
  kr_one(c, c->tmp, c->ky, c->phys[7]);
  kr_one(c, c->tmp, c->kz, c->phys[8]);
  double **p = c->phys;
  for (long i = 0; i < c->g.nr; i++) {
    double tu = p[3][i], tv = p[4][i], tw = p[5][i], tx = p[0][i], te = p[1][i], tz = p[2][i];
    double nx = tv * tz - tw * te, ny = tw * tx - tu * tz, nz = tu * te - tv * tx;
    p[3][i] = -nx; p[4][i] = -ny; p[5][i] = -nz;
    p[6][i] = tu * p[6][i] + tv * p[7][i] + tw * p[8][i];
  }
  double in = 1.0 / (double)c->g.nr;
  const int src[4] = {3, 4, 5, 6};
  for (int f = 0; f < 4; f++) {
    or_fft3_r2c(p[src[f]], c->FF, c->g.nx, c->g.ny, c->g.nz);
    for (long ik = 0; ik < c->g.nkt; ik++) {
      orc v = c->FF[c->ikn[ik]];
      c->uk[f][ik].re = v.re * in; c->uk[f][ik].im = v.im * in;
    }
  }
}

static inline void vel_from_s0(orb_ctx *c, long ik, double wki) {
  const double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
  orc x0 = c->s0[0][ik], e0 = c->s0[1][ik], z0 = c->s0[2][ik];
  c->uk[0][ik].re = -(kky * z0.im - kkz * e0.im) * wki;
  c->uk[0][ik].im = -(kkz * e0.re - kky * z0.re) * wki;
  c->uk[1][ik].re = -(kkz * x0.im - kkx * z0.im) * wki;
  c->uk[1][ik].im = -(kkx * z0.re - kkz * x0.re) * wki;
  c->uk[2][ik].re = -(kkx * e0.im - kky * x0.im) * wki;
  c->uk[2][ik].im = -(kky * x0.re - kkx * e0.re) * wki;
}

/* the spectral right-hand side terms shared by Step1/Step2/Step/RHS;
 * mask = 1 applies LC inside (Step1, Step2, RHS), 0 leaves it to the caller (Step) */
static inline void rhs_terms(const orb_ctx *c, long ik, int mask, double *C /* 8: xi,et,zt,rho re/im */) {
  const double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
  const double UKX = c->uk[0][ik].re, UKY = c->uk[0][ik].im, VKX = c->uk[1][ik].re, VKY = c->uk[1][ik].im;
  const double WKX = c->uk[2][ik].re, WKY = c->uk[2][ik].im;
  const double ROX = c->s0[3][ik].re, ROY = c->s0[3][ik].im;
  const double wk = kkx * kkx + kky * kky + kkz * kkz;
  const double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
  double rXR = -kkx * ROY, rXC = kkx * ROX, rYR = -kky * ROY, rYC = kky * ROX, rZR = -kkz * ROY, rZC = kkz * ROX;
  const double LC = mask ? (double)c->L[ik] : 1.0;
  const double ri = c->ri, s = c->sinth, co = c->costh;
  C[0] = LC * (kky * WKY - kkz * VKY - ri * (rYR * s - rZR * co));
  C[1] = LC * (kkz * VKX - kky * WKX - ri * (rYC * s - rZC * co));
  C[2] = LC * (kkz * UKY - kkx * WKY + ri * rXR * s);
  C[3] = LC * (kkx * WKX - kkz * UKX + ri * rXC * s);
  C[4] = LC * (kkx * VKY - kky * UKY - ri * rXR * co);
  C[5] = LC * (kky * UKX - kkx * VKX - ri * rXC * co);
  orc x0 = c->s0[0][ik], e0 = c->s0[1][ik], z0 = c->s0[2][ik];
  C[6] = -LC * (c->uk[3][ik].re + (kkz * x0.im - kkx * z0.im) * wki * co + (kkx * e0.im - kky * x0.im) * wki * s);
  C[7] = -LC * (c->uk[3][ik].im + (kkx * z0.re - kkz * x0.re) * wki * co + (kky * x0.re - kkx * e0.re) * wki * s);
  if (kkx == 0.0 && kkz == 0.0 && kky == c->kf) C[4] += c->ampfor;
}

static void step_stage(orb_ctx *c, orc *S[4], int stage) {
  const double dt = c->dt, v2 = c->v2, sc = c->sc;
  for (long ik = 0; ik < c->g.nkt; ik++) {
    const double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    const double wk = kkx * kkx + kky * kky + kkz * kkz;
    const double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
    double C[8];
    if (stage < 3) {
      rhs_terms(c, ik, 1, C);
      const double q = stage < 2 ? 0.25 : 0.5, h = stage < 2 ? 0.5 : 1.0;
      const double damp1 = (1. - q * dt * v2 * wk) / (1. + q * dt * v2 * wk);
      const double damp2 = (1. - q * dt * v2 * wk / sc) / (1. + q * dt * v2 * wk / sc);
      for (int f = 0; f < 4; f++) {
        orc *Xs = c->st[stage][f];
        Xs[ik].re = dt * C[2 * f]; Xs[ik].im = dt * C[2 * f + 1];
        const double d = f < 3 ? damp1 : damp2;
        c->s0[f][ik].re = d * (S[f][ik].re + h * dt * C[2 * f]);
        c->s0[f][ik].im = d * (S[f][ik].im + h * dt * C[2 * f + 1]);
      }
      vel_from_s0(c, ik, wki);
    } else {
      rhs_terms(c, ik, 0, C);
      const double LC = (double)c->L[ik];
      const double damp1 = 1. - 0.5 * dt * v2 * wk, damp2 = 1. - 0.5 * dt * v2 * wk / sc;
      const double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
      for (int f = 0; f < 4; f++) {
        const double d = f < 3 ? damp1 : damp2;
        const double den = f < 3 ? (1. + 0.5 * dt * v2 * wk) : (1. + 0.5 * dt * v2 * wk / sc);
        orc *x1 = c->st[0][f], *x2 = c->st[1][f], *x3 = c->st[2][f];
        S[f][ik].re = LC * (d * S[f][ik].re + sixth * (x1[ik].re + dt * C[2 * f]) + third * (x2[ik].re + x3[ik].re)) / den;
        S[f][ik].im = LC * (d * S[f][ik].im + sixth * (x1[ik].im + dt * C[2 * f + 1]) + third * (x2[ik].im + x3[ik].im)) / den;
        c->s0[f][ik] = S[f][ik];
      }
      /* Leray projection of the state vorticity; the stage copy keeps the unprojected one */
      const double phiR = (kkx * S[0][ik].im + kky * S[1][ik].im + kkz * S[2][ik].im) * wki;
      const double phiC = -(kkx * S[0][ik].re + kky * S[1][ik].re + kkz * S[2][ik].re) * wki;
      S[0][ik].re -= -kkx * phiC; S[0][ik].im -= kkx * phiR;
      S[1][ik].re -= -kky * phiC; S[1][ik].im -= kky * phiR;
      S[2][ik].re -= -kkz * phiC; S[2][ik].im -= kkz * phiR;
      vel_from_s0(c, ik, wki);
    }
  }
}

static void set_velocity(orb_ctx *c, orc *S[4]) {
  for (long ik = 0; ik < c->g.nkt; ik++) {
    for (int f = 0; f < 4; f++) c->s0[f][ik] = S[f][ik];
    const double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    const double wk = kkx * kkx + kky * kky + kkz * kkz;
    double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
    if (c->L[ik] == 0) wki = 0.0;
    vel_from_s0(c, ik, wki);
  }
}

void orb_run(orb_ctx *c, orc *xii, orc *eta, orc *zet, orc *rho, int nsteps) {
  orc *S[4] = {xii, eta, zet, rho};
  set_velocity(c, S);
  for (int nt = 0; nt < nsteps; nt++)
    for (int stage = 0; stage < 4; stage++) {
      rhs_fft(c);
      step_stage(c, S, stage);
    }
}

/* dz/dt = -nu k^2 z + N(z) + F for the Newton period constraint */
void orb_rhs(orb_ctx *c, const orc *xii, const orc *eta, const orc *zet, const orc *rho, orc *dxii, orc *deta,
             orc *dzet, orc *drho) {
  orc *S[4] = {(orc *)xii, (orc *)eta, (orc *)zet, (orc *)rho};
  orc *D[4] = {dxii, deta, dzet, drho};
  set_velocity(c, S);
  rhs_fft(c);
  for (long ik = 0; ik < c->g.nkt; ik++) {
    const double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    const double wk = kkx * kkx + kky * kky + kkz * kkz;
    double C[8];
    rhs_terms(c, ik, 1, C);
    const double d1 = -c->v2 * wk, d2 = -c->v2 * wk / c->sc;
    for (int f = 0; f < 4; f++) {
      const double d = f < 3 ? d1 : d2;
      D[f][ik].re = d * S[f][ik].re + C[2 * f];
      D[f][ik].im = d * S[f][ik].im + C[2 * f + 1];
    }
  }
}
