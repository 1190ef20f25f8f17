/*
 * Oracle 3D Navier-Stokes (Heun + Crank-Nicolson) and 3D Euler (Heun)
 * steppers: a step-by-step FP64 restatement of the reference time loop
 * (test infrastructure only; see oracle.h).
 *
 *   loop        3D_NavierStokes/src/cuda.cu:60-112 ; 3D_Euler/src/cuda.cu:49-101
 *   KR_FFT_ALL  3D_NavierStokes/src/functions.h:26-75  (conj -> setFFN -> Z2D)
 *   conj        3D_NavierStokes/src/kernels.h:113-163
 *   setFFN      3D_NavierStokes/src/kernels.h:51-68
 *   multReal    3D_NavierStokes/src/kernels.h:281-306
 *   RK_FFT      3D_NavierStokes/src/functions.h:79-108 (D2Z -> normFF1)
 *   normFF1     3D_NavierStokes/src/kernels.h:266-277
 *   setVelocity 3D_NavierStokes/src/kernels.h:168-205
 *   preStep     3D_NavierStokes/src/kernels.h:309-386 ; GMRES forcing
 *               3D_NavierStokes/src_gmres/kernels.h:234-238 ;
 *               Euler 3D_Euler/src/kernels.h:296-366
 *   corStep     3D_NavierStokes/src/kernels.h:389-464 ; Euler 3D_Euler/src/kernels.h:369-440
 *   diagnostics 3D_NavierStokes/src/functions.h:128-161
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or3_ctx {
  or_grid g;
  int flavour, force;
  double v2, dt, kf, ampfor;
  double *kx, *ky, *kz, *nuzn, *nu1;
  int *L, *ikf, *ikn;
  orc *uk[3], *rh[3], *tmp, *FF;
  double *phys[6];
};

or3_ctx *or3_create(int nx, int ny, int nz, double alpha, int flavour,
                    double v2, double dt, int force_mode, double kf, double ampfor) {
  or3_ctx *c = (or3_ctx *)calloc(1, sizeof(or3_ctx));
  or_grid3(&c->g, nx, ny, nz, alpha);
  c->flavour = flavour; c->force = force_mode;
  c->v2 = v2; c->dt = dt; c->kf = kf; c->ampfor = ampfor;
  long nkt = c->g.nkt;
  c->kx = malloc(sizeof(double) * nkt); c->ky = malloc(sizeof(double) * nkt);
  c->kz = malloc(sizeof(double) * nkt);
  c->nuzn = malloc(sizeof(double) * nkt); c->nu1 = malloc(sizeof(double) * nkt);
  c->L = malloc(sizeof(int) * nkt);
  c->ikf = malloc(sizeof(int) * c->g.n2); c->ikn = malloc(sizeof(int) * nkt);
  or_tables(&c->g, c->kx, c->ky, c->kz, c->L);
  or_nutab(&c->g, v2, dt, NULL, c->nuzn, c->nu1);
  or_ikf_ikn(&c->g, c->ikf, c->ikn);
  for (int i = 0; i < 3; i++) {
    c->uk[i] = calloc(nkt, sizeof(orc));
    c->rh[i] = calloc(nkt, sizeof(orc));
  }
  c->tmp = calloc(nkt, sizeof(orc));
  c->FF = calloc(c->g.n2, sizeof(orc));
  for (int i = 0; i < 6; i++) c->phys[i] = calloc(c->g.nr, sizeof(double));
  return c;
}

void or3_destroy(or3_ctx *c) {
  if (!c) return;
  free(c->kx); free(c->ky); free(c->kz); free(c->nuzn); free(c->nu1);
  free(c->L); free(c->ikf); free(c->ikn);
  for (int i = 0; i < 3; i++) { free(c->uk[i]); free(c->rh[i]); }
  free(c->tmp); free(c->FF);
  for (int i = 0; i < 6; i++) free(c->phys[i]);
  free(c);
}

const or_grid *or3_grid(const or3_ctx *c) { return &c->g; }

/* conj: Hermitian completion of the kx=0 plane into tmp (kernels.h:113-163).
 * The reference compares the ky index with KTZ; KTY is used here (identical
 * for the NY=NZ boxes the reference runs). */
static void conj_fill(const or3_ctx *c, const orc *ZK, orc *ZZ) {
  const or_grid *g = &c->g;
#pragma omp parallel for schedule(static)
  for (long ik = 0; ik < g->nkt; ik++) {
    if (c->kx[ik] == 0.0) {
      long ikt = ik / g->iktx;
      int ikz = (int)(ikt / g->ikty);
      int iky = (int)(ikt - (long)ikz * g->ikty);
      if (iky == g->kty) {
        if (ikz == g->ktz) { ZZ[ik].re = 0.0; ZZ[ik].im = 0.0; }
        else if (ikz < g->ktz) {
          long ikk = ((long)(g->iktz - 1 - ikz) * g->ikty + iky) * g->iktx;
          ZZ[ik].re = ZK[ikk].re; ZZ[ik].im = -ZK[ikk].im;
        } else ZZ[ik] = ZK[ik];
      } else if (iky < g->kty) {
        long ikk = ((long)(g->iktz - 1 - ikz) * g->ikty + (g->ikty - 1 - iky)) * g->iktx;
        ZZ[ik].re = ZK[ikk].re; ZZ[ik].im = -ZK[ikk].im;
      } else ZZ[ik] = ZK[ik];
    } else ZZ[ik] = ZK[ik];
  }
}

static void set_ffn(const or3_ctx *c, const orc *V, orc *F1) {
#pragma omp parallel for schedule(static)
  for (long ip = 0; ip < c->g.n2; ip++) {
    int ikk = c->ikf[ip];
    if (ikk < 0) { F1[ip].re = 0.0; F1[ip].im = 0.0; } else F1[ip] = V[ikk];
  }
}

static void kr_one(or3_ctx *c, const orc *V, double *out) {
  conj_fill(c, V, c->tmp);
  set_ffn(c, c->tmp, c->FF);
  or_fft3_c2r(c->FF, out, c->g.nx, c->g.ny, c->g.nz);
}

/* KR_FFT_ALL: phys[0..2] = vorticity, phys[3..5] = velocity */
static void kr_fft_all(or3_ctx *c, const orc *w[3]) {
  for (int i = 0; i < 3; i++) kr_one(c, w[i], c->phys[i]);
  for (int i = 0; i < 3; i++) kr_one(c, c->uk[i], c->phys[3 + i]);
}

/* multReal: (U,V,W) <- -(u x w) */
static void mult_real(or3_ctx *c) {
  double *X = c->phys[0], *E = c->phys[1], *Z = c->phys[2];
  double *U = c->phys[3], *V = c->phys[4], *W = c->phys[5];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < c->g.nr; i++) {
    double tu = U[i], tv = V[i], tw = W[i], tx = X[i], te = E[i], tz = Z[i];
    double nx = tv * tz - tw * te;
    double ny = tw * tx - tu * tz;
    double nz = tu * te - tv * tx;
    U[i] = -nx; V[i] = -ny; W[i] = -nz;
  }
}

/* RK_FFT: D2Z then normFF1 gather scaled by 1/NR into uk[] */
static void rk_fft(or3_ctx *c) {
  double in = 1.0 / (double)c->g.nr;
  for (int f = 0; f < 3; f++) {
    or_fft3_r2c(c->phys[3 + f], c->FF, c->g.nx, c->g.ny, c->g.nz);
    orc *o = c->uk[f];
#pragma omp parallel for schedule(static)
    for (long ik = 0; ik < c->g.nkt; ik++) {
      orc v = c->FF[c->ikn[ik]];
      o[ik].re = v.re * in; o[ik].im = v.im * in;
    }
  }
}

static void set_velocity(or3_ctx *c, const orc *xii, const orc *eta, const orc *zet) {
#pragma omp parallel for schedule(static)
  for (long ik = 0; ik < c->g.nkt; ik++) {
    double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    double wk = kkx * kkx + kky * kky + kkz * kkz;
    double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
    if (c->L[ik] == 0) wki = 0.0;
    double xiX = xii[ik].re, xiY = xii[ik].im, etX = eta[ik].re, etY = eta[ik].im;
    double ztX = zet[ik].re, ztY = zet[ik].im;
    c->uk[0][ik].re = -(kky * ztY - kkz * etY) * wki;
    c->uk[0][ik].im = -(kkz * etX - kky * ztX) * wki;
    c->uk[1][ik].re = -(kkz * xiY - kkx * ztY) * wki;
    c->uk[1][ik].im = -(kkx * ztX - kkz * xiX) * wki;
    c->uk[2][ik].re = -(kkx * etY - kky * xiY) * wki;
    c->uk[2][ik].im = -(kky * xiX - kkx * etX) * wki;
  }
}

/* forcing added to the curl result at the forcing wavenumbers */
static inline void add_forcing(const or3_ctx *c, double kkx, double kky, double kkz,
                               double *xiX, double *xiY, double *etX, double *etY,
                               double *ztX, double *ztY) {
  double kf = c->kf, a = c->ampfor;
  (void)xiY; (void)etX;
  switch (c->force) {
  case OR_FORCE_DNS:    /* 3D_NavierStokes/src/kernels.h:342-343 */
    if (kkx == 0.0 && kkz == kf && kky == kf) *xiX += a;
    if (kkx == 0.0 && kkz == -kf && kky == kf) *xiX -= a;
    break;
  case OR_FORCE_GMRES:  /* 3D_NavierStokes/src_gmres/kernels.h:234-238 */
    if (kkx == 0.0 && kkz == kf && kky == kf) *etY += a;
    if (kkx == 0.0 && kkz == kf && kky == kf) *ztY -= a;
    if (kkx == 0.0 && kkz == -kf && kky == kf) *etY += a;
    if (kkx == 0.0 && kkz == -kf && kky == kf) *ztY += a;
    break;
  case OR_FORCE_KOLMO:  /* f=(sin kf y,0,0): zeta(0,kf,0) real part, cf. 3D_Boussinesq/src/kernels.h:382-384 */
    if (kkx == 0.0 && kkz == 0.0 && kky == kf) *ztX += a;
    break;
  default: break;
  }
}

#define CURL_FROM_UK                                                  \
  double UKX = c->uk[0][ik].re, UKY = c->uk[0][ik].im;                \
  double VKX = c->uk[1][ik].re, VKY = c->uk[1][ik].im;                \
  double WKX = c->uk[2][ik].re, WKY = c->uk[2][ik].im;                \
  double wk = kkx * kkx + kky * kky + kkz * kkz;                      \
  double wki = 1.0 / (wk > 0.001 ? wk : 0.001);                       \
  double xiX = kky * WKY - kkz * VKY, xiY = kkz * VKX - kky * WKX;   \
  double etX = kkz * UKY - kkx * WKY, etY = kkx * WKX - kkz * UKX;   \
  double ztX = kkx * VKY - kky * UKY, ztY = kky * UKX - kkx * VKX;

#define STORE_STATE_AND_VELOCITY                                      \
  c->uk[0][ik].re = -(kky * ztY - kkz * etY) * wki;                   \
  c->uk[0][ik].im = -(kkz * etX - kky * ztX) * wki;                   \
  c->uk[1][ik].re = -(kkz * xiY - kkx * ztY) * wki;                   \
  c->uk[1][ik].im = -(kkx * ztX - kkz * xiX) * wki;                   \
  c->uk[2][ik].re = -(kkx * etY - kky * xiY) * wki;                   \
  c->uk[2][ik].im = -(kky * xiX - kkx * etX) * wki;                   \
  xii[ik].re = xiX; xii[ik].im = xiY;                                 \
  eta[ik].re = etX; eta[ik].im = etY;                                 \
  zet[ik].re = ztX; zet[ik].im = ztY;

static void pre_step(or3_ctx *c, orc *xii, orc *eta, orc *zet) {
#pragma omp parallel for schedule(static)
  for (long ik = 0; ik < c->g.nkt; ik++) {
    double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    CURL_FROM_UK
    if (c->flavour == 1) {           /* Euler */
      double ll = (double)c->L[ik], dt = c->dt;
      c->rh[0][ik].re = ll * xiX; c->rh[0][ik].im = ll * xiY;
      c->rh[1][ik].re = ll * etX; c->rh[1][ik].im = ll * etY;
      c->rh[2][ik].re = ll * ztX; c->rh[2][ik].im = ll * ztY;
      xiX = xii[ik].re + dt * xiX; xiY = xii[ik].im + dt * xiY;
      etX = eta[ik].re + dt * etX; etY = eta[ik].im + dt * etY;
      ztX = zet[ik].re + dt * ztX; ztY = zet[ik].im + dt * ztY;
    } else {
      double NUZN = c->nuzn[ik];
      double NU1 = (double)c->L[ik] * c->nu1[ik];
      add_forcing(c, kkx, kky, kkz, &xiX, &xiY, &etX, &etY, &ztX, &ztY);
      c->rh[0][ik].re = xiX; c->rh[0][ik].im = xiY;
      xiX = NU1 * (NUZN * xii[ik].re + xiX); xiY = NU1 * (NUZN * xii[ik].im + xiY);
      c->rh[1][ik].re = etX; c->rh[1][ik].im = etY;
      etX = NU1 * (NUZN * eta[ik].re + etX); etY = NU1 * (NUZN * eta[ik].im + etY);
      c->rh[2][ik].re = ztX; c->rh[2][ik].im = ztY;
      ztX = NU1 * (NUZN * zet[ik].re + ztX); ztY = NU1 * (NUZN * zet[ik].im + ztY);
    }
    STORE_STATE_AND_VELOCITY
  }
}

static void cor_step(or3_ctx *c, orc *xii, orc *eta, orc *zet) {
#pragma omp parallel for schedule(static)
  for (long ik = 0; ik < c->g.nkt; ik++) {
    double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    CURL_FROM_UK
    double ll = (double)c->L[ik];
    double NU2;
    if (c->flavour == 1) NU2 = 0.5 * c->dt;
    else { NU2 = 0.5 * c->nu1[ik]; add_forcing(c, kkx, kky, kkz, &xiX, &xiY, &etX, &etY, &ztX, &ztY); }
    xiX = ll * (xii[ik].re + NU2 * (xiX - c->rh[0][ik].re));
    xiY = ll * (xii[ik].im + NU2 * (xiY - c->rh[0][ik].im));
    etX = ll * (eta[ik].re + NU2 * (etX - c->rh[1][ik].re));
    etY = ll * (eta[ik].im + NU2 * (etY - c->rh[1][ik].im));
    ztX = ll * (zet[ik].re + NU2 * (ztX - c->rh[2][ik].re));
    ztY = ll * (zet[ik].im + NU2 * (ztY - c->rh[2][ik].im));
    STORE_STATE_AND_VELOCITY
    c->rh[0][ik].re = c->rh[0][ik].im = 0.0;
    c->rh[1][ik].re = c->rh[1][ik].im = 0.0;
    c->rh[2][ik].re = c->rh[2][ik].im = 0.0;
  }
}

void or3_run(or3_ctx *c, orc *xii, orc *eta, orc *zet, int nsteps) {
  const orc *w[3] = {xii, eta, zet};
  set_velocity(c, xii, eta, zet);
  for (int nt = 0; nt < nsteps; nt++) {
    kr_fft_all(c, w);
    mult_real(c);
    rk_fft(c);
    pre_step(c, xii, eta, zet);
    kr_fft_all(c, w);
    mult_real(c);
    rk_fft(c);
    cor_step(c, xii, eta, zet);
  }
}

void or3_physical(or3_ctx *c, const orc *xii, const orc *eta, const orc *zet, double *out6) {
  const orc *w[3] = {xii, eta, zet};
  set_velocity(c, xii, eta, zet);
  kr_fft_all(c, w);
  for (int i = 0; i < 6; i++) memcpy(out6 + (long)i * c->g.nr, c->phys[i], sizeof(double) * c->g.nr);
}

/* masked nonlinear term curl(u x w) + forcing, the convol part of
 * 3D_NavierStokes/src_gmres/timederiv.F90:2-44 */
void or3_nonlin(or3_ctx *c, const orc *xii, const orc *eta, const orc *zet,
                orc *nx_, orc *ny_, orc *nz_) {
  const orc *w[3] = {xii, eta, zet};
  set_velocity(c, xii, eta, zet);
  kr_fft_all(c, w);
  mult_real(c);
  rk_fft(c);
  for (long ik = 0; ik < c->g.nkt; ik++) {
    double kkx = c->kx[ik], kky = c->ky[ik], kkz = c->kz[ik];
    CURL_FROM_UK
    (void)wki;
    if (c->flavour != 1) add_forcing(c, kkx, kky, kkz, &xiX, &xiY, &etX, &etY, &ztX, &ztY);
    double ll = (double)c->L[ik];
    nx_[ik].re = ll * xiX; nx_[ik].im = ll * xiY;
    ny_[ik].re = ll * etX; ny_[ik].im = ll * etY;
    nz_[ik].re = ll * ztX; nz_[ik].im = ll * ztY;
  }
}

void or3_diag(const or3_ctx *c, const orc *xii, const orc *eta, const orc *zet,
              double *energy, double *enstrophy, double *spec) {
  /* the sums of computeStats in compensated arithmetic */
  or_ksum e = {0, 0}, z = {0, 0};
  const int nb = c->g.kty + 2;
  or_ksum *bins = (or_ksum *)calloc((size_t)nb, sizeof(or_ksum));
  for (long ik = 0; ik < c->g.nkt; ik++) {
    if (c->L[ik] != 1) continue;
    double kk = c->kx[ik] * c->kx[ik] + c->ky[ik] * c->ky[ik] + c->kz[ik] * c->kz[ik];
    double wk = kk > 0.001 ? kk : 0.001;
    double vz = xii[ik].re * xii[ik].re + xii[ik].im * xii[ik].im +
                eta[ik].re * eta[ik].re + eta[ik].im * eta[ik].im +
                zet[ik].re * zet[ik].re + zet[ik].im * zet[ik].im;
    or_kadd(&z, vz);
    or_kadd(&e, vz / wk);
    int j = (int)(sqrt(kk) + 0.5);
    if (j >= 0 && j < nb) or_kadd(&bins[j], vz / wk);
  }
  if (spec)
    for (int j = 0; j < nb; j++) spec[j] = or_kval(&bins[j]);
  free(bins);
  *energy = or_kval(&e);
  *enstrophy = or_kval(&z);
}
