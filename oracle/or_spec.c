/*
 * Oracle grid sizes, wavenumber tables, masks, CN coefficient tables,
 * padded<->truncated index maps, reference RNG and initial conditions
 * (test infrastructure only; see oracle.h).
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 3D sizes: 3D_NavierStokes/src/global_params.F90:13-26 (NY=NZ, NX=NZ/alpha
 * in the reference; here all three are explicit so 3D_Boussinesq's
 * NX=NZ/alpha1 boxes are covered, 3D_Boussinesq/src/global_params.F90:15-30) */
int or_grid3(or_grid *g, int nx, int ny, int nz, double alpha) {
  memset(g, 0, sizeof(*g));
  g->dim = 3; g->nx = nx; g->ny = ny; g->nz = nz; g->alpha = alpha;
  g->ktx = nx / 3; g->kty = ny / 3; g->ktz = nz / 3;
  g->iktx = g->ktx + 1; g->ikty = 2 * g->kty + 1; g->iktz = 2 * g->ktz + 1;
  g->nx2 = nx / 2 + 1;
  g->nkt = (long)g->iktx * g->ikty * g->iktz;
  g->nr = (long)nx * ny * nz;
  g->n2 = (long)g->nx2 * ny * nz;
  return 0;
}

/* 2D sizes: 2D_NavierStokes/src_RK4/src_gmres/global_params.F90:14-21 */
int or_grid2(or_grid *g, int nx, int ny, double alpha) {
  memset(g, 0, sizeof(*g));
  g->dim = 2; g->nx = nx; g->ny = ny; g->nz = 1; g->alpha = alpha;
  g->ktx = nx / 3; g->kty = ny / 3; g->ktz = 0;
  g->iktx = g->ktx + 1; g->ikty = 2 * g->kty + 1; g->iktz = 1;
  g->nx2 = nx / 2 + 1;
  g->nkt = (long)g->iktx * g->ikty;
  g->nr = (long)nx * ny;
  g->n2 = (long)g->nx2 * ny;
  return 0;
}

/* 3D: 3D_NavierStokes/src/init.F90:25-64 ; 2D: 2D_NavierStokes/src_RK4/init.F90:25-38 */
void or_tables(const or_grid *g, double *kx, double *ky, double *kz, int *L) {
  for (int ikz = 0; ikz < g->iktz; ikz++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikx = 0; ikx < g->iktx; ikx++) {
        long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        double kkx = (double)ikx * g->alpha;
        double kky = (double)(iky - g->kty);
        double kkz = g->dim == 3 ? (double)(ikz - g->ktz) : 0.0;
        double wk = sqrt(kkx * kkx + kky * kky + kkz * kkz);
        int act = 1;
        if (kkx < 0.0) act = 0;
        if (g->dim == 3) {
          if (kkx == 0.0 && kky < 0.0) act = 0;
          if (kkx == 0.0 && kky == 0.0 && kkz <= 0.0) act = 0;
        } else {
          if (kkx == 0.0 && kky <= 0.0) act = 0;
        }
        if (wk > (double)g->kty - 0.5) act = 0;
        if (kx) kx[ik] = kkx;
        if (ky) ky[ik] = kky;
        if (kz) kz[ik] = kkz;
        if (L) L[ik] = act;
      }
}

long or_count_active(const or_grid *g) {
  int *L = (int *)malloc(sizeof(int) * g->nkt);
  or_tables(g, NULL, NULL, NULL, L);
  long n = 0;
  for (long i = 0; i < g->nkt; i++) n += L[i];
  free(L);
  return n;
}

/* 3D_NavierStokes/src/init.F90:194-198 (and src_gmres/main.F90:759-768) */
void or_nutab(const or_grid *g, double v2, double dt, double *nu, double *nuzn, double *nu1) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt),
         *kz = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  or_tables(g, kx, ky, kz, L);
  for (long ik = 0; ik < g->nkt; ik++) {
    if (!L[ik]) { if (nu) nu[ik] = 0.0; nuzn[ik] = 0.0; nu1[ik] = 0.0; continue; }
    double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
    double r3 = log10(v2) + 2.0 * log10(wk);
    double n = pow(10.0, r3);
    if (nu) nu[ik] = n;
    nuzn[ik] = (1.0 - 0.5 * n * dt) / dt;
    nu1[ik] = dt / (1.0 + 0.5 * n * dt);
  }
  free(kx); free(ky); free(kz); free(L);
}

/* 3D: 3D_NavierStokes/src/main.F90:100-199 ; 2D: 2D_NavierStokes/src_RK4/main.F90:169-202.
 * ikf has n2 entries (padded index -> truncated index or -1), ikn has nkt. */
void or_ikf_ikn(const or_grid *g, int *ikf, int *ikn) {
  int nx2 = g->nx2, ny = g->ny, nz = g->nz, kty = g->kty, ktz = g->ktz;
  for (long i = 0; i < g->n2; i++) ikf[i] = -1;
  for (int iz = 0; iz < nz; iz++) {
    int kz;
    if (g->dim == 2) kz = 0;
    else if (iz <= ktz) kz = iz;
    else if (iz >= nz - ktz) kz = iz - nz;
    else continue;
    for (int iy = 0; iy < ny; iy++) {
      int ky;
      if (iy <= kty) ky = iy;
      else if (iy >= ny - kty) ky = iy - ny;
      else continue;
      for (int ix = 0; ix < g->iktx; ix++) {
        long ip = ((long)iz * ny + iy) * nx2 + ix;
        ikf[ip] = ix + g->iktx * ((ky + kty) + g->ikty * (kz + ktz));
      }
    }
  }
  for (int ikz = 0; ikz < g->iktz; ikz++) {
    int kz = g->dim == 3 ? ikz - ktz : 0;
    int iz = kz < 0 ? nz + kz : kz;
    for (int iky = 0; iky < g->ikty; iky++) {
      int ky = iky - kty;
      int iy = ky < 0 ? ny + ky : ky;
      for (int ikx = 0; ikx < g->iktx; ikx++) {
        long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        ikn[ik] = (int)(((long)iz * ny + iy) * nx2 + ikx);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* ran1 (Numerical Recipes 2nd ed. 7.1: Park-Miller minimal standard with
 * Bays-Durham shuffle) and ranno, as 3D_NavierStokes/src/init.F90:237-294.
 * ranno(0) decrements the stored seed and re-initialises ran1 with it on
 * every draw (init.F90:256-259), so each draw is a fresh shuffle table.  */
#define PM_IA 16807
#define PM_IM 2147483647
#define PM_IQ 127773
#define PM_IR 2836
#define NTAB 32
#define NDIV (1 + (PM_IM - 1) / NTAB)

static int g_junk, g_iv[NTAB], g_iy;

int or_pm_next(int idum) {
  int k = idum / PM_IQ;
  idum = PM_IA * (idum - k * PM_IQ) - PM_IR * k;
  if (idum < 0) idum += PM_IM;
  return idum;
}

static double ran1(int *idum) {
  const double am = 1.0 / PM_IM, eps = 1.2e-7, rnmx = 1.0 - eps;
  if (*idum <= 0 || g_iy == 0) {
    *idum = -(*idum) > 1 ? -(*idum) : 1;
    for (int j = NTAB + 8; j >= 1; j--) {
      *idum = or_pm_next(*idum);
      if (j <= NTAB) g_iv[j - 1] = *idum;
    }
    g_iy = g_iv[0];
  }
  *idum = or_pm_next(*idum);
  int j = g_iy / NDIV;                  /* 0-based: 1+iy/ndiv - 1 */
  g_iy = g_iv[j];
  g_iv[j] = *idum;
  double r = am * g_iy;
  return r < rnmx ? r : rnmx;
}

void or_rng_seed(int iseed) {
  int i = iseed > 0 ? -iseed : iseed;
  g_junk = i;
  (void)ran1(&i);
}

double or_ranno0(void) {
  const double twopi = 4.0 * asin(1.0);
  g_junk = g_junk - 1;
  int ihold = g_junk;
  return (ran1(&ihold) - 0.5) * twopi;
}

/* Reference 3D random IC: 3D_NavierStokes/src/init.F90:151-230.
 * Loop order IKX, IKY, IKZ (IKZ innermost); three phase draws per active mode;
 * unit modulus on 2.5<|k|<=9.5; enstrophy 2*sum|w|^2 normalised to AMPV. */
void or_ic_ns3d_ref(const or_grid *g, int iseed, double ampv, orc *xii, orc *eta, orc *zet) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt),
         *kz = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  or_tables(g, kx, ky, kz, L);
  or_rng_seed(iseed);
  memset(xii, 0, sizeof(orc) * g->nkt);
  memset(eta, 0, sizeof(orc) * g->nkt);
  memset(zet, 0, sizeof(orc) * g->nkt);
  double enstro = 0.0;
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikz = 0; ikz < g->iktz; ikz++) {
        long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        if (!L[ik]) continue;
        double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
        double ek = (wk > 2.5 && wk <= 9.5) ? 1.0 : 0.0;
        double s, c, ph;
        ph = or_ranno0(); sincos(ph, &s, &c); xii[ik].re = ek * c; xii[ik].im = ek * s;
        ph = or_ranno0(); sincos(ph, &s, &c); eta[ik].re = ek * c; eta[ik].im = ek * s;
        ph = or_ranno0(); sincos(ph, &s, &c); zet[ik].re = ek * c; zet[ik].im = ek * s;
        double r1 = xii[ik].re * xii[ik].re + eta[ik].re * eta[ik].re + zet[ik].re * zet[ik].re;
        double r2 = xii[ik].im * xii[ik].im + eta[ik].im * eta[ik].im + zet[ik].im * zet[ik].im;
        enstro = enstro + r1 + r2;
      }
  enstro *= 2.0;
  double f = sqrt(ampv / enstro);
  for (long ik = 0; ik < g->nkt; ik++) {
    if (L[ik]) {
      xii[ik].re *= f; xii[ik].im *= f; eta[ik].re *= f; eta[ik].im *= f;
      zet[ik].re *= f; zet[ik].im *= f;
    } else {
      xii[ik].re = xii[ik].im = eta[ik].re = eta[ik].im = zet[ik].re = zet[ik].im = 0.0;
    }
  }
  free(kx); free(ky); free(kz); free(L);
}

/* Reference 3D Euler IC: 3D_Euler/src/init.F90:137-190.  Loop order IKX, IKY,
 * IKZ over every stored mode (no mask); one phase draw per mode;
 * zet = EK + PK e^{i phase} with the Taylor-Green pair EK = +-1/4 at
 * (kx = alpha, ky = +-1, kz = 0) and PK = AMPV on 1.5 < |k| <= 3.5, kz = 1;
 * xii = eta = 0; no normalisation. */
void or_ic_euler_ref(const or_grid *g, int iseed, double ampv, orc *xii, orc *eta, orc *zet) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt),
         *kz = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  or_tables(g, kx, ky, kz, L);
  or_rng_seed(iseed);
  memset(xii, 0, sizeof(orc) * g->nkt);
  memset(eta, 0, sizeof(orc) * g->nkt);
  memset(zet, 0, sizeof(orc) * g->nkt);
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikz = 0; ikz < g->iktz; ikz++) {
        long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
        double pk = (wk > 1.5 && wk <= 3.5 && kz[ik] == 1.0) ? ampv : 0.0;
        double ek = 0.0;
        if (ikx == 1 && ky[ik] == 1.0 && kz[ik] == 0.0) ek = 0.25;
        else if (ikx == 1 && ky[ik] == -1.0 && kz[ik] == 0.0) ek = -0.25;
        double s, c, ph = or_ranno0();
        sincos(ph, &s, &c);
        zet[ik].re = ek + pk * c;
        zet[ik].im = pk * s;
      }
  free(kx); free(ky); free(kz); free(L);
}

/* Reference 3D Boussinesq IC: 3D_Boussinesq/src/init.F90:234-305.  Loop order
 * IKX, IKY, IKZ over every stored mode; three phase draws per mode (the third is
 * drawn and unused); xii, eta = 0.1 e^{i phase} on 2.5 < |k| <= 9.5,
 * zet = -(kx xii + ky eta)/kz (all three zero where kz = 0); the enstrophy of
 * every stored mode, 2 sum|w|^2, scales the active modes to AMPV; rho = 0. */
void or_ic_bous_ref(const or_grid *g, int iseed, double ampv, orc *xii, orc *eta, orc *zet, orc *rho) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt),
         *kz = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  or_tables(g, kx, ky, kz, L);
  or_rng_seed(iseed);
  memset(rho, 0, sizeof(orc) * g->nkt);
  double enstro = 0.0;
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikz = 0; ikz < g->iktz; ikz++) {
        long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
        double ek = (wk > 2.5 && wk <= 9.5) ? 0.1 : 0.0;
        double s, c;
        sincos(or_ranno0(), &s, &c);
        xii[ik].re = ek * c; xii[ik].im = ek * s;
        sincos(or_ranno0(), &s, &c);
        eta[ik].re = ek * c; eta[ik].im = ek * s;
        (void)or_ranno0();
        if (kz[ik] != 0.0) {
          zet[ik].re = -(kx[ik] * xii[ik].re + ky[ik] * eta[ik].re) / kz[ik];
          zet[ik].im = -(kx[ik] * xii[ik].im + ky[ik] * eta[ik].im) / kz[ik];
        } else {
          xii[ik].re = xii[ik].im = eta[ik].re = eta[ik].im = zet[ik].re = zet[ik].im = 0.0;
        }
        double r1 = xii[ik].re * xii[ik].re + eta[ik].re * eta[ik].re + zet[ik].re * zet[ik].re;
        double r2 = xii[ik].im * xii[ik].im + eta[ik].im * eta[ik].im + zet[ik].im * zet[ik].im;
        enstro = enstro + r1 + r2;
      }
  enstro *= 2.0;
  if (enstro != 0.0) {
    double f = sqrt(ampv / enstro);
    for (long ik = 0; ik < g->nkt; ik++) {
      if (L[ik]) {
        xii[ik].re *= f; xii[ik].im *= f; eta[ik].re *= f; eta[ik].im *= f;
        zet[ik].re *= f; zet[ik].im *= f;
      } else {
        xii[ik].re = xii[ik].im = eta[ik].re = eta[ik].im = zet[ik].re = zet[ik].im = 0.0;
      }
    }
  }
  free(kx); free(ky); free(kz); free(L);
}

/* Reference 2D random IC: 2D_NavierStokes/src_RK4/init.F90:138-208.
 * EK=1/|k| on 2.5<|k|<=9.5 with kx>=0.1, VK=|k|^2 EK, Z=sqrt(VK/|k|) e^{i phase};
 * one phase draw per active mode (IKX outer, IKY inner); 2*sum|Z|^2 = AMPV. */
void or_ic_ns2d_ref(const or_grid *g, int iseed, double ampv, orc *z) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  or_tables(g, kx, ky, NULL, L);
  or_rng_seed(iseed);
  memset(z, 0, sizeof(orc) * g->nkt);
  double enstro = 0.0;
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++) {
      long ik = (long)iky * g->iktx + ikx;
      if (!L[ik]) continue;
      double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik]);
      double ek = (wk > 2.5 && wk <= 9.5) ? 1.0 / wk : 0.0;
      if (kx[ik] < 0.1) ek = 0.0;
      double vk = pow(wk, 2.0) * ek;
      double ph = or_ranno0(), s, c;
      sincos(ph, &s, &c);
      double a = sqrt(vk / wk);
      z[ik].re = a * c; z[ik].im = a * s;
      enstro = enstro + z[ik].re * z[ik].re + z[ik].im * z[ik].im;
    }
  enstro *= 2.0;
  double f = sqrt(ampv / enstro);
  for (long ik = 0; ik < g->nkt; ik++) {
    if (L[ik]) { z[ik].re *= f; z[ik].im *= f; } else { z[ik].re = z[ik].im = 0.0; }
  }
  free(kx); free(ky); free(L);
}

/* ------------------------------------------------------------------ */
/* Benchmark initial conditions (no reference counterpart; the law is the
 * benchmark's): SplitMix64 (Steele, Lea, Flood 2014) counter-based uniforms,
 * Box-Muller normals.  Identical code lives in the product's host layer; the
 * two are compared bit-for-bit in the tests. */
uint64_t or_splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double unif(uint64_t key, uint64_t idx) {
  uint64_t s = key * 0x100000001B3ULL ^ idx;
  return (double)(or_splitmix64(&s) >> 11) * 0x1.0p-53;
}

double or_gauss(uint64_t key, uint64_t idx) {
  double u1 = 1.0 - unif(key, 2 * idx), u2 = unif(key, 2 * idx + 1);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Random solenoidal velocity with shell spectrum E(k) ~ k^4 exp(-2 (k/4)^2):
 * per active mode g = complex Gaussian 3-vector, projected normal to k,
 * scaled by |k| exp(-(|k|/4)^2) (mode density ~ k^2), then normalised so that
 * sum_L |u|^2 = energy; vorticity w = i k x u.  Optional scalar rho with the
 * same amplitude law, normalised to sum_L |rho|^2 = amp_scale^2 * energy. */
void or_ic_solenoidal(const or_grid *g, uint64_t seed, double energy, double amp_scale,
                      orc *xii, orc *eta, orc *zet, orc *rho) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt),
         *kz = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  orc *u = malloc(sizeof(orc) * 3 * g->nkt);
  or_tables(g, kx, ky, kz, L);
  double esum = 0.0, rsum = 0.0;
  for (long ik = 0; ik < g->nkt; ik++) {
    orc *uk = u + 3 * ik;
    if (!L[ik]) { memset(uk, 0, sizeof(orc) * 3); if (rho) rho[ik].re = rho[ik].im = 0.0; continue; }
    double k[3] = {kx[ik], ky[ik], kz[ik]};
    double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2], km = sqrt(k2);
    double amp = km * exp(-(km / 4.0) * (km / 4.0));
    double gr[3], gi[3];
    for (int c = 0; c < 3; c++) {
      gr[c] = or_gauss(seed, 6 * (uint64_t)ik + 2 * c);
      gi[c] = or_gauss(seed, 6 * (uint64_t)ik + 2 * c + 1);
    }
    double dr = (k[0] * gr[0] + k[1] * gr[1] + k[2] * gr[2]) / k2;
    double di = (k[0] * gi[0] + k[1] * gi[1] + k[2] * gi[2]) / k2;
    for (int c = 0; c < 3; c++) {
      uk[c].re = amp * (gr[c] - k[c] * dr);
      uk[c].im = amp * (gi[c] - k[c] * di);
      esum += uk[c].re * uk[c].re + uk[c].im * uk[c].im;
    }
    if (rho) {
      uint64_t rk = seed ^ 0xD1B54A32D192ED03ULL;
      rho[ik].re = amp * or_gauss(rk, 2 * (uint64_t)ik);
      rho[ik].im = amp * or_gauss(rk, 2 * (uint64_t)ik + 1);
      rsum += rho[ik].re * rho[ik].re + rho[ik].im * rho[ik].im;
    }
  }
  double f = esum > 0.0 ? sqrt(energy / esum) : 0.0;
  double fr = rsum > 0.0 ? amp_scale * sqrt(energy / rsum) : 0.0;
  for (long ik = 0; ik < g->nkt; ik++) {
    orc *uk = u + 3 * ik;
    double k[3] = {kx[ik], ky[ik], kz[ik]};
    for (int c = 0; c < 3; c++) { uk[c].re *= f; uk[c].im *= f; }
    /* w = i k x u : (k x u)_x = ky uz - kz uy ; i*(a+ib) = -b + i a */
    double cxr = k[1] * uk[2].re - k[2] * uk[1].re, cxi = k[1] * uk[2].im - k[2] * uk[1].im;
    double cyr = k[2] * uk[0].re - k[0] * uk[2].re, cyi = k[2] * uk[0].im - k[0] * uk[2].im;
    double czr = k[0] * uk[1].re - k[1] * uk[0].re, czi = k[0] * uk[1].im - k[1] * uk[0].im;
    xii[ik].re = -cxi; xii[ik].im = cxr;
    eta[ik].re = -cyi; eta[ik].im = cyr;
    zet[ik].re = -czi; zet[ik].im = czr;
    if (rho) { rho[ik].re *= fr; rho[ik].im *= fr; }
  }
  free(kx); free(ky); free(kz); free(L); free(u);
}

/* 2D benchmark IC: |w| ~ |k| exp(-(|k|/4)^2) on 1<=|k|<=8, uniform phases
 * 2*pi*U[0,1) from the same counter RNG, 2*sum_L|w|^2 = ampv. */
void or_ic_ns2d_cfg(const or_grid *g, uint64_t seed, double ampv, orc *z) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  or_tables(g, kx, ky, NULL, L);
  double s2 = 0.0;
  for (long ik = 0; ik < g->nkt; ik++) {
    z[ik].re = z[ik].im = 0.0;
    if (!L[ik]) continue;
    double km = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik]);
    if (km < 1.0 || km > 8.0) continue;
    double amp = km * exp(-(km / 4.0) * (km / 4.0));
    double ph = 2.0 * M_PI * unif(seed, (uint64_t)ik), s, c;
    sincos(ph, &s, &c);
    z[ik].re = amp * c; z[ik].im = amp * s;
    s2 += z[ik].re * z[ik].re + z[ik].im * z[ik].im;
  }
  double f = s2 > 0.0 ? sqrt(ampv / (2.0 * s2)) : 0.0;
  for (long ik = 0; ik < g->nkt; ik++) { z[ik].re *= f; z[ik].im *= f; }
  free(kx); free(ky); free(L);
}
