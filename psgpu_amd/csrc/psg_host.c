/*
 * Host layer of psgpu-mi355x in plain C11: the duties the reference's
 * Fortran drivers perform before and after each timestep_cuda_ call.
 * Declarations and reference citations are in include/psg_host.h.
 */
#define _GNU_SOURCE
#include "../../include/psg_host.h"
#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ sizes */
int psgh_grid3(psgh_grid *g, int nx, int ny, int nz, double alpha) {
  if (nx < 6 || ny < 6 || nz < 6) return -1;
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

int psgh_grid2(psgh_grid *g, int nx, int ny, double alpha) {
  if (nx < 6 || ny < 6) return -1;
  memset(g, 0, sizeof(*g));
  g->dim = 2; g->nx = nx; g->ny = ny; g->nz = 1; g->alpha = alpha;
  g->ktx = nx / 3; g->kty = ny / 3;
  g->iktx = g->ktx + 1; g->ikty = 2 * g->kty + 1; g->iktz = 1;
  g->nx2 = nx / 2 + 1;
  g->nkt = (long)g->iktx * g->ikty;
  g->nr = (long)nx * ny;
  g->n2 = (long)g->nx2 * ny;
  return 0;
}

/* mask: half space kx>0 | (kx=0, ky>0) | (kx=ky=0, kz>0) [2D: kx>0 | (kx=0, ky>0)]
 * inside the sphere |k| <= KTY - 1/2 */
void psgh_tables(const psgh_grid *g, double *kx, double *ky, double *kz, int *L) {
  long ik = 0;
  for (int ikz = 0; ikz < g->iktz; ikz++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikx = 0; ikx < g->iktx; ikx++, ik++) {
        const double a = (double)ikx * g->alpha;
        const double b = (double)(iky - g->kty);
        const double c = g->dim == 3 ? (double)(ikz - g->ktz) : 0.0;
        int on = !(a < 0.0);
        if (a == 0.0) {
          if (b < 0.0) on = 0;
          if (g->dim == 3 && b == 0.0 && c <= 0.0) on = 0;
          if (g->dim == 2 && b <= 0.0) on = 0;
        }
        if (sqrt(a * a + b * b + c * c) > (double)g->kty - 0.5) on = 0;
        if (kx) kx[ik] = a;
        if (ky) ky[ik] = b;
        if (kz) kz[ik] = c;
        if (L) L[ik] = on;
      }
}

long psgh_count_active(const psgh_grid *g) {
  int *L = malloc(sizeof(int) * g->nkt);
  psgh_tables(g, NULL, NULL, NULL, L);
  long n = 0;
  for (long i = 0; i < g->nkt; i++) n += L[i];
  free(L);
  return n;
}

void psgh_nutab(const psgh_grid *g, double v2, double dt, double *nu, double *nuzn, double *nu1) {
  double *kx = malloc(sizeof(double) * g->nkt), *ky = malloc(sizeof(double) * g->nkt),
         *kz = malloc(sizeof(double) * g->nkt);
  int *L = malloc(sizeof(int) * g->nkt);
  psgh_tables(g, kx, ky, kz, L);
  for (long ik = 0; ik < g->nkt; ik++) {
    double n = 0.0, zn = 0.0, n1 = 0.0;
    if (L[ik]) {
      const double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
      n = pow(10.0, log10(v2) + 2.0 * log10(wk));
      zn = (1.0 - 0.5 * n * dt) / dt;
      n1 = dt / (1.0 + 0.5 * n * dt);
    }
    if (nu) nu[ik] = n;
    if (nuzn) nuzn[ik] = zn;
    if (nu1) nu1[ik] = n1;
  }
  free(kx); free(ky); free(kz); free(L);
}

void psgh_ikf_ikn(const psgh_grid *g, int *ikf, int *ikn) {
  const int ny = g->ny, nz = g->nz, nx2 = g->nx2;
  for (long i = 0; i < g->n2; i++) ikf[i] = -1;
  for (int ikz = 0; ikz < g->iktz; ikz++) {
    const int kz = g->dim == 3 ? ikz - g->ktz : 0;
    const int iz = kz < 0 ? kz + nz : kz;
    for (int iky = 0; iky < g->ikty; iky++) {
      const int ky = iky - g->kty;
      const int iy = ky < 0 ? ky + ny : ky;
      for (int ikx = 0; ikx < g->iktx; ikx++) {
        const long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        const long ip = ((long)iz * ny + iy) * nx2 + ikx;
        ikn[ik] = (int)ip;
        ikf[ip] = (int)ik;
      }
    }
  }
}

void psgh_twiddles(int n, double *w) {
  for (int k = 0; k < n; k++) {
    double s, c;
    sincos(2.0 * M_PI * (double)k / (double)n, &s, &c);
    w[2 * k] = c;
    w[2 * k + 1] = -s;
  }
}

/* ------------------------------------------------------------------ RNG */
static int h_seed, h_tab[32], h_y;

static int pm_step(int s) {
  /* Schrage's factorisation of 16807*s mod (2^31-1) */
  const int q = 127773, r = 2836;
  int k = s / q;
  s = 16807 * (s - k * q) - r * k;
  return s < 0 ? s + 2147483647 : s;
}

static double ran1_draw(int *idum) {
  const int ndiv = 1 + 2147483646 / 32;
  if (*idum <= 0 || h_y == 0) {
    *idum = -*idum < 1 ? 1 : -*idum;
    for (int j = 39; j >= 0; j--) {
      *idum = pm_step(*idum);
      if (j < 32) h_tab[j] = *idum;
    }
    h_y = h_tab[0];
  }
  *idum = pm_step(*idum);
  const int j = h_y / ndiv;
  h_y = h_tab[j];
  h_tab[j] = *idum;
  const double r = (1.0 / 2147483647) * h_y;
  const double top = 1.0 - 1.2e-7;
  return r < top ? r : top;
}

void psgh_rng_seed(int iseed) {
  int i = iseed > 0 ? -iseed : iseed;
  h_seed = i;
  (void)ran1_draw(&i);
}

double psgh_ranno0(void) {
  h_seed -= 1;
  int t = h_seed;
  return (ran1_draw(&t) - 0.5) * (4.0 * asin(1.0));
}

/* ------------------------------------------------------------------ ICs */
void psgh_ic_ns3d_ref(const psgh_grid *g, int iseed, double ampv, double *xii, double *eta, double *zet) {
  const long n = g->nkt;
  double *kx = malloc(sizeof(double) * n), *ky = malloc(sizeof(double) * n), *kz = malloc(sizeof(double) * n);
  int *L = malloc(sizeof(int) * n);
  psgh_tables(g, kx, ky, kz, L);
  memset(xii, 0, sizeof(double) * 2 * n);
  memset(eta, 0, sizeof(double) * 2 * n);
  memset(zet, 0, sizeof(double) * 2 * n);
  psgh_rng_seed(iseed);
  double ens = 0.0;
  double *f[3] = {xii, eta, zet};
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikz = 0; ikz < g->iktz; ikz++) {
        const long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        if (!L[ik]) continue;
        const double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
        const double ek = (wk > 2.5 && wk <= 9.5) ? 1.0 : 0.0;
        for (int c = 0; c < 3; c++) {
          double s, co;
          sincos(psgh_ranno0(), &s, &co);
          f[c][2 * ik] = ek * co;
          f[c][2 * ik + 1] = ek * s;
        }
        const double r1 = xii[2 * ik] * xii[2 * ik] + eta[2 * ik] * eta[2 * ik] + zet[2 * ik] * zet[2 * ik];
        const double r2 = xii[2 * ik + 1] * xii[2 * ik + 1] + eta[2 * ik + 1] * eta[2 * ik + 1] +
                          zet[2 * ik + 1] * zet[2 * ik + 1];
        ens = ens + r1 + r2;
      }
  ens *= 2.0;
  const double sc = sqrt(ampv / ens);
  for (long ik = 0; ik < n; ik++)
    for (int c = 0; c < 3; c++) {
      f[c][2 * ik] = L[ik] ? f[c][2 * ik] * sc : 0.0;
      f[c][2 * ik + 1] = L[ik] ? f[c][2 * ik + 1] * sc : 0.0;
    }
  free(kx); free(ky); free(kz); free(L);
}

/* 3D_Euler/src/init.F90:137-190: every stored mode draws one phase;
 * zet = EK + PK e^{i phase}, EK = +-1/4 at (kx = alpha, ky = +-1, kz = 0) (the Taylor-Green pair),
 * PK = AMPV on 1.5 < |k| <= 3.5 with kz = 1; xii = eta = 0; no mask and no normalisation */
void psgh_ic_euler_ref(const psgh_grid *g, int iseed, double ampv, double *xii, double *eta, double *zet) {
  const long n = g->nkt;
  double *kx = malloc(sizeof(double) * n), *ky = malloc(sizeof(double) * n), *kz = malloc(sizeof(double) * n);
  psgh_tables(g, kx, ky, kz, NULL);
  memset(xii, 0, sizeof(double) * 2 * n);
  memset(eta, 0, sizeof(double) * 2 * n);
  psgh_rng_seed(iseed);
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikz = 0; ikz < g->iktz; ikz++) {
        const long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        const double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
        const double pk = (wk > 1.5 && wk <= 3.5 && kz[ik] == 1.0) ? ampv : 0.0;
        const double ek = (ikx == 1 && kz[ik] == 0.0) ? (ky[ik] == 1.0 ? 0.25 : ky[ik] == -1.0 ? -0.25 : 0.0) : 0.0;
        double s, co;
        sincos(psgh_ranno0(), &s, &co);
        zet[2 * ik] = ek + pk * co;
        zet[2 * ik + 1] = pk * s;
      }
  free(kx); free(ky); free(kz);
}

/* 3D_Boussinesq/src/init.F90:234-305: every stored mode draws three phases (the third unused);
 * xii, eta = 0.1 e^{i phase} on 2.5 < |k| <= 9.5, zet = -(kx xii + ky eta)/kz, all three zero
 * where kz = 0; 2 sum|w|^2 over every stored mode normalised to AMPV on the mask; rho = 0 */
void psgh_ic_bous_ref(const psgh_grid *g, int iseed, double ampv, double *xii, double *eta, double *zet,
                      double *rho) {
  const long n = g->nkt;
  double *kx = malloc(sizeof(double) * n), *ky = malloc(sizeof(double) * n), *kz = malloc(sizeof(double) * n);
  int *L = malloc(sizeof(int) * n);
  psgh_tables(g, kx, ky, kz, L);
  memset(rho, 0, sizeof(double) * 2 * n);
  psgh_rng_seed(iseed);
  double ens = 0.0;
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++)
      for (int ikz = 0; ikz < g->iktz; ikz++) {
        const long ik = ((long)ikz * g->ikty + iky) * g->iktx + ikx;
        const double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik] + kz[ik] * kz[ik]);
        const double ek = (wk > 2.5 && wk <= 9.5) ? 0.1 : 0.0;
        double s1, c1, s2, c2;
        sincos(psgh_ranno0(), &s1, &c1);
        sincos(psgh_ranno0(), &s2, &c2);
        (void)psgh_ranno0();
        double *w[3] = {xii + 2 * ik, eta + 2 * ik, zet + 2 * ik};
        if (kz[ik] != 0.0) {
          w[0][0] = ek * c1; w[0][1] = ek * s1;
          w[1][0] = ek * c2; w[1][1] = ek * s2;
          w[2][0] = -(kx[ik] * w[0][0] + ky[ik] * w[1][0]) / kz[ik];
          w[2][1] = -(kx[ik] * w[0][1] + ky[ik] * w[1][1]) / kz[ik];
        } else {
          for (int c = 0; c < 3; c++) w[c][0] = w[c][1] = 0.0;
        }
        const double r1 = w[0][0] * w[0][0] + w[1][0] * w[1][0] + w[2][0] * w[2][0];
        const double r2 = w[0][1] * w[0][1] + w[1][1] * w[1][1] + w[2][1] * w[2][1];
        ens = ens + r1 + r2;
      }
  ens *= 2.0;
  if (ens != 0.0) {
    const double sc = sqrt(ampv / ens);
    double *f[3] = {xii, eta, zet};
    for (long ik = 0; ik < n; ik++)
      for (int c = 0; c < 3; c++) {
        f[c][2 * ik] = L[ik] ? f[c][2 * ik] * sc : 0.0;
        f[c][2 * ik + 1] = L[ik] ? f[c][2 * ik + 1] * sc : 0.0;
      }
  }
  free(kx); free(ky); free(kz); free(L);
}

void psgh_ic_ns2d_ref(const psgh_grid *g, int iseed, double ampv, double *z) {
  const long n = g->nkt;
  double *kx = malloc(sizeof(double) * n), *ky = malloc(sizeof(double) * n);
  int *L = malloc(sizeof(int) * n);
  psgh_tables(g, kx, ky, NULL, L);
  memset(z, 0, sizeof(double) * 2 * n);
  psgh_rng_seed(iseed);
  double ens = 0.0;
  for (int ikx = 0; ikx < g->iktx; ikx++)
    for (int iky = 0; iky < g->ikty; iky++) {
      const long ik = (long)iky * g->iktx + ikx;
      if (!L[ik]) continue;
      const double wk = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik]);
      double ek = (wk > 2.5 && wk <= 9.5) ? 1.0 / wk : 0.0;
      if (kx[ik] < 0.1) ek = 0.0;
      const double vk = pow(wk, 2.0) * ek;
      double s, co;
      sincos(psgh_ranno0(), &s, &co);
      const double a = sqrt(vk / wk);
      z[2 * ik] = a * co;
      z[2 * ik + 1] = a * s;
      ens = ens + z[2 * ik] * z[2 * ik] + z[2 * ik + 1] * z[2 * ik + 1];
    }
  ens *= 2.0;
  const double sc = sqrt(ampv / ens);
  for (long ik = 0; ik < n; ik++) {
    z[2 * ik] = L[ik] ? z[2 * ik] * sc : 0.0;
    z[2 * ik + 1] = L[ik] ? z[2 * ik + 1] * sc : 0.0;
  }
  free(kx); free(ky); free(L);
}

/* SplitMix64 (Steele, Lea & Flood 2014) as a counter-based generator */
static uint64_t sm64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
static double u01(uint64_t key, uint64_t i) { return (double)(sm64(key * 0x100000001B3ULL ^ i) >> 11) * 0x1.0p-53; }
static double gauss(uint64_t key, uint64_t i) {
  const double a = 1.0 - u01(key, 2 * i), b = u01(key, 2 * i + 1);
  return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * b);
}

void psgh_ic_solenoidal(const psgh_grid *g, uint64_t seed, double energy, double rho_scale,
                        double *xii, double *eta, double *zet, double *rho) {
  const long n = g->nkt;
  double *kx = malloc(sizeof(double) * n), *ky = malloc(sizeof(double) * n), *kz = malloc(sizeof(double) * n);
  int *L = malloc(sizeof(int) * n);
  double *u = calloc(6 * n, sizeof(double));
  psgh_tables(g, kx, ky, kz, L);
  double eu = 0.0, er = 0.0;
  const uint64_t rkey = seed ^ 0xD1B54A32D192ED03ULL;
  for (long ik = 0; ik < n; ik++) {
    if (rho) { rho[2 * ik] = 0.0; rho[2 * ik + 1] = 0.0; }
    if (!L[ik]) continue;
    const double k[3] = {kx[ik], ky[ik], kz[ik]};
    const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2], km = sqrt(k2);
    const double amp = km * exp(-(km / 4.0) * (km / 4.0));
    double re[3], im[3];
    for (int c = 0; c < 3; c++) {
      re[c] = gauss(seed, 6 * (uint64_t)ik + 2 * c);
      im[c] = gauss(seed, 6 * (uint64_t)ik + 2 * c + 1);
    }
    const double pr = (k[0] * re[0] + k[1] * re[1] + k[2] * re[2]) / k2;
    const double pi = (k[0] * im[0] + k[1] * im[1] + k[2] * im[2]) / k2;
    for (int c = 0; c < 3; c++) {
      u[6 * ik + 2 * c] = amp * (re[c] - k[c] * pr);
      u[6 * ik + 2 * c + 1] = amp * (im[c] - k[c] * pi);
      eu += u[6 * ik + 2 * c] * u[6 * ik + 2 * c] + u[6 * ik + 2 * c + 1] * u[6 * ik + 2 * c + 1];
    }
    if (rho) {
      rho[2 * ik] = amp * gauss(rkey, 2 * (uint64_t)ik);
      rho[2 * ik + 1] = amp * gauss(rkey, 2 * (uint64_t)ik + 1);
      er += rho[2 * ik] * rho[2 * ik] + rho[2 * ik + 1] * rho[2 * ik + 1];
    }
  }
  const double su = eu > 0.0 ? sqrt(energy / eu) : 0.0;
  const double sr = er > 0.0 ? rho_scale * sqrt(energy / er) : 0.0;
  for (long ik = 0; ik < n; ik++) {
    double *q = u + 6 * ik;
    for (int c = 0; c < 6; c++) q[c] *= su;
    const double a = kx[ik], b = ky[ik], c = kz[ik];
    /* w = i (k x u) */
    const double cxr = b * q[4] - c * q[2], cxi = b * q[5] - c * q[3];
    const double cyr = c * q[0] - a * q[4], cyi = c * q[1] - a * q[5];
    const double czr = a * q[2] - b * q[0], czi = a * q[3] - b * q[1];
    xii[2 * ik] = -cxi; xii[2 * ik + 1] = cxr;
    eta[2 * ik] = -cyi; eta[2 * ik + 1] = cyr;
    zet[2 * ik] = -czi; zet[2 * ik + 1] = czr;
    if (rho) { rho[2 * ik] *= sr; rho[2 * ik + 1] *= sr; }
  }
  free(kx); free(ky); free(kz); free(L); free(u);
}

void psgh_ic_ns2d_cfg(const psgh_grid *g, uint64_t seed, double ampv, double *z) {
  const long n = g->nkt;
  double *kx = malloc(sizeof(double) * n), *ky = malloc(sizeof(double) * n);
  int *L = malloc(sizeof(int) * n);
  psgh_tables(g, kx, ky, NULL, L);
  double s2 = 0.0;
  for (long ik = 0; ik < n; ik++) {
    z[2 * ik] = 0.0; z[2 * ik + 1] = 0.0;
    if (!L[ik]) continue;
    const double km = sqrt(kx[ik] * kx[ik] + ky[ik] * ky[ik]);
    if (km < 1.0 || km > 8.0) continue;
    const double amp = km * exp(-(km / 4.0) * (km / 4.0));
    double s, c;
    sincos(2.0 * M_PI * u01(seed, (uint64_t)ik), &s, &c);
    z[2 * ik] = amp * c; z[2 * ik + 1] = amp * s;
    s2 += z[2 * ik] * z[2 * ik] + z[2 * ik + 1] * z[2 * ik + 1];
  }
  const double f = s2 > 0.0 ? sqrt(ampv / (2.0 * s2)) : 0.0;
  for (long ik = 0; ik < 2 * n; ik++) z[ik] *= f;
  free(kx); free(ky); free(L);
}

/* ------------------------------------------------------------------ params */
typedef struct { const char *label; char kind; } lab_t;

static const lab_t LAB_NS3D_DNS[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"KF = ", 'I'},
  {"Re = ", 'F'}, {"TSTEP = ", 'F'}, {"ResidualThreshold = ", 'F'}, {"outfreq1 = ", 'F'},
  {"outfreq2 = ", 'F'}, {"DELT = ", 'F'}, {"AMPV = ", 'F'}, {"STATSFLAG = ", 'I'}, {"RCFLAG = ", 'I'},
  {"UPOflag = ", 'I'}, {"ISTOP = ", 'I'}, {NULL, 0}};
static const lab_t LAB_NS3D_GMRES[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"Re = ", 'F'},
  {"TSTEP = ", 'F'}, {"outfreq1 = ", 'F'}, {"outfreq2 = ", 'F'}, {"DELT = ", 'F'}, {"KF = ", 'I'},
  {"ResidualThreshold = ", 'F'}, {"STATSFLAG = ", 'I'}, {"RCFLAG = ", 'I'}, {"ICFLAG = ", 'I'},
  {"WNFLAG = ", 'I'}, {"initial x shift = ", 'F'}, {"initial y shift = ", 'F'},
  {"initial z shift = ", 'F'}, {"lamFLAG = ", 'I'}, {"UPOflag = ", 'I'}, {NULL, 0}};
static const lab_t LAB_EULER[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"TSTEP = ", 'F'},
  {"outfreq1 = ", 'F'}, {"outfreq2 = ", 'F'}, {"DELT = ", 'F'}, {"AMPV = ", 'F'}, {"STATSFLAG = ", 'I'},
  {NULL, 0}};
static const lab_t LAB_BOUS_DNS[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"KF = ", 'I'},
  {"Re = ", 'F'}, {"Ri = ", 'F'}, {"Theta = ", 'F'}, {"Sc = ", 'F'}, {"TSTEP = ", 'F'},
  {"ResidualThreshold = ", 'F'}, {"outfreq1 = ", 'F'}, {"outfreq2 = ", 'F'}, {"DELT = ", 'F'},
  {"AMPV = ", 'F'}, {"STATSFLAG = ", 'I'}, {"RCFLAG = ", 'I'}, {"UPOflag = ", 'I'}, {"ADAPTflag = ", 'I'},
  {"ISTOP = ", 'I'}, {"Dtarget = ", 'F'}, {NULL, 0}};
static const lab_t LAB_BOUS_GMRES[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"KF = ", 'I'},
  {"Re = ", 'F'}, {"Ri = ", 'F'}, {"Theta = ", 'F'}, {"Sc = ", 'F'}, {"TSTEP = ", 'F'},
  {"ResidualThreshold = ", 'F'}, {"outfreq1 = ", 'F'}, {"outfreq2 = ", 'F'}, {"DELT = ", 'F'},
  {"AMPV = ", 'F'}, {"OUTPLANE = ", 'I'}, {"STATSFLAG = ", 'I'}, {"RCFLAG = ", 'I'}, {"UPOflag = ", 'I'},
  {"ADAPTflag = ", 'I'}, {"ISTOP = ", 'I'}, {NULL, 0}};
static const lab_t LAB_NS2D_DNS[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"Re = ", 'F'},
  {"TSTEP = ", 'F'}, {"outfreq1 = ", 'F'}, {"outfreq2 = ", 'F'}, {"DELT = ", 'F'}, {"AMPFOR = ", 'F'},
  {"SCALE = ", 'F'}, {"ResidualThreshold = ", 'F'}, {"STFLAG = ", 'I'}, {"TDFCFLAG = ", 'I'},
  {"ICFLAG = ", 'I'}, {"WNFLAG = ", 'I'}, {"initial x shift = ", 'F'}, {"initial y shift = ", 'F'},
  {"lamFLAG = ", 'I'}, {"UPOflag = ", 'I'}, {NULL, 0}};
/* read only when TDFCFLAG == 1 (2D_NavierStokes/src_RK4/parameters.F90:168-172) */
static const lab_t LAB_NS2D_TDFC[] = {{"G = ", 'F'}, {"GAMMA = ", 'F'}, {"R = ", 'F'}, {NULL, 0}};
static const lab_t LAB_NS2D_GMRES[] = {
  {"ISEED = ", 'I'}, {"IREST = ", 'I'}, {"infile = ", 'A'}, {"infile type = ", 'A'}, {"Re = ", 'F'},
  {"TSTEP = ", 'F'}, {"outfreq1 = ", 'F'}, {"outfreq2 = ", 'F'}, {"DELT = ", 'F'}, {"AMPFOR = ", 'F'},
  {"ResidualThreshold = ", 'F'}, {"STFLAG = ", 'I'}, {"RCFLAG = ", 'I'}, {"ICFLAG = ", 'I'},
  {"WNFLAG = ", 'I'}, {"initial x shift = ", 'F'}, {"initial y shift = ", 'F'}, {"lamFLAG = ", 'I'},
  {"UPOflag = ", 'I'}, {NULL, 0}};
static const lab_t LAB_EXTRA_NS2D[] = {
  {"IGUESSEND = ", 'I'}, {"maxPERIOD = ", 'F'}, {"minPERIOD = ", 'F'}, {"gmres ResidualThreshold = ", 'F'},
  {"arclcFLAG = ", 'L'}, {"1stOrder = ", 'L'}, {"2ndOrder PredOnly = ", 'L'}, {"Full 2ndOrder = ", 'L'},
  {"drN_reset = ", 'L'}, {"Fixed Period = ", 'L'}, {"Fixed Period Size = ", 'F'}, {"delta_v2 = ", 'F'},
  {"ArnoldiFLAG = ", 'L'}, {NULL, 0}};
static const lab_t LAB_EXTRA_NS3D[] = {
  {"IGUESSEND = ", 'I'}, {"maxPERIOD = ", 'F'}, {"gmres ResidualThreshold = ", 'F'}, {"arclcFLAG = ", 'L'},
  {"1stOrder = ", 'L'}, {"2ndOrder PredOnly = ", 'L'}, {"Full 2ndOrder = ", 'L'}, {"drN_reset = ", 'L'},
  {"Fixed Period = ", 'L'}, {"Fixed Period Size = ", 'F'}, {"delta_v2 = ", 'F'}, {"ArnoldiFLAG = ", 'L'},
  {NULL, 0}};
static const lab_t LAB_EXTRA_BOUS[] = {
  {"IGUESSEND = ", 'I'}, {"maxPERIOD = ", 'F'}, {"gmres ResidualThreshold = ", 'F'}, {"arclcFLAG = ", 'L'},
  {"1stOrder = ", 'L'}, {"2ndOrder PredOnly = ", 'L'}, {"Full 2ndOrder = ", 'L'}, {"drN_reset = ", 'L'},
  {"Fixed Period = ", 'L'}, {"Fixed Period Size = ", 'F'}, {"delta_v2 = ", 'F'}, {"ArnoldiFLAG = ", 'L'},
  {"ParFLAG = ", 'I'}, {NULL, 0}};

static const lab_t *layout_table(int layout) {
  switch (layout) {
  case PSGH_PARAMS_NS3D_DNS: return LAB_NS3D_DNS;
  case PSGH_PARAMS_NS3D_GMRES: return LAB_NS3D_GMRES;
  case PSGH_PARAMS_EULER: return LAB_EULER;
  case PSGH_PARAMS_BOUS_DNS: return LAB_BOUS_DNS;
  case PSGH_PARAMS_BOUS_GMRES: return LAB_BOUS_GMRES;
  case PSGH_PARAMS_NS2D_DNS: return LAB_NS2D_DNS;
  case PSGH_PARAMS_NS2D_GMRES: return LAB_NS2D_GMRES;
  case PSGH_EXTRA_NS2D: return LAB_EXTRA_NS2D;
  case PSGH_EXTRA_NS3D: return LAB_EXTRA_NS3D;
  case PSGH_EXTRA_BOUS: return LAB_EXTRA_BOUS;
  default: return NULL;
  }
}

/* Fortran list-free formatted input of one Fw.d field: blanks ignored, an
 * explicit decimal point or exponent wins, otherwise the last d digits are
 * the fraction. */
double psgh_fortran_f(const char *field, int w, int d) {
  char buf[64];
  int n = 0, has_point = 0, has_exp = 0;
  for (int i = 0; i < w && field[i] && field[i] != '\n' && field[i] != '\r'; i++) {
    char ch = field[i];
    if (ch == ' ' || ch == '\t') continue;
    if (ch == '.') has_point = 1;
    if (ch == 'e' || ch == 'E' || ch == 'd' || ch == 'D') { has_exp = 1; ch = 'e'; }
    if (n < 63) buf[n++] = ch;
  }
  buf[n] = 0;
  if (n == 0) return 0.0;
  double v = strtod(buf, NULL);
  if (!has_point) {
    if (has_exp) {
      /* mantissa digits carry the implied fraction, exponent applies after */
      char *e = strchr(buf, 'e');
      *e = 0;
      v = strtod(buf, NULL) * pow(10.0, -d) * pow(10.0, strtod(e + 1, NULL));
    } else {
      v = v * pow(10.0, -d);
    }
  }
  return v;
}

static void set_msg(char *msg, int len, const char *a, const char *b) {
  if (msg && len > 0) snprintf(msg, (size_t)len, "%s%s", a, b ? b : "");
}

int psgh_params_read(const char *path, int layout, psgh_kv *kv, char *msg, int msglen) {
  const lab_t *tab = layout_table(layout);
  if (!tab) { set_msg(msg, msglen, "unknown parameter layout", NULL); return -1; }
  FILE *f = fopen(path, "r");
  if (!f) { set_msg(msg, msglen, "cannot open ", path); return -1; }
  memset(kv, 0, sizeof(*kv));
  char line[1024];
  const lab_t *parts[2] = {tab, NULL};
  for (int pass = 0; pass < 2 && parts[pass]; pass++) {
    for (const lab_t *t = parts[pass]; t->label; t++) {
      if (!fgets(line, sizeof line, f)) { fclose(f); set_msg(msg, msglen, "bad label ", t->label); return -1; }
      const size_t ll = strlen(t->label);
      if (strncmp(line, t->label, ll) != 0) { fclose(f); set_msg(msg, msglen, "bad label ", t->label); return -1; }
      const char *val = line + ll;
      int i = kv->n++;
      if (i >= PSGH_MAXKEYS) { fclose(f); set_msg(msg, msglen, "too many keys", NULL); return -1; }
      size_t kl = ll - 3;  /* drop " = " */
      memcpy(kv->key[i], t->label, kl);
      kv->key[i][kl] = 0;
      kv->kind[i] = t->kind;
      if (t->kind == 'I') {
        char fld[11] = {0};
        int n = 0;
        for (int c = 0; c < 10 && val[c] && val[c] != '\n' && val[c] != '\r'; c++)
          if (val[c] != ' ') fld[n++] = val[c];
        kv->ival[i] = strtol(fld, NULL, 10);
      } else if (t->kind == 'F') {
        kv->fval[i] = psgh_fortran_f(val, 12, 4);
      } else if (t->kind == 'L') {
        const char *p = val;
        while (*p == ' ' || *p == '.') p++;
        kv->ival[i] = (*p == 'T' || *p == 't') ? 1 : 0;
      } else {
        int n = 0;
        for (int c = 0; c < 256 && val[c] && val[c] != '\n' && val[c] != '\r'; c++) kv->sval[i][n++] = val[c];
        while (n > 0 && kv->sval[i][n - 1] == ' ') n--;
        kv->sval[i][n] = 0;
      }
    }
    if (pass == 0 && layout == PSGH_PARAMS_NS2D_DNS && psgh_kv_int(kv, "TDFCFLAG", 0) == 1) parts[1] = LAB_NS2D_TDFC;
  }
  fclose(f);
  return 0;
}

static int kv_find(const psgh_kv *kv, const char *key) {
  for (int i = 0; i < kv->n; i++)
    if (strcmp(kv->key[i], key) == 0) return i;
  return -1;
}
long psgh_kv_int(const psgh_kv *kv, const char *key, long dflt) {
  int i = kv_find(kv, key);
  return i < 0 ? dflt : kv->ival[i];
}
double psgh_kv_float(const psgh_kv *kv, const char *key, double dflt) {
  int i = kv_find(kv, key);
  return i < 0 ? dflt : (kv->kind[i] == 'F' ? kv->fval[i] : (double)kv->ival[i]);
}
const char *psgh_kv_str(const psgh_kv *kv, const char *key) {
  int i = kv_find(kv, key);
  return i < 0 ? NULL : kv->sval[i];
}

/* ------------------------------------------------------------------ records */
int psgh_rec_write(FILE *f, const void *const *parts, const long *sizes, int nparts) {
  long tot = 0;
  for (int i = 0; i < nparts; i++) tot += sizes[i];
  if (tot > 2147483639L) return -1;
  int32_t m = (int32_t)tot;
  if (fwrite(&m, 4, 1, f) != 1) return -1;
  for (int i = 0; i < nparts; i++)
    if (sizes[i] && fwrite(parts[i], 1, (size_t)sizes[i], f) != (size_t)sizes[i]) return -1;
  if (fwrite(&m, 4, 1, f) != 1) return -1;
  return 0;
}

long psgh_rec_read(FILE *f, void *buf, long cap) {
  int32_t m1, m2;
  if (fread(&m1, 4, 1, f) != 1 || m1 < 0) return -1;
  if ((long)m1 > cap) {
    if (fread(buf, 1, (size_t)cap, f) != (size_t)cap) return -1;
    if (fseek(f, (long)m1 - cap, SEEK_CUR) != 0) return -1;
  } else if (fread(buf, 1, (size_t)m1, f) != (size_t)m1) return -1;
  if (fread(&m2, 4, 1, f) != 1 || m2 != m1) return -1;
  return m1;
}

int psgh_zk_header(FILE *f, int nrec, int iktx, int ikty, int iktz) {
  int32_t h[4] = {nrec, iktx, ikty, iktz};
  const void *p[1] = {h};
  long s[1] = {iktz > 0 ? 16 : 12};
  return psgh_rec_write(f, p, s, 1);
}

int psgh_zk_record(FILE *f, const double *scalars, int nscalars, const double *fields, long ncomplex) {
  const void *p[2] = {scalars, fields};
  long s[2] = {8L * nscalars, 16L * ncomplex};
  return psgh_rec_write(f, p, s, 2);
}

int psgh_raw_write(const char *path, const double *r, long nr) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  size_t w = fwrite(r, sizeof(double), (size_t)nr, f);
  fclose(f);
  return w == (size_t)nr ? 0 : -1;
}

/* ---------------- restart records ----------------
 * One layout per "infile type": the scalars ahead of the fields as a string of
 * slot letters (T time, V v2, G Ri, A alpha, P period, X/Y/Z shifts, N NER,
 * M NER2, U timeUNDER, R RsymFLAG, - a value the reference reads into
 * `dummy`), whether the file is a C stream (no record markers) and whether a
 * header record (jw, IKTX, IKTY[, IKTZ]) comes first. */
typedef struct {
  char type;
  int stream, header;
  const char *slots, *first;   /* first: slots of record 1 when they differ (2D type B) */
} rec_layout;

static const rec_layout REC_NS2D[] = {   /* 2D_NavierStokes/src_RK4/src_gmres/init.F90:66-106 */
  {'A', 0, 1, "T", NULL},        {'B', 0, 1, "TPXNM", "T"},    {'C', 0, 1, "TPXNM", NULL},
  {'D', 0, 1, "TVPXNM", NULL},   {'E', 0, 1, "T-PXNM", NULL},  {'G', 0, 1, "TPXYN", NULL},
  {'H', 0, 1, "TVPXYN", NULL},   {'I', 0, 1, "T-PXYN", NULL},  {'J', 0, 1, "TVPXYNU", NULL},
  {'K', 0, 1, "T-PXYNU", NULL},  {'L', 0, 1, "T-PXYRNU", NULL}, {0, 0, 0, NULL, NULL}};
static const rec_layout REC_NS3D[] = {   /* 3D_NavierStokes/src/init.F90:76-129 */
  {'A', 1, 0, "TVPXYZN", NULL},  {'B', 1, 0, "T-PXYZN", NULL}, {'C', 0, 0, "", NULL},
  {'D', 0, 1, "TVPXYZN", NULL},  {0, 0, 0, NULL, NULL}};
static const rec_layout REC_EULER[] = {  /* 3D_Euler/src/init.F90:73-118 */
  {'A', 1, 0, "T", NULL},        {'B', 0, 0, "", NULL},          {'C', 0, 1, "T", NULL},
  {0, 0, 0, NULL, NULL}};
static const rec_layout REC_BOUS[] = {   /* 3D_Boussinesq/src/init.F90:46-182 */
  {'A', 1, 0, "TVPXYZN", NULL},  {'B', 1, 0, "T-PXYZN", NULL},   {'C', 0, 0, "", NULL},
  {'D', 0, 1, "TVPXYZN", NULL},  {'E', 0, 1, "T-PXYZN", NULL},   {'F', 0, 0, "TVGAPXYZN", NULL},
  {'G', 0, 1, "TVGAPXYZN", NULL}, {'H', 0, 1, "TVG-PXYZN", NULL}, {0, 0, 0, NULL, NULL}};

static const rec_layout *rec_find(int family, const char *type) {
  const rec_layout *t = family == PSGH_REC_NS2D ? REC_NS2D : family == PSGH_REC_NS3D ? REC_NS3D :
                        family == PSGH_REC_BOUS ? REC_BOUS : family == PSGH_REC_EULER ? REC_EULER : NULL;
  if (!t || !type || !type[0] || type[1]) return NULL;
  for (; t->type; t++)
    if (t->type == type[0]) return t;
  return NULL;
}

int psgh_restart_known(int family, const char *type) { return rec_find(family, type) != NULL; }

static void rec_assign(const char *slots, const double *sc, psgh_restart *r) {
  for (int i = 0; slots[i]; i++) {
    const double v = sc[i];
    switch (slots[i]) {
      case 'T': r->time = v; break;
      case 'V': r->v2 = v; r->has_v2 = 1; break;
      case 'G': r->ri = v; r->has_ri = 1; break;
      case 'A': r->alpha = v; r->has_alpha = 1; break;
      case 'P': r->period = v; break;
      case 'X': r->shiftx = v; break;
      case 'Y': r->shifty = v; break;
      case 'Z': r->shiftz = v; break;
      case 'N': r->ner = v; break;
      case 'M': r->ner2 = v; break;
      case 'U': r->tunder = v; break;
      case 'R': r->rsym = v; break;
      default: break;
    }
  }
}

struct psgh_restart_file {
  FILE *f;
  const rec_layout *t;
  int family, nf;
  long nkt, nread;
  double sc[16];
};

psgh_restart_file *psgh_restart_open(const char *path, int family, const char *type, int iktx, int ikty, int iktz,
                                     int nf, long nkt, char *msg, int msglen) {
  const rec_layout *t = rec_find(family, type);
  if (!t) { snprintf(msg, (size_t)msglen, "infile type %s unknown", type ? type : "(none)"); return NULL; }
  FILE *f = fopen(path, "rb");
  if (!f) { snprintf(msg, (size_t)msglen, "cannot open %s", path); return NULL; }
  if (t->header) {
    int32_t h[4] = {0, 0, 0, 0};
    const long want = family == PSGH_REC_NS2D ? 12 : 16;
    if (psgh_rec_read(f, h, sizeof h) != want || h[1] != iktx || h[2] != ikty ||
        (family != PSGH_REC_NS2D && h[3] != iktz)) {
      snprintf(msg, (size_t)msglen, "infile header does not match the grid (IKTX, IKTY%s)",
               family == PSGH_REC_NS2D ? "" : ", IKTZ");
      fclose(f);
      return NULL;
    }
  }
  psgh_restart_file *rf = calloc(1, sizeof *rf);
  if (!rf) { fclose(f); snprintf(msg, (size_t)msglen, "out of memory"); return NULL; }
  rf->f = f; rf->t = t; rf->family = family; rf->nf = nf; rf->nkt = nkt;
  return rf;
}

int psgh_restart_next(psgh_restart_file *rf, double *state, psgh_restart *r, char *msg, int msglen) {
  const rec_layout *t = rf->t;
  const char *slots = (rf->nread == 0 && t->first) ? t->first : t->slots;
  const long ns = (long)strlen(slots), fbytes = 16L * rf->nf * rf->nkt, want = 8L * ns + fbytes;
  int bad = 0;
  if (t->stream) {
    bad = (ns && fread(rf->sc, 8, (size_t)ns, rf->f) != (size_t)ns) ||
          fread(state, 1, (size_t)fbytes, rf->f) != (size_t)fbytes;
  } else {
    /* one record: the scalars, then the fields, from the one payload */
    int32_t m1 = 0, m2 = 0;
    bad = fread(&m1, 4, 1, rf->f) != 1 || (long)m1 != want ||
          (ns && fread(rf->sc, 8, (size_t)ns, rf->f) != (size_t)ns) ||
          fread(state, 1, (size_t)fbytes, rf->f) != (size_t)fbytes || fread(&m2, 4, 1, rf->f) != 1 || m2 != m1;
  }
  if (bad) {
    snprintf(msg, (size_t)msglen, "infile record %ld: missing, or its length does not match type %c on this grid",
             rf->nread + 1, t->type);
    return -1;
  }
  rec_assign(slots, rf->sc, r);
  rf->nread++;
  return 0;
}

void psgh_restart_close(psgh_restart_file *rf) {
  if (!rf) return;
  fclose(rf->f);
  free(rf);
}

int psgh_restart_read(const char *path, int family, const char *type, long irest, int iktx, int ikty, int iktz,
                      int nf, long nkt, double *state, psgh_restart *r, char *msg, int msglen) {
  if (irest < 1) { snprintf(msg, (size_t)msglen, "IREST must be >= 1"); return -1; }
  psgh_restart_file *rf = psgh_restart_open(path, family, type, iktx, ikty, iktz, nf, nkt, msg, msglen);
  if (!rf) return -1;
  int rc = 0;
  for (long i = 1; i <= irest && rc == 0; i++) rc = psgh_restart_next(rf, state, r, msg, msglen);
  psgh_restart_close(rf);
  return rc;
}
