/*
 * FP64 FFT for the oracle (test infrastructure only; see oracle.h).
 *
 * Plain iterative radix-2 decimation-in-time with a per-length FP64
 * twiddle table (glibc sincos of 2*pi*k/n).  Real transforms use the
 * textbook half-length packing (even/odd samples as real/imag parts).
 * The 3D/2D wrappers restate what cufftPlan3d(NZ,NY,NX,CUFFT_Z2D/D2Z)
 * compute on the padded half-complex array used by the reference
 * (3D_NavierStokes/src/init.h:100-101, functions.h:51,95;
 *  2D_NavierStokes/src_RK4/src_gmres/init.h:82-84).
 * Pinned against numpy.fft golden vectors in tests/golden/.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <assert.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXLOG 16
static orc *g_tw[MAXLOG + 1];       /* g_tw[l][k] = e^{-2 pi i k / 2^l}, k < 2^(l-1) */

static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

static const orc *twiddles(int n) {
  int l = ilog2(n);
  assert((1 << l) == n && l <= MAXLOG);
  orc *t = __atomic_load_n(&g_tw[l], __ATOMIC_ACQUIRE);
  if (t) return t;
#pragma omp critical(or_fft_tw)
  {
    if (!g_tw[l]) {
      int h = n / 2 > 0 ? n / 2 : 1;
      orc *w = (orc *)malloc(sizeof(orc) * h);
      for (int k = 0; k < h; k++) {
        double s, c;
        sincos(2.0 * M_PI * (double)k / (double)n, &s, &c);
        w[k].re = c; w[k].im = -s;
      }
      __atomic_store_n(&g_tw[l], w, __ATOMIC_RELEASE);
    }
  }
  return g_tw[l];
}

void or_fft_c2c(orc *x, int n, int sign) {
  if (n <= 1) return;
  const orc *w = twiddles(n);
  /* bit reversal */
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { orc t = x[i]; x[i] = x[j]; x[j] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < half; k++) {
        orc t = w[k * step];
        double wr = t.re, wi = sign < 0 ? t.im : -t.im;
        orc *a = &x[i + k], *b = &x[i + k + half];
        double br = b->re * wr - b->im * wi;
        double bi = b->re * wi + b->im * wr;
        b->re = a->re - br; b->im = a->im - bi;
        a->re += br;        a->im += bi;
      }
    }
  }
}

void or_fft_many(orc *x, int n, int sign, long es, long nb, long bs) {
#pragma omp parallel
  {
    orc *buf = (orc *)malloc(sizeof(orc) * n);
#pragma omp for schedule(static)
    for (long b = 0; b < nb; b++) {
      orc *p = x + b * bs;
      for (int j = 0; j < n; j++) buf[j] = p[j * es];
      or_fft_c2c(buf, n, sign);
      for (int j = 0; j < n; j++) p[j * es] = buf[j];
    }
    free(buf);
  }
}

/* r2c via half-length complex FFT of z_m = r_{2m} + i r_{2m+1} */
static void r2c_buf(const double *r, orc *out, int n, orc *z) {
  int m = n / 2;
  for (int j = 0; j < m; j++) { z[j].re = r[2 * j]; z[j].im = r[2 * j + 1]; }
  or_fft_c2c(z, m, -1);
  const orc *w = twiddles(n);
  for (int k = 0; k <= m; k++) {
    orc zk = z[k % m], zc = z[(m - k) % m];
    double er = 0.5 * (zk.re + zc.re), ei = 0.5 * (zk.im - zc.im);
    /* O_k = (Z_k - conj Z_{m-k}) / (2i) */
    double dr = 0.5 * (zk.re - zc.re), di = 0.5 * (zk.im + zc.im);
    double orr = di, oi = -dr;
    double wr, wi;
    if (k < m) { wr = w[k].re; wi = w[k].im; } else { wr = -1.0; wi = 0.0; }
    out[k].re = er + (orr * wr - oi * wi);
    out[k].im = ei + (orr * wi + oi * wr);
  }
}

/* c2r: Z_k = (X_k + conj X_{m-k}) + i W^{-k} (X_k - conj X_{m-k}); z = IFFT_m(Z) */
static void c2r_buf(const orc *in, double *r, int n, orc *z) {
  int m = n / 2;
  const orc *w = twiddles(n);
  for (int k = 0; k < m; k++) {
    orc a = in[k], b = in[m - k];
    if (k == 0) { a.im = 0.0; b.im = 0.0; }       /* Im X_0, Im X_{n/2} ignored */
    double sr = a.re + b.re, si = a.im - b.im;    /* X_k + conj X_{m-k} */
    double dr = a.re - b.re, di = a.im + b.im;    /* X_k - conj X_{m-k} */
    double wr = w[k].re, wi = -w[k].im;           /* W^{-k} */
    double tr = dr * wr - di * wi, ti = dr * wi + di * wr;
    z[k].re = sr - ti; z[k].im = si + tr;         /* + i*(tr + i ti) */
  }
  or_fft_c2c(z, m, +1);
  for (int j = 0; j < m; j++) { r[2 * j] = z[j].re; r[2 * j + 1] = z[j].im; }
}

void or_fft_r2c(const double *r, orc *out, int n) {
  orc *z = (orc *)malloc(sizeof(orc) * (n / 2 + 1));
  r2c_buf(r, out, n, z);
  free(z);
}

void or_fft_c2r(const orc *in, double *r, int n) {
  orc *z = (orc *)malloc(sizeof(orc) * (n / 2 + 1));
  c2r_buf(in, r, n, z);
  free(z);
}

void or_fft3_c2r(orc *F, double *r, int nx, int ny, int nz) {
  int nx2 = nx / 2 + 1;
  /* z then y (complex, inverse), then x (C2R) */
#pragma omp parallel
  {
    int nmax = nz > ny ? nz : ny;
    orc *buf = (orc *)malloc(sizeof(orc) * (nmax > nx ? nmax : nx));
#pragma omp for schedule(static)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx2; ix++) {
        orc *p = F + (long)iy * nx2 + ix;
        long st = (long)ny * nx2;
        for (int j = 0; j < nz; j++) buf[j] = p[j * st];
        or_fft_c2c(buf, nz, +1);
        for (int j = 0; j < nz; j++) p[j * st] = buf[j];
      }
#pragma omp for schedule(static)
    for (int iz = 0; iz < nz; iz++)
      for (int ix = 0; ix < nx2; ix++) {
        orc *p = F + (long)iz * ny * nx2 + ix;
        for (int j = 0; j < ny; j++) buf[j] = p[(long)j * nx2];
        or_fft_c2c(buf, ny, +1);
        for (int j = 0; j < ny; j++) p[(long)j * nx2] = buf[j];
      }
#pragma omp for schedule(static)
    for (long row = 0; row < (long)ny * nz; row++)
      c2r_buf(F + row * nx2, r + row * nx, nx, buf);
    free(buf);
  }
}

void or_fft3_r2c(const double *r, orc *F, int nx, int ny, int nz) {
  int nx2 = nx / 2 + 1;
#pragma omp parallel
  {
    int nmax = nz > ny ? nz : ny;
    orc *buf = (orc *)malloc(sizeof(orc) * (nmax > nx ? nmax : nx));
#pragma omp for schedule(static)
    for (long row = 0; row < (long)ny * nz; row++)
      r2c_buf(r + row * nx, F + row * nx2, nx, buf);
#pragma omp for schedule(static)
    for (int iz = 0; iz < nz; iz++)
      for (int ix = 0; ix < nx2; ix++) {
        orc *p = F + (long)iz * ny * nx2 + ix;
        for (int j = 0; j < ny; j++) buf[j] = p[(long)j * nx2];
        or_fft_c2c(buf, ny, -1);
        for (int j = 0; j < ny; j++) p[(long)j * nx2] = buf[j];
      }
#pragma omp for schedule(static)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx2; ix++) {
        orc *p = F + (long)iy * nx2 + ix;
        long st = (long)ny * nx2;
        for (int j = 0; j < nz; j++) buf[j] = p[j * st];
        or_fft_c2c(buf, nz, -1);
        for (int j = 0; j < nz; j++) p[j * st] = buf[j];
      }
    free(buf);
  }
}

void or_fft2_c2r(orc *F, double *r, int nx, int ny) { or_fft3_c2r(F, r, nx, ny, 1); }
void or_fft2_r2c(const double *r, orc *F, int nx, int ny) { or_fft3_r2c(r, F, nx, ny, 1); }

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
