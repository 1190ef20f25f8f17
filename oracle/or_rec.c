/*
 * Oracle: recurrence search over a DNS history (TEST INFRASTRUCTURE ONLY).
 *
 * A sequential restatement of the reference's near-recurrence check, with a
 * host ring of full states as the reference keeps it:
 *   computeNorm      3D_NavierStokes/src/functions.h:111-125 (masked),
 *                    3D_Boussinesq/src/functions.h:190-208 (unmasked, ri*x*x + y*y
 *                    for the density), 2D_NavierStokes/src_RK4/functions.h:59-67
 *   computeZdash     3D_NavierStokes/src/functions.h:213-251, 2D :105-130
 *   recurrence_check 3D_NavierStokes/src/functions.h:253-389, 3D_Boussinesq
 *                    :459-619 (periods < 0.5 skipped, residual over both
 *                    norms, ri-weighted density), 2D :194-309 (timeUNDER)
 *   the NOUT block   3D_NavierStokes/src/cuda.cu:114-139, 3D_Boussinesq/src/out_high.h:8-33
 *   set-up           3D_NavierStokes/src/init.h:165-213, 3D_Boussinesq/src/init.h:275-310
 * The 2D reference has no caller of its search in the tree; its subset is
 * taken as in 3D (0 <= ikx < W, KTY-W <= iky < KTY+W) and its y shifts as
 * 2 pi i/NSHIFTY.  The reference's abs() of the period test is read as the
 * floating-point absolute value.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct or_rec {
  int rule, nf, iktx, ikty, iktz, kty, ktz;
  long nkt;
  double alpha, dt0, tstart, thr, ri, v2, freq2;
  int storeSize, normWidth, NSX, NSY, NSZ, nOut, nNorm;
  int *iNorm;
  int8_t *LL;
  double *kx, *ky, *kz;
  double *shiftX, *shiftY, *shiftZ, *normZStore;
  orc *store;      /* [storeSize][nf][nkt] */
  orc *dash;       /* [nf][nNorm * NSX * NSY * NSZ] */
  int nStore, nCopy, iStart, RecOUTflag, minSTOREouter, izkout;
  double minRESIDUALouter, minTIMEouter, RecPERIOD, minTIMEinner;
  double minSHIFTXouter, minSHIFTYouter, minSHIFTZouter, minSHIFTXinner, minSHIFTYinner, minSHIFTZinner;
  /* the last check's residual table */
  double *last_res;
  int *last_slots;
  int last_nj;
};

static orc *slot_field(or_rec *r, int slot, int f) { return r->store + ((long)slot * r->nf + f) * r->nkt; }

static double compute_norm(or_rec *r, int iStore) {
  if (r->rule == 1) {                      /* 3D_Boussinesq/src/functions.h:190-208 */
    double normX = 0.0, normE = 0.0, normZ = 0.0, normR = 0.0;
    const orc *X = slot_field(r, iStore, 0), *E = slot_field(r, iStore, 1), *Z = slot_field(r, iStore, 2),
              *R = slot_field(r, iStore, 3);
    for (int i = 0; i < r->nNorm; i++) {
      const int ik = r->iNorm[i];
      normX += X[ik].re * X[ik].re + X[ik].im * X[ik].im;
      normE += E[ik].re * E[ik].re + E[ik].im * E[ik].im;
      normZ += Z[ik].re * Z[ik].re + Z[ik].im * Z[ik].im;
      normR += r->ri * R[ik].re * R[ik].re + R[ik].im * R[ik].im;
    }
    normZ += normX + normE + normR;
    return normZ;
  }
  if (r->rule == 2) {                      /* 2D functions.h:59-67 */
    double normZ = 0.0;
    const orc *Z = slot_field(r, iStore, 0);
    for (int i = 0; i < r->nNorm; i++) {
      const int ik = r->iNorm[i];
      normZ += Z[ik].re * Z[ik].re + Z[ik].im * Z[ik].im;
    }
    return normZ;
  }
  double normX = 0.0, normE = 0.0, normZ = 0.0;   /* 3D_NavierStokes/src/functions.h:111-125 */
  const orc *X = slot_field(r, iStore, 0), *E = slot_field(r, iStore, 1), *Z = slot_field(r, iStore, 2);
  for (int i = 0; i < r->nNorm; i++) {
    const int ik = r->iNorm[i];
    if (r->LL[ik] == 1) {
      normX += X[ik].re * X[ik].re + X[ik].im * X[ik].im;
      normE += E[ik].re * E[ik].re + E[ik].im * E[ik].im;
      normZ += Z[ik].re * Z[ik].re + Z[ik].im * Z[ik].im;
    }
  }
  normZ += normX + normE;
  return normZ;
}

static void compute_zdash(or_rec *r, int iStore) {
  const long nsh = (long)r->nNorm * r->NSX * r->NSY * r->NSZ;
  for (int iShiftY = 0; iShiftY < r->NSY; iShiftY++)
    for (int iShiftZ = 0; iShiftZ < r->NSZ; iShiftZ++)
      for (int iShiftX = 0; iShiftX < r->NSX; iShiftX++)
        for (int i = 0; i < r->nNorm; i++) {
          const int ik = r->iNorm[i];
          const long iks = i + (long)r->nNorm * (iShiftX + r->NSX * (iShiftY + r->NSY * iShiftZ));
          const double shift = r->kx[ik] * r->shiftX[iShiftX] + r->ky[ik] * r->shiftY[iShiftY] +
                               r->kz[ik] * r->shiftZ[iShiftZ];
          for (int f = 0; f < r->nf; f++) {
            const orc z = slot_field(r, iStore, f)[ik];
            orc *d = r->dash + f * nsh + iks;
            d->re = cos(shift) * z.re + sin(shift) * z.im;
            d->im = cos(shift) * z.im - sin(shift) * z.re;
          }
        }
}

or_rec *or_rec_new(int rule, int nf, int iktx, int ikty, int iktz, int kty, int ktz, double alpha, const int8_t *L,
                   int store_size, int norm_width, int nsx, int nsy, int nsz, int nout, double dt0, double tstart,
                   double thr, double ri, double v2, const double *state0) {
  or_rec *r = calloc(1, sizeof *r);
  if (!r) return NULL;
  r->rule = rule; r->nf = nf; r->iktx = iktx; r->ikty = ikty; r->iktz = iktz; r->kty = kty; r->ktz = ktz;
  r->alpha = alpha; r->dt0 = dt0; r->tstart = tstart; r->thr = thr; r->ri = ri; r->v2 = v2;
  r->storeSize = store_size; r->normWidth = norm_width; r->NSX = nsx; r->NSY = nsy; r->NSZ = nsz; r->nOut = nout;
  r->nkt = (long)iktx * ikty * (rule == 2 ? 1 : iktz);
  r->freq2 = 2 * nout * dt0;
  const int W = norm_width;
  r->nNorm = rule == 2 ? 2 * W * W : 4 * W * W * W;
  r->iNorm = malloc(sizeof(int) * r->nNorm);
  r->LL = malloc(r->nkt);
  r->kx = malloc(sizeof(double) * r->nkt);
  r->ky = malloc(sizeof(double) * r->nkt);
  r->kz = malloc(sizeof(double) * r->nkt);
  r->shiftX = malloc(sizeof(double) * nsx);
  r->shiftY = malloc(sizeof(double) * nsy);
  r->shiftZ = malloc(sizeof(double) * nsz);
  r->normZStore = calloc(store_size, sizeof(double));
  r->store = calloc((size_t)store_size * nf * r->nkt, sizeof(orc));
  r->dash = malloc(sizeof(orc) * (size_t)nf * r->nNorm * nsx * nsy * nsz);
  r->last_res = malloc(sizeof(double) * (size_t)store_size * nsx * nsy * nsz);
  r->last_slots = malloc(sizeof(int) * store_size);
  if (!r->iNorm || !r->LL || !r->kx || !r->ky || !r->kz || !r->shiftX || !r->shiftY || !r->shiftZ ||
      !r->normZStore || !r->store || !r->dash || !r->last_res || !r->last_slots) {
    or_rec_free(r);
    return NULL;
  }
  memcpy(r->LL, L, r->nkt);
  for (long ik = 0; ik < r->nkt; ik++) {   /* init.F90 KX, KY, KZ of each stored mode */
    const long ikx = ik % iktx, iky = (ik / iktx) % ikty, ikz = rule == 2 ? 0 : ik / ((long)iktx * ikty);
    r->kx[ik] = ikx * alpha;
    r->ky[ik] = (double)(iky - kty);
    r->kz[ik] = rule == 2 ? 0.0 : (double)(ikz - ktz);
  }
  if (rule == 2) {
    for (int iy = 0; iy < 2 * W; iy++)
      for (int ix = 0; ix < W; ix++) r->iNorm[iy * W + ix] = (kty - W + iy) * iktx + ix;
  } else {                                  /* init.h:184-191 */
    for (int iz = 0; iz < 2 * W; iz++)
      for (int iy = 0; iy < 2 * W; iy++)
        for (int ix = 0; ix < W; ix++) {
          const int ik = ((ktz - W + iz) * ikty + ktz - W + iy) * iktx + ix;
          r->iNorm[(iz * 2 * W + iy) * W + ix] = ik;
        }
  }
  const double twopi = 4.0 * asin(1.0);
  for (int i = 0; i < nsy; i++) r->shiftY[i] = twopi * i / (nsy);
  for (int i = 0; i < nsx; i++) r->shiftX[i] = twopi * i / (nsx * (alpha));
  for (int i = 0; i < nsz; i++) r->shiftZ[i] = twopi * i / (nsz);
  /* init.h:165-170, 211-212: entry 0 is the initial state */
  memcpy(r->store, state0, sizeof(orc) * nf * r->nkt);
  r->normZStore[0] = compute_norm(r, 0);
  return r;
}

void or_rec_free(or_rec *r) {
  if (!r) return;
  free(r->iNorm); free(r->LL); free(r->kx); free(r->ky); free(r->kz);
  free(r->shiftX); free(r->shiftY); free(r->shiftZ); free(r->normZStore);
  free(r->store); free(r->dash); free(r->last_res); free(r->last_slots);
  free(r);
}

/* one NOUT block: recurrence_check (when nStore >= 1), then the push of `state`.
 * guess (10): index, store, time, v2, period, sx, sy, sz, residual, time_under;
 * chk (9): checked, compared, store, count, residual, sx, sy, sz, time.
 * Returns 1 when a guess record was produced. */
int or_rec_update(or_rec *r, long NT, double delt, const double *state, double *guess, double *guess_state,
                  double *chk) {
  int produced = 0;
  const int nOut = r->nOut, storeSize = r->storeSize;
  const long nsh = (long)r->nNorm * r->NSX * r->NSY * r->NSZ;
  if (chk) memset(chk, 0, sizeof(double) * 9);
  r->last_nj = 0;
  if (r->nStore >= 1) {
    const int nCopy = r->nCopy;
    double minSx = 0, minSy = 0, minSz = 0, resPrev;
    double minRESIDUALinner = 10000.0;
    double minPERIOD = 10000.0;
    double res = 0.0;
    double normFN = 0.0;
    double normFNN = 0.0;
    int minPflag = 1;
    int minSTOREinner = 0;
    int minCOUNTinner = 0;
    compute_zdash(r, nCopy);
    const double tme = r->tstart + (NT - nOut) * delt;
    for (int iCount = (r->nStore - 1); iCount >= 1; iCount--) {
      int iStore;
      const double PERIODinner = ((r->nStore - iCount) * nOut) * (delt);
      iStore = iCount + r->iStart;
      iStore -= (int)floor(iStore / storeSize) * storeSize;
      if (r->rule == 1 && PERIODinner < 0.5) continue;
      if (PERIODinner > 20.0 && fabs(PERIODinner - r->freq2 * floor(PERIODinner / r->freq2 + 0.5)) > 0.0001) continue;
      resPrev = res;
      res = 10000.0;
      const int jj = r->last_nj++;
      r->last_slots[jj] = iStore;
      for (int iShiftY = 0; iShiftY < r->NSY; iShiftY++)
        for (int iShiftZ = 0; iShiftZ < r->NSZ; iShiftZ++)
          for (int iShiftX = 0; iShiftX < r->NSX; iShiftX++) {
            const int sidx = iShiftX + r->NSX * (iShiftY + r->NSY * iShiftZ);
            const long is = (long)r->nNorm * sidx;
            double nZdash = 0.0;
            normFNN = 0.0;
            for (int i = 0; i < r->nNorm; i++) {
              const int ik = r->iNorm[i];
              for (int f = 0; f < r->nf; f++) {
                const orc d = r->dash[f * nsh + is + i];
                const orc s = slot_field(r, iStore, f)[ik];
                if (r->rule == 1 && f == 3) {
                  normFNN -= r->ri * (2.0 * (d.re * s.re + d.im * s.im));
                  nZdash += r->ri * (d.re * d.re + d.im * d.im);
                } else {
                  normFNN -= 2.0 * (d.re * s.re + d.im * s.im);
                  nZdash += d.re * d.re + d.im * d.im;
                }
              }
            }
            normFN = normFNN + nZdash + r->normZStore[iStore];
            const double den = r->rule == 1 ? r->normZStore[iStore] + r->normZStore[nCopy] : r->normZStore[nCopy];
            const double rr = sqrt(normFN / den);
            r->last_res[(long)jj * r->NSX * r->NSY * r->NSZ + sidx] = rr;
            if (rr < res) {
              res = rr;
              minSx = r->shiftX[iShiftX];
              minSy = r->shiftY[iShiftY];
              minSz = r->shiftZ[iShiftZ];
            }
          }
      if (res > resPrev && minPflag == 1) {
        minPERIOD = PERIODinner;
      } else {
        minPflag = 0;
      }
      if (res < minRESIDUALinner && PERIODinner > minPERIOD) {
        minRESIDUALinner = res;
        minSTOREinner = iStore;
        minCOUNTinner = r->nStore - iCount;
        r->minSHIFTXinner = minSx;
        r->minSHIFTYinner = minSy;
        r->minSHIFTZinner = minSz;
      }
    }
    if (chk) {
      chk[0] = 1; chk[1] = r->last_nj; chk[4] = minRESIDUALinner; chk[8] = tme;
      if (minRESIDUALinner < 10000.0) {
        chk[2] = minSTOREinner; chk[3] = minCOUNTinner;
        chk[5] = r->minSHIFTXinner; chk[6] = r->minSHIFTYinner; chk[7] = r->minSHIFTZinner;
      }
    }
    if (minRESIDUALinner < r->thr && r->RecOUTflag == 0) {
      r->minRESIDUALouter = r->thr;
      if (r->rule == 2) r->minTIMEinner = tme + minCOUNTinner * nOut * delt;
      r->RecOUTflag = 1;
    }
    if (minRESIDUALinner < r->minRESIDUALouter && r->RecOUTflag == 1) {
      r->minRESIDUALouter = minRESIDUALinner;
      r->minSTOREouter = minSTOREinner;
      r->minTIMEouter = tme + minCOUNTinner * nOut * delt;
      r->minSHIFTXouter = r->minSHIFTXinner;
      r->minSHIFTYouter = r->minSHIFTYinner;
      r->minSHIFTZouter = r->minSHIFTZinner;
      r->RecPERIOD = ((minCOUNTinner) * nOut) * (delt);
    }
    if (minRESIDUALinner > r->thr && r->RecOUTflag == 1) {
      double timeUNDER = 0.0;
      r->RecOUTflag = 0;
      if (r->rule == 2) {
        r->minTIMEinner = tme + minCOUNTinner * nOut * delt - r->minTIMEinner;
        timeUNDER = sqrt(r->minTIMEinner * r->minTIMEinner / (r->RecPERIOD * r->RecPERIOD));
      }
      r->izkout = r->izkout + 1;
      if (guess) {
        guess[0] = r->izkout; guess[1] = r->minSTOREouter; guess[2] = r->minTIMEouter; guess[3] = r->v2;
        guess[4] = r->RecPERIOD; guess[5] = r->minSHIFTXouter; guess[6] = r->minSHIFTYouter;
        guess[7] = r->rule == 2 ? 0.0 : r->minSHIFTZouter; guess[8] = r->minRESIDUALouter; guess[9] = timeUNDER;
      }
      if (guess_state) memcpy(guess_state, slot_field(r, r->minSTOREouter, 0), sizeof(orc) * r->nf * r->nkt);
      produced = 1;
    }
  }
  /* cuda.cu:120-138 */
  r->nStore++;
  if (r->nStore >= storeSize) {
    r->nStore = storeSize;
    r->nCopy++;
    r->nCopy -= (int)floor(r->nCopy / storeSize) * storeSize;
    r->iStart = r->nCopy;
  } else {
    r->nCopy = r->nStore;
  }
  memcpy(slot_field(r, r->nCopy, 0), state, sizeof(orc) * r->nf * r->nkt);
  r->normZStore[r->nCopy] = compute_norm(r, r->nCopy);
  return produced;
}

/* residual table of the last check: res[j * nshift + s], slots[j] */
int or_rec_last(const or_rec *r, double *res, int *slots, int cap) {
  if (r->last_nj > cap) return -1;
  const long ns = (long)r->NSX * r->NSY * r->NSZ;
  if (res) memcpy(res, r->last_res, sizeof(double) * r->last_nj * ns);
  if (slots) memcpy(slots, r->last_slots, sizeof(int) * r->last_nj);
  return r->last_nj;
}
