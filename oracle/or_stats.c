/*
 * Oracle: run statistics of the DNS callers (TEST INFRASTRUCTURE ONLY).
 * Sequential restatements of the reference's host loops:
 *   or_stats_ns3d   3D_NavierStokes/src/functions.h:128-180 (computeStats sums)
 *   or_stats_bous   3D_Boussinesq/src/functions.h:275-360 (computeStats sums)
 *   or_rig          3D_Boussinesq/src/out_low.h:18-76 (RiG plane sums; the
 *                   theta < pi/4 branch's index -1 read through the flat index,
 *                   wrapped modulo the grid, as the GPU path defines it)
 *   or_spectra      3D_Boussinesq/src/finalise.h:24-60
 * Raw sums, in the order of include/psg_stats.h, before the reference's
 * normalisations.  Time averages are updated in place with count.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

static double dmax(double a, double b) { return a > b ? a : b; }

void or_stats_ns3d(long nkt, const double *kx, const double *ky, const double *kz, const int *LL, const orc *X,
                   const orc *E, const orc *Z, orc *Xt, orc *Et, orc *Zt, double cnt, double ampfor, double kfy,
                   double *out) {
  double diss = 0, energy = 0, eturb = 0, dturb = 0, input = 0, XIa = 0, ETa = 0, ZTa = 0, Ua = 0, Va = 0, Wa = 0;
  for (long i = 0; i < nkt; i++) {
    if (LL[i] != 1) continue;
    const double wk = dmax(kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i], 0.001);
    double VZX = X[i].re * X[i].re + X[i].im * X[i].im;
    double VZY = E[i].re * E[i].re + E[i].im * E[i].im;
    double VZZ = Z[i].re * Z[i].re + Z[i].im * Z[i].im;
    double VZ = VZX + VZY + VZZ;
    diss += VZ;
    energy += VZ / wk;
    XIa += VZX;
    ETa += VZY;
    ZTa += VZZ;
    Ua += pow((ky[i] * Z[i].im - kz[i] * E[i].im) / wk, 2) + pow((kz[i] * E[i].re - ky[i] * Z[i].re) / wk, 2);
    Va += pow((kz[i] * X[i].im - kx[i] * Z[i].im) / wk, 2) + pow((kx[i] * Z[i].re - kz[i] * X[i].re) / wk, 2);
    Wa += pow((kx[i] * E[i].im - ky[i] * X[i].im) / wk, 2) + pow((ky[i] * X[i].re - kx[i] * E[i].re) / wk, 2);
    Xt[i].re = (Xt[i].re * cnt + X[i].re) / (cnt + 1.0);
    Xt[i].im = (Xt[i].im * cnt + X[i].im) / (cnt + 1.0);
    Et[i].re = (Et[i].re * cnt + E[i].re) / (cnt + 1.0);
    Et[i].im = (Et[i].im * cnt + E[i].im) / (cnt + 1.0);
    Zt[i].re = (Zt[i].re * cnt + Z[i].re) / (cnt + 1.0);
    Zt[i].im = (Zt[i].im * cnt + Z[i].im) / (cnt + 1.0);
    VZX = Xt[i].re * Xt[i].re + Xt[i].im * Xt[i].im;
    VZY = Et[i].re * Et[i].re + Et[i].im * Et[i].im;
    VZZ = Zt[i].re * Zt[i].re + Zt[i].im * Zt[i].im;
    VZ = VZX + VZY + VZZ;
    eturb += VZ / wk;
    dturb += VZ;
    if (kx[i] == 0.0 && ky[i] == kfy && kz[i] == kfy) input += ampfor * X[i].re;
    if (kx[i] == 0.0 && ky[i] == kfy && kz[i] == -kfy) input -= ampfor * X[i].re;
  }
  const double v[11] = {diss, energy, eturb, dturb, input, XIa, ETa, ZTa, Ua, Va, Wa};
  memcpy(out, v, sizeof v);
}

void or_stats_bous(long nkt, const double *kx, const double *ky, const double *kz, const int *LL, const orc *X,
                   const orc *E, const orc *Z, const orc *R, orc *Xt, orc *Et, orc *Zt, orc *Rt, double cnt,
                   double ampfor, double kfy, double ri, double theta, double *out) {
  const double sinTh = sin(theta), cosTh = cos(theta);
  double diss = 0, energy = 0, eturb = 0, dturb = 0, input = 0, pe = 0, ped = 0, ped1 = 0, ped2 = 0;
  double peturb = 0, pedturb = 0, bflux = 0;
  for (long i = 0; i < nkt; i++) {
    if (LL[i] != 1) continue;
    const double wk = dmax(kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i], 0.001);
    double VZX = X[i].re * X[i].re + X[i].im * X[i].im;
    double VZY = E[i].re * E[i].re + E[i].im * E[i].im;
    double VZZ = Z[i].re * Z[i].re + Z[i].im * Z[i].im;
    double VZ = VZX + VZY + VZZ;
    diss += VZ;
    energy += VZ / wk;
    if (kx[i] == 0.0 && ky[i] == kfy && kz[i] == 0.0) input -= ampfor * Z[i].re / kfy;
    double VR = R[i].re * R[i].re + R[i].im * R[i].im;
    pe += VR;
    ped += VR * wk;
    ped1 += kx[i] * kx[i] * VR;
    ped2 += ky[i] * ky[i] * VR;
    Xt[i].re = (Xt[i].re * cnt + X[i].re) / (cnt + 1.0);
    Xt[i].im = (Xt[i].im * cnt + X[i].im) / (cnt + 1.0);
    Et[i].re = (Et[i].re * cnt + E[i].re) / (cnt + 1.0);
    Et[i].im = (Et[i].im * cnt + E[i].im) / (cnt + 1.0);
    Zt[i].re = (Zt[i].re * cnt + Z[i].re) / (cnt + 1.0);
    Zt[i].im = (Zt[i].im * cnt + Z[i].im) / (cnt + 1.0);
    Rt[i].re = (Rt[i].re * cnt + R[i].re) / (cnt + 1.0);
    Rt[i].im = (Rt[i].im * cnt + R[i].im) / (cnt + 1.0);
    VZX = Xt[i].re * Xt[i].re + Xt[i].im * Xt[i].im;
    VZY = Et[i].re * Et[i].re + Et[i].im * Et[i].im;
    VZZ = Zt[i].re * Zt[i].re + Zt[i].im * Zt[i].im;
    VZ = VZX + VZY + VZZ;
    VR = Rt[i].re * Rt[i].re + Rt[i].im * Rt[i].im;
    eturb += VZ / wk;
    dturb += VZ;
    peturb += VR;
    pedturb += VR * wk;
    const double vpR = (kz[i] * X[i].im - kx[i] * Z[i].im) / wk;
    const double vpC = (kx[i] * Z[i].re - kz[i] * X[i].re) / wk;
    const double wpR = (kx[i] * E[i].im - ky[i] * X[i].im) / wk;
    const double wpC = (ky[i] * X[i].re - kx[i] * E[i].re) / wk;
    const double ubpR = vpR * sinTh + wpR * cosTh;
    const double ubpC = vpC * sinTh + wpC * cosTh;
    const double phR = atan2(R[i].im, R[i].re);
    const double phW = atan2(ubpC, ubpR);
    bflux += ri * sqrt(ubpR * ubpR + ubpC * ubpC) * sqrt(R[i].re * R[i].re + R[i].im * R[i].im) * cos(phR - phW);
  }
  const double v[12] = {diss, energy, eturb, dturb, input, pe, ped, ped1, ped2, peturb, pedturb, bflux};
  memcpy(out, v, sizeof v);
}

/* U, V, W, RO: physical fields, x fastest.  navg/savg: one entry per plane. */
void or_rig(int nx, int ny, int nz, double theta, const double *U, const double *V, const double *W,
            const double *RO, double *navg, double *savg) {
  const double twopi = 4.0 * asin(1.0);
  const long nr = (long)nx * ny * nz;
  if (theta < twopi / 8.0) {
    const double *xir = U;
    (void)W;
    for (int iy = 0; iy < ny; iy++) {
      double Navg = 0.0, Savg = 0.0;
      int iyp = iy + 1;
      if (iy == ny - 1) iyp = 1;
      int iym = iy - 1;
      if (iy == 0) iyp = ny - 1;
      for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
          const long ikp = ((long)iz * ny + iyp) * nx + ix;
          long ikm = ((long)iz * ny + iym) * nx + ix;
          if (ikm < 0) ikm += nr;
          Navg += 0.5 * ny * (RO[ikp] - RO[ikm]) / twopi + 1.0;
          Savg += pow(0.5 * ny * (xir[ikp] - xir[ikm]) / twopi, 2);
        }
      navg[iy] = Navg;
      savg[iy] = Savg;
    }
  } else {
    const double *ztr = V, *xir = U;
    for (int iz = 0; iz < nz; iz++) {
      double Navg = 0.0, Savg = 0.0;
      int izp = iz + 1;
      if (iz == nz - 1) izp = 0;
      int izm = iz - 1;
      if (iz == 0) izm = nz - 1;
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
          const long ikp = ((long)izp * ny + iy) * nx + ix;
          const long ikm = ((long)izm * ny + iy) * nx + ix;
          Navg += 0.5 * nz * (RO[ikp] - RO[ikm]) / twopi + 1.0;
          Savg += pow(0.5 * nz * (ztr[ikp] - ztr[ikm]) / twopi, 2) + pow(0.5 * nz * (xir[ikp] - xir[ikm]) / twopi, 2);
        }
      navg[iz] = Navg;
      savg[iz] = Savg;
    }
  }
}

/* spch/spcv: nx + 1 entries, zeroed here; returns size */
int or_spectra(long nkt, int nx, const double *kx, const double *ky, const double *kz, const orc *Xt, const orc *Et,
               const orc *Zt, double theta, double *spch, double *spcv) {
  const double twopi = 4.0 * asin(1.0);
  int size = 0;
  for (int i = 0; i <= nx; i++) spch[i] = spcv[i] = 0.0;
  for (long i = 0; i < nkt; i++) {
    double wkh, wkv;
    if (theta < twopi / 8.0) {
      wkh = dmax(sqrt(kx[i] * kx[i] + kz[i] * kz[i]), 0.001);
      wkv = dmax(ky[i], 0.001);
    } else {
      wkh = dmax(sqrt(kx[i] * kx[i] + ky[i] * ky[i]), 0.001);
      wkv = dmax(kz[i], 0.001);
    }
    const int jh = (int)(wkh + 0.5), jv = (int)(wkv + 0.5);
    if (jh > nx || jv > nx) continue;
    const double wk = dmax(kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i], 0.001);
    const double VZX = Xt[i].re * Xt[i].re + Xt[i].im * Xt[i].im;
    const double VZY = Et[i].re * Et[i].re + Et[i].im * Et[i].im;
    const double VZZ = Zt[i].re * Zt[i].re + Zt[i].im * Zt[i].im;
    const double VZ = (VZX + VZY + VZZ) / wk;
    spch[jh] += VZ;
    spcv[jv] += VZ;
    size = size > jh ? size : jh;
    size = size > jv ? size : jv;
  }
  return size;
}
