// Newton-GMRES-hookstep periodic-orbit solver on the persistent GPU context
// (algorithm and citations in include/psg_newton.h).  Krylov vectors and all
// state-sized algebra live in HBM as FP64; every inner product is the masked
// real product sum_L Re(conj(a) b) reduced on the GPU (block partials summed on
// the host in a fixed order, so results are run-to-run deterministic).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <complex>
#include <coroutine>
#include <vector>
#include "psg_launch.hpp"
#include "../../include/psg_newton.h"

namespace psg {

int timederiv_dev(Ctx *c, const cpx<double> *in, cpx<double> *out);
int tl_flow_map(Ctx *c, const cpx<double> *z0, const cpx<double> *dz0, cpx<double> *z1, cpx<double> *dz1, double T,
                double dt0, int policy, bool reuse_base);

namespace {

constexpr int RED_BLOCKS = 240;

__global__ void k_dot(const cpx<double> *__restrict__ a, const cpx<double> *__restrict__ b,
                      const int8_t *__restrict__ L, long nkt, long n, double *__restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (L[i % nkt]) s += a[i].x * b[i].x + a[i].y * b[i].y;
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// masked real product over the modes with kx^2 + ky^2 + kz^2 <= kmax2 only
// (the low-wavenumber residual NER2, 2D_NavierStokes/src_RK4/src_gmres/main.F90:737-751)
__global__ void k_dot_low(const cpx<double> *__restrict__ a, const cpx<double> *__restrict__ b, Grid g,
                          double kmax2, const int8_t *__restrict__ L, long n, double *__restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long ik = i % g.nkt;
    if (!L[ik]) continue;
    const int ikx = (int)(ik % g.iktx);
    const long t = ik / g.iktx;
    const int iky = (int)(t % g.ikty), ikz = (int)(t / g.ikty);
    const double kx = ikx * g.alpha, ky = iky - g.kty, kz = g.dim == 3 ? (double)(ikz - g.ktz) : 0.0;
    if (kx * kx + ky * ky + kz * kz > kmax2) continue;
    s += a[i].x * b[i].x + a[i].y * b[i].y;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// y = a*x + b*y  (masked: off-mask entries set to 0)
__global__ void k_axpby(double a, const cpx<double> *__restrict__ x, double b, cpx<double> *__restrict__ y,
                        const int8_t *__restrict__ L, long nkt, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (!L[i % nkt]) { y[i] = cpx<double>{0.0, 0.0}; continue; }
    // b == 0 never reads y (it may hold uninitialised memory)
    const cpx<double> yy = b == 0.0 ? cpx<double>{0.0, 0.0} : y[i];
    const cpx<double> xx = a == 0.0 ? cpx<double>{0.0, 0.0} : x[i];
    y[i] = cpx<double>{a * xx.x + b * yy.x, a * xx.y + b * yy.y};
  }
}

// out = exp(-i(kx sx + ky sy + kz sz)) in, masked (the reference's ZNSHIFT loops)
__global__ void k_shift(const cpx<double> *__restrict__ in, cpx<double> *__restrict__ out, Grid g, double sx,
                        double sy, double sz, const int8_t *__restrict__ L, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long ik = i % g.nkt;
    if (!L[ik]) { out[i] = cpx<double>{0.0, 0.0}; continue; }
    const int ikx = (int)(ik % g.iktx);
    const long t = ik / g.iktx;
    const int iky = (int)(t % g.ikty), ikz = (int)(t / g.ikty);
    const double kx = ikx * g.alpha, ky = iky - g.kty, kz = g.dim == 3 ? (double)(ikz - g.ktz) : 0.0;
    double s, co;
    sincos(-(kx * sx + ky * sy + kz * sz), &s, &co);
    const cpx<double> a = in[i];
    out[i] = cpx<double>{a.x * co - a.y * s, a.x * s + a.y * co};
  }
}

// out = -i k_dir in, masked (dZ/dS)
__global__ void k_ddir(const cpx<double> *__restrict__ in, cpx<double> *__restrict__ out, Grid g, int dir,
                       const int8_t *__restrict__ L, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long ik = i % g.nkt;
    const int ikx = (int)(ik % g.iktx);
    const long t = ik / g.iktx;
    const int iky = (int)(t % g.ikty), ikz = (int)(t / g.ikty);
    const double k = dir == 0 ? ikx * g.alpha : (dir == 1 ? (double)(iky - g.kty) : (double)(ikz - g.ktz));
    const cpx<double> a = in[i];
    out[i] = L[ik] ? cpx<double>{k * a.y, -k * a.x} : cpx<double>{0.0, 0.0};
  }
}

}  // namespace

// ------------------------------------------------------------------ vector helpers
struct VecOps {
  Ctx *c;
  long n;  // complex entries per vector (nf*nkt)
  double *dpart = nullptr;
  std::vector<double> hpart;
  unsigned grid() const { return (unsigned)std::min<long>(4096, (n + 255) / 256); }
  explicit VecOps(Ctx *cc) : c(cc), n((long)cc->nf * cc->g.nkt), hpart(RED_BLOCKS) {
    if (hipMalloc(&dpart, sizeof(double) * RED_BLOCKS) != hipSuccess) dpart = nullptr;
  }
  ~VecOps() { if (dpart) (void)hipFree(dpart); }
  double dot(const cpx<double> *a, const cpx<double> *b) {
    hipLaunchKernelGGL(k_dot, dim3(RED_BLOCKS), dim3(256), 0, c->stream, a, b, c->mask, c->g.nkt, n, dpart);
    (void)hipMemcpyAsync(hpart.data(), dpart, sizeof(double) * RED_BLOCKS, hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    double s = 0.0;
    for (int i = 0; i < RED_BLOCKS; i++) s += hpart[i];
    return s;
  }
  double dot_low(const cpx<double> *a, const cpx<double> *b, double kmax2) {
    hipLaunchKernelGGL(k_dot_low, dim3(RED_BLOCKS), dim3(256), 0, c->stream, a, b, c->g, kmax2, c->mask, n, dpart);
    (void)hipMemcpyAsync(hpart.data(), dpart, sizeof(double) * RED_BLOCKS, hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    double s = 0.0;
    for (int i = 0; i < RED_BLOCKS; i++) s += hpart[i];
    return s;
  }
  void axpby(double a, const cpx<double> *x, double b, cpx<double> *y) {
    hipLaunchKernelGGL(k_axpby, dim3(grid()), dim3(256), 0, c->stream, a, x, b, y, c->mask, c->g.nkt, n);
  }
  void copy(const cpx<double> *x, cpx<double> *y) { axpby(1.0, x, 0.0, y); }
  void shift(const cpx<double> *in, cpx<double> *out, double sx, double sy, double sz) {
    hipLaunchKernelGGL(k_shift, dim3(grid()), dim3(256), 0, c->stream, in, out, c->g, sx, sy, sz, c->mask, n);
  }
  void ddir(const cpx<double> *in, cpx<double> *out, int dir) {
    hipLaunchKernelGGL(k_ddir, dim3(grid()), dim3(256), 0, c->stream, in, out, c->g, dir, c->mask, n);
  }
};

// Phi_T on device buffers: out = flow map of in over time T
static int flow(Ctx *c, const cpx<double> *in, cpx<double> *out, double T, double dt0, int policy) {
  const size_t bytes = sizeof(cpx<double>) * c->nf * c->g.nkt;
  PSG_HIP(hipMemcpyAsync(c->state, in, bytes, hipMemcpyDeviceToDevice, c->stream));
  int nstop = (int)(T / dt0);
  if (nstop < 1) nstop = 1;
  int r = 0;
  if (policy == PSG_TIME_UNIFORM) {
    const double dt = T / (double)nstop;
    if (c->cfg.dt != dt && (r = psg_set_dt(static_cast<psg_ctx *>(c), dt))) return r;
    if ((r = psg_run(static_cast<psg_ctx *>(c), nstop))) return r;
  } else {
    nstop = (int)(T / dt0);
    if (nstop > 0) {
      if (c->cfg.dt != dt0 && (r = psg_set_dt(static_cast<psg_ctx *>(c), dt0))) return r;
      if ((r = psg_run(static_cast<psg_ctx *>(c), nstop))) return r;
    }
    const double last = last_step(T, dt0, nstop);
    if (last != 0.0) {
      if ((r = psg_set_dt(static_cast<psg_ctx *>(c), last))) return r;
      if ((r = psg_run(static_cast<psg_ctx *>(c), 1))) return r;
    }
  }
  PSG_HIP(hipMemcpyAsync(out, c->state, bytes, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// ------------------------------------------------------------------ dense algebra
// One-sided Jacobi (Hestenes) SVD of column-major A (m x n, m >= n).
static int jacobi_svd(int m, int n, const double *A, double *U, double *s, double *V) {
  std::vector<double> W(A, A + (size_t)m * n);
  for (int i = 0; i < n * n; i++) V[i] = 0.0;
  for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 80; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < m; i++) {
          const double a = W[(size_t)p * m + i], b = W[(size_t)q * m + i];
          alpha += a * a; beta += b * b; gamma += a * b;
        }
        if (gamma == 0.0) continue;
        const double rel = fabs(gamma) / sqrt(alpha * beta);
        if (!(rel > 1e-15)) continue;
        off = fmax(off, rel);
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < m; i++) {
          const double a = W[(size_t)p * m + i], b = W[(size_t)q * m + i];
          W[(size_t)p * m + i] = cs * a - sn * b;
          W[(size_t)q * m + i] = sn * a + cs * b;
        }
        for (int i = 0; i < n; i++) {
          const double a = V[(size_t)p * n + i], b = V[(size_t)q * n + i];
          V[(size_t)p * n + i] = cs * a - sn * b;
          V[(size_t)q * n + i] = sn * a + cs * b;
        }
      }
    if (off <= 1e-15) break;
  }
  // singular values = column norms, sorted descending
  std::vector<int> ord(n);
  std::vector<double> nrm(n);
  for (int j = 0; j < n; j++) {
    double t = 0;
    for (int i = 0; i < m; i++) t += W[(size_t)j * m + i] * W[(size_t)j * m + i];
    nrm[j] = sqrt(t);
    ord[j] = j;
  }
  for (int a = 0; a < n; a++)
    for (int b = a + 1; b < n; b++)
      if (nrm[ord[b]] > nrm[ord[a]]) std::swap(ord[a], ord[b]);
  std::vector<double> Vc(V, V + (size_t)n * n);
  for (int k = 0; k < n; k++) {
    const int j = ord[k];
    s[k] = nrm[j];
    for (int i = 0; i < m; i++) U[(size_t)k * m + i] = nrm[j] > 0 ? W[(size_t)j * m + i] / nrm[j] : 0.0;
    for (int i = 0; i < n; i++) V[(size_t)k * n + i] = Vc[(size_t)j * n + i];
  }
  return 0;
}

// trust-region search of main.F90:1283-1333
static double hook_mu(int k, const double *S, const double *P, double delta, double *X) {
  const double r2 = sqrt(2.0);
  auto xnorm = [&](double mu) {
    double t = 0;
    for (int i = 0; i < k; i++) {
      X[i] = mu == 0.0 ? P[i] / S[i] : P[i] * S[i] / (mu + S[i] * S[i]);
      t += X[i] * X[i];
    }
    return sqrt(t);
  };
  double normx = xnorm(0.0), mumid = 0.0;
  if (normx > r2 * delta) {
    double mu1 = 0.0, mu2 = 0.0;
    for (int i = 0; i < k; i++) mu2 = fmax(mu2, S[i]);
    normx = xnorm(mu2);
    while (!(normx < delta)) {
      mu1 = mu2;
      mu2 = 2.0 * mu2;
      normx = xnorm(mu2);
    }
    mumid = mu2;
    for (int it = 0; it < 200 && !(normx < r2 * delta && normx > delta / r2); it++) {
      mumid = 0.5 * (mu2 + mu1);
      normx = xnorm(mumid);
      if (normx > delta) mu1 = mumid;
      else mu2 = mumid;
    }
  }
  return mumid;
}

}  // namespace psg

using namespace psg;

// ------------------------------------------------------------------ Hessenberg eigenvalues
// Householder reflector P = I - beta v v^T with v[0] = 1 mapping x (length m,
// 2 or 3) onto a multiple of e_1 (Golub & Van Loan, Matrix Computations,
// Alg. 5.1.1).
static void house(const double *x, int m, double *v, double &beta) {
  double sigma = 0.0;
  for (int i = 1; i < m; i++) sigma += x[i] * x[i];
  v[0] = 1.0;
  for (int i = 1; i < m; i++) v[i] = x[i];
  if (sigma == 0.0) { beta = 0.0; return; }
  const double mu = sqrt(x[0] * x[0] + sigma);
  const double v0 = x[0] <= 0.0 ? x[0] - mu : -sigma / (x[0] + mu);
  beta = 2.0 * v0 * v0 / (sigma + v0 * v0);
  for (int i = 1; i < m; i++) v[i] /= v0;
}

// Eigenvalues of an n x n upper Hessenberg matrix h (row-major, overwritten)
// by Francis double-shift QR steps with bulge chasing (Golub & Van Loan,
// Alg. 7.5.1-7.5.2), deflating at negligible subdiagonal entries and taking
// the exceptional shifts of EISPACK's HQR after 10 and 20 fruitless steps.
// Complex pairs are stored adjacently, positive imaginary part first.
static int hess_eig(int n, double *h, double *wr, double *wi) {
  auto H = [&](int i, int j) -> double & { return h[(size_t)i * n + j]; };
  double anorm = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = (i > 0 ? i - 1 : 0); j < n; j++) anorm += fabs(H(i, j));
  const double eps = 2.220446049250313e-16;
  if (!(anorm == anorm) || anorm > 1e300) return -1;   // NaN or overflow in the input
  double tshift = 0.0;   // accumulated exceptional shifts, added back to every later eigenvalue
  int hi = n - 1, iter = 0, total = 0;
  const int itmax = 30 * (n > 10 ? n : 10);               // the whole reduction's budget, as LAPACK's dlahqr
  while (hi >= 0) {
    int lo = hi;
    for (; lo > 0; lo--) {
      double s = fabs(H(lo - 1, lo - 1)) + fabs(H(lo, lo));
      if (s == 0.0) s = anorm;
      if (fabs(H(lo, lo - 1)) <= eps * s) { H(lo, lo - 1) = 0.0; break; }
    }
    if (lo == hi) {                       // 1 x 1 block
      wr[hi] = H(hi, hi) + tshift;
      wi[hi] = 0.0;
      hi--;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {                   // 2 x 2 block: its two roots, without cancellation
      const double a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      const double p = 0.5 * (a - d), bc = b * c, disc = p * p + bc;
      if (disc >= 0.0) {
        const double z = p + copysign(sqrt(disc), p);
        wr[hi - 1] = d + z + tshift;
        wr[hi] = (z != 0.0 ? d - bc / z : d) + tshift;
        wi[hi - 1] = wi[hi] = 0.0;
      } else {
        wr[hi - 1] = wr[hi] = d + p + tshift;
        wi[hi - 1] = sqrt(-disc);
        wi[hi] = -wi[hi - 1];
      }
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++total > itmax) return -1;
    // the double shift as the sum s and product t of the two shifts; an
    // exceptional one after every 10 steps without a deflation
    double s, t;
    if (iter > 0 && iter % 10 == 0) {
      const double x = H(hi, hi);
      tshift += x;
      for (int i = 0; i <= hi; i++) H(i, i) -= x;
      const double e = fabs(H(hi, hi - 1)) + fabs(H(hi - 1, hi - 2));
      s = 1.5 * e;
      t = e * e;
    } else {
      s = H(hi - 1, hi - 1) + H(hi, hi);
      t = H(hi - 1, hi - 1) * H(hi, hi) - H(hi - 1, hi) * H(hi, hi - 1);
    }
    iter++;
    // first column of (H - s1)(H - s2) on the active block, then chase the bulge
    double x[3] = {H(lo, lo) * H(lo, lo) + H(lo, lo + 1) * H(lo + 1, lo) - s * H(lo, lo) + t,
                   H(lo + 1, lo) * (H(lo, lo) + H(lo + 1, lo + 1) - s), H(lo + 1, lo) * H(lo + 2, lo + 1)};
    for (int k = lo; k <= hi - 2; k++) {
      double v[3], beta;
      house(x, 3, v, beta);
      const int c0 = k > lo ? k - 1 : lo;
      for (int j = c0; j <= hi; j++) {
        const double dd = beta * (H(k, j) + v[1] * H(k + 1, j) + v[2] * H(k + 2, j));
        H(k, j) -= dd;
        H(k + 1, j) -= dd * v[1];
        H(k + 2, j) -= dd * v[2];
      }
      const int r1 = k + 3 < hi ? k + 3 : hi;
      for (int i = lo; i <= r1; i++) {
        const double dd = beta * (H(i, k) + v[1] * H(i, k + 1) + v[2] * H(i, k + 2));
        H(i, k) -= dd;
        H(i, k + 1) -= dd * v[1];
        H(i, k + 2) -= dd * v[2];
      }
      x[0] = H(k + 1, k);
      x[1] = H(k + 2, k);
      x[2] = k < hi - 2 ? H(k + 3, k) : 0.0;
    }
    {
      double v[2], beta;
      house(x, 2, v, beta);
      for (int j = hi - 2; j <= hi; j++) {
        const double dd = beta * (H(hi - 1, j) + v[1] * H(hi, j));
        H(hi - 1, j) -= dd;
        H(hi, j) -= dd * v[1];
      }
      for (int i = lo; i <= hi; i++) {
        const double dd = beta * (H(i, hi - 1) + v[1] * H(i, hi));
        H(i, hi - 1) -= dd;
        H(i, hi) -= dd * v[1];
      }
    }
    // the reflectors zero the bulge exactly in exact arithmetic: keep the form exact
    for (int i = lo + 2; i <= hi; i++)
      for (int j = lo; j < i - 1; j++) H(i, j) = 0.0;
  }
  return 0;
}

// Eigenvector y (H y = mu y, |y| = 1) of the n x n Hessenberg matrix h
// (row-major) for the eigenvalue mu, by inverse iteration on H - (mu + d) I
// with complex LU and partial pivoting.
static void hess_evec(int n, const double *h, std::complex<double> mu, std::vector<std::complex<double>> &y) {
  typedef std::complex<double> Z;
  double hn = 0.0;
  for (int i = 0; i < n * n; i++) hn = std::max(hn, fabs(h[i]));
  const Z sh = mu + Z(1e-10 * (std::abs(mu) + hn), 0.0);
  std::vector<Z> A((size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A[(size_t)i * n + j] = Z(h[(size_t)i * n + j], 0.0) - (i == j ? sh : Z(0.0, 0.0));
  std::vector<int> piv(n);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::abs(A[(size_t)i * n + k]) > std::abs(A[(size_t)p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
    Z d = A[(size_t)k * n + k];
    if (std::abs(d) == 0.0) d = A[(size_t)k * n + k] = Z(1e-300, 0.0);
    for (int i = k + 1; i < n; i++) {
      const Z f = A[(size_t)i * n + k] / d;
      A[(size_t)i * n + k] = f;
      if (f != Z(0.0, 0.0))
        for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
    }
  }
  y.assign(n, Z(1.0, 0.0));
  for (int it = 0; it < 3; it++) {
    for (int k = 0; k < n; k++) std::swap(y[k], y[piv[k]]);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < i; j++) y[i] -= A[(size_t)i * n + j] * y[j];
    for (int i = n - 1; i >= 0; i--) {
      for (int j = i + 1; j < n; j++) y[i] -= A[(size_t)i * n + j] * y[j];
      y[i] /= A[(size_t)i * n + i];
    }
    double nr = 0.0;
    for (auto &a : y) nr += std::norm(a);
    nr = sqrt(nr);
    for (auto &a : y) a /= nr;
  }
}

extern "C" {

int psg_hess_eig(int n, const double *Hc, double *wr, double *wi) {
  if (n < 1 || !Hc || !wr || !wi) { set_error("psg_hess_eig: bad argument"); return PSG_EINVAL; }
  std::vector<double> h((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++)
    for (int i = 0; i <= std::min(j + 1, n - 1); i++) h[(size_t)i * n + j] = Hc[(size_t)j * n + i];
  if (hess_eig(n, h.data(), wr, wi)) { set_error("psg_hess_eig: QR iteration did not converge"); return -1; }
  return 0;
}

int psg_svd(int m, int n, const double *A, double *U, double *s, double *V) {
  if (m < n || n < 1) return PSG_EINVAL;
  return jacobi_svd(m, n, A, U, s, V);
}

double psg_hook_mu(int k, const double *s, const double *p, double delta, double *x) {
  return hook_mu(k, s, p, delta, x);
}

void psg_newton_defaults(int flavour, psg_newton_opts *o) {
  memset(o, 0, sizeof(*o));
  const bool is3 = flavour != PSG_NS2D;
  o->max_newton = is3 ? 75 : 100;
  o->max_gmres = 40;
  o->max_hook = 50;
  o->netol = is3 ? 1e-10 : 1e-12;
  o->gmtol = flavour == PSG_BOUS3D ? 1e-7 : 1e-3;
  o->eps = 1e-7;
  o->delta0 = flavour == PSG_BOUS3D ? 0.05 : 0.1;
  o->nshift = flavour == PSG_BOUS3D ? 2 : 1;
  o->fixed_period = 0;
  o->jv = 0;
  o->time_policy = is3 ? PSG_TIME_LASTSTEP : PSG_TIME_UNIFORM;
  o->dt0 = 0.005;
  o->sy = 0.0;
  o->stall_count = is3 ? 5 : 3;
  o->tl_record = 1;
}

int psg_flow_map(psg_ctx *c, double *z, double T, double dt0, int policy) {
  const size_t bytes = sizeof(cpx<double>) * c->nf * c->g.nkt;
  cpx<double> *d = nullptr;
  PSG_HIP(hipMalloc(&d, bytes));
  PSG_HIP(hipMemcpyAsync(d, z, bytes, hipMemcpyHostToDevice, c->stream));
  int r = flow(c, d, d, T, dt0, policy);
  if (!r) {
    PSG_HIP(hipMemcpyAsync(z, d, bytes, hipMemcpyDeviceToHost, c->stream));
    PSG_HIP(hipStreamSynchronize(c->stream));
  }
  (void)hipFree(d);
  return r;
}

int psg_tangent(psg_ctx *c, const double *z, const double *dz0, double T, double dt0, int policy, double *z1,
                double *dz1) {
  if (!c || !z || !dz0 || !dz1) { set_error("null argument"); return PSG_EINVAL; }
  const size_t bytes = sizeof(cpx<double>) * c->nf * c->g.nkt;
  cpx<double> *d = nullptr;
  PSG_HIP(hipMalloc(&d, 4 * bytes));
  cpx<double> *a = d, *b = d + c->nf * c->g.nkt, *o1 = b + c->nf * c->g.nkt, *o2 = o1 + c->nf * c->g.nkt;
  int r = 0;
  if (hipMemcpyAsync(a, z, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(b, dz0, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) { set_error("copy"); r = PSG_EHIP; }
  if (!r) r = tl_flow_map(c, a, b, o1, o2, T, dt0, policy, false);
  if (!r) {
    if ((z1 && hipMemcpyAsync(z1, o1, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        hipMemcpyAsync(dz1, o2, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) { set_error("copy"); r = PSG_EHIP; }
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  return r;
}

int psg_tangent_many(psg_ctx *c, const double *z, int ndir, const double *dz0, double T, double dt0, int policy,
                     int fp32, double *z1, double *dz1) {
  if (!c || !z || !dz0 || !dz1 || ndir < 1) { set_error("null argument"); return PSG_EINVAL; }
  if (c->g.dim == 3 && c->cfg.flavour != PSG_NS3D) { set_error("tangent linear: 2D and 3D Navier-Stokes only"); return PSG_EINVAL; }
  const long nn = c->nf * c->g.nkt;
  const size_t bytes = sizeof(cpx<double>) * nn;
  cpx<double> *d = nullptr;
  PSG_HIP(hipMalloc(&d, 4 * bytes));
  cpx<double> *a = d, *b = d + nn, *o1 = b + nn, *o2 = o1 + nn;
  const int rp = (fp32 || c->cfg.precision == PSG_FP32BULK) ? 4 : 8;
  const bool old = c->tl_fp32;
  if (fp32 && !old) { c->tl_fp32 = true; c->graph_gen++; }
  int r = hipMemcpyAsync(a, z, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess ? 0 : PSG_EHIP;
  const bool rec = !r && rec_usable(c, T, dt0, policy, rp);
  if (!r) r = rec ? rec_flow(c, a, o1, T, dt0, policy) : flow(c, a, o1, T, dt0, policy);
  if (!r && z1 && hipMemcpyAsync(z1, o1, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) r = PSG_EHIP;
  for (int i = 0; i < ndir && !r; i++) {
    if (hipMemcpyAsync(b, dz0 + 2 * nn * i, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) { r = PSG_EHIP; break; }
    r = rec ? tl_replay(c, b, o2, T, dt0, policy) : tl_flow_map(c, a, b, o1, o2, T, dt0, policy, false);
    if (!r && hipMemcpyAsync(dz1 + 2 * nn * i, o2, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) r = PSG_EHIP;
  }
  if (r == PSG_EHIP) set_error("psg_tangent_many: copy");
  (void)hipStreamSynchronize(c->stream);
  if (c->tl_fp32 != old) { c->tl_fp32 = old; c->graph_gen++; }
  (void)hipFree(d);
  return r;
}

// jv = 2: the tangent-linear flows of the solve run their passes in FP32-bulk
// (the residual flow stays in the context's precision); captured step graphs
// are rebuilt on each switch
struct TlPrecision {
  Ctx *c;
  bool old;
  TlPrecision(Ctx *ctx, bool fp32) : c(ctx), old(ctx->tl_fp32) {
    if (fp32 != old) { c->tl_fp32 = fp32; c->graph_gen++; }
  }
  ~TlPrecision() {
    if (c->tl_fp32 != old) { c->tl_fp32 = old; c->graph_gen++; }
  }
};

// The solver; with alc, v2 is a further unknown and the pseudo-arclength
// condition <dX/dr, X - X_r> = ds a further row (the reference's arclcFLAG2
// branches, 2D_NavierStokes/src_RK4/src_gmres/main.F90:805-812, 1049-1062,
// 1123-1155, 1350-1366, 1441-1530, 1678, 1720-1726).
}  // extern "C"

// The solver is a coroutine that suspends at every flow map it needs (a residual,
// a finite-difference Jv, a hookstep candidate); its driver serves the request,
// either with one flow on the context (psg_newton_solve) or, for many guesses at
// once, with one batched run of all waiting members (psg_newton_batch).
struct FlowReq {
  const cpx<double> *in = nullptr;
  cpx<double> *out = nullptr;
  double T = 0.0, dt0 = 0.0;
  int policy = 0, rc = 0;
  bool rec = false;   // record the base for the tangent-linear replays (one-guess solves only)
};
struct NTask {
  struct promise_type {
    FlowReq req;
    int result = 0;
    NTask get_return_object() { return NTask{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    std::suspend_always final_suspend() noexcept { return {}; }
    void return_value(int v) { result = v; }
    void unhandled_exception() { result = PSG_EINVAL; }
  };
  std::coroutine_handle<promise_type> h;
  explicit NTask(std::coroutine_handle<promise_type> hh) : h(hh) {}
  NTask(NTask &&o) noexcept : h(o.h) { o.h = {}; }
  NTask(const NTask &) = delete;
  ~NTask() { if (h) h.destroy(); }
};
struct FlowAwait {
  FlowReq r;
  std::coroutine_handle<NTask::promise_type> hp;
  bool await_ready() const noexcept { return false; }
  void await_suspend(std::coroutine_handle<NTask::promise_type> h) noexcept {
    hp = h;
    h.promise().req = r;
  }
  int await_resume() const noexcept { return hp.promise().req.rc; }
};

NTask newton_core(psg_ctx *c, double *zh, double *Tio, double *sxio, double *szio, double *v2io,
                       const psg_alc *alc, const psg_newton_opts *o, psg_newton_result *res) {
  if (!c || !zh || !Tio || !o || !res) { set_error("null argument"); co_return PSG_EINVAL; }
  if (alc && (!v2io || !alc->dzdr || !alc->zr || !(*v2io > 0.0))) { set_error("newton: continuation needs v2, dzdr and zr"); co_return PSG_EINVAL; }
  if (o->jv < 0 || o->jv > 2) { set_error("newton: jv %d is not 0, 1 or 2", o->jv); co_return PSG_EINVAL; }
  if (o->jv == 2 && c->cfg.flavour != PSG_NS2D && c->cfg.flavour != PSG_NS3D) {
    set_error("newton: tangent-linear Jv is 2D/3D Navier-Stokes only");
    co_return PSG_EINVAL;
  }
  TlPrecision tlp(c, o->jv == 2);
  const auto t_start = std::chrono::steady_clock::now();
  memset(res, 0, sizeof(*res));
  VecOps V(c);
  if (!V.dpart) { set_error("newton: reduction buffer"); co_return PSG_ENOMEM; }
  const long n = V.n;
  const size_t bytes = sizeof(cpx<double>) * n;
  const int K = o->max_gmres;
  const int ns = o->nshift;                    // shift unknowns (x[, z])
  const int sdir[2] = {0, 2};
  // device buffers
  std::vector<cpx<double> *> bufs;
  auto dalloc = [&](cpx<double> **p) -> bool {
    if (hipMalloc(p, bytes) != hipSuccess) return false;
    bufs.push_back(*p);
    return true;
  };
  cpx<double> *ZNEWT, *ZN, *ZNN, *ZNSH, *FN, *dZNdt, *dzdt, *dZNdS[2], *dzdS[2], *v, *DIFF, *Zplus, *FNplus, *Zplus2,
      *ZN2, *FNplus2, *work, *Q;
  cpx<double> *dZrdr = nullptr, *Zr = nullptr, *dFdv = nullptr, *work2 = nullptr;
  bool ok = dalloc(&ZNEWT) && dalloc(&ZN) && dalloc(&ZNN) && dalloc(&ZNSH) && dalloc(&FN) && dalloc(&dZNdt) &&
            dalloc(&dzdt) && dalloc(&dZNdS[0]) && dalloc(&dZNdS[1]) && dalloc(&dzdS[0]) && dalloc(&dzdS[1]) &&
            dalloc(&v) && dalloc(&DIFF) && dalloc(&Zplus) && dalloc(&FNplus) && dalloc(&Zplus2) && dalloc(&ZN2) &&
            dalloc(&FNplus2) && dalloc(&work);
  if (ok && alc) ok = dalloc(&dZrdr) && dalloc(&Zr) && dalloc(&dFdv) && dalloc(&work2);
  if (ok) {
    if (hipMalloc(&Q, bytes * (size_t)(K + 1)) != hipSuccess) ok = false;
    else bufs.push_back(Q);
  }
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    for (auto p : bufs) (void)hipFree(p);
  };
  if (!ok) { cleanup(); set_error("newton: device allocation"); co_return PSG_ENOMEM; }
  auto q = [&](int i) { return Q + (size_t)i * n; };
  constexpr int NQ = 4;   // scalar parts of a Krylov vector: [T, sx, sz, v2]
  std::vector<double> qN((size_t)(K + 1) * NQ, 0.0);
  auto qn = [&](int i, int a) -> double & { return qN[(size_t)i * NQ + a]; };
  std::vector<double> H((size_t)(K + 1) * K, 0.0);   // column-major (K+1) x K
  auto h = [&](int i, int j) -> double & { return H[(size_t)j * (K + 1) + i]; };

  double T = *Tio, sh[2] = {sxio ? *sxio : 0.0, szio ? *szio : 0.0};
  const double sy = o->sy;
  double dt0 = o->dt0;
  const double dt0_orig = o->dt0;
  long phi = 0;
  int rc = 0;
#define NTRY(x)              \
  do {                       \
    if ((rc = (x))) {        \
      cleanup();             \
      co_return rc;          \
    }                        \
  } while (0)
#define NFLOW(in, out, TT) NTRY((co_await FlowAwait{FlowReq{(in), (out), (TT), dt0, o->time_policy, 0}}))
#define NFLOWR(in, out, TT) NTRY((co_await FlowAwait{FlowReq{(in), (out), (TT), dt0, o->time_policy, 0, true}}))
  // tangent-linear Jv against a recorded base (2D, one member): the record is
  // taken by the residual flows of the iterate and of the hookstep candidates,
  // and is valid while it holds the flow of the current (ZNEWT, T, dt0)
  const int rec_rp = (o->jv == 2 || c->cfg.precision == PSG_FP32BULK) ? 4 : 8;
  const bool try_rec = o->jv >= 1 && o->tl_record && c->g.dim == 2 && c->batch == 1;
  bool rec_valid = false;
  auto can_rec = [&](double TT) { return try_rec && rec_usable(c, TT, dt0, o->time_policy, rec_rp); };
  auto shift_of = [&](const cpx<double> *in, cpx<double> *out, double sx, double sz) {
    V.shift(in, out, sx, sy, ns > 1 ? sz : 0.0);
  };
  auto norm = [&](const cpx<double> *a) { return sqrt(V.dot(a, a)); };
  // continuation: the context's viscosity follows the flow being computed
  double v2N = v2io ? *v2io : c->cfg.v2;
  auto at_v2 = [&](double vv) { return vv == c->cfg.v2 ? 0 : psg_set_v2(static_cast<psg_ctx *>(c), vv); };
  auto narc_of = [&](const cpx<double> *Zz, double Tz, double Sz, double vz) {
    V.copy(Zz, work2);
    V.axpby(-1.0, Zr, 1.0, work2);
    double r = V.dot(dZrdr, work2) + alc->dsdr * (Sz - alc->sr) + alc->dv2dr * (vz - alc->v2r) - alc->ds;
    if (!o->fixed_period) r += alc->dTdr * (Tz - alc->Tr);
    return r;
  };

  NTRY(hipMemcpyAsync(work, zh, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess ? 0 : PSG_EHIP);
  V.copy(work, ZNEWT);
  if (alc) {
    NTRY(hipMemcpyAsync(work, alc->dzdr, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess ? 0 : PSG_EHIP);
    V.copy(work, dZrdr);
    NTRY(hipMemcpyAsync(work, alc->zr, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess ? 0 : PSG_EHIP);
    V.copy(work, Zr);
  }
  NTRY(at_v2(v2N));
  if (can_rec(T)) {
    NFLOWR(ZNEWT, ZN, T);
    rec_valid = true;
  } else {
    NFLOW(ZNEWT, ZN, T);
  }
  phi++;
  V.copy(ZN, ZNN);
  shift_of(ZN, ZNSH, sh[0], sh[1]);
  // the Boussinesq solver normalises by ||(Phi_T(z), z)|| (3D_Boussinesq/src/src_gmres/main.F90:775,
  // 1698); the Navier-Stokes solvers by ||Phi_T(z)|| (3D main.F90:813,1755; 2D main.F90:734,1696)
  const bool two_norm = c->cfg.flavour == PSG_BOUS3D;
  auto normz_of = [&](const cpx<double> *zn, const cpx<double> *z0) {
    return two_norm ? sqrt(V.dot(zn, zn) + V.dot(z0, z0)) : norm(zn);
  };
  double NORMZ = normz_of(ZN, ZNEWT);
  V.copy(ZNSH, FN);
  V.axpby(-1.0, ZNEWT, 1.0, FN);
  double NORMFN = norm(FN);
  double NER = NORMFN / NORMZ;
  // second residual over |k|^2 <= KYmax2 = ((IKTY-KTY-1)^2)/4 (2D main.F90:124-125); the 3D
  // solvers define none
  const double kymax2 = (double)c->g.kty * (double)c->g.kty / 4.0;
  auto ner2_of = [&](const cpx<double> *F) { return c->g.dim == 2 ? sqrt(V.dot_low(F, F, kymax2)) / NORMZ : -1.0; };
  double NER2 = ner2_of(FN);
  double Delta = o->delta0, NERold = 0.0;
  int NERcount = 0;
  int status = PSG_NEWTON_FAILED, INE = 0;
  long gm_total = 0, hook_total = 0;
  if (o->verbose) { fprintf(stderr, "newton: start T=%.10g sx=%.6g NER=%.3e\n", T, sh[0], NER); fflush(stderr); }
  if (NER <= o->netol) status = PSG_NEWTON_CONVERGED;
  double Narc = alc ? narc_of(ZNEWT, T, sh[0], v2N) : 0.0;
  for (INE = 1; INE <= o->max_newton && status != PSG_NEWTON_CONVERGED; INE++) {
    const double NORMb = alc ? sqrt(NORMFN * NORMFN + Narc * Narc) : NORMFN;
    V.axpby(-1.0 / NORMb, FN, 0.0, q(0));
    for (int a = 0; a < NQ; a++) qn(0, a) = 0.0;
    if (alc) qn(0, 3) = -Narc / NORMb;
    NTRY(at_v2(v2N));
    if (alc && o->jv >= 1) {
      // S dPhi/dv2 for the tangent-linear Jv, by a central difference in v2
      const double hv = 1e-4 * v2N;
      NTRY(at_v2(v2N + hv));
      NFLOW(ZNEWT, dFdv, T);
      NTRY(at_v2(v2N - hv));
      NFLOW(ZNEWT, work2, T);
      phi += 2;
      V.axpby(-1.0, work2, 1.0, dFdv);
      shift_of(dFdv, dFdv, sh[0], sh[1]);
      V.axpby(0.0, dFdv, 0.5 / hv, dFdv);
      NTRY(at_v2(v2N));
    }
    if (try_rec && !rec_valid && can_rec(T)) {
      // the accepted step was not the last recorded flow: record the iterate's
      NFLOWR(ZNEWT, work, T);
      phi++;
      rec_valid = true;
    }
    if (!o->fixed_period) {
      NTRY(timederiv_dev(c, ZNSH, dZNdt));
      NTRY(timederiv_dev(c, ZNEWT, dzdt));
    }
    for (int s = 0; s < ns; s++) {
      V.ddir(ZNSH, dZNdS[s], sdir[s]);
      V.ddir(ZNEWT, dzdS[s], sdir[s]);
    }
    std::fill(H.begin(), H.end(), 0.0);
    int k = 0;
    std::vector<double> U, S, VV, P, X, Y;
    double GMR = 1.0;
    // J applied to (dz, dT, ds): rows 1..N into out, scalar rows into vn.  In three
    // parts around the flow a finite-difference Jv needs: jv_pre perturbs the state
    // (into work), the driver flows it (APPLY_J), jv_post finishes
    auto jv_pre = [&](const cpx<double> *dz, const double *dsc) -> int {
      if (o->jv >= 1) return 0;
      V.copy(ZNEWT, work);
      V.axpby(o->eps, dz, 1.0, work);
      // the finite difference perturbs v2 along with z (main.F90:1049-1062)
      return alc ? at_v2(v2N + o->eps * dsc[3]) : 0;
    };
    auto jv_post = [&](const cpx<double> *dz, const double *dsc, cpx<double> *out, double *vn) -> int {
      int r;
      if (o->jv >= 1) {
        r = rec_valid ? tl_replay(c, dz, out, T, dt0, o->time_policy)
                      : tl_flow_map(c, ZNEWT, dz, work, out, T, dt0, o->time_policy, true);
        if (r) return r;
        shift_of(out, out, sh[0], sh[1]);   // S dPhi dz
        if (alc && dsc[3] != 0.0) V.axpby(dsc[3], dFdv, 1.0, out);
      } else {
        if (alc && (r = at_v2(v2N))) return r;
        V.axpby(-1.0, ZNN, 1.0, out);
        shift_of(out, out, sh[0], sh[1]);
        V.axpby(0.0, out, 1.0 / o->eps, out);
      }
      phi++;
      V.axpby(-1.0, dz, 1.0, out);
      vn[3] = 0.0;
      if (alc) {   // the arclength row (main.F90:1123-1126)
        vn[3] = V.dot(dZrdr, dz) + alc->dsdr * dsc[1] + alc->dv2dr * dsc[3];
        if (!o->fixed_period) vn[3] += alc->dTdr * dsc[0];
      }
      int a = 0;
      if (!o->fixed_period) { V.axpby(dsc[0], dZNdt, 1.0, out); vn[0] = V.dot(dzdt, dz); a = 1; }
      else vn[0] = 0.0;
      for (int s = 0; s < ns; s++) {
        V.axpby(dsc[1 + s], dZNdS[s], 1.0, out);
        vn[1 + s] = V.dot(dzdS[s], dz);
      }
      (void)a;
      return 0;
    };
#define APPLY_J(dz, dsc, out, vn)                              \
  do {                                                         \
    NTRY(jv_pre((dz), (dsc)));                                 \
    if (o->jv == 0) NFLOW(work, (out), T);                     \
    NTRY(jv_post((dz), (dsc), (out), (vn)));                   \
  } while (0)
    for (k = 1; k <= K; k++) {
      double vn[NQ] = {0, 0, 0, 0};
      double dsc[NQ] = {qn(k - 1, 0), qn(k - 1, 1), qn(k - 1, 2), qn(k - 1, 3)};
      APPLY_J(q(k - 1), dsc, v, vn);
      for (int i = 0; i < k; i++) {
        double hij = V.dot(q(i), v);
        if (!o->fixed_period) hij += qn(i, 0) * vn[0];
        for (int s = 0; s < ns; s++) hij += qn(i, 1 + s) * vn[1 + s];
        if (alc) hij += qn(i, 3) * vn[3];
        h(i, k - 1) = hij;
        V.axpby(-hij, q(i), 1.0, v);
        for (int a = 0; a < NQ; a++) vn[a] -= hij * qn(i, a);
      }
      double hn = V.dot(v, v);
      if (!o->fixed_period) hn += vn[0] * vn[0];
      for (int s = 0; s < ns; s++) hn += vn[1 + s] * vn[1 + s];
      if (alc) hn += vn[3] * vn[3];
      hn = sqrt(hn);
      h(k, k - 1) = hn;
      V.axpby(1.0 / hn, v, 0.0, q(k));
      for (int a = 0; a < NQ; a++) qn(k, a) = vn[a] / hn;
      // SVD of H(0:k, 0:k-1)
      std::vector<double> Hk((size_t)(k + 1) * k);
      for (int j = 0; j < k; j++)
        for (int i = 0; i <= k; i++) Hk[(size_t)j * (k + 1) + i] = h(i, j);
      U.assign((size_t)(k + 1) * k, 0.0);
      S.assign(k, 0.0);
      VV.assign((size_t)k * k, 0.0);
      jacobi_svd(k + 1, k, Hk.data(), U.data(), S.data(), VV.data());
      P.assign(k, 0.0);
      X.assign(k, 0.0);
      Y.assign(k, 0.0);
      for (int i = 0; i < k; i++) P[i] = NORMb * U[(size_t)i * (k + 1)];
      for (int i = 0; i < k; i++) X[i] = P[i] / S[i];
      for (int i = 0; i < k; i++) {
        double t = 0;
        for (int j = 0; j < k; j++) t += VV[(size_t)j * k + i] * X[j];
        Y[i] = t;
      }
      double r2 = 0;
      for (int i = 0; i <= k; i++) {
        double t = i == 0 ? -NORMb : 0.0;
        for (int j = 0; j < k; j++) t += h(i, j) * Y[j];
        r2 += t * t;
      }
      GMR = sqrt(r2) / NORMb;
      if (o->verbose > 1) { fprintf(stderr, "  gmres %d  residual %.3e\n", k, GMR); fflush(stderr); }
      if (GMR <= o->gmtol) break;
    }
    if (k > K) k = K;   // Krylov cap reached: hookstep in the space built so far
    gm_total += k;
    if (GMR > o->gmtol && o->max_gmres >= 2000) { status = PSG_NEWTON_GMRES_FAILED; break; }
    // ---------------- hookstep
    int IHOOK = 0;
    bool halve = false, dbl = false, hook_failed = false, rec_last = false;
    double NORMFNplus = 0, NORMFNpred = 0, NORMFNplus2 = 0, Tplus = T, Tplus2 = T, Splus[2] = {sh[0], sh[1]},
           Splus2[2] = {sh[0], sh[1]}, v2plus = v2N, v2plus2 = v2N;
    for (;;) {
      IHOOK++;
      if (IHOOK > o->max_hook) { hook_failed = true; break; }
      const double mu = hook_mu(k, S.data(), P.data(), Delta, X.data());
      for (int i = 0; i < k; i++) {
        double t = 0;
        for (int j = 0; j < k; j++) t += VV[(size_t)j * k + i] * X[j];
        Y[i] = t;
      }
      V.axpby(0.0, DIFF, 0.0, DIFF);
      double dsc[NQ] = {0, 0, 0, 0};
      for (int i = 0; i < k; i++) {
        V.axpby(Y[i], q(i), 1.0, DIFF);
        for (int a = 0; a < NQ; a++) dsc[a] += qn(i, a) * Y[i];
      }
      if (o->fixed_period) dsc[0] = 0.0;
      if (!alc) dsc[3] = 0.0;
      // predicted residual F + J DIFF.  A tangent-linear J is linear in DIFF = Q_k Y,
      // so J DIFF = Q_{k+1} (H_k Y) by the Arnoldi relation; the finite-difference
      // Jv is integrated as the reference does
      if (o->jv >= 1) {
        V.axpby(0.0, v, 0.0, v);
        for (int jj = 0; jj <= k; jj++) {
          double cj = 0.0;
          for (int i = (jj > 0 ? jj - 1 : 0); i < k; i++) cj += h(jj, i) * Y[i];
          if (cj != 0.0) V.axpby(cj, q(jj), 1.0, v);
        }
      } else {
        double vn[NQ];
        APPLY_J(DIFF, dsc, v, vn);
      }
      V.axpby(1.0, FN, 1.0, v);
      NORMFNpred = norm(v);
      // const = 0 in the reference (main.F90:135), so the bound is |b| itself (main.F90:1435-1442)
      const double NORMbpredC = alc ? sqrt(NORMFN * NORMFN + Narc * Narc) : NORMFN;
      // evaluate the candidate
      V.copy(ZNEWT, Zplus);
      V.axpby(1.0, DIFF, 1.0, Zplus);
      Tplus = T + dsc[0];
      for (int s = 0; s < ns; s++) Splus[s] = sh[s] + dsc[1 + s];
      v2plus = v2N + dsc[3];
      if (!(Tplus > 0.0)) {
        // a step to a non-positive period has no flow to evaluate: the trust region
        // is halved as for a candidate whose integration blew up (the reference would
        // step a negative INT(T/dt) count, i.e. only its shortened last step)
        Delta /= 2.0;
        halve = true;
        continue;
      }
      NTRY(at_v2(v2plus));
      // the Jv of this Newton step are done: a candidate may take the record
      rec_valid = false;
      if (v2plus == v2N && can_rec(Tplus)) {
        NFLOWR(Zplus, ZN, Tplus);
        rec_last = true;
      } else {
        NFLOW(Zplus, ZN, Tplus);
        rec_last = false;
      }
      NTRY(at_v2(v2N));
      phi++;
      shift_of(ZN, ZNSH, Splus[0], Splus[1]);
      V.copy(ZNSH, FNplus);
      V.axpby(-1.0, Zplus, 1.0, FNplus);
      NORMFNplus = norm(FNplus);
      double NORMbplus = NORMFNplus;
      if (alc) {   // main.F90:1511-1518
        const double na = narc_of(Zplus, Tplus, Splus[0], v2plus);
        NORMbplus = sqrt(NORMFNplus * NORMFNplus + na * na);
      }
      if (o->verbose > 1) {
        fprintf(stderr, "  hook %d mu=%.3e Delta=%.3e |F+|=%.3e pred=%.3e |F|=%.3e\n", IHOOK, mu, Delta,
                NORMFNplus, NORMFNpred, NORMFN);
        fflush(stderr);
      }
      if (NORMFNplus != NORMFNplus || NORMFNpred != NORMFNpred) {
        Delta /= 2.0;
        halve = true;
        continue;
      }
      if (NORMbplus >= NORMbpredC) {
        if (dbl) {
          V.copy(Zplus2, Zplus); Tplus = Tplus2; Splus[0] = Splus2[0]; Splus[1] = Splus2[1]; v2plus = v2plus2;
          V.copy(ZN2, ZN); V.copy(FNplus2, FNplus); NORMFNplus = NORMFNplus2;
          rec_last = false;
          Delta /= 2.0;
          break;
        }
        Delta /= 2.0;
        halve = true;
      } else {
        if (halve) break;
        if (dbl && NORMFNplus > NORMFNplus2) {
          V.copy(Zplus2, Zplus); Tplus = Tplus2; Splus[0] = Splus2[0]; Splus[1] = Splus2[1]; v2plus = v2plus2;
          V.copy(ZN2, ZN); V.copy(FNplus2, FNplus); NORMFNplus = NORMFNplus2;
          rec_last = false;
          Delta /= 2.0;
          break;
        }
        if (mu == 0.0) break;
        if (fabs(NORMFNpred - NORMFNplus) < 0.1 * fabs(NORMFNplus - NORMFN) || NORMFNplus < NORMFNpred) {
          Delta *= 2.0;
          dbl = true;
          V.copy(Zplus, Zplus2); Tplus2 = Tplus; Splus2[0] = Splus[0]; Splus2[1] = Splus[1]; v2plus2 = v2plus;
          V.copy(ZN, ZN2); V.copy(FNplus, FNplus2); NORMFNplus2 = NORMFNplus;
        } else {
          break;
        }
      }
    }
    hook_total += IHOOK;
    if (NORMFNpred != NORMFNpred || NORMFNplus != NORMFNplus) {
      // time stepping blew up: halve dt and retry this Newton step (main.F90:1636-1665)
      if (dt0 < 0.1 * dt0_orig) { status = PSG_NEWTON_TSTEP_FAILED; break; }
      dt0 /= 2.0;
      rec_valid = false;
      shift_of(ZNN, ZNSH, sh[0], sh[1]);
      continue;
    }
    if (hook_failed) { status = PSG_NEWTON_HSTEP_FAILED; break; }
    // accept (the record is the accepted candidate's when it was the last flowed)
    rec_valid = rec_last;
    V.copy(ZN, ZNN);
    V.copy(Zplus, ZNEWT);
    T = Tplus;
    sh[0] = Splus[0];
    sh[1] = Splus[1];
    v2N = v2plus;
    NTRY(at_v2(v2N));
    shift_of(ZNN, ZNSH, sh[0], sh[1]);
    NORMZ = normz_of(ZNN, ZNEWT);
    V.copy(FNplus, FN);
    NORMFN = NORMFNplus;
    NER = NORMFN / NORMZ;
    if (alc) Narc = narc_of(ZNEWT, T, sh[0], v2N);   // main.F90:1720-1726
    NER2 = ner2_of(FN);   // main.F90:1700-1712
    if (o->verbose) { fprintf(stderr, "newton %d: NER=%.3e T=%.10g sx=%.6g gmres=%d hook=%d\n", INE, NER, T, sh[0], k, IHOOK); fflush(stderr); }
    const double relch = fabs((NER - NERold) / NER);
    if (relch < 0.001) {
      if (++NERcount > o->stall_count) { status = PSG_NEWTON_STALLED; break; }
    } else {
      NERcount = 0;
    }
    NERold = NER;
    if (NER <= o->netol) { status = PSG_NEWTON_CONVERGED; break; }
    if (T < 0.1) { status = PSG_NEWTON_PERIOD_SMALL; break; }
  }
  NTRY(hipMemcpyAsync(zh, ZNEWT, bytes, hipMemcpyDeviceToHost, c->stream) == hipSuccess ? 0 : PSG_EHIP);
  (void)hipStreamSynchronize(c->stream);
  *Tio = T;
  if (sxio) *sxio = sh[0];
  if (szio) *szio = sh[1];
  if (v2io) *v2io = v2N;
  res->T = T; res->sx = sh[0]; res->sz = sh[1];
  res->ner = NER;
  res->ner2 = NER2;
  res->newton_iters = INE > o->max_newton ? o->max_newton : INE;
  res->gmres_iters = (int)gm_total;
  res->hook_iters = (int)hook_total;
  res->status = status;
  res->phi_evals = phi;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  cleanup();
#undef NTRY
#undef NFLOW
#undef APPLY_J
  co_return 0;
}

extern "C" {

// serve each flow request of one solver with a flow on the context (member 0)
static int run_direct(Ctx *c, NTask &t) {
  auto h = t.h;
  h.resume();
  while (!h.done()) {
    FlowReq &rq = h.promise().req;
    rq.rc = rq.rec ? rec_flow(c, rq.in, rq.out, rq.T, rq.dt0, rq.policy)
                   : flow(c, rq.in, rq.out, rq.T, rq.dt0, rq.policy);
    h.resume();
  }
  return h.promise().result;
}

int psg_newton_solve(psg_ctx *c, double *zh, double *Tio, double *sxio, double *szio, const psg_newton_opts *o,
                     psg_newton_result *res) {
  if (c && c->batch != 1) { set_error("psg_newton_solve needs a batch-1 context (psg_newton_batch: several)"); return PSG_EINVAL; }
  NTask t = newton_core(c, zh, Tio, sxio, szio, nullptr, nullptr, o, res);
  return run_direct(c, t);
}

int psg_newton_solve_alc(psg_ctx *c, double *zh, double *Tio, double *sxio, double *v2io, const psg_alc *alc,
                         const psg_newton_opts *o, psg_newton_result *res) {
  if (!alc) { set_error("null argument"); return PSG_EINVAL; }
  if (o && o->nshift > 1) { set_error("newton: continuation takes the x shift only"); return PSG_EINVAL; }
  if (c && c->batch != 1) { set_error("psg_newton_solve_alc needs a batch-1 context"); return PSG_EINVAL; }
  NTask t = newton_core(c, zh, Tio, sxio, nullptr, v2io, alc, o, res);
  return run_direct(c, t);
}

int psg_newton_batch(psg_ctx *c, int nguess, double *z, double *T, double *sx, double *sz, const psg_newton_opts *o,
                     psg_newton_result *res) {
  if (!c || !z || !T || !o || !res || nguess < 1) { set_error("psg_newton_batch: bad argument"); return PSG_EINVAL; }
  if (c->g.dim != 2) { set_error("psg_newton_batch: 2D Navier-Stokes contexts only (per-member time steps)"); return PSG_EINVAL; }
  if (o->jv != 0) { set_error("psg_newton_batch: finite-difference Jv only (jv = 0)"); return PSG_EINVAL; }
  if (o->time_policy != PSG_TIME_UNIFORM) { set_error("psg_newton_batch: uniform steps only"); return PSG_EINVAL; }
  const int B = c->batch;
  const long n = (long)c->nf * c->g.nkt;
  const size_t bytes = sizeof(cpx<double>) * n;
  using H = std::coroutine_handle<NTask::promise_type>;
  std::vector<H> lane(B);
  int next = 0, rc = 0;
  // start guesses in lane m until one waits on a flow (or none are left)
  auto start = [&](int m) {
    while (next < nguess) {
      const int gi = next++;
      NTask t = newton_core(c, z + (size_t)gi * 2 * n, T + gi, sx ? sx + gi : nullptr, sz ? sz + gi : nullptr,
                            nullptr, nullptr, o, res + gi);
      H h = t.h;
      t.h = {};
      h.resume();
      if (!h.done()) { lane[m] = h; return; }
      if (h.promise().result && !rc) rc = h.promise().result;
      h.destroy();
    }
    lane[m] = H{};
  };
  for (int m = 0; m < B; m++) start(m);
  std::vector<double> dts(B, 0.0), dtm(B, 0.0);
  std::vector<int> nst(B, 0);
  int e = 0;
  for (;;) {
    // one batched flow for every lane that waits on one: member m steps with
    // dt_m = T_m / N_m for N_m = INT(T_m / dt0) steps, then stands still (dt 0)
    std::vector<int> bounds;
    bool any = false;
    for (int m = 0; m < B; m++) {
      nst[m] = 0;
      dtm[m] = 0.0;
      if (!lane[m]) continue;
      const FlowReq &rq = lane[m].promise().req;
      int ns = (int)(rq.T / rq.dt0);
      if (ns < 1) ns = 1;
      nst[m] = ns;
      dtm[m] = rq.T / (double)ns;
      bounds.push_back(ns);
      any = true;
      if (hipMemcpyAsync(c->state + (size_t)m * n, rq.in, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        set_error("psg_newton_batch: state copy");
        e = PSG_EHIP;
      }
    }
    if (!any || e) break;
    std::sort(bounds.begin(), bounds.end());
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
    int cur = 0;
    for (int b : bounds) {
      for (int m = 0; m < B; m++) dts[m] = nst[m] > cur ? dtm[m] : 0.0;
      if ((e = set_member_dt(c, dts.data())) || (e = psg_run(static_cast<psg_ctx *>(c), b - cur))) break;
      cur = b;
    }
    if (e) break;
    for (int m = 0; m < B; m++) {
      if (!lane[m]) continue;
      FlowReq &rq = lane[m].promise().req;
      rq.rc = hipMemcpyAsync(rq.out, c->state + (size_t)m * n, bytes, hipMemcpyDeviceToDevice, c->stream) == hipSuccess
                  ? 0 : PSG_EHIP;
    }
    for (int m = 0; m < B; m++) {
      if (!lane[m]) continue;
      lane[m].resume();
      if (lane[m].done()) {
        if (lane[m].promise().result && !rc) rc = lane[m].promise().result;
        lane[m].destroy();
        start(m);
      }
    }
  }
  for (int m = 0; m < B; m++)
    if (lane[m]) lane[m].destroy();
  const int e2 = set_member_dt(c, nullptr);
  if (e) return e;
  return rc ? rc : e2;
}

void psg_floquet_defaults(psg_floquet_opts *o) {
  memset(o, 0, sizeof(*o));
  o->nev = 10;            // main.F90:63-67
  o->ncv = 200;
  o->max_restarts = 20;
  o->tol = 1e-7;
  o->which = PSG_FLOQUET_LR;   // main.F90:388
  o->jv = 1;
  o->eps = 1e-7;
  o->time_policy = PSG_TIME_UNIFORM;
  o->dt0 = 0.005;
  o->seed = 1;
}

int psg_floquet(psg_ctx *c, const double *zh, double T, double sx, double sz, const psg_floquet_opts *o,
                double *mu_re, double *mu_im, double *err, double *vecs, psg_floquet_result *res) {
  typedef std::complex<double> Z;
  if (!c || !zh || !o || !mu_re || !mu_im || !err || !res) { set_error("null argument"); return PSG_EINVAL; }
  if (c->batch != 1) { set_error("psg_floquet needs a batch-1 context"); return PSG_EINVAL; }
  const int nev = o->nev, m = o->ncv;
  if (nev < 1 || m < nev + 2 || m > 400) { set_error("psg_floquet: need 1 <= nev and nev + 2 <= ncv <= 400"); return PSG_EINVAL; }
  if (o->jv < 0 || o->jv > 2) { set_error("psg_floquet: jv %d is not 0, 1 or 2", o->jv); return PSG_EINVAL; }
  if (o->jv >= 1 && c->cfg.flavour != PSG_NS2D && c->cfg.flavour != PSG_NS3D) {
    set_error("psg_floquet: tangent-linear Jv is 2D/3D Navier-Stokes only");
    return PSG_EINVAL;
  }
  TlPrecision tlp(c, o->jv == 2);
  const auto t_start = std::chrono::steady_clock::now();
  memset(res, 0, sizeof(*res));
  VecOps V(c);
  if (!V.dpart) { set_error("psg_floquet: reduction buffer"); return PSG_ENOMEM; }
  const long n = V.n;
  const size_t bytes = sizeof(cpx<double>) * n;
  cpx<double> *buf = nullptr;
  // Q (m + 1 basis vectors), the orbit, its image, and three work vectors
  if (hipMalloc(&buf, bytes * (size_t)(m + 6)) != hipSuccess) { set_error("psg_floquet: device allocation"); return PSG_ENOMEM; }
  auto q = [&](int i) { return buf + (size_t)i * n; };
  cpx<double> *Z0 = q(m + 1), *ZN = q(m + 2), *w = q(m + 3), *t1 = q(m + 4), *t2 = q(m + 5);
  int rc = 0;
  auto done = [&](int r) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return r;
  };
  if (hipMemcpyAsync(t1, zh, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) { set_error("copy"); return done(PSG_EHIP); }
  V.copy(t1, Z0);
  // with the tangent linear on a one-member 2D context the orbit's flow is
  // recorded once and every matvec integrates the perturbation alone
  const int rec_rp = (o->jv == 2 || c->cfg.precision == PSG_FP32BULK) ? 4 : 8;
  const bool rec = o->jv >= 1 && rec_usable(c, T, o->dt0, o->time_policy, rec_rp);
  if (o->jv == 0) {
    if ((rc = flow(c, Z0, ZN, T, o->dt0, o->time_policy))) return done(rc);
  } else if (rec) {
    if ((rc = rec_flow(c, Z0, ZN, T, o->dt0, o->time_policy))) return done(rc);
  }
  // A q = S_s dPhi_T(z) q
  auto apply_A = [&](const cpx<double> *in, cpx<double> *out) -> int {
    int r;
    if (o->jv >= 1) {
      r = rec ? tl_replay(c, in, out, T, o->dt0, o->time_policy)
              : tl_flow_map(c, Z0, in, t2, out, T, o->dt0, o->time_policy, true);
      if (r) return r;
    } else {
      V.copy(Z0, t2);
      V.axpby(o->eps, in, 1.0, t2);
      if ((r = flow(c, t2, out, T, o->dt0, o->time_policy))) return r;
      V.axpby(-1.0, ZN, 1.0, out);
      V.axpby(0.0, out, 1.0 / o->eps, out);
    }
    V.shift(out, out, sx, o->sy, sz);
    res->matvecs++;
    return 0;
  };
  // seeded start vector on the active modes
  {
    std::vector<cpx<double>> h0(n);
    unsigned long long st = o->seed * 0x9E3779B97F4A7C15ull + 12345u;
    auto rnd = [&]() {
      st ^= st << 13; st ^= st >> 7; st ^= st << 17;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    };
    for (long i = 0; i < n; i++) h0[i] = cpx<double>{rnd(), rnd()};
    if (hipMemcpyAsync(t1, h0.data(), bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) { set_error("copy"); return done(PSG_EHIP); }
  }
  const double n0 = sqrt(V.dot(t1, t1));
  V.axpby(1.0 / n0, t1, 0.0, q(0));
  std::vector<double> Hm((size_t)(m + 1) * m);
  auto h = [&](int i, int j) -> double & { return Hm[(size_t)i * m + j]; };
  std::vector<int> order;
  std::vector<Z> mus;
  std::vector<std::vector<Z>> ys;
  std::vector<double> errs;
  int nconv = 0, kdim = m;
  for (int cyc = 0; cyc <= o->max_restarts; cyc++) {
    res->cycles = cyc + 1;
    std::fill(Hm.begin(), Hm.end(), 0.0);
    kdim = m;
    for (int j = 0; j < m; j++) {
      if ((rc = apply_A(q(j), w))) return done(rc);
      const double wn0 = sqrt(V.dot(w, w));
      if (!(wn0 == wn0)) { set_error("psg_floquet: the Jacobian action is not finite"); return done(PSG_EINVAL); }
      // modified Gram-Schmidt, twice
      for (int pass = 0; pass < 2; pass++)
        for (int i = 0; i <= j; i++) {
          const double hij = V.dot(q(i), w);
          h(i, j) += hij;
          V.axpby(-hij, q(i), 1.0, w);
        }
      const double hn = sqrt(V.dot(w, w));
      h(j + 1, j) = hn;
      // invariant subspace: what is left after orthogonalisation is rounding
      // (operators with repeated eigenvalues, e.g. symmetry pairs, reach it
      // before the space is exhausted)
      if (!(hn > 1e-12 * wn0)) { h(j + 1, j) = 0.0; kdim = j + 1; break; }
      V.axpby(1.0 / hn, w, 0.0, q(j + 1));
    }
    // Ritz values of the kdim x kdim Hessenberg matrix
    std::vector<double> hh((size_t)kdim * kdim), wr(kdim), wi(kdim);
    for (int i = 0; i < kdim; i++)
      for (int j = 0; j < kdim; j++) hh[(size_t)i * kdim + j] = h(i, j);
    std::vector<double> hq = hh;
    if (hess_eig(kdim, hq.data(), wr.data(), wi.data())) { set_error("psg_floquet: Hessenberg QR failed"); return done(-1); }
    order.resize(kdim);
    for (int i = 0; i < kdim; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const double ka = o->which == PSG_FLOQUET_LM ? std::hypot(wr[a], wi[a]) : wr[a];
      const double kb = o->which == PSG_FLOQUET_LM ? std::hypot(wr[b], wi[b]) : wr[b];
      if (ka != kb) return ka > kb;
      return wi[a] > wi[b];   // conjugate pairs: positive imaginary part first
    });
    const int nw = std::min(nev, kdim);
    mus.assign(nw, Z(0, 0));
    ys.assign(nw, {});
    errs.assign(nw, 0.0);
    nconv = 0;
    const double beta = kdim < m || h(m, m - 1) == 0.0 ? 0.0 : h(m, m - 1);
    for (int i = 0; i < nw; i++) {
      mus[i] = Z(wr[order[i]], wi[order[i]]);
      hess_evec(kdim, hh.data(), mus[i], ys[i]);
      errs[i] = (kdim < m ? 0.0 : beta) * std::abs(ys[i][kdim - 1]);
      if (errs[i] <= o->tol * std::max(std::abs(mus[i]), 1e-300)) nconv++;
    }
    if (o->verbose) {
      fprintf(stderr, "floquet cycle %d: %d/%d converged, leading mu = %.10g%+.10gi (err %.2e)\n", cyc, nconv, nw,
              mus[0].real(), mus[0].imag(), errs[0]);
      fflush(stderr);
    }
    if (nconv >= nw || cyc == o->max_restarts) break;
    // explicit restart from the sum of the wanted Ritz vectors' real and imaginary parts
    V.axpby(0.0, w, 0.0, w);
    for (int i = 0; i < nw; i++)
      for (int j = 0; j < kdim; j++) V.axpby(ys[i][j].real() + ys[i][j].imag(), q(j), 1.0, w);
    const double wn = sqrt(V.dot(w, w));
    V.axpby(1.0 / wn, w, 0.0, q(0));
  }
  const int nw = (int)mus.size();
  for (int i = 0; i < nev; i++) {
    mu_re[i] = i < nw ? mus[i].real() : 0.0;
    mu_im[i] = i < nw ? mus[i].imag() : 0.0;
    err[i] = i < nw ? errs[i] : -1.0;
  }
  res->nconv = nconv;
  if (vecs) {
    // Ritz vectors; a conjugate pair gives its real then its imaginary part
    for (int i = 0; i < nw; i++) {
      const bool second = i > 0 && mus[i].imag() < 0.0 && mus[i - 1].imag() > 0.0 &&
                          mus[i] == std::conj(mus[i - 1]);
      const std::vector<Z> &y = second ? ys[i - 1] : ys[i];
      V.axpby(0.0, w, 0.0, w);
      for (int j = 0; j < kdim; j++) V.axpby(second ? y[j].imag() : y[j].real(), q(j), 1.0, w);
      if (hipMemcpyAsync(vecs + (size_t)i * 2 * n, w, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
        set_error("copy");
        return done(PSG_EHIP);
      }
    }
  }
  return done(0);
}

}  // extern "C"
