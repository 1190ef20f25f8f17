// The device pencil FFT (psg_fft.hpp: Stockham stages, packed LDS twiddles,
// row-team swizzles) run on the host through the emulation header, for every
// supported length 16..1024, 4/8/16 points per thread, both precisions, both
// directions and every swizzle the kernels use, against a direct DFT in long
// double.  Also checks that each swizzle permutes 0..N-1.  Exit status 0 =
// all passed.  TEST INFRASTRUCTURE ONLY.
#include <math.h>
#include <stdio.h>
#include <complex>
#include <vector>
#include "psg_passes.hpp"

using namespace psg;
static int g_fail = 0, g_pass = 0;

template <typename P, int N, int E, int SW, bool INV>
__global__ void k_fft_test(const cpx<P> *in, cpx<P> *out, const cpx<P> *tw) {
  PSG_LDS_TWIDDLES(P, N, E);
  constexpr int TP = N / E;
  const int team = threadIdx.x / TP, j = threadIdx.x % TP;
  cpx<P> *tl = lds + team * row_len<N>();
  cpx<P> v[E];
  for (int m = 0; m < E; m++) v[m] = in[team * N + j + TP * m];
  fft_pencil<P, N, E, INV, (TP <= 64), SW>(v, tl, j, 1, 0, tw);
  for (int m = 0; m < E; m++) out[team * N + j + TP * m] = v[m];
}

template <int SW, int N> static bool bijective() {
  std::vector<int> seen(N, 0);
  for (int p = 0; p < N; p++) {
    const int q = lpos<SW>(p);
    if (q < 0 || q >= N || seen[q]++) return false;
  }
  return true;
}

template <typename P, int N, int E, int SW, bool INV> static void one() {
  constexpr int TEAMS = 2;
  std::vector<cpx<P>> tw(N), in(TEAMS * N), out(TEAMS * N);
  for (int k = 0; k < N; k++) {
    double s, c;
    sincos(2.0 * M_PI * (double)k / (double)N, &s, &c);
    tw[k] = cpx<P>{(P)c, (P)-s};
  }
  unsigned x = 12345u + N * 7u + E;
  for (auto &z : in) {
    x = x * 1664525u + 1013904223u;
    z.x = (P)((x >> 8) * (1.0 / 16777216.0) - 0.5);
    x = x * 1664525u + 1013904223u;
    z.y = (P)((x >> 8) * (1.0 / 16777216.0) - 0.5);
  }
  const size_t lds = (size_t)(row_len<N>() * TEAMS + N) * sizeof(cpx<P>);
  hipLaunchKernelGGL((k_fft_test<P, N, E, SW, INV>), dim3(1), dim3(TEAMS * (N / E)), lds, nullptr, in.data(),
                     out.data(), tw.data());
  double err = 0, mag = 0;
  const long double sg = INV ? 1.0L : -1.0L;
  for (int t = 0; t < TEAMS; t++)
    for (int k = 0; k < N; k++) {
      std::complex<long double> acc = 0;
      for (int n = 0; n < N; n++) {
        const long double a = sg * 2.0L * M_PIl * (long double)((long)n * k % N) / N;
        acc += std::complex<long double>(in[t * N + n].x, in[t * N + n].y) * std::polar(1.0L, a);
      }
      err = fmax(err, (double)std::abs(acc - std::complex<long double>(out[t * N + k].x, out[t * N + k].y)));
      mag = fmax(mag, (double)std::abs(acc));
    }
  const double rel = err / mag, tol = sizeof(P) == 8 ? 1e-14 : 2e-6;
  const bool ok = rel < tol && bijective<SW, N>();
  printf("%s fft %s N=%-5d E=%-2d SW=%-3d %s %.3e\n", ok ? "PASS" : "FAIL", sizeof(P) == 8 ? "fp64" : "fp32", N, E, SW,
         INV ? "inv" : "fwd", rel);
  fflush(stdout);
  (ok ? g_pass : g_fail)++;
}

template <typename P, int N, int E> static void sizes() {
  one<P, N, E, 0, false>();
  one<P, N, E, 0, true>();
  one<P, N, E, row_swizzle<E>(), false>();
  one<P, N, E, row_swizzle<E>(), true>();
}
template <typename P, int E> static void lengths() {
  if constexpr (E <= 16) sizes<P, 16, E>();
  sizes<P, 32, E>();
  sizes<P, 64, E>();
  sizes<P, 128, E>();
  sizes<P, 256, E>();
  sizes<P, 512, E>();
  sizes<P, 1024, E>();
}

int main() {
  lengths<double, 4>();
  lengths<double, 8>();
  lengths<double, 16>();
  lengths<float, 4>();
  lengths<float, 8>();
  lengths<float, 16>();
  printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail ? 1 : 0;
}
