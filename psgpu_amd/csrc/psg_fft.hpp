// Device FFT building blocks for gfx950 (CDNA4).
//
// A pencil of N points (N a power of two) is transformed by a team of
// TP = N/E threads, each holding E points in registers: thread j holds the
// points j + TP*m, m < E, both before and after the transform (self-sorting
// Stockham formulation: every stage reads with stride N/R and writes to
// (b/Ns)*Ns*R + b%Ns + r*Ns).  Stages are radix E until fewer than E points
// per sub-transform remain, then one smaller radix; intermediate stages
// exchange through LDS, the first reads registers and the last writes
// registers, so a 256-point FP64 pencil with E=8 costs two LDS round trips.
//
// LDS addressing is lds[pos*stride + off]: column tiles use stride = tile
// width and off = column (consecutive lanes hit consecutive columns), row
// teams use stride = 1 and off = team base, with positions permuted by an XOR
// swizzle (lpos) that keeps both the contiguous reads and the strided
// Stockham writes of a wavefront on distinct LDS banks.
//
// Twiddles come from a per-length table W_N^k = e^{-2 pi i k/N} built on the
// host in FP64 (glibc sincos) and rounded once for the FP32 table.  Kernels
// stage them into LDS per stage and thread (tw_base): the R-1 twiddles a
// butterfly needs are adjacent, so lanes with different k read different
// banks instead of the bank-aligned strides r*k*N/(NS*R) of the natural
// table.
#pragma once
#include <hip/hip_runtime.h>

namespace psg {

// Aligned to its full size, so that one LDS or global access moves a whole
// point (ds_read_b128 / global_load_dwordx4 in FP64) instead of two halves.
template <typename T> struct alignas(2 * sizeof(T)) cpx { T x, y; };

template <typename T> __device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <typename T> __device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by conj(b)
template <typename T> __device__ __forceinline__ cpx<T> mulc(cpx<T> a, cpx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __device__ __forceinline__ cpx<T> rot90(cpx<T> a) {
  return INV ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

template <typename T> struct K;
template <> struct K<double> { static constexpr double r2 = 0.70710678118654752440084436210485; };
template <> struct K<float> { static constexpr float r2 = 0.70710678118654752440084436210485f; };

template <typename T, bool INV> __device__ __forceinline__ void dft2(cpx<T> &a, cpx<T> &b) {
  cpx<T> t = a; a = add(t, b); b = sub(t, b);
}

template <typename T, bool INV> __device__ __forceinline__ void dft4(cpx<T> *v, int s) {
  // v[0], v[s], v[2s], v[3s]
  cpx<T> a0 = add(v[0], v[2 * s]), a1 = sub(v[0], v[2 * s]);
  cpx<T> a2 = add(v[s], v[3 * s]), a3 = sub(v[s], v[3 * s]);
  a3 = rot90<T, INV>(a3);
  v[0] = add(a0, a2); v[2 * s] = sub(a0, a2);
  v[s] = add(a1, a3); v[3 * s] = sub(a1, a3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft8(cpx<T> *v) {
  // even/odd split: E = dft4(v0,v2,v4,v6), O = dft4(v1,v3,v5,v7)
  dft4<T, INV>(v, 2);
  dft4<T, INV>(v + 1, 2);
  const T c = K<T>::r2;
  // twiddle odd part by W8^k (forward e^{-i pi k/4})
  cpx<T> o1 = v[3], o2 = v[5], o3 = v[7];
  o1 = INV ? cpx<T>{c * (o1.x - o1.y), c * (o1.x + o1.y)} : cpx<T>{c * (o1.x + o1.y), c * (o1.y - o1.x)};
  o2 = rot90<T, INV>(o2);
  o3 = INV ? cpx<T>{-c * (o3.x + o3.y), c * (o3.x - o3.y)} : cpx<T>{c * (o3.y - o3.x), -c * (o3.x + o3.y)};
  cpx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = add(e0, o0); v[4] = sub(e0, o0);
  v[1] = add(e1, o1); v[5] = sub(e1, o1);
  v[2] = add(e2, o2); v[6] = sub(e2, o2);
  v[3] = add(e3, o3); v[7] = sub(e3, o3);
}

// 16-point DFT as 4 x 4: n = n1 + 4 n2, k = 4 k1 + k2,
// X[4k1+k2] = sum_n1 W4^{n1 k1} W16^{n1 k2} sum_n2 W4^{n2 k2} v[n1+4n2]
template <typename T> struct K16 {
  static constexpr T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  static constexpr T r2 = (T)0.70710678118654752440;
};
template <typename T, bool INV> __device__ __forceinline__ void dft16(cpx<T> *v) {
#pragma unroll
  for (int n1 = 0; n1 < 4; n1++) dft4<T, INV>(v + n1, 4);   // v[n1+4k2] = Y[n1][k2]
  // W16^m = (cos 2 pi m/16, -sin 2 pi m/16) for m = n1*k2; inverse uses the conjugate
  const T c1 = K16<T>::c1, s1 = K16<T>::s1, r2 = K16<T>::r2;
  const cpx<T> w1{c1, -s1}, w2{r2, -r2}, w3{s1, -c1}, w6{-r2, -r2}, w9{-c1, s1};
  auto tw = [](cpx<T> a, cpx<T> w) { return INV ? mulc(a, w) : mul(a, w); };
  v[5] = tw(v[5], w1);    // n1=1,k2=1
  v[9] = tw(v[9], w2);    // n1=1,k2=2
  v[13] = tw(v[13], w3);  // n1=1,k2=3
  v[6] = tw(v[6], w2);    // n1=2,k2=1
  v[10] = rot90<T, INV>(v[10]);  // n1=2,k2=2: W16^4 = -i (forward)
  v[14] = tw(v[14], w6);  // n1=2,k2=3
  v[7] = tw(v[7], w3);    // n1=3,k2=1
  v[11] = tw(v[11], w6);  // n1=3,k2=2
  v[15] = tw(v[15], w9);  // n1=3,k2=3
#pragma unroll
  for (int k2 = 0; k2 < 4; k2++) dft4<T, INV>(v + 4 * k2, 1);  // v[4k2+k1] = X[4k1+k2]
  // 4x4 transpose to natural order
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = a + 1; b < 4; b++) {
      cpx<T> t = v[4 * a + b];
      v[4 * a + b] = v[4 * b + a];
      v[4 * b + a] = t;
    }
}

template <typename T, bool INV, int R> __device__ __forceinline__ void dftR(cpx<T> *v) {
  if constexpr (R == 2) dft2<T, INV>(v[0], v[1]);
  else if constexpr (R == 4) dft4<T, INV>(v, 1);
  else if constexpr (R == 8) dft8<T, INV>(v);
  else if constexpr (R == 16) dft16<T, INV>(v);
  else static_assert(R == 1, "unsupported radix");
}

// One Stockham stage.  NS = product of the previous radices.
// Synchronisation/layout policy of a pencil team:
//   WAVE: the team lives inside one wavefront, so LDS exchanges only need a
//         wavefront-scope fence (no block barrier);
//   SW  : row-team swizzle of LDS positions (lpos); 0 for column tiles,
//         whose lanes already address consecutive columns.
template <bool WAVE> __device__ __forceinline__ void team_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}
// Row-team LDS position of point p under swizzle SW (0: none): bit T+i of p
// is flipped by bit S+i, i < 4, SW = S | T << 4.  With S > T every flipped
// bit is below the bit that flips it, so this permutes 0..N-1 for any
// power-of-two N (the inverse restores bits from the top down).
template <int SW> __device__ __forceinline__ int lpos(int p) {
  if constexpr (SW == 0) return p;
  else return p ^ (((p >> (SW & 15)) & 15) << (SW >> 4));
}
// Swizzle of a row team by points per thread.  gfx950 serves an LDS read of
// a wavefront in groups of 16 (128-bit) or 32 (64-bit) lanes over 64 banks
// and a write in groups of 8 or 16 lanes over 32 banks; these swizzles keep
// the reads, the radix-E Stockham writes and the output exchange of both
// precisions within one pass per group.
template <int E> constexpr int row_swizzle() { return E == 4 ? 2 : (E == 16 ? 4 : 3); }

// Stage-major packed twiddles: stage NS (radix R, not the first) stores
// W_N^{r k N/(NS R)}, r = 1..R-1, at tw_base + k (R-1) + r - 1.  Fewer than N
// entries in total.
constexpr int radix_at(int N, int E, int NS) { return (N / NS >= E) ? E : N / NS; }
constexpr int tw_base(int N, int E, int NS) {
  int base = 0;
  for (int ns = radix_at(N, E, 1); ns < NS; ns *= radix_at(N, E, ns)) base += ns * (radix_at(N, E, ns) - 1);
  return base;
}
constexpr int tw_count(int N, int E) { return tw_base(N, E, N); }

template <typename T, int N, int E, int NS, bool INV, bool WAVE = false, int SW = 0>
__device__ __forceinline__ void fft_stage(cpx<T> (&v)[E], cpx<T> *lds, int j, int stride, int off,
                                          const cpx<T> *__restrict__ tw) {
  constexpr int TP = N / E;
  constexpr int R = (N / NS >= E) ? E : N / NS;
  constexpr int G = E / R;            // butterflies per thread
  constexpr int NB = N / R;
  constexpr bool FIRST = (NS == 1);
  constexpr bool LAST = (NS * R == N);
  if constexpr (!FIRST) {
    team_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = lds[lpos<SW>(j + TP * m) * stride + off];
  }
#pragma unroll
  for (int t = 0; t < G; t++) {
    cpx<T> u[R];
#pragma unroll
    for (int r = 0; r < R; r++) u[r] = v[t + r * G];
    const int b = j + t * TP;
    const int k = b % NS;
    if constexpr (!FIRST) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        cpx<T> w = tw[tw_base(N, E, NS) + k * (R - 1) + r - 1];
        u[r] = INV ? mulc(u[r], w) : mul(u[r], w);
      }
    }
    dftR<T, INV, R>(u);
#pragma unroll
    for (int r = 0; r < R; r++) v[t + r * G] = u[r];
  }
  (void)NB;
  if constexpr (!LAST) {
    team_sync<WAVE>();
#pragma unroll
    for (int t = 0; t < G; t++) {
      const int b = j + t * TP;
      const int k = b % NS;
      const int d = (b / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) lds[lpos<SW>(d + r * NS) * stride + off] = v[t + r * G];
    }
  }
}

template <typename T, int N, int E, int NS, bool INV, bool WAVE, int SW>
__device__ __forceinline__ void fft_stages(cpx<T> (&v)[E], cpx<T> *lds, int j, int stride, int off,
                                           const cpx<T> *__restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int R = (N / NS >= E) ? E : N / NS;
    fft_stage<T, N, E, NS, INV, WAVE, SW>(v, lds, j, stride, off, tw);
    fft_stages<T, N, E, NS * R, INV, WAVE, SW>(v, lds, j, stride, off, tw);
  }
}

// Full N-point transform of the pencil held by a team (see header comment).
// Every thread of the block must call this (it contains block barriers).
template <typename T, int N, int E, bool INV, bool WAVE = false, int SW = 0>
__device__ __forceinline__ void fft_pencil(cpx<T> (&v)[E], cpx<T> *lds, int j, int stride, int off,
                                           const cpx<T> *__restrict__ tw) {
  static_assert((N & (N - 1)) == 0 && N >= E, "power-of-two pencil");
  fft_stages<T, N, E, 1, INV, WAVE, SW>(v, lds, j, stride, off, tw);
}

// LDS complex slots per row-team row of N points
template <int N> constexpr int row_len() { return N; }

}  // namespace psg
