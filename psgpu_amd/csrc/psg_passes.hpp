// Pencil-pass kernels of the pseudospectral right-hand side on gfx950.
//
// One right-hand side (3D) is five HBM passes over pruned pencils:
//   1 zinv : truncated spectral state -> Hermitian fill of the kx=0 plane,
//            Biot-Savart velocity, zero-pad in z, inverse z-FFT
//            state [m][f][nkt] (FP64)          -> A [m][F][NZ][IKTY][XP]
//   2 yinv : zero-pad in y, inverse y-FFT        A -> B [m][F][NZ][NY][XP]
//   3 xpass: per (y,z) row: C2R of the fields (two real fields per complex
//            FFT), pointwise product in registers, R2C, truncation to IKTX;
//            physical space never reaches HBM    B -> C [m][FO][NZ][NY][XP]
//   4 yfwd : forward y-FFT, truncation in y      C -> A' [m][FO][NZ][IKTY][XP]
//   5 zfwd : forward z-FFT, truncation in z, x 1/NR (FP64), then the scheme's
//            update functor at every stored mode (curl, forcing, mask, CN).
// 2D uses passes 2-5 shape-wise: y-inverse straight from the state, the same
// x-pass, and a y-forward pass that ends in the update functor.
//
// Column passes use a tile of T consecutive kx columns per block (coalesced
// along the contiguous kx axis) and N/8 threads per column; the x-pass uses
// one team of N/8 threads per row.  Every pass is templated on the storage
// precision P of the pass buffers (double: plain form, float: FP32-bulk);
// state, tables, update arithmetic and diagnostics stay FP64.
#pragma once
#include <type_traits>

#include "psg_internal.hpp"

namespace psg {

constexpr int EPT = 8;  // points per thread in every pencil transform

// Every transform kernel keeps the twiddles of its pencil length at the start
// of its dynamic LDS, repacked from the natural table W_N^k in HBM into the
// per-stage order the Stockham stages read (tw_base in psg_fft.hpp), so
// they are read at LDS latency and without bank conflicts.  The packed table
// has fewer than N entries; the kernel's data region starts N entries in
// (launches add N*sizeof(cpx<P>) to the LDS size).
// packed entry i of the stages from NS on, whose entries start at BASE:
// stage NS (radix R) holds w^(r k N/(NS R)) at BASE + k (R-1) + r-1
template <typename P, int N, int E, int NS, int BASE>
__device__ __forceinline__ void stage_twiddle(const cpx<P> *__restrict__ tw, cpx<P> *lds_tw, int i) {
  if constexpr (NS < N) {
    constexpr int R = radix_at(N, E, NS), CNT = NS * (R - 1);
    if (i < BASE + CNT) {
      const int q = i - BASE, k = q / (R - 1), r = q % (R - 1) + 1;
      lds_tw[i] = tw[(r * k) * (N / (NS * R))];
    } else {
      stage_twiddle<P, N, E, NS * R, BASE + CNT>(tw, lds_tw, i);
    }
  }
}
template <typename P, int N, int E>
__device__ __forceinline__ const cpx<P> *stage_twiddles(const cpx<P> *__restrict__ tw, cpx<P> *lds_tw) {
  constexpr int M = tw_count(N, E);
  static_assert(M < N, "packed twiddles fit the reserved N entries");
  for (int i = threadIdx.x; i < M; i += blockDim.x) stage_twiddle<P, N, E, radix_at(N, E, 1), 0>(tw, lds_tw, i);
  __syncthreads();
  return lds_tw;
}
// XCD-aware tile order.  The dispatcher hands workgroups to the 8 XCDs round
// robin (workgroup b runs on XCD b % 8, x fastest, then y, then z), and each
// XCD has its own L2.  Column tiles of neighbouring kx ranges share cache lines
// of the NKT-ordered state and tables (and of FP32 pass buffers), so the
// logical tile order gives XCD x one contiguous range of tiles: shared lines
// are then fetched once per XCD range boundary instead of once per tile pair.
struct TileId { int x, y, z; };
__device__ __forceinline__ TileId xcd_tile() {
  constexpr unsigned X = 8;
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned per = nb / X, tail = nb - per * X;
  const unsigned x = b % X, q = b / X;
  const unsigned l = x < tail ? x * (per + 1) + q : tail * (per + 1) + (x - tail) * per + q;
  return TileId{(int)(l % gx), (int)((l / gx) % gy), (int)(l / (gx * gy))};
}

// dynamic LDS of a kernel (a CPU build for bounds checking supplies its own)
// occupancy request of a kernel (a CPU build for bounds checking drops it)
#ifndef PSG_WAVES_PER_EU
#define PSG_WAVES_PER_EU(w) __attribute__((amdgpu_waves_per_eu(w)))
#endif
#ifndef PSG_DYN_SMEM
#define PSG_DYN_SMEM(name) extern __shared__ __attribute__((aligned(16))) char name[]
#endif
#define PSG_LDS_TWIDDLES(P, N, E)                                       \
  PSG_DYN_SMEM(smem_raw);                                              \
  cpx<P> *lds = reinterpret_cast<cpx<P> *>(smem_raw) + (N);            \
  tw = stage_twiddles<P, N, E>(tw, reinterpret_cast<cpx<P> *>(smem_raw))

template <typename P> __device__ __forceinline__ cpx<P> to(cpx<double> a) { return {(P)a.x, (P)a.y}; }
template <typename P> __device__ __forceinline__ cpx<P> to(cpx<float> a) { return {(P)a.x, (P)a.y}; }
__device__ __forceinline__ cpx<double> up(cpx<double> a) { return a; }
__device__ __forceinline__ cpx<double> up(cpx<float> a) { return {(double)a.x, (double)a.y}; }

// Pass-buffer element access.  Every pass-buffer element is written once by
// one pass and read once by the next, so these accesses may bypass the caches
// as non-temporal loads and stores.  Which passes gain was measured on MI355X
// (256^3, per-kernel HIP event times): in FP64 streaming the stores of the
// z-inverse, y-inverse and x-pass took 8% off a step and streaming the loads
// of the y-inverse and z-forward a little more, while the y-forward and the
// x-pass's loads ran slower streamed; in FP32 the z-inverse's stores did.
enum { PASS_ZINV, PASS_YINV, PASS_XPASS, PASS_YFWD, PASS_ZFWD };
template <typename P, int PASS> struct NT {
  static constexpr bool f64 = sizeof(P) == 8;
  static constexpr bool st = PASS == PASS_YINV || PASS == PASS_XPASS || (PASS == PASS_ZINV && f64) ||
                             (PASS == PASS_YFWD && !f64);
  static constexpr bool ld = PASS == PASS_YINV || PASS == PASS_ZFWD;
};
template <bool STREAM, typename P> __device__ __forceinline__ void st_pass(cpx<P> *p, cpx<P> v) {
#if defined(__clang__)
  if constexpr (STREAM) {
    typedef P vec2 __attribute__((ext_vector_type(2)));
    vec2 t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<vec2 *>(p));
    return;
  }
#endif
  *p = v;
}
template <bool STREAM, typename P> __device__ __forceinline__ cpx<P> ld_pass(const cpx<P> *p) {
#if defined(__clang__)
  if constexpr (STREAM) {
    typedef P vec2 __attribute__((ext_vector_type(2)));
    const vec2 t = __builtin_nontemporal_load(reinterpret_cast<const vec2 *>(p));
    return cpx<P>{t.x, t.y};
  }
#endif
  return *p;
}

// kz (or ky) of padded position p, or INT_MIN when p is in the zero band
__device__ __forceinline__ int wrapk(int p, int n, int kt) {
  return p <= kt ? p : (p >= n - kt ? p - n : -0x40000000);
}
#define PSG_NOK (-0x40000000)

// Pass-buffer layouts, per field.  kx columns are grouped in tiles of LW<P>
// (xp/LW tiles).  A column pass's workgroup owns one tile of one plane, so
// the y passes read and write whole tiles contiguously (21.9 KB of ky and
// 32 KB of y at 256^3 FP64) instead of 128-byte pieces of rows; the z passes
// see 128-byte pieces either way, ordered so that the tiles of neighbouring
// workgroups (consecutive kx tiles of a plane) are adjacent.
//   Z, Z' (z-inverse -> y-inverse, y-forward -> z-forward):  [z][tile][iky][LW]
//   Y     (y-inverse -> x-pass):                            [z][tile][y][LW]
//   C     (x-pass -> y-forward): row-major [z][y][xp] (the x-pass writes rows)
// The tile width is LW<P>: 8 points (128 B) in FP64; in FP32 PSG_LW32 points
// (8: 64 B, 16: a whole 128-byte line).  XP is a multiple of 16, so either
// tiles the padded kx range.
#ifndef PSG_LW32
#define PSG_LW32 8
#endif
template <typename P> struct Lay {
  static constexpr int W = sizeof(P) == 4 ? PSG_LW32 : 8;
  static constexpr int SH = W == 16 ? 4 : 3;
  static_assert(W == 8 || W == 16, "tile width");
};
template <typename P> __device__ __forceinline__ long zoff(const Grid &g, int ikx, int iky, int z) {
  constexpr int W = Lay<P>::W, SH = Lay<P>::SH;
  return ((long)z * (g.xp / W) + (ikx >> SH)) * ((long)g.ikty * W) + iky * W + (ikx & (W - 1));
}
template <typename P> __device__ __forceinline__ long yoff(const Grid &g, int ikx, int y, int z) {
  constexpr int W = Lay<P>::W, SH = Lay<P>::SH;
  return ((long)z * (g.xp / W) + (ikx >> SH)) * ((long)g.ny * W) + y * W + (ikx & (W - 1));
}
// x row rr = z*ny + y of a Y buffer: element k sits at yrow(rr)[ycol(k)]
template <typename P> __device__ __forceinline__ long yrow(const Grid &g, long rr) {
  return yoff<P>(g, 0, (int)(rr % g.ny), (int)(rr / g.ny));
}
template <typename P> __device__ __forceinline__ int ycol(const Grid &g, int k) {
  constexpr int W = Lay<P>::W, SH = Lay<P>::SH;
  return (k >> SH) * (g.ny * W) + (k & (W - 1));
}

// Biot-Savart u = i k x w / max(|k|^2,1e-3): 3D_NavierStokes/src/kernels.h:187-201
__device__ __forceinline__ void biot_savart(double kx, double ky, double kz, cpx<double> xi, cpx<double> et,
                                            cpx<double> zt, double wki, cpx<double> &u, cpx<double> &v,
                                            cpx<double> &w) {
  u.x = -(ky * zt.y - kz * et.y) * wki;
  u.y = -(kz * et.x - ky * zt.x) * wki;
  v.x = -(kz * xi.y - kx * zt.y) * wki;
  v.y = -(kx * zt.x - kz * xi.x) * wki;
  w.x = -(kx * et.y - ky * xi.y) * wki;
  w.y = -(ky * xi.x - kx * et.x) * wki;
}

// ---------------------------------------------------------------- pass 1 (3D)
// out fields: 0..2 vorticity, 3..5 velocity (if BS).  umask: velocity masked
// by L as setVelocity does before the first step (kernels.h:191).
// Each output field re-reads the state components it needs (one for a
// vorticity component, two for a velocity component); the block's state tile
// stays cache-resident, so HBM sees each state value once while the kernel
// keeps only one pencil's worth of registers live.
// Storage layouts of a 3D state (per member; field f of mode (ikx, iky, ikz)
// at at() + f fs(), and per-mode tables at tab()):
//  LayNKT: the reference order, field-major, ik = (ikz IKTY + iky) IKTX + ikx
//  LayW:   the stepper's working copy: runs along z of 8-column tiles with the
//          three fields of a mode adjacent, [tile][iky][ikz][field][8].  A
//          z-pass block (one tile, one iky) then reads and writes one
//          contiguous run of IKTZ x 3 x 8 points instead of three 8-point
//          pieces per kz spread over the three fields.
struct LayNKT {
  static __device__ __forceinline__ long at(const Grid &g, int ikx, int iky, int ikz) {
    return ((long)ikz * g.ikty + iky) * g.iktx + ikx;
  }
  static __device__ __forceinline__ long tab(const Grid &g, int ikx, int iky, int ikz) { return at(g, ikx, iky, ikz); }
  static __device__ __forceinline__ long fs(const Grid &g) { return g.nkt; }
};
struct LayW {
  static __device__ __forceinline__ long run(const Grid &g, int ikx, int iky, int ikz) {
    return ((long)(ikx >> 3) * g.ikty + iky) * g.iktz + ikz;
  }
  static __device__ __forceinline__ long at(const Grid &g, int ikx, int iky, int ikz) {
    return run(g, ikx, iky, ikz) * 24 + (ikx & 7);
  }
  static __device__ __forceinline__ long tab(const Grid &g, int ikx, int iky, int ikz) {
    return run(g, ikx, iky, ikz) * 8 + (ikx & 7);
  }
  static __device__ __forceinline__ long fs(const Grid &) { return 8; }
};

struct FillIdx {
  long idx;     // field-0 index to read (mirror for the filled half-plane), -1 = zero
  long tidx;    // the same mode's index in the per-mode tables
  bool conj;    // conjugate on read
};
template <class LAY = LayNKT>
__device__ __forceinline__ FillIdx fill_index(const Grid &g, int ikx, int iky, int ky, int kz) {
  FillIdx r{-1, -1, false};
  if (kz == PSG_NOK) return r;
  if (ikx == 0 && ky == 0 && kz == 0) return r;
  const bool mirror = (ikx == 0) && (ky < 0 || (ky == 0 && kz < 0));
  const int sx = mirror ? 0 : ikx, sy = mirror ? -ky + g.kty : iky, sz = mirror ? -kz + g.ktz : kz + g.ktz;
  r.idx = LAY::at(g, sx, sy, sz);
  r.tidx = LAY::tab(g, sx, sy, sz);
  r.conj = mirror;
  return r;
}
template <typename S>
__device__ __forceinline__ cpx<S> ld_fill(const cpx<S> *__restrict__ f, const FillIdx &q) {
  if (q.idx < 0) return cpx<S>{0, 0};
  cpx<S> a = f[q.idx];
  if (q.conj) a.y = -a.y;
  return a;
}

// S: the state's storage type (double; float for the FP32 working state of
// an FP32-bulk run)
// LAY: the state's layout (st_fs its field stride); mask is in LAY's table order
// PIPE: ky rows per block, one after the other; the next row's state values
// are requested before this row's transforms, so that their loads are in
// flight while the row computes (the grid's y extent is IKTY/PIPE rounded up)
template <typename P, int N, int T, bool BS, typename S = double, class LAY = LayNKT, int PIPE = 1>
__global__ void __launch_bounds__(T *(N / EPT))
k_zinv(const cpx<S> *__restrict__ st, long st_ms, long st_fs, cpx<P> *__restrict__ out, long out_ms, Grid g,
       const cpx<P> *__restrict__ tw, const int8_t *__restrict__ mask, int umask) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, iky0 = tid.y * PIPE, m = tid.z;
  const bool valid = ikx < g.iktx;
  const double kx = (double)ikx * g.alpha;
  const cpx<S> *s0 = st + (long)m * st_ms;
  // every state value of the pencil is requested before the first transform;
  // it is kept in the pass precision (in FP32-bulk the Biot-Savart product then
  // starts from the rounded vorticity, and the registers halve)
  auto load = [&](int iky, cpx<P>(&w)[3][EPT], P(&lm)[EPT]) {
    const bool ok = valid && iky < g.ikty;
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) {
      const int kz = ok ? wrapk(j + TP * mm, g.nz, g.ktz) : PSG_NOK;
      const FillIdx q = fill_index<LAY>(g, ikx, iky, iky - g.kty, kz);
      w[0][mm] = to<P>(ld_fill(s0, q));
      w[1][mm] = to<P>(ld_fill(s0 + st_fs, q));
      w[2][mm] = to<P>(ld_fill(s0 + 2 * st_fs, q));
      lm[mm] = (umask && q.idx >= 0) ? (P)mask[q.tidx] : (P)1;
    }
  };
  cpx<P> w[3][EPT];
  P lm[EPT];
  load(iky0, w, lm);
  constexpr int NFO = BS ? 6 : 3;
#pragma unroll
  for (int r = 0; r < PIPE; r++) {
    const int iky = iky0 + r;
    const double kyd = (double)(iky - g.kty);
    cpx<P> wn[3][EPT];
    P lmn[EPT];
    if (r + 1 < PIPE) load(iky + 1, wn, lmn);
#pragma unroll
    for (int f = 0; f < NFO; f++) {
      cpx<P> v[EPT];
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) {
        if (f < 3) {
          v[mm] = w[f][mm];
        } else {
          const int kz = wrapk(j + TP * mm, g.nz, g.ktz);
          const double kzd = kz == PSG_NOK ? 0.0 : (double)kz;
          const double wk = kx * kx + kyd * kyd + kzd * kzd;
          const double wki = (double)lm[mm] / (wk > 0.001 ? wk : 0.001);
          cpx<double> u, vv, ww;
          biot_savart(kx, kyd, kzd, up(w[0][mm]), up(w[1][mm]), up(w[2][mm]), wki, u, vv, ww);
          v[mm] = to<P>(f == 3 ? u : (f == 4 ? vv : ww));
        }
      }
      fft_pencil<P, N, EPT, true>(v, lds, j, T, col, tw);
      if (valid && iky < g.ikty) {
        cpx<P> *o = out + (long)m * out_ms + (long)f * g.zsz;
#pragma unroll
        for (int mm = 0; mm < EPT; mm++) st_pass<NT<P, PASS_ZINV>::st>(&o[zoff<P>(g, ikx, iky, j + TP * mm)], v[mm]);
      }
    }
    if (r + 1 < PIPE) {
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) {
        w[0][mm] = wn[0][mm];
        w[1][mm] = wn[1][mm];
        w[2][mm] = wn[2][mm];
        lm[mm] = lmn[mm];
      }
    }
  }
}

// ---------------------------------------------------------------- pass 2 (3D)
template <typename P, int N, int T>
__global__ void __launch_bounds__(T *(N / EPT))
k_yinv(const cpx<P> *__restrict__ in, cpx<P> *__restrict__ out, Grid g, const cpx<P> *__restrict__ tw, int z0) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, z = z0 + tid.y, mf = tid.z;
  const bool valid = ikx < g.iktx;
  const cpx<P> *src = in + (long)mf * g.zsz;
  cpx<P> v[EPT];
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) {
    const int y = j + TP * mm;
    const int ky = wrapk(y, g.ny, g.kty);
    v[mm] = (valid && ky != PSG_NOK) ? ld_pass<NT<P, PASS_YINV>::ld>(&src[zoff<P>(g, ikx, ky + g.kty, z)]) : cpx<P>{0, 0};
  }
  fft_pencil<P, N, EPT, true>(v, lds, j, T, col, tw);
  if (valid) {
    cpx<P> *o = out + (long)mf * g.ysz;
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) st_pass<NT<P, PASS_YINV>::st>(&o[yoff<P>(g, ikx, j + TP * mm, z)], v[mm]);
  }
}

// ---------------------------------------------------------------- pass 4 (3D)
template <typename P, int N, int T>
__global__ void __launch_bounds__(T *(N / EPT))
k_yfwd(const cpx<P> *__restrict__ in, cpx<P> *__restrict__ out, Grid g, const cpx<P> *__restrict__ tw, int z0) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, z = z0 + tid.y, mf = tid.z;
  const bool valid = ikx < g.iktx;
  const cpx<P> *src = in + (long)mf * g.ysz + (long)z * g.ny * g.xp + ikx;
  cpx<P> v[EPT];
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) v[mm] = valid ? ld_pass<NT<P, PASS_YFWD>::ld>(&src[(long)(j + TP * mm) * g.xp]) : cpx<P>{0, 0};
  fft_pencil<P, N, EPT, false>(v, lds, j, T, col, tw);
  if (valid) {
    cpx<P> *o = out + (long)mf * g.zsz;
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) {
      const int ky = wrapk(j + TP * mm, g.ny, g.kty);
      if (ky != PSG_NOK) st_pass<NT<P, PASS_YFWD>::st>(&o[zoff<P>(g, ikx, ky + g.kty, z)], v[mm]);
    }
  }
}

// ---------------------------------------------------------------- pass 5 (3D)
// UPD::operator()(m, ik, ikx, kx, ky, kz, N[FO]) with N already scaled by 1/NR.
// FRONT: every input is requested before the first transform (the default up
// to 4 fields); otherwise each field is loaded just before its transform,
// which frees 4*(FO-1)*EPT registers (the 6-field tangent-linear/adjoint
// updates always do this).
// PIPE (with FRONT): ky rows per block, the next row's inputs requested
// before this row's transforms (the grid's y extent is IKTY/PIPE rounded up)
template <typename P, int N, int T, int FO, class UPD, bool FRONT = (FO <= 4), int PIPE = 1>
__global__ void __launch_bounds__(T *(N / EPT))
k_zfwd(const cpx<P> *__restrict__ in, Grid g, const cpx<P> *__restrict__ tw, UPD upd) {
  static_assert(PIPE == 1 || FRONT, "row pipelining needs front-loaded inputs");
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  constexpr int NV = FRONT ? FO : 1;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, iky0 = tid.y * PIPE, m = tid.z;
  const bool valid = ikx < g.iktx;
  const double kx = (double)ikx * g.alpha;
  auto load = [&](int f, int iky, cpx<P>(&v)[EPT]) {
    const cpx<P> *src = in + ((long)m * FO + f) * g.zsz;
    const bool ok = valid && iky < g.ikty;
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) v[mm] = ok ? ld_pass<NT<P, PASS_ZFWD>::ld>(&src[zoff<P>(g, ikx, iky, j + TP * mm)]) : cpx<P>{0, 0};
  };
  cpx<P> vin[NV][EPT];
  if constexpr (FRONT) {
#pragma unroll
    for (int f = 0; f < FO; f++) load(f, iky0, vin[f]);
  }
#pragma unroll
  for (int r = 0; r < PIPE; r++) {
    const int iky = iky0 + r;
    cpx<P> vnext[PIPE > 1 ? NV : 1][EPT];
    if constexpr (PIPE > 1) {
      if (r + 1 < PIPE) {
#pragma unroll
        for (int f = 0; f < FO; f++) load(f, iky + 1, vnext[f]);
      }
    }
    cpx<double> res[FO][EPT];
#pragma unroll
    for (int f = 0; f < FO; f++) {
      cpx<P> v[EPT];
      if constexpr (FRONT) {
#pragma unroll
        for (int mm = 0; mm < EPT; mm++) v[mm] = vin[f][mm];
      } else {
        load(f, iky, v);
      }
      fft_pencil<P, N, EPT, false>(v, lds, j, T, col, tw);
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) {
        cpx<double> t = up(v[mm]);
        res[f][mm] = cpx<double>{t.x * g.inr, t.y * g.inr};
      }
    }
    // rows past IKTY and columns past IKTX run the transforms (the block's
    // barriers) but update nothing
    if (valid && iky < g.ikty) {
      const double ky = (double)(iky - g.kty);
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) {
        const int kz = wrapk(j + TP * mm, g.nz, g.ktz);
        if (kz == PSG_NOK) continue;
        const long ik = ((long)(kz + g.ktz) * g.ikty + iky) * g.iktx + ikx;
        cpx<double> nn[FO];
#pragma unroll
        for (int f = 0; f < FO; f++) nn[f] = res[f][mm];
        upd(m, ik, ikx, kx, ky, (double)kz, nn);
      }
    }
    if constexpr (PIPE > 1) {
      if (r + 1 < PIPE) {
#pragma unroll
        for (int f = 0; f < FO; f++)
#pragma unroll
          for (int mm = 0; mm < EPT; mm++) vin[f][mm] = vnext[f][mm];
      }
    }
  }
}

// ---------------------------------------------------------------- pass 3 (x)
// Row teams: TEAMS rows per block, N/E threads per row.  Fields are paired
// into complex transforms z = a + i b.  PROD::apply(double in[FI], double out[FO]).
// A team that fits in one wavefront synchronises with wavefront fences only;
// its LDS row is XOR-swizzled (lpos, row_swizzle) so that the strided radix-E
// writes and the contiguous reads are bank-conflict free.
// FRONT: every input of the row is requested before the first transform
// (latency hiding for the 6-field products); without it each pair is loaded
// just before its transform, which keeps the 9- and 12-field products of the
// tangent-linear and adjoint passes within the register file.
// kept kx modes of an N-point x transform under the 2/3 rule: 0..N/3
// (IKTX = NX/3 + 1 for every grid the context builds)
template <int N> constexpr int xband() { return N / 3 + 1; }

template <typename P, int N, int E> struct XRow {
  // pair p of a row: z = a + i b from the half spectra of fields fa, fb, whose
  // row starts ra, rb hold element k at ycol<P>(g, k)
  __device__ __forceinline__ static void load(cpx<P> (&v)[E], const cpx<P> *ra, const cpx<P> *rb, bool hasb,
                                              bool valid, int j, const Grid &g) {
    constexpr int TPX = N / E;
    constexpr int ix = xband<N>();
#pragma unroll
    for (int mm = 0; mm < E; mm++) {
      const int k = j + TPX * mm;
      cpx<P> a{0, 0}, b{0, 0};
      cpx<P> z{0, 0};
      if (valid) {
        if (k < ix) {
          a = ld_pass<NT<P, PASS_XPASS>::ld>(&ra[ycol<P>(g, k)]);
          if (hasb) b = ld_pass<NT<P, PASS_XPASS>::ld>(&rb[ycol<P>(g, k)]);
          if (k == 0) { a.y = 0; b.y = 0; }
          z = cpx<P>{a.x - b.y, a.y + b.x};
        } else if (k > N - ix) {
          a = ld_pass<NT<P, PASS_XPASS>::ld>(&ra[ycol<P>(g, N - k)]);
          if (hasb) b = ld_pass<NT<P, PASS_XPASS>::ld>(&rb[ycol<P>(g, N - k)]);
          z = cpx<P>{a.x + b.y, b.x - a.y};
        }
      }
      v[mm] = z;
    }
  }
};

// LAZY: the product is re-evaluated per output pair instead of being kept
// for all outputs, which frees 2*FO*E registers for occupancy.
// WPE > 0 asks the compiler for at least WPE waves per SIMD (a register cap
// of 512/WPE VGPRs), trading spills or rematerialisation for occupancy.
template <typename P, int N, int TEAMS, int FI, int FO, class PROD, int E = EPT, bool FRONT = true,
          bool LAZY = false, int WPE = 0>
__global__ void __launch_bounds__(TEAMS *(N / E)) PSG_WAVES_PER_EU(WPE > 0 ? WPE : 1)
k_xpass(const cpx<P> *__restrict__ in, cpx<P> *__restrict__ out, long nrow_m, long nrows, Grid g,
        const cpx<P> *__restrict__ tw, PROD prod, int zs0) {
  PSG_LDS_TWIDDLES(P, N, E);
  constexpr int TPX = N / E;
  constexpr bool WAVE = TPX <= 64;
  constexpr int PI = (FI + 1) / 2, PO = (FO + 1) / 2;
  constexpr int NV = FRONT ? PI : 1;
  const int team = threadIdx.x / TPX, j = threadIdx.x % TPX;
  // row = (m, z, y) with y fastest over the launch's planes zs0 .. zs0 +
  // nrow_m/ny - 1 of each member: the block's first row is split once per
  // block (uniform), each team steps y from it (launches keep nrows < 2^31)
  const unsigned row0 = blockIdx.x * (unsigned)TEAMS, nrm = (unsigned)nrow_m, ny = (unsigned)g.ny;
  const unsigned m0 = row0 / nrm, rr0 = row0 - m0 * nrm, z0 = rr0 / ny, nzr = nrm / ny;
  unsigned y = rr0 - z0 * ny + team, z = z0, m = m0;
  while (y >= ny) { y -= ny; z++; }
  while (z >= nzr) { z -= nzr; m++; }
  const bool valid = row0 + team < (unsigned long)nrows;
  if (!valid) m = 0, z = 0, y = 0;
  z += (unsigned)zs0;
  const long rr = (long)z * ny + y;
  cpx<P> *tl = lds + team * row_len<N>();
  constexpr int SW = row_swizzle<E>();
  constexpr int ix = xband<N>();
  const long ro = yoff<P>(g, 0, (int)y, (int)z);
  auto rowp = [&](int f) { return in + ((long)m * FI + f) * g.ysz + ro; };
  cpx<P> vin[NV][E];
  if constexpr (FRONT) {
#pragma unroll
    for (int p = 0; p < PI; p++)
      XRow<P, N, E>::load(vin[p], rowp(2 * p), rowp(2 * p + 1 < FI ? 2 * p + 1 : 2 * p), 2 * p + 1 < FI, valid, j,
                          g);
  }
  P phys[2 * PI][E];
#pragma unroll
  for (int p = 0; p < PI; p++) {
    const int fa = 2 * p, fb = 2 * p + 1;
    cpx<P> v[E];
    if constexpr (FRONT) {
#pragma unroll
      for (int mm = 0; mm < E; mm++) v[mm] = vin[p][mm];
    } else {
      XRow<P, N, E>::load(v, rowp(fa), rowp(fb < FI ? fb : fa), fb < FI, valid, j, g);
    }
    fft_pencil<P, N, E, true, WAVE, SW>(v, tl, j, 1, 0, tw);
#pragma unroll
    for (int mm = 0; mm < E; mm++) { phys[fa][mm] = v[mm].x; phys[fb][mm] = v[mm].y; }
  }
  constexpr int NR = LAZY ? 1 : 2 * PO;
  P res[NR][E];
  auto product = [&](int mm, P (&pout)[2 * PO]) {
    P pin[2 * PI];
#pragma unroll
    for (int f = 0; f < 2 * PI; f++) pin[f] = phys[f][mm];
#pragma unroll
    for (int f = 0; f < 2 * PO; f++) pout[f] = 0;
    prod(pin, pout);
  };
  if constexpr (!LAZY) {
#pragma unroll
    for (int mm = 0; mm < E; mm++) {
      P pout[2 * PO];
      product(mm, pout);
#pragma unroll
      for (int f = 0; f < 2 * PO; f++) res[f][mm] = pout[f];
    }
  }
  (void)res;
#pragma unroll
  for (int q = 0; q < PO; q++) {
    cpx<P> v[E];
#pragma unroll
    for (int mm = 0; mm < E; mm++) {
      if constexpr (LAZY) {
        P pout[2 * PO];
        product(mm, pout);
        v[mm] = cpx<P>{pout[2 * q], pout[2 * q + 1]};
      } else {
        v[mm] = cpx<P>{res[2 * q][mm], res[2 * q + 1][mm]};
      }
    }
    fft_pencil<P, N, E, false, WAVE, SW>(v, tl, j, 1, 0, tw);
    team_sync<WAVE>();
#pragma unroll
    for (int mm = 0; mm < E; mm++) {
      const int k = j + TPX * mm;   // only the kept modes and their mirrors are read back
      if (k < ix || k > N - ix) tl[lpos<SW>(k)] = v[mm];
    }
    team_sync<WAVE>();
    if (valid) {
      cpx<P> *oa = out + ((long)m * FO + 2 * q) * g.ysz + rr * g.xp;
      cpx<P> *ob = out + ((long)m * FO + 2 * q + 1) * g.ysz + rr * g.xp;
      for (int k = j; k < ix; k += TPX) {
        cpx<P> zk = tl[lpos<SW>(k)], zc = tl[lpos<SW>((N - k) & (N - 1))];
        const P h = (P)0.5;
        cpx<P> A{h * (zk.x + zc.x), h * (zk.y - zc.y)};
        cpx<P> B{h * (zk.y + zc.y), h * (zc.x - zk.x)};
        st_pass<NT<P, PASS_XPASS>::st>(&oa[k], A);
        if (2 * q + 1 < FO) st_pass<NT<P, PASS_XPASS>::st>(&ob[k], B);
      }
    }
  }
}

// Split-row x-pass for grids with few rows (2D): one row per block and one
// team per field pair, so that a row's transforms run side by side instead of
// one after another in a single team; the kernel time is then about one
// inverse and one forward transform.  Phase 1: team p inverse-transforms
// input pair p and leaves the physical values in LDS (phys[f][n]).  Phase 2:
// team q forms the product at its points and forward-transforms output pair
// q.  Every team runs both phases (on zeros where it has no pair) so that the
// synchronisation is the same in every thread.
// A product that needs the row and column of each point (the recorded-base
// products of the 2D tangent linear) declares ROW = true and provides
// bind(member, row, team) returning an object called as (n, in, out).
template <class T, class = void> struct prod_row : std::false_type {};
template <class T> struct prod_row<T, std::void_t<decltype(T::ROW)>> : std::bool_constant<T::ROW> {};
// An update that must act once per launch (the step counter of the recorded
// base) provides once(), called by one thread of the first tile.
template <class T, class = void> struct upd_once : std::false_type {};
template <class T> struct upd_once<T, std::void_t<decltype(T::ONCE)>> : std::bool_constant<T::ONCE> {};

template <int FI, int FO> constexpr int xsplit_teams() { return (FI + 1) / 2 > (FO + 1) / 2 ? (FI + 1) / 2 : (FO + 1) / 2; }
template <typename P, int N, int FI, int FO>
constexpr size_t xsplit_lds() {
  return (size_t)(xsplit_teams<FI, FO>() * row_len<N>() + N) * sizeof(cpx<P>) +
         (size_t)2 * ((FI + 1) / 2) * N * sizeof(P);
}

template <typename P, int N, int FI, int FO, class PROD, int E = EPT>
__global__ void __launch_bounds__((xsplit_teams<FI, FO>() * (N / E)))
k_xpass_split(const cpx<P> *__restrict__ in, cpx<P> *__restrict__ out, long nrow_m, long nrows, Grid g,
              const cpx<P> *__restrict__ tw, PROD prod) {
  PSG_LDS_TWIDDLES(P, N, E);
  constexpr int TPX = N / E;
  constexpr bool WAVE = TPX <= 64;
  constexpr int PI = (FI + 1) / 2, PO = (FO + 1) / 2, NTEAM = xsplit_teams<FI, FO>();
  constexpr int SW = row_swizzle<E>();
  constexpr int ix = xband<N>();
  const int team = threadIdx.x / TPX, j = threadIdx.x % TPX;
  const unsigned row = blockIdx.x, nrm = (unsigned)nrow_m, ny = (unsigned)g.ny;
  const bool valid = row < (unsigned long)nrows;
  const unsigned m = valid ? row / nrm : 0, rr = valid ? row - m * nrm : 0, z = rr / ny, y = rr - z * ny;
  cpx<P> *tl = lds + team * row_len<N>();
  P *phys = reinterpret_cast<P *>(lds + NTEAM * row_len<N>());
  const long ro = yoff<P>(g, 0, (int)y, (int)z);
  auto rowp = [&](int f) { return in + ((long)m * FI + f) * g.ysz + ro; };
  {
    const int p = team < PI ? team : PI - 1, fa = 2 * p, fb = 2 * p + 1;
    cpx<P> v[E];
    XRow<P, N, E>::load(v, rowp(fa), rowp(fb < FI ? fb : fa), fb < FI, valid && team < PI, j, g);
    fft_pencil<P, N, E, true, WAVE, SW>(v, tl, j, 1, 0, tw);
    if (team < PI) {
#pragma unroll
      for (int mm = 0; mm < E; mm++) {
        const int n = j + TPX * mm;
        phys[fa * N + n] = v[mm].x;
        phys[fb * N + n] = v[mm].y;
      }
    }
  }
  __syncthreads();
  const int q = team < PO ? team : PO - 1;
  cpx<P> v[E];
  auto bound = [&]() {
    if constexpr (prod_row<PROD>::value) return prod.bind(m, rr, team);
    else return 0;
  }();
#pragma unroll
  for (int mm = 0; mm < E; mm++) {
    const int n = j + TPX * mm;
    P pin[2 * PI], pout[2 * PO];
#pragma unroll
    for (int f = 0; f < 2 * PI; f++) pin[f] = phys[f * N + n];
#pragma unroll
    for (int f = 0; f < 2 * PO; f++) pout[f] = 0;
    if constexpr (prod_row<PROD>::value) bound(n, pin, pout);
    else prod(pin, pout);
    v[mm] = cpx<P>{pout[2 * q], pout[2 * q + 1]};
  }
  (void)bound;
  fft_pencil<P, N, E, false, WAVE, SW>(v, tl, j, 1, 0, tw);
  team_sync<WAVE>();
#pragma unroll
  for (int mm = 0; mm < E; mm++) {
    const int k = j + TPX * mm;   // only the kept modes and their mirrors are read back
    if (k < ix || k > N - ix) tl[lpos<SW>(k)] = v[mm];
  }
  team_sync<WAVE>();
  if (valid && team < PO) {
    cpx<P> *oa = out + ((long)m * FO + 2 * q) * g.ysz + (long)rr * g.xp;
    cpx<P> *ob = out + ((long)m * FO + 2 * q + 1) * g.ysz + (long)rr * g.xp;
    for (int k = j; k < ix; k += TPX) {
      cpx<P> zk = tl[lpos<SW>(k)], zc = tl[lpos<SW>((N - k) & (N - 1))];
      const P h = (P)0.5;
      cpx<P> A{h * (zk.x + zc.x), h * (zk.y - zc.y)};
      cpx<P> B{h * (zk.y + zc.y), h * (zc.x - zk.x)};
      st_pass<false>(&oa[k], A);   // 2D grids: the y-forward reads C back from the caches
      if (2 * q + 1 < FO) st_pass<false>(&ob[k], B);
    }
  }
}

// ---------------------------------------------------------------- products
// 3D rotational form: (U,V,W) <- -(u x w)  (3D_NavierStokes/src/kernels.h:281-306)
// inputs: 0..2 vorticity, 3..5 velocity
template <typename P> struct ProdNS {
  __device__ __forceinline__ void operator()(const P *in, P *out) const {
    const P tx = in[0], te = in[1], tz = in[2], tu = in[3], tv = in[4], tw = in[5];
    const P nx = tv * tz - tw * te;
    const P ny = tw * tx - tu * tz;
    const P nz = tu * te - tv * tx;
    out[0] = -nx; out[1] = -ny; out[2] = -nz;
  }
};

// 2D tangent linear: base (v w, u w) and their linearisation
// (dv w + v dw, du w + u dw); inputs w,u,v,dw,du,dv
template <typename P> struct Prod2DTL {
  __device__ __forceinline__ void operator()(const P *in, P *out) const {
    out[0] = in[2] * in[0];
    out[1] = in[1] * in[0];
    out[2] = in[5] * in[0] + in[2] * in[3];
    out[3] = in[4] * in[0] + in[1] * in[3];
  }
};

// 2D: NZR = (v w, u w)  (2D_NavierStokes/src_RK4/kernels.h:187-201); inputs w,u,v
template <typename P> struct Prod2D {
  __device__ __forceinline__ void operator()(const P *in, P *out) const {
    out[0] = in[2] * in[0];
    out[1] = in[1] * in[0];
  }
};

// ---------------------------------------------------------------- 2D passes
// y-inverse straight from the state: Hermitian fill of kx=0 (in place in the
// reference, 2D_NavierStokes/src_RK4/kernels.h:82-96), mask, Biot-Savart
// (kernels.h:112-147), zero-pad, inverse y-FFT.  out fields w,u,v.
// SPLIT: blockIdx.y picks one of the 3*NSRC output fields (3*NSRC times the
// blocks, for grids too small to fill the GPU with column tiles alone)
template <typename P, int N, int T, int NSRC = 1, bool SPLIT = false>
__global__ void __launch_bounds__(T *(N / EPT))
k2_yinv(const cpx<double> *__restrict__ st, long st_ms, cpx<P> *__restrict__ out, long out_ms, Grid g,
        const cpx<P> *__restrict__ tw, const int8_t *__restrict__ mask,
        const cpx<double> *__restrict__ st2 = nullptr) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, m = tid.z, sel = tid.y;
  const bool valid = ikx < g.iktx;
  const double kx = (double)ikx * g.alpha;
#pragma unroll 1
  for (int src = 0; src < NSRC; src++) {
    if (SPLIT && src != sel / 3) continue;
    const cpx<double> *s0 = (src == 0 ? st : st2) + (long)m * st_ms;
    cpx<double> w[EPT];
    double kyv[EPT];
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) {
      const int y = j + TP * mm;
      const int ky = wrapk(y, g.ny, g.kty);
      cpx<double> a{0, 0};
      kyv[mm] = ky == PSG_NOK ? 0.0 : (double)ky;
      if (valid && ky != PSG_NOK) {
        if (ikx == 0 && ky < 0) {
          long idx = (long)(-ky + g.kty) * g.iktx;
          if (mask[idx]) { a = s0[idx]; a.y = -a.y; }
        } else {
          long idx = (long)(ky + g.kty) * g.iktx + ikx;
          if (mask[idx]) a = s0[idx];
        }
      }
      w[mm] = a;
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
      if (SPLIT && f != sel % 3) continue;
      cpx<P> v[EPT];
#pragma unroll
      for (int mm = 0; mm < EPT; mm++) {
        if (f == 0) v[mm] = to<P>(w[mm]);
        else {
          const double ky = kyv[mm];
          const double wk = kx * kx + ky * ky;
          const double wki = 1.0 / (wk > 0.001 ? wk : 0.001);
          cpx<double> r;
          if (f == 1) r = cpx<double>{-ky * w[mm].y * wki, ky * w[mm].x * wki};
          else r = cpx<double>{kx * w[mm].y * wki, -kx * w[mm].x * wki};
          v[mm] = to<P>(r);
        }
      }
      fft_pencil<P, N, EPT, true>(v, lds, j, T, col, tw);
      if (valid) {
        cpx<P> *o = out + (long)m * out_ms + (long)(3 * src + f) * g.ysz;
#pragma unroll
        for (int mm = 0; mm < EPT; mm++) st_pass<false>(&o[yoff<P>(g, ikx, j + TP * mm, 0)], v[mm]);   // 2D: the next pass reads it from cache
      }
    }
  }
}

// y-forward + update (2D).  UPD::operator()(m, ik, ikx, kx, ky, N[2]) with
// N[0] = FFT(v w)/NR, N[1] = FFT(u w)/NR  (the reference's NZK halves).
// NG > 1: the NF fields form NG independent groups, blockIdx.y picks one and
// its update is UPD::part(group, m, ik, ikx, kx, ky, N[NF/NG]).
template <typename P, int N, int T, class UPD, int NF = 2, int NG = 1>
__global__ void __launch_bounds__(T *(N / EPT))
k2_yfwd(const cpx<P> *__restrict__ in, long in_ms, Grid g, const cpx<P> *__restrict__ tw, UPD upd) {
  PSG_LDS_TWIDDLES(P, N, EPT);
  constexpr int TP = N / EPT, NFG = NF / NG;
  const int col = threadIdx.x % T, j = threadIdx.x / T;
  const TileId tid = xcd_tile();
  const int ikx = tid.x * T + col, m = tid.z, grp = NG > 1 ? tid.y : 0;
  const bool valid = ikx < g.iktx;
  cpx<double> res[NFG][EPT];
#pragma unroll
  for (int f = 0; f < NFG; f++) {
    const cpx<P> *src = in + (long)m * in_ms + (long)(grp * NFG + f) * g.ysz + ikx;
    cpx<P> v[EPT];
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) v[mm] = valid ? src[(long)(j + TP * mm) * g.xp] : cpx<P>{0, 0};
    fft_pencil<P, N, EPT, false>(v, lds, j, T, col, tw);
#pragma unroll
    for (int mm = 0; mm < EPT; mm++) {
      cpx<double> t = up(v[mm]);
      res[f][mm] = cpx<double>{t.x * g.inr, t.y * g.inr};
    }
  }
  if constexpr (upd_once<UPD>::value) {
    if (tid.x == 0 && tid.y == 0 && tid.z == 0 && threadIdx.x == 0) upd.once();
  }
  if (!valid) return;
  const double kx = (double)ikx * g.alpha;
#pragma unroll
  for (int mm = 0; mm < EPT; mm++) {
    const int ky = wrapk(j + TP * mm, g.ny, g.kty);
    if (ky == PSG_NOK) continue;
    const long ik = (long)(ky + g.kty) * g.iktx + ikx;
    cpx<double> nn[NFG];
#pragma unroll
    for (int f = 0; f < NFG; f++) nn[f] = res[f][mm];
    if constexpr (NG > 1) upd.part(grp, m, ik, ikx, kx, (double)ky, nn);
    else upd(m, ik, ikx, kx, (double)ky, nn);
  }
}

}  // namespace psg
