// 16 x 16 tile building blocks shared by the diagonal-block kernels
// (diag.hip) and the batched factorisation (batched.hip).
//
// * TileAcc / tile_mma: a 16 x 16 accumulator in the v_mfma_f64_16x16x4_f64
//   result layout (lane l, register r holds element (l / 16 + 4 r, l % 16))
//   and C += s X op(Y) over a 16-deep k range, complex as four real MFMAs.
// * leaf_factor_rows / leaf_inverse_cols: one wave factors a 16 x 16
//   Hermitian positive definite tile held in registers (lane r + 16 q holds
//   row r) and inverts the factor, moving values between lanes with DPP row
//   broadcasts, so the 16 sequential pivots need no workgroup barrier and no
//   LDS round trip.
#pragma once

#include <type_traits>

#include "internal.h"

namespace hc {
namespace t16 {

constexpr int LEAF = 16;

typedef double v4d __attribute__((ext_vector_type(4)));

// Last argument of __builtin_amdgcn_mfma_f64_*: on gfx950 the FP64 MFMA's
// BLGP field holds negate modifiers (bit 0: A, bit 1: B, bit 2: C), so
// acc - A B is one instruction with no VALU negation of the operand.
constexpr int kNegA = 1;

__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ zd sel(bool c, zd a, zd b) { return make_zd(c ? a.x : b.x, c ? a.y : b.y); }

template <typename T> struct TileAcc;
template <> struct TileAcc<double> { v4d re; };
template <> struct TileAcc<zd> { v4d re, im; };

__device__ __forceinline__ void acc_zero(TileAcc<double> &c) { c.re = (v4d){0.0, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ void acc_zero(TileAcc<zd> &c) {
  c.re = (v4d){0.0, 0.0, 0.0, 0.0};
  c.im = c.re;
}

__device__ __forceinline__ double acc_get(const TileAcc<double> &c, int r) { return c.re[r]; }
__device__ __forceinline__ zd acc_get(const TileAcc<zd> &c, int r) { return make_zd(c.re[r], c.im[r]); }
__device__ __forceinline__ void acc_set(TileAcc<double> &c, int r, double v) { c.re[r] = v; }
__device__ __forceinline__ void acc_set(TileAcc<zd> &c, int r, zd v) {
  c.re[r] = v.x;
  c.im[r] = v.y;
}

// Tile at C (column-major, leading dimension ls) <-> accumulator.
template <typename T>
__device__ __forceinline__ void acc_load(TileAcc<T> &c, const T *C, int ls) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int r = 0; r < 4; r++) acc_set(c, r, C[li * ls + lk + 4 * r]);
}
template <typename T>
__device__ __forceinline__ void acc_store(const TileAcc<T> &c, T *C, int ls) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int r = 0; r < 4; r++) C[li * ls + lk + 4 * r] = acc_get(c, r);
}

// acc += sx * X op(Y), 16-deep, X and Y column-major tiles; op(Y) = Y^H when
// YH.  MFMA operands: X fragment lane l = X(l % 16, k0 + l / 16), op(Y)
// fragment lane l = op(Y)(k0 + l / 16, l % 16).
template <bool YH>
__device__ __forceinline__ void tile_mma(TileAcc<double> &c, const double *X, int ldx, const double *Y,
                                         int ldy, double sx) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int k0 = 0; k0 < LEAF; k0 += 4) {
    const double a = sx * X[(k0 + lk) * ldx + li];
    const double b = YH ? Y[(k0 + lk) * ldy + li] : Y[li * ldy + k0 + lk];
    c.re = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c.re, 0, 0, 0);
  }
}
template <bool YH>
__device__ __forceinline__ void tile_mma(TileAcc<zd> &c, const zd *X, int ldx, const zd *Y, int ldy,
                                         double sx) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int k0 = 0; k0 < LEAF; k0 += 4) {
    const zd a = sx * X[(k0 + lk) * ldx + li];
    const zd b = YH ? conjx(Y[(k0 + lk) * ldy + li]) : Y[li * ldy + k0 + lk];
    c.re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, c.re, 0, 0, 0);
    c.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, c.re, 0, 0, 0);
    c.im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, c.im, 0, 0, 0);
    c.im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, c.im, 0, 0, 0);
  }
}

// Orders this wave's LDS writes before its following LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// ---------------------------------------------------------------------------
// Row layout: lane r + 16 q holds row r of the leaf in a[0..15] (the four
// 16-lane rows of the wave, q = 0..3, hold the same matrix row; entries above
// the diagonal are not used).  A pivot's own row element is then in the lane's
// registers, and the other elements it needs come from DPP row_newbcast (lane
// m of each 16-lane row to the whole row: a VGPR move, no LDS round trip and
// no scalar register).
// ---------------------------------------------------------------------------

// f(integral_constant<int, I>) for I = B .. E - 1: loop indices that are
// compile-time constants without relying on the unroller (which declines
// large bodies, leaving register arrays indexed at run time, i.e. in scratch).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// v from lane M of this lane's 16-lane row.
template <int M>
__device__ __forceinline__ int row_bcast_b32(int v) {
  // every lane reads a valid source lane, so no "old" value is needed
  return __builtin_amdgcn_mov_dpp(v, 0x150 + M, 0xf, 0xf, true);
}
// A 64-bit value moves in one v_mov_b64_dpp (gfx950 has 64-bit DPP moves
// with row_newbcast), not as two 32-bit halves.
template <int M>
__device__ __forceinline__ double row_bcast(double v) {
  return __builtin_amdgcn_mov_dpp(v, 0x150 + M, 0xf, 0xf, true);
}
template <int M>
__device__ __forceinline__ zd row_bcast(zd v) {
  return make_zd(row_bcast<M>(v.x), row_bcast<M>(v.y));
}

// 1 / sqrt(d) for a normal d > 0: the hardware estimate (v_rsq_f64) refined by
// two Newton-Raphson steps y <- y + y (1 - d y^2) / 2, each of which squares
// the relative error; the result is within a few ulp of the correctly rounded
// value, at a third of the dependent operations of sqrt followed by a divide.
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = __builtin_fma(-d * y, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
  }
  return y;
}

// Factor the leaf (row layout) in place: on return a[c], c <= r, is L(r, c).
// Returns the first non-positive / NaN pivot (0-based, the same in every
// lane) or -1; *logsum accumulates log(d_k) of the pivots k < kmax.  The
// rank-1 update uses A(r, k) / d and conj(A(c, k)) unscaled, so the pivot's
// square root is off the update chain.  There is no branch on the pivot: a
// failed pivot is recorded and replaced by 1, and the caller discards the
// leaf (a data-dependent exit per pivot costs nested exec-mask regions).
template <typename T>
__device__ __forceinline__ int leaf_factor_rows(T a[LEAF], int kmax, double *logsum) {
  int r = threadIdx.x & 15;
  asm volatile("" : "+v"(r));
  int bad = -1;
  static_for<0, LEAF>([&](auto K) {
    constexpr int k = decltype(K)::value;
    const double d0 = realx(row_bcast<k>(a[k]));  // A(k, k) after k rank-1 updates
    const bool ok = d0 > 0.0;                     // false for NaN
    bad = (bad < 0 && !ok) ? k : bad;
    const double d = ok ? d0 : 1.0;
    if (k < kmax) *logsum += log(d);
    const double rs = rsqrt_nr(d);
    const T ld = (rs * rs) * a[k];                // A(r, k) / d
    static_for<k + 1, LEAF>([&](auto C) {         // rows r < c only touch unused entries
      constexpr int c = decltype(C)::value;
      a[c] = a[c] - ld * conjx(row_bcast<c>(a[k]));
    });
    a[k] = sel(r == k, from_real<T>(d * rs), sel(r > k, rs * a[k], a[k]));
  });
  return bad;
}

// Columns q, q + 4, q + 8, q + 12 of X = inv(L) for a leaf factored by
// leaf_factor_rows (lane r + 16 q: x[j] = X(r, q + 4 j), zeros above the
// diagonal; the wave's four rows share the sixteen columns).  Forward
// substitution X(k, c) = (delta_kc - sum_{j < k} L(k, j) X(j, c)) / L(k, k):
// step k broadcasts row k of X inside each 16-lane row and every lane adds
// its L(r, k) X(k, c); with the strictly lower part of L the sums of rows
// already finished stop growing, so X is formed once after the loop.
template <typename T>
__device__ __forceinline__ void leaf_inverse_cols(const T a[LEAF], T x[4]) {
  int r = threadIdx.x & 15, q = (threadIdx.x >> 4) & 3;
  // opaque here, so that what is derived from the lane index is computed
  // here rather than hoisted out of the caller's loop and kept live (and
  // spilled) through the factorisation
  asm volatile("" : "+v"(r), "+v"(q));
  T dg = a[0];
#pragma unroll
  for (int c = 1; c < LEAF; c++) dg = sel(r == c, a[c], dg);
  const T dinv = from_real<T>(1.0) / dg;  // every lane its own row's pivot, in parallel
  T e[4], acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    e[j] = sel(r == q + 4 * j, from_real<T>(1.0), from_real<T>(0.0));
    acc[j] = from_real<T>(0.0);
  }
  static_for<0, LEAF>([&](auto K) {
    constexpr int k = decltype(K)::value;
    const T lk = sel(k < r, a[k], from_real<T>(0.0));  // strictly lower part of L
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const T xk = row_bcast<k>((e[j] - acc[j]) * dinv);  // X(k, q + 4 j), final in lane k
      acc[j] = acc[j] + lk * xk;
    }
  });
#pragma unroll
  for (int j = 0; j < 4; j++) x[j] = (e[j] - acc[j]) * dinv;
}

// leaf_factor_rows followed by the inverse, with fewer instructions: the leaf
// wave issues every one of them for all 64 lanes, so the count, not the
// pivot chain's latency, sets the leaf's time.  With L = L1 D (L1 unit lower,
// D = diag(L_kk)), inv(L) = D^-1 inv(L1), and the rows of Y = inv(L1) follow
// from Y(k, c) = delta_kc - sum_{j<k} L1(k, j) Y(j, c) with no division:
// at step k, row k of Y is final in lane k, is broadcast, and every lane
// r > k subtracts L1(r, k) = L(r, k) / L_kk times it (two instructions per
// column, against four for the substitution with L).  X = D^-1 Y and L1 use
// 1 / L_kk = rsqrt(d_k), which the factorisation computes anyway, so no
// divide remains.  On return a[c] (c <= r) is L(r, c) as from
// leaf_factor_rows, x[j] = X(r, q + 4 j) (zeros above the diagonal) and
// rpiv = 1 / L(r, r) of the lane's row r.  Returns the first non-positive /
// NaN pivot (0-based) or -1; on failure x and rpiv are unspecified.
template <typename T>
__device__ __forceinline__ int leaf_factor_invert(T a[LEAF], T x[4], double &rpiv) {
  int r = threadIdx.x & 15, q = (threadIdx.x >> 4) & 3;
  asm volatile("" : "+v"(r), "+v"(q));
  int bad = -1;
  double rs_own = 1.0;
  static_for<0, LEAF>([&](auto K) {
    constexpr int k = decltype(K)::value;
    const double d0 = realx(row_bcast<k>(a[k]));  // A(k, k) after k rank-1 updates
    const bool ok = d0 > 0.0;                     // false for NaN
    bad = (bad < 0 && !ok) ? k : bad;
    const double d = ok ? d0 : 1.0;
    const double rs = rsqrt_nr(d);
    const T ld = (rs * rs) * a[k];                // A(r, k) / d
    static_for<k + 1, LEAF>([&](auto C) {         // rows r < c only touch unused entries
      constexpr int c = decltype(C)::value;
      a[c] = a[c] - ld * conjx(row_bcast<c>(a[k]));
    });
    rs_own = r == k ? rs : rs_own;
    a[k] = sel(r == k, from_real<T>(d * rs), rs * a[k]);  // entries r < k are unused
  });
  T y[4];
#pragma unroll
  for (int j = 0; j < 4; j++) y[j] = sel(r == q + 4 * j, from_real<T>(1.0), from_real<T>(0.0));
  static_for<0, LEAF>([&](auto K) {
    constexpr int k = decltype(K)::value;
    const T l1 = sel(r > k, row_bcast<k>(rs_own) * a[k], from_real<T>(0.0));  // L1(r, k)
#pragma unroll
    for (int j = 0; j < 4; j++) y[j] = y[j] - l1 * row_bcast<k>(y[j]);  // row k of Y is final in lane k
  });
#pragma unroll
  for (int j = 0; j < 4; j++) x[j] = rs_own * y[j];
  rpiv = rs_own;
  return bad;
}

}  // namespace t16
}  // namespace hc
