// 16 x 16 tile building blocks shared by the diagonal-block kernels
// (diag.hip) and the batched factorisation (batched.hip).
//
// * TileAcc / tile_mma: a 16 x 16 accumulator in the v_mfma_f64_16x16x4_f64
//   result layout (lane l, register r holds element (l / 16 + 4 r, l % 16))
//   and C += s X op(Y) over a 16-deep k range, complex as four real MFMAs.
// * leaf_factor / leaf_inverse: one wave factors a 16 x 16 Hermitian
//   positive definite tile held in registers (lane r + 16 q holds row r,
//   columns q + 4 j, j = 0..3) and inverts the factor, moving pivots and
//   column values with cross-lane shuffles, so the 16 sequential pivots need
//   no workgroup barrier.
#pragma once

#include "internal.h"

namespace hc {
namespace t16 {

constexpr int LEAF = 16;

typedef double v4d __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ double lane_shfl(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ zd lane_shfl(zd v, int src) {
  return make_zd(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// Orders this wave's LDS writes before its following LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Factor the leaf in registers in place: on return the lower part of `a` is
// L.  Returns the first non-positive / NaN pivot (0-based, the same value in
// every lane) or -1.  *logsum accumulates log(d_k) of the pivots k < kmax
// (d_k = L_kk^2), in pivot order.
template <typename T>
__device__ __forceinline__ int leaf_factor(T a[4], int kmax, double *logsum) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int k = 0; k < LEAF; k++) {
    const int jk = k >> 2, qk = k & 3;  // column k: lanes (*, qk), register jk
    const T ak = a[jk];                 // column k before this step's update
    const double d = realx(lane_shfl(ak, k + 16 * qk));
    if (!(d > 0.0)) return k;           // uniform: every lane holds the same d
    if (k < kmax) *logsum += log(d);
    const double sd = sqrt(d), rs = 1.0 / sd;
    const T lrk = rs * lane_shfl(ak, r + 16 * qk);
    const T fin = sel(r == k, from_real<T>(sd), lrk);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int c = q + 4 * j;
      const T lck = conjx(rs * lane_shfl(ak, c + 16 * qk));
      const T upd = a[j] - lrk * lck;
      a[j] = sel(c > k, upd, sel(c == k, fin, a[j]));
    }
  }
  return -1;
}

// x = inv(L) of the factored leaf (lower, exact zeros above), by forward
// substitution over rows.
template <typename T>
__device__ __forceinline__ void leaf_inverse(const T a[4], T x[4]) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int j = 0; j < 4; j++) x[j] = from_real<T>(0.0);
#pragma unroll
  for (int k = 0; k < LEAF; k++) {
    const int jk = k >> 2, qk = k & 3;
    const T dk = from_real<T>(1.0) / lane_shfl(a[jk], k + 16 * qk);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int c = q + 4 * j;
      const T fin = (sel(c == k, from_real<T>(1.0), from_real<T>(0.0)) - x[j]) * dk;
      x[j] = sel(r == k, fin, x[j]);
    }
    const T lrk = lane_shfl(a[jk], r + 16 * qk);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const T xkc = lane_shfl(x[j], k + 16 * q);
      x[j] = sel(r > k, x[j] + lrk * xkc, x[j]);
    }
  }
}

// ---------------------------------------------------------------------------
// Row layout: lane r + 16 q holds row r of the leaf in a[0..15] (the four lane
// groups q hold the same row; entries above the diagonal are not used).  A
// pivot's own row element is then in the lane's registers, and everything
// else a pivot needs comes from v_readlane at a compile-time lane, with no LDS
// round trip: one wave factors and inverts a 16 x 16 leaf in about a fifth of
// the time of the shuffle-based column layout above.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double readlane(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ zd readlane(zd v, int lane) { return make_zd(readlane(v.x, lane), readlane(v.y, lane)); }

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
// lane) or -1; *logsum accumulates log(d_k) of the pivots k < kmax.
template <typename T>
__device__ __forceinline__ int leaf_factor_rows(T a[LEAF], int kmax, double *logsum) {
  const int r = threadIdx.x & 15;
#pragma unroll
  for (int k = 0; k < LEAF; k++) {
    const double d = realx(readlane(a[k], k));  // A(k, k) after k rank-1 updates
    if (!(d > 0.0)) return k;                   // uniform
    if (k < kmax) *logsum += log(d);
    const double rs = rsqrt_nr(d);
    const T lr = rs * a[k];                     // L(r, k) for r > k
#pragma unroll
    for (int c = k + 1; c < LEAF; c++) {
      // A(r, c) -= L(r, k) conj(L(c, k)); rows r < c only touch unused entries
      const T lc = conjx(rs * readlane(a[k], c));
      a[c] = a[c] - lr * lc;
    }
    a[k] = sel(r == k, from_real<T>(d * rs), sel(r > k, lr, a[k]));
  }
  return -1;
}

// x = inv(L) (row layout: x[c], c <= r, is X(r, c); zeros above) of a leaf
// factored by leaf_factor_rows, by forward substitution over rows:
// X(k, :) = (e_k - sum_{j < k} L(k, j) X(j, :)) / L(k, k).
template <typename T>
__device__ __forceinline__ void leaf_inverse_rows(const T a[LEAF], T x[LEAF]) {
  const int r = threadIdx.x & 15;
  T dg = a[0];
#pragma unroll
  for (int c = 1; c < LEAF; c++) dg = sel(r == c, a[c], dg);
  const T dinv = from_real<T>(1.0) / dg;  // every lane its own row's pivot, in parallel
  T acc[LEAF];
#pragma unroll
  for (int c = 0; c < LEAF; c++) {
    acc[c] = from_real<T>(0.0);
    x[c] = from_real<T>(0.0);
  }
#pragma unroll
  for (int k = 0; k < LEAF; k++) {
    // row k of X, finished by lane k (other lanes compute and discard it)
#pragma unroll
    for (int c = 0; c <= k; c++) {
      const T t = (sel(r == c, from_real<T>(1.0), from_real<T>(0.0)) - acc[c]) * dinv;
      x[c] = sel(r == k, t, x[c]);
      const T xk = readlane(t, k);
      acc[c] = acc[c] + a[k] * xk;  // rows r > k: sum_{j <= k} L(r, j) X(j, c)
    }
  }
}

}  // namespace t16
}  // namespace hc
