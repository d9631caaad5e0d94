// Diagonal-block kernels and reductions, double and double complex.
//
// A diagonal block of order n <= 128 (double) or n <= 64 (complex) is held in
// LDS (and, for the factor and inverse, in registers) by one workgroup and
// processed in "lower form": for an
// upper-triangle caller the block is loaded conjugate-transposed, so
// U^H U = L L^H, inv(U) = inv(L)^H and U U^H = L^H L reduce every case to a
// lower-triangular kernel, and stored back conjugate-transposed.
//
//   diag_potrf   factor + inverse (reference lapack/dpotrf.cu:23-148 and the
//                host dpotrf of lapack/dpotrf.c:268; here it never leaves the GPU)
//   diag_trtri   triangular inverse (reference lapack/dtrtri.cu:23-167)
//   diag_lauum   L^H L / U U^H      (reference lapack/dlauum.cu:23-91)
//   trsm_diag_inverses  batched inverses of op(A)'s diagonal blocks
//   logdet       2 sum log(Re x_i)  (reference lapack/dlogdet.cu:53-118)
#include "internal.h"
#include "tile16.h"
#include "publish.h"

#include <climits>

namespace hc {
namespace {

// Factor / inverse kernels: ten waves, one per 32 x 32 (double) or 16 x 16
// (complex) region of the lower triangle of a 4 x 4 grid of regions.
constexpr int DIAG_THREADS = 640;
// L^H L kernel: 512 threads in groups of TPR per row.
constexpr int LAUUM_THREADS = 512;

template <typename T> __device__ __forceinline__ T shfl_xor_t(T v, int m);
template <> __device__ __forceinline__ double shfl_xor_t<double>(double v, int m) {
  return __shfl_xor(v, m, 64);
}
template <> __device__ __forceinline__ zd shfl_xor_t<zd>(zd v, int m) {
  return make_zd(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

// Sum over groups of G consecutive lanes (G a power of two dividing 64).
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v = v + shfl_xor_t(v, off);
  return v;
}

// Visits the n x n block at A (column-major, 1 <= n <= NMAX) as
// use(i, k, A(i, k)), i the contiguous index: thread t takes row t % NMAX of
// columns t / NMAX + j * (THREADS / NMAX), U loads in flight together from
// clamped (always valid) addresses.  Nothing divides by the runtime n: a
// 32-bit division is a ~40-instruction sequence here, and per-element
// divisions used to cost the diagonal-block kernels more than their loads.
template <int NMAX, int THREADS, int U, typename T, typename Use>
__device__ __forceinline__ void gather_block(int n, const T *A, long lda, Use use) {
  static_assert(THREADS % NMAX == 0, "whole rows per sweep");
  constexpr int STEP = THREADS / NMAX, PER = (NMAX + STEP - 1) / STEP;
  const int i = (int)threadIdx.x % NMAX, k0 = (int)threadIdx.x / NMAX;
  const int ic = min(i, n - 1);
#pragma unroll
  for (int j0 = 0; j0 < PER; j0 += U) {
    if (k0 + j0 * STEP >= n) break;
    T v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = A[(long)min(k0 + (j0 + u) * STEP, n - 1) * lda + ic];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = k0 + (j0 + u) * STEP;
      if (i < n && k < n) use(i, k, v[u]);
    }
  }
}

// fn(i, k) for every element of an n x n block, with the thread mapping of
// gather_block (stores to column-major memory stay contiguous across lanes).
template <int NMAX, int THREADS, typename Fn>
__device__ __forceinline__ void for_block(int n, Fn fn) {
  static_assert(THREADS % NMAX == 0, "whole rows per sweep");
  constexpr int STEP = THREADS / NMAX;
  const int i = (int)threadIdx.x % NMAX, k0 = (int)threadIdx.x / NMAX;
  if (i < n)
    for (int k = k0; k < n; k += STEP) fn(i, k);
}

// S is column-major with leading dimension NMAX + 1: S(r, c) = S[c * LS + r].
template <typename T, int NMAX>
struct Lds {
  static constexpr int LS = NMAX + 1;
  static constexpr int TPR = LAUUM_THREADS / NMAX;  // threads per row in sweeps
};

// Load the lower form of the n x n block at A: S(r, c) = A(r, c) (lower) or
// conj(A(c, r)) (upper), r >= c.  Memory is read along its contiguous index.
template <typename T, int NMAX, int THREADS, typename ST = T>
__device__ __forceinline__ void load_lower_form(T *S, bool lower, int n, const ST *A, long lda) {
  constexpr int LS = Lds<T, NMAX>::LS;
  // a: contiguous index in memory; the unreferenced triangle is read but not used
  gather_block<NMAX, THREADS, 16>(n, A, lda, [&](int a, int b, ST x) {
    const T v = convert<T>(x);
    if (lower) {
      if (a >= b) S[b * LS + a] = v;
    } else {
      // A(a, b) with a <= b is upper; it is conj of lower-form (b, a)
      if (a <= b) S[a * LS + b] = conjx(v);
    }
  });
}

// Store the lower form back into the caller's triangle (diagonal included
// unless skip_diag).
template <typename T, int NMAX, int THREADS, typename ST = T>
__device__ __forceinline__ void store_lower_form(const T *S, bool lower, int n, ST *A, long lda, bool skip_diag) {
  constexpr int LS = Lds<T, NMAX>::LS;
  for_block<NMAX, THREADS>(n, [&](int a, int b) {
    if (skip_diag && a == b) return;
    if (lower) {
      if (a >= b) A[(long)b * lda + a] = convert<ST>(S[b * LS + a]);
    } else {
      if (a <= b) A[(long)b * lda + a] = convert<ST>(conjx(S[a * LS + b]));
    }
  });
}

// W(r, c) = S(r, c) for r >= c, zero above; unit -> ones on the diagonal.
// With upper, W is the conjugate transpose instead (upper triangular).
template <typename T, int NMAX, int THREADS, typename ST = T>
__device__ __forceinline__ void store_dense(const T *S, bool upper, int n, bool unit, ST *W, long ldw) {
  constexpr int LS = Lds<T, NMAX>::LS;
  for_block<NMAX, THREADS>(n, [&](int r, int c) {
    T v = from_real<T>(0.0);
    if (r == c) v = unit ? from_real<T>(1.0) : (upper ? conjx(S[c * LS + r]) : S[c * LS + r]);
    else if (!upper && r > c) v = S[c * LS + r];
    else if (upper && r < c) v = conjx(S[r * LS + c]);
    W[(long)c * ldw + r] = convert<ST>(v);
  });
}

// In-place R = L^H L of the lower form (lower part), columns left to right:
// R(r, c) = sum_{k >= r} conj(L(k, r)) L(k, c).
template <typename T, int NMAX>
__device__ __forceinline__ void lds_lauum(T *S, int n) {
  constexpr int LS = Lds<T, NMAX>::LS, TPR = Lds<T, NMAX>::TPR;
  const int tid = threadIdx.x, row = tid / TPR, part = tid % TPR;
  for (int c = 0; c < n; c++) {
    const int r = c + row;
    T sum = from_real<T>(0.0);
    if (r < n)
      for (int k = r + part; k < n; k += TPR) sum = sum + conjx(S[r * LS + k]) * S[c * LS + k];
    sum = group_sum<TPR>(sum);
    // the diagonal of L^H L is real; fused rounding of conj(a) a may leave an
    // imaginary part of order eps |a|^2, which a Hermitian result must not carry
    if (r == c) sum = from_real<T>(realx(sum));
    __syncthreads();
    if (r < n && part == 0) S[c * LS + r] = sum;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Register-resident factor and inverse.  Each of ten waves owns one region
// of the lower triangle (a 4 x 4 grid of NMAX/4-square regions); each lane
// owns an RB x CB block of its wave's region in registers.  A column step
// needs one broadcast vector, published through a double-buffered LDS array,
// and one barrier; a wave whose region the step cannot change skips it, so
// the work per step follows the shrinking trailing triangle.
// ---------------------------------------------------------------------------
template <typename T, int NMAX>
struct Reg {
  static constexpr int RS = NMAX / 4;          // region side (one wave)
  static constexpr int RB = RS / 8, CB = RS / 8;  // lane block: 8 x 8 lanes per region
  // region (I, J), I >= J, of wave w in the order (0,0) (1,0) (1,1) (2,0) ...
  __device__ static int reg_i() {
    const int w = threadIdx.x >> 6;
    return w >= 6 ? 3 : (w >= 3 ? 2 : (w >= 1 ? 1 : 0));
  }
  __device__ static int reg_j() {
    const int i = reg_i();
    return (threadIdx.x >> 6) - i * (i + 1) / 2;
  }
  __device__ static int r0() { return reg_i() * RS + ((threadIdx.x & 63) >> 3) * RB; }
  __device__ static int c0() { return reg_j() * RS + (threadIdx.x & 7) * CB; }
};
template <typename T, int NMAX>
using RegTile = T[NMAX / 32][NMAX / 32];

// Lane-wise select written per component, so a complex value never becomes
// an aggregate select (which the compiler lowers through private memory).
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ zd sel(bool c, zd a, zd b) { return make_zd(c ? a.x : b.x, c ? a.y : b.y); }

// X = inv(L) of the lower form L held in LDS (S, leading dimension NMAX + 1),
// forward substitution by rows: row k of X is final once the updates of rows
// 0..k-1 have been applied; it is published through xrow[2][NMAX] and the
// rows below take its contribution.  dinv[k] = 1 / L(k, k) (unused for a unit
// diagonal).  The strict upper part of X is exactly zero throughout (it starts
// at zero and only ever gains products with published zeros), so neither the
// finalisation nor the update needs a column condition.  On return x holds
// the thread's block of X.
template <typename T, int NMAX>
__device__ __forceinline__ void reg_trtri(const T *S, RegTile<T, NMAX> &x, T *xrow, const T *dinv, int n, bool unit) {
  constexpr int LS = NMAX + 1, RB = Reg<T, NMAX>::RB, CB = Reg<T, NMAX>::CB;
  const int R0 = Reg<T, NMAX>::r0(), C0 = Reg<T, NMAX>::c0();
#pragma unroll
  for (int ri = 0; ri < RB; ri++)
#pragma unroll
    for (int ci = 0; ci < CB; ci++) x[ri][ci] = from_real<T>(0.0);
  const T one = from_real<T>(1.0), zero = from_real<T>(0.0);
  for (int k = 0; k < n; k++) {
    if (k >= R0 && k < R0 + RB) {  // this thread holds part of row k: finalise it
      const T dk = unit ? one : dinv[k];
      const int kr = k - R0;
      T v[CB];
#pragma unroll
      for (int ci = 0; ci < CB; ci++) {
        const T delta = sel(C0 + ci == k, one, zero);
        v[ci] = zero;
#pragma unroll
        for (int ri = 0; ri < RB; ri++) {
          const T fin = (delta - x[ri][ci]) * dk;
          x[ri][ci] = sel(ri == kr, fin, x[ri][ci]);
          v[ci] = sel(ri == kr, fin, v[ci]);
        }
        xrow[(k & 1) * NMAX + C0 + ci] = v[ci];
      }
    }
    __syncthreads();
    if ((Reg<T, NMAX>::reg_i() + 1) * Reg<T, NMAX>::RS > k + 1 &&
        Reg<T, NMAX>::reg_j() * Reg<T, NMAX>::RS <= k) {  // regions with rows below k, columns <= k
      const T *xr = xrow + (k & 1) * NMAX;
      T lk[RB], xk[CB];
#pragma unroll
      for (int ri = 0; ri < RB; ri++) lk[ri] = S[k * LS + R0 + ri];
#pragma unroll
      for (int ci = 0; ci < CB; ci++) xk[ci] = xr[C0 + ci];
#pragma unroll
      for (int ri = 0; ri < RB; ri++) {
        const bool below = R0 + ri > k;
#pragma unroll
        for (int ci = 0; ci < CB; ci++) {
          const T upd = x[ri][ci] + lk[ri] * xk[ci];
          x[ri][ci] = sel(below, upd, x[ri][ci]);
        }
      }
    }
  }
}

// Registers <-> LDS copies of the thread's block (lower part, inside n x n).
template <typename T, int NMAX>
__device__ __forceinline__ void reg_to_lds(const RegTile<T, NMAX> &v, T *S, int n) {
  constexpr int LS = NMAX + 1, RB = Reg<T, NMAX>::RB, CB = Reg<T, NMAX>::CB;
  const int R0 = Reg<T, NMAX>::r0(), C0 = Reg<T, NMAX>::c0();
#pragma unroll
  for (int ri = 0; ri < RB; ri++)
#pragma unroll
    for (int ci = 0; ci < CB; ci++) {
      const int r = R0 + ri, c = C0 + ci;
      if (r < n && c < n && r >= c) S[c * LS + r] = v[ri][ci];
    }
}
template <typename T, int NMAX>
__device__ __forceinline__ void lds_to_reg(const T *S, RegTile<T, NMAX> &v, int n) {
  constexpr int LS = NMAX + 1, RB = Reg<T, NMAX>::RB, CB = Reg<T, NMAX>::CB;
  const int R0 = Reg<T, NMAX>::r0(), C0 = Reg<T, NMAX>::c0();
#pragma unroll
  for (int ri = 0; ri < RB; ri++)
#pragma unroll
    for (int ci = 0; ci < CB; ci++) {
      const int r = R0 + ri, c = C0 + ci;
      v[ri][ci] = (r < n && c < n && r >= c) ? S[c * LS + r] : from_real<T>(0.0);
    }
}

// dinv[k] = 1 / S(k, k) for k < n.
template <typename T, int NMAX>
__device__ __forceinline__ void diag_reciprocals(const T *S, T *dinv, int n) {
  constexpr int LS = NMAX + 1;
  for (int k = threadIdx.x; k < n; k += blockDim.x) dinv[k] = from_real<T>(1.0) / S[k * LS + k];
}

// ---------------------------------------------------------------------------
// Blocked factor + inverse of one diagonal block (diag_potrf).
//
// The block is split into 16-column leaves.  Leaf p is factored and inverted
// by one wave entirely in registers (lane r + 16 q holds row r, columns
// q + 4 j) with cross-lane shuffles, so its 16 sequential pivots need no
// workgroup barrier.  Everything else is 16 x 16 x 16 tile products on the
// matrix cores spread over the eight waves: the panel below the leaf
// (L_Ip = A_Ip inv(L_pp)^H), the trailing update (A_IJ -= L_Ip L_Jp^H), and
// the inverse X = inv(L) row block by row block as the factor's rows
// complete (X_IJ = -X_II sum_{K=J..I-1} L_IK X_KJ, from L X = I), formed by
// the waves that are not factoring the next leaf.  Two barriers per leaf.
// ---------------------------------------------------------------------------
using t16::LEAF;
using t16::TileAcc;
using t16::acc_zero;
using t16::acc_load;
using t16::acc_store;
using t16::tile_mma;
using t16::wave_lds_fence;

constexpr int BLK_THREADS = 512;

// One wave factors the 16 x 16 leaf at S (leading dimension ls), writes the
// factor straight to its place in the global matrix (rows and columns
// g0 .. g0 + 15 of the n x n block at A, lower form: A(r, c) when lower,
// conj into A(c, r) when upper; padding beyond n is not written) and
// overwrites the leaf with its inverse (lower, zeros above).  Returns the
// first non-positive / NaN pivot (0-based) or -1; on failure nothing is
// written.
template <typename T, typename ST>
__device__ __forceinline__ int leaf_factor_inverse(T *S, int ls, bool lower, int n, int g0, ST *A,
                                                long lda) {
  int l = threadIdx.x & 63;
  asm volatile("" : "+v"(l));  // lane-derived values computed here, not kept live across the kernel's loop
  const int r = l & 15, q = l >> 4;
  T a[LEAF];  // row r (row layout: the four lane groups hold the same row)
#pragma unroll
  for (int c = 0; c < LEAF; c++) a[c] = S[c * ls + r];
  T x[4];
  int bad;
  if constexpr (is_complex<T>::value) {
    // complex: factor and invert one after the other (the fused pass and the
    // inverse-row accumulator of the caller do not fit the registers together)
    double unused = 0.0;
    bad = t16::leaf_factor_rows<T>(a, 0, &unused);
    if (bad < 0) t16::leaf_inverse_cols<T>(a, x);
  } else {
    double rpiv;
    bad = t16::leaf_factor_invert<T>(a, x, rpiv);
  }
  if (bad >= 0) return bad;
  // lane group q writes columns q + 4 j of its row
  auto pick = [&](const T *v, int j) {
    return sel(q == 0, v[4 * j], sel(q == 1, v[4 * j + 1], sel(q == 2, v[4 * j + 2], v[4 * j + 3])));
  };
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int c = q + 4 * j, gr = g0 + r, gc = g0 + c;
    const T v = pick(a, j);
    if (r >= c && gr < n) {
      if (lower) A[(long)gc * lda + gr] = convert<ST>(v);
      else A[(long)gr * lda + gc] = convert<ST>(conjx(v));
    }
  }
#pragma unroll
  for (int j = 0; j < 4; j++) S[(q + 4 * j) * ls + r] = x[j];
  return -1;
}

// LDS holds only the lower-triangle tiles, packed: tile (I, J), I >= J, at
// index I (I + 1) / 2 + J, each 16 x 16 column-major with leading dimension
// 17 (78 KB for a 128 x 128 real block), so the kernel fits beside a
// trailing-update workgroup on the same CU.  Once leaf p is factored its tile
// holds inv(L_pp); its factor has already gone to global memory.
// T is the compute type (the LDS copy), ST the storage type of A and W
// (float for the mixed-precision factorisation, which computes the diagonal
// blocks in double).
// One workgroup per launch, run beside the trailing update's workgroups
// (256 registers per wave, up to two per SIMD): with at most 128 registers per
// wave (WPE = 4 waves per SIMD) the real-type kernel fits next to one of them
// and starts at once; the complex one needs more registers and is left free.
template <typename T>
constexpr int diag_waves_per_eu() { return is_complex<T>::value ? 2 : 4; }

// Factor and invert the lower-form block held in S (the packed tiles above,
// already loaded, identity padding beyond n; `flag` an LDS word).  The factor
// goes to A (lower form: A(r, c), or conj into A(c, r) when !lower) and
// inv(L) densely to W (ld ldw, zeros above the diagonal, n x n only).
// Returns -1, or the first non-positive / NaN pivot (0-based, the same value
// in every thread; the factor is then only partly written).
// EARLY: wave 0 has written tile (0, 0) itself and set flag = -1, and the
// other waves may still be filling the rest of S (leaf 0 overlaps their loads;
// the barrier after it completes S).
template <typename T, typename ST, int NMAX, bool EARLY = false>
__device__ __forceinline__ int diag_block_factor(T *S, int &flag, int lower, int n, ST *A, long lda, ST *W,
                                                 long ldw) {
  constexpr int NPMAX = NMAX / LEAF, LL = LEAF + 1, TS = LEAF * LL, NW = BLK_THREADS / 64;
  static_assert(NPMAX <= NW, "each wave finishes at most one tile of a row of the inverse");
  const int tid = threadIdx.x, wave = tid >> 6;
  const int np = (n + LEAF - 1) / LEAF;
  auto tile = [&](int I, int J) { return S + (I * (I + 1) / 2 + J) * TS; };
  auto elem = [&](int r, int c) -> T & { return tile(r / LEAF, c / LEAF)[(c % LEAF) * LL + r % LEAF]; };
  if constexpr (!EARLY) {
    if (tid == 0) flag = -1;
    __syncthreads();
  }
  // One leaf of look-ahead: while wave 0 updates tile (p+1, p+1) and factors
  // leaf p + 1, waves 1 .. NW-1 apply the rest of step p's trailing update
  // and form S_{p+1,J} = sum_{K=J..p} L_{p+1,K} X_KJ (wave 1 + J) for row
  // p + 1 of X = inv(L).  At the start of step p + 1 (phase A) the same wave
  // writes L_{p+1,J} to A and finishes X_{p+1,J} = -X_{p+1,p+1} S_{p+1,J} in
  // its tile (L X = I row by row), beside the panel solves below leaf p + 1:
  // the inverse is complete when the last leaf is, with no pass of its own.
  // Step p = -1 only factors leaf 0 (one call site keeps the leaf inlined once).
  TileAcc<T> sacc;  // S_{p+1, wave-1}, carried from phase B to the next phase A
  acc_zero(sacc);
  for (int p = -1; p < np; p++) {
    if (p >= 0) {
      // phase A: panel item k (L_Ip = A_Ip X_pp^H, I = p + 1 + k) on wave 0
      // (k = 0) or wave p + k; row p of the inverse on waves 1 .. p
      const int k = wave == 0 ? 0 : (wave > p ? wave - p : -1);
      if (k >= 0 && p + 1 + k < np) {
        const int I = p + 1 + k;
        TileAcc<T> acc;
        acc_zero(acc);
        tile_mma<true>(acc, tile(I, p), LL, tile(p, p), LL, 1.0);
        acc_store(acc, tile(I, p), LL);
      }
      if (wave >= 1 && wave <= p) {
        const int J = wave - 1;
        T *tpj = tile(p, J);
        {  // L_pJ to its place in A (lower form; padding rows are not written)
          const int l = tid & 63, rr = l & 15, cq = l >> 4, r = p * LEAF + rr;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int c = J * LEAF + cq + 4 * j;
            const T v = tpj[(cq + 4 * j) * LL + rr];
            if (r < n) {
              if (lower) A[(long)c * lda + r] = convert<ST>(v);
              else A[(long)r * lda + c] = convert<ST>(conjx(v));
            }
          }
        }
        wave_lds_fence();
        acc_store(sacc, tpj, LL);
        wave_lds_fence();
        TileAcc<T> acc;
        acc_zero(acc);
        tile_mma<false>(acc, tile(p, p), LL, tpj, LL, -1.0);
        wave_lds_fence();
        acc_store(acc, tpj, LL);
      }
      __syncthreads();
    }
    if (p + 1 < np) {
      if (wave == 0) {  // A_{p+1,p+1} -= L_{p+1,p} L_{p+1,p}^H, then leaf p + 1
        if (p >= 0) {
          TileAcc<T> acc;
          acc_load(acc, tile(p + 1, p + 1), LL);
          tile_mma<true>(acc, tile(p + 1, p), LL, tile(p + 1, p), LL, -1.0);
          acc_store(acc, tile(p + 1, p + 1), LL);
          wave_lds_fence();
        }
        const int bad = leaf_factor_inverse<T, ST>(tile(p + 1, p + 1), LL, lower, n, (p + 1) * LEAF, A, lda);
        if (bad >= 0 && (tid & 63) == 0) flag = (p + 1) * LEAF + bad;
      } else if (p >= 0) {
        int t = 0;
        for (int J = p + 1; J < np; J++)
          for (int I = (J == p + 1) ? J + 1 : J; I < np; I++, t++)
            if (t % (NW - 1) == wave - 1) {  // A_IJ -= L_Ip L_Jp^H
              TileAcc<T> acc;
              acc_load(acc, tile(I, J), LL);
              tile_mma<true>(acc, tile(I, p), LL, tile(J, p), LL, -1.0);
              acc_store(acc, tile(I, J), LL);
            }
        if (wave - 1 <= p) {  // S_{p+1,J}, J = wave - 1: row p + 1 of L is complete
          const int J = wave - 1;
          acc_zero(sacc);
          for (int K = J; K <= p; K++) tile_mma<false>(sacc, tile(p + 1, K), LL, tile(K, J), LL, 1.0);
        }
      }
    }
    __syncthreads();
    const int f = flag;
    if (f >= 0) return f;  // leaf p + 1 failed (padding never fails, so f < n)
  }
  // W = inv(L) dense, zeros above the diagonal
  for_block<NMAX, BLK_THREADS>(n, [&](int r, int c) {
    W[(long)c * ldw + r] = convert<ST>(r >= c ? elem(r, c) : from_real<T>(0.0));
  });
  return -1;
}

// Bytes of the packed lower-triangle tiles of a diagonal block of order NMAX.
template <typename T, int NMAX>
constexpr int diag_tiles_bytes() {
  return (NMAX / LEAF) * (NMAX / LEAF + 1) / 2 * LEAF * (LEAF + 1) * (int)sizeof(T);
}

template <typename T, typename ST, int NMAX>
__global__ void __launch_bounds__(BLK_THREADS)
__attribute__((amdgpu_waves_per_eu(diag_waves_per_eu<T>(), diag_waves_per_eu<T>())))
diag_potrf_kernel(int lower, int n, ST *A, long lda, ST *W, long ldw, int *info, long info_off) {
  if (*info != 0) return;
  // This block is on the critical path of the factorisation while trailing-
  // update waves share its SIMDs: take instruction issue ahead of them.
  __builtin_amdgcn_s_setprio(3);
  constexpr int NPMAX = NMAX / LEAF, LL = LEAF + 1, TS = LEAF * LL;
  __shared__ T S[NPMAX * (NPMAX + 1) / 2 * TS];
  __shared__ int flag;
  const int tid = threadIdx.x;
  const int np = (n + LEAF - 1) / LEAF, npad = np * LEAF;
  auto tile = [&](int I, int J) { return S + (I * (I + 1) / 2 + J) * TS; };
  auto elem = [&](int r, int c) -> T & { return tile(r / LEAF, c / LEAF)[(c % LEAF) * LL + r % LEAF]; };
  if (tid < 64) {
    // wave 0: tile (0, 0) only (lower form, identity padding), then leaf 0
    // at once while the other waves load the rest of the block
    const int rr = tid & 15, cq = tid >> 4, rc = min(rr, n - 1);
    T v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int cc = min(cq + 4 * j, n - 1);
      const ST x = lower ? A[(long)cc * lda + rc] : A[(long)rc * lda + cc];
      v[j] = lower ? convert<T>(x) : conjx(convert<T>(x));
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int c = cq + 4 * j;
      if (rr >= c) tile(0, 0)[c * LL + rr] = (rr < n && c < n) ? v[j] : from_real<T>(rr == c ? 1.0 : 0.0);
    }
    if (tid == 0) flag = -1;
    wave_lds_fence();
  } else {
    // waves 1 .. 7: identity padding and the lower form of every other tile,
    // read along memory's contiguous index (i), every load of a thread in
    // flight at once (clamped, always valid addresses)
    constexpr int T2 = BLK_THREADS - 64, U = (NMAX * NMAX + T2 - 1) / T2;
    const int t2 = tid - 64;
    for (int c = t2; c < npad; c += T2)
      for (int r = max(n, c); r < npad; r++)
        if (r >= LEAF || c >= LEAF) elem(r, c) = from_real<T>(r == c ? 1.0 : 0.0);
    // in batches that fit the registers (one for 4-byte data, two for 8-byte)
    constexpr int NB = sizeof(ST) > 4 ? 2 : 1, UB = (U + NB - 1) / NB;
#pragma unroll
    for (int b = 0; b < NB; b++) {
      ST x[UB];
#pragma unroll
      for (int u = 0; u < UB; u++) {
        const int e = min(t2 + T2 * (b * UB + u), NMAX * NMAX - 1), i = e % NMAX, k = e / NMAX;
        x[u] = A[(long)min(k, n - 1) * lda + min(i, n - 1)];
      }
#pragma unroll
      for (int u = 0; u < UB; u++) {
        const int e = t2 + T2 * (b * UB + u), i = e % NMAX, k = e / NMAX;
        const int r = lower ? i : k, c = lower ? k : i;  // lower form, r >= c wanted
        if (b * UB + u < U && e < NMAX * NMAX && i < n && k < n && r >= c && (r >= LEAF || c >= LEAF))
          elem(r, c) = lower ? convert<T>(x[u]) : conjx(convert<T>(x[u]));
      }
    }
  }
  const int bad = diag_block_factor<T, ST, NMAX, true>(S, flag, lower, n, A, lda, W, ldw);
  if (bad >= 0 && tid == 0) *info = (int)(info_off + bad + 1);
}

// ---------------------------------------------------------------------------
// A whole panel of the blocked factorisation in one launch (lower, real):
// the m x jb panel at A, its top jb x jb block on the diagonal, becomes L,
// left-looking by 128-column blocks J = 0 .. C-1 (C = ceil(jb / 128)).
// The reference factors a panel as a chain of diagonal-block, panel-solve
// and panel-update calls (lapack/dpotrf.c:268-326); here one workgroup owns
// each 128-row block: row block b < C is the diagonal part (rows 128 b ..),
// row block b >= C the rows jb + 128 (b - C) .. below it.  For every column
// block J left of its rows, workgroup b
//   * once row block J has published its off-diagonal blocks L_Jk (k < J),
//     accumulates S = sum_{k<J} L_bk L_Jk^T (K = 128 J, its own L_bk
//     re-read from memory);
//   * once L_JJ and inv(L_JJ) are published, stores X = A_bJ - S in place and
//     then L_bJ = X inv(L_JJ)^T.
// A diagonal workgroup b then publishes its off-diagonal blocks, factors and
// inverts A_bb - sum_k L_bk L_bk^T with diag_block_factor (FP64 for FP32
// storage; the partial sum over k < b - 1 is formed ahead of the wait for
// block b - 1 and parked in the workspace) and publishes L_bb and inv(L_bb).
// The critical path per column block is one flag hand-off, two 128^3
// products and the diagonal block; the other workgroups' accumulations run
// while a diagonal block is factored.
//
// Dependencies only point to lower workgroup indices, which the hardware
// dispatches first, so every awaited workgroup is resident or finished.
// Flags are set to the launch's epoch (a per-handle counter, so they never
// need resetting); a wait gives up when the info word becomes nonzero (a
// failed pivot anywhere) or after about four seconds, which it reports as
// kInfoAbandoned.  Publishing: every wave waits for its own stores, then one
// wave writes the L2 back once and sets the flag; waiting: one wave polls,
// then acquires with one agent-scope fence before the workgroup barrier.
// ---------------------------------------------------------------------------
typedef float pv4f __attribute__((ext_vector_type(4)));
typedef double pv4d __attribute__((ext_vector_type(4)));
typedef double pv2d __attribute__((ext_vector_type(2)));

constexpr int PK_THREADS = 512;
constexpr int PK_B = 128;          // rows and columns of a block
constexpr int PK_FLAG_STRIDE = 32; // ints between flags (one 128-byte line each)

// Matrix-core arithmetic of the panel products per storage type: K step KT
// staged in LDS as rows of LK elements (K contiguous), a 16-byte fragment of
// FV consecutive k per lane, and the result column of register r in lane
// group g (the row is lane % 16).
template <typename ST> struct PanelMma;
template <> struct PanelMma<float> {
  using V = pv4f;
  using F = pv4f;
  static constexpr int KT = 32, LK = KT + 4, FV = 4;
  __device__ static V mma(float b, float a, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c, 0, 0, 0); }
  __device__ static int col(int g, int r) { return 4 * g + r; }
};
template <> struct PanelMma<double> {
  using V = pv4d;
  using F = pv2d;
  static constexpr int KT = 16, LK = KT + 2, FV = 2;
  __device__ static V mma(double b, double a, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c, 0, 0, 0); }
  __device__ static int col(int g, int r) { return g + 4 * r; }
};

// 128 x 128 accumulator of the workgroup: 8 waves in 2 x 4, each 64 x 32 as
// 4 x 2 blocks of 16 x 16.
template <typename ST> struct PanelAcc {
  typename PanelMma<ST>::V v[4][2];
  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) v[i][j][r] = ST(0);
  }
  // element <- fn(row, col, element) over the thread's 32 elements
  template <typename Fn> __device__ void map(Fn fn) {
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // indices formed per call, not hoisted by the caller
    const int lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lg = lane >> 4, wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 32;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) v[i][j][r] = fn(wm0 + 16 * i + lc, wn0 + 16 * j + PanelMma<ST>::col(lg, r), v[i][j][r]);
  }
  // fn(row, col, C(row, col) - element) where ok(row, col), C(r, c) at
  // Cp[c * ldc + r]; C is read in two halves with all loads of a half in
  // flight before its results are used (fn may store into C)
  template <typename Ok, typename Fn> __device__ void sub_from(const ST *Cp, long ldc, Ok ok, Fn fn) {
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lg = lane >> 4, wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 32;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      ST cv[16];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int row = wm0 + 16 * (2 * h + i) + lc, col = wn0 + 16 * j + PanelMma<ST>::col(lg, r);
            cv[8 * i + 4 * j + r] = Cp[ok(row, col) ? (long)col * ldc + row : 0];
          }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int row = wm0 + 16 * (2 * h + i) + lc, col = wn0 + 16 * j + PanelMma<ST>::col(lg, r);
            fn(row, col, cv[8 * i + 4 * j + r] - v[2 * h + i][j][r]);
          }
    }
  }
};

// acc += P Q^T over k < K: P(r, k) = Pp[k * ldp + r] (r < rp), Q(c, k) =
// Qp[k * ldq + c] (c < rq); k >= klim reads as zero.  K is a multiple of KT.
// Both operands are staged through LDS (double-buffered, the next step's
// global loads in registers during this step's products); returns with the
// LDS free (after a barrier).
template <typename ST>
__device__ __forceinline__ void panel_nt(PanelAcc<ST> &acc, const ST *Pp, long ldp, int rp, const ST *Qp, long ldq,
                                         int rq, int K, int klim, ST *lds) {
  using M = PanelMma<ST>;
  constexpr int KT = M::KT, LK = M::LK, FV = M::FV, E = KT / 4;
  using F = typename M::F;
  ST *sP = lds, *sQ = lds + 2 * PK_B * LK;
  int tid = threadIdx.x;
  // thread-derived values (and the addresses built from them) are formed here
  // at every call, not hoisted out of the caller's loop and kept live across it
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = tid & (PK_B - 1), kq = tid >> 7;  // loader: row lr, k = kq E .. kq E + E - 1
  const int lc = lane & 15, lg = lane >> 4, wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 32;
  ST rp_[E], rq_[E];
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int k = k0 + kq * E + e;
      const bool okp = lr < rp && k < klim, okq = lr < rq && k < klim;
      const ST a = Pp[okp ? (long)k * ldp + lr : 0];
      const ST b = Qp[okq ? (long)k * ldq + lr : 0];
      rp_[e] = okp ? a : ST(0);
      rq_[e] = okq ? b : ST(0);
    }
  };
  auto store = [&](int buf) {
    ST *dp = sP + buf * PK_B * LK + lr * LK + kq * E, *dq = sQ + buf * PK_B * LK + lr * LK + kq * E;
#pragma unroll
    for (int e = 0; e < E; e += FV) {
      F a, b;
#pragma unroll
      for (int t = 0; t < FV; t++) { a[t] = rp_[e + t]; b[t] = rq_[e + t]; }
      *reinterpret_cast<F *>(dp + e) = a;
      *reinterpret_cast<F *>(dq + e) = b;
    }
  };
  const int nk = K / KT;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const bool more = kt + 1 < nk;
    if (more) load((kt + 1) * KT);
    const ST *cP = sP + (kt & 1) * PK_B * LK, *cQ = sQ + (kt & 1) * PK_B * LK;
#pragma unroll
    for (int h = 0; h < KT; h += 4 * FV) {
      F fp[4], fq[2];
#pragma unroll
      for (int i = 0; i < 4; i++) fp[i] = *reinterpret_cast<const F *>(cP + (wm0 + 16 * i + lc) * LK + h + FV * lg);
#pragma unroll
      for (int j = 0; j < 2; j++) fq[j] = *reinterpret_cast<const F *>(cQ + (wn0 + 16 * j + lc) * LK + h + FV * lg);
#pragma unroll
      for (int t = 0; t < FV; t++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc.v[i][j] = M::mma(fq[j][t], fp[i][t], acc.v[i][j]);
    }
    if (more) store((kt + 1) & 1);
    __syncthreads();
  }
}

// One wave polls flag f until it holds `epoch` (or info becomes nonzero, or
// the wait times out), then acquires; the result (true: proceed) is the
// same in every thread.
__device__ __forceinline__ bool panel_wait(const int *f, int epoch, int *info, int &sh) {
  if (threadIdx.x < 64) {
    int st = 0;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
      const int v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      const int iv = __builtin_amdgcn_readfirstlane(__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (iv != 0) { st = 2; break; }
      if (v == epoch) { st = 1; break; }
      if (__builtin_amdgcn_s_memrealtime() - t0 > 400000000ull) {  // ~4 s at 100 MHz
        if (threadIdx.x == 0) {
          int expected = 0;
          __hip_atomic_compare_exchange_strong(info, &expected, kInfoAbandoned, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
        }
        st = 2;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (threadIdx.x == 0) sh = st;
  }
  __syncthreads();
  return sh == 1;
}

__device__ __forceinline__ void panel_publish(int *f, int epoch) {
  publish_wait_own();
  __syncthreads();
  if (threadIdx.x < 64) {
    publish_release();
    if (threadIdx.x == 0) __hip_atomic_store(f, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename ST>
constexpr int panel_lds_bytes() {
  constexpr int g = 4 * PK_B * PanelMma<ST>::LK * (int)sizeof(ST);
  constexpr int d = diag_tiles_bytes<double, PK_B>();
  return g > d ? g : d;
}

// ws: C inverses inv(L_JJ) (128 x 128, ld 128), then C partial sums (128 x 128).
template <typename ST>
__global__ void __launch_bounds__(PK_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4)))
panel_potrf_kernel(int m, int jb, ST *A, long lda, ST *ws, int *flags, int epoch, int *info, long info_off) {
  using T = double;  // the diagonal blocks are factored in FP64 for both storage types
  if (*info != 0) return;
  __builtin_amdgcn_s_setprio(3);
  __shared__ __attribute__((aligned(16))) char lds_raw[panel_lds_bytes<ST>()];
  __shared__ int sh_wait, sh_flag;
  ST *lds = reinterpret_cast<ST *>(lds_raw);
  const int C = (jb + PK_B - 1) / PK_B, b = blockIdx.x;
  const bool dg = b < C;
  const int r0 = dg ? PK_B * b : jb + PK_B * (b - C);
  const int ra = dg ? min(PK_B, jb - r0) : min(PK_B, m - r0);
  const int nJ = dg ? b : C;
  constexpr long BB = (long)PK_B * PK_B;
  ST *Winv = ws, *Psum = ws + C * BB;
  int *const off_flag = flags, *const diag_flag = flags + PK_FLAG_STRIDE;
  auto fo = [&](int J) { return off_flag + 2 * PK_FLAG_STRIDE * J; };
  auto fd = [&](int J) { return diag_flag + 2 * PK_FLAG_STRIDE * J; };
  PanelAcc<ST> acc;
  for (int J = 0; J < nJ; J++) {
    const int cw = min(PK_B, jb - PK_B * J);
    ST *AbJ = A + (long)PK_B * J * lda + r0;
    if (dg && J == b - 1 && b >= 2) {  // P = sum_{k < b-1} L_bk L_bk^T (own rows, no wait)
      acc.zero();
      panel_nt<ST>(acc, A + r0, lda, ra, A + r0, lda, ra, PK_B * J, PK_B * J, lds);
      ST *P = Psum + b * BB;
      acc.map([&](int r, int c, ST x) { P[(long)c * PK_B + r] = x; return x; });
    }
    acc.zero();
    if (J > 0) {
      if (!panel_wait(fo(J), epoch, info, sh_wait)) return;
      panel_nt<ST>(acc, A + r0, lda, ra, A + PK_B * J, lda, cw, PK_B * J, PK_B * J, lds);
    }
    if (!panel_wait(fd(J), epoch, info, sh_wait)) return;
    // X = A_bJ - S in place (C and the sum kept apart: they nearly cancel)
    acc.sub_from(AbJ, lda, [&](int r, int c) { return r < ra && c < cw; }, [&](int r, int c, ST v) {
      if (r < ra && c < cw) AbJ[(long)c * lda + r] = v;
    });
    publish_wait_own();
    __syncthreads();
    // L_bJ = X inv(L_JJ)^T
    acc.zero();
    panel_nt<ST>(acc, AbJ, lda, ra, Winv + J * BB, PK_B, cw, PK_B, cw, lds);
    acc.map([&](int r, int c, ST x) {
      if (r < ra && c < cw) AbJ[(long)c * lda + r] = x;
      return x;
    });
    publish_wait_own();
    __syncthreads();
  }
  if (!dg) return;
  // the diagonal workgroup's row block left of the diagonal is final
  if (b > 0) panel_publish(fo(b), epoch);
  // A_bb - sum_{k<b} L_bk L_bk^T into the packed FP64 tiles
  ST *Abb = A + (long)r0 * lda + r0;
  acc.zero();
  if (b >= 2) {
    const ST *P = Psum + b * BB;
    acc.map([&](int r, int c, ST) { return P[(long)c * PK_B + r]; });
  }
  if (b >= 1) {
    const ST *Lp = A + (long)PK_B * (b - 1) * lda + r0;
    panel_nt<ST>(acc, Lp, lda, ra, Lp, lda, ra, PK_B, PK_B, lds);
  }
  {
    constexpr int LL = LEAF + 1, TS = LEAF * LL;
    T *S = reinterpret_cast<T *>(lds_raw);
    const int npad = (ra + LEAF - 1) / LEAF * LEAF;
    acc.sub_from(Abb, lda, [&](int r, int c) { return r < ra && c < ra && r >= c; }, [&](int r, int c, ST v) {
      if (r >= c && r < npad && c < npad) {
        T *tl = S + ((r / LEAF) * (r / LEAF + 1) / 2 + c / LEAF) * TS;
        tl[(c % LEAF) * LL + r % LEAF] = (r < ra && c < ra) ? convert<T>(v) : from_real<T>(r == c ? 1.0 : 0.0);
      }
    });
  }
  const int bad = diag_block_factor<T, ST, PK_B>(reinterpret_cast<T *>(lds_raw), sh_flag, 1, ra, Abb, lda,
                                                 Winv + b * BB, PK_B);
  if (bad >= 0) {
    if (threadIdx.x == 0) {
      int expected = 0;
      __hip_atomic_compare_exchange_strong(info, &expected, (int)(info_off + r0 + bad + 1), __ATOMIC_RELAXED,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  panel_publish(fd(b), epoch);
}

// Workgroup b inverts the diagonal block of order min(nb, ntot - b nb) at
// (b nb, b nb) of A into the same place of B (one workgroup: nb = ntot).
// T is the compute type, ST the storage type of A, B and W.
template <typename T, typename ST, int NMAX>
__global__ void __launch_bounds__(DIAG_THREADS)
diag_trtri_kernel(int lower, int unit, int ntot, int nb, const ST *A, long lda, ST *B, long ldb, ST *W,
                  long ldw, const int *abort_flag) {
  if (abort_flag != nullptr && *abort_flag != 0) return;
  const long off = (long)blockIdx.x * nb;
  const int n = (int)min((long)nb, (long)ntot - off);
  A += off * lda + off;
  B += off * ldb + off;
  constexpr int RB = Reg<T, NMAX>::RB, CB = Reg<T, NMAX>::CB;
  __shared__ T S[NMAX * (NMAX + 1)];
  __shared__ T vec[2 * NMAX];
  __shared__ T dinv[NMAX];
  load_lower_form<T, NMAX, DIAG_THREADS, ST>(S, lower, n, A, lda);
  __syncthreads();
  if (!unit) diag_reciprocals<T, NMAX>(S, dinv, n);
  __syncthreads();
  T x[RB][CB];
  reg_trtri<T, NMAX>(S, x, vec, dinv, n, unit != 0);
  __syncthreads();
  reg_to_lds<T, NMAX>(x, S, n);
  __syncthreads();
  store_lower_form<T, NMAX, DIAG_THREADS, ST>(S, lower, n, B, ldb, unit != 0);
  if (W != nullptr) store_dense<T, NMAX, DIAG_THREADS, ST>(S, false, n, unit != 0, W, ldw);
}

// Workgroup b squares the diagonal block of order min(nb, ntot - b nb) at
// (b nb, b nb) of A into B + b bstride (ld ldb); one workgroup in place:
// nb = ntot, B = A, ldb = lda.
template <typename T, typename ST, int NMAX>
__global__ void __launch_bounds__(LAUUM_THREADS)
diag_lauum_kernel(int lower, int ntot, int nb, const ST *A, long lda, ST *B, long ldb, long bstride) {
  __shared__ T S[NMAX * (NMAX + 1)];
  const long off = (long)blockIdx.x * nb;
  const int n = (int)min((long)nb, (long)ntot - off);
  load_lower_form<T, NMAX, LAUUM_THREADS, ST>(S, lower, n, A + off * lda + off, lda);
  __syncthreads();
  lds_lauum<T, NMAX>(S, n);
  store_lower_form<T, NMAX, LAUUM_THREADS, ST>(S, lower, n, B + blockIdx.x * bstride, ldb, false);
}

// D <- S on the triangle uplo (diagonal included) of an n x n block
template <typename T>
__global__ void tri_copy_kernel(int lower, int n, const T *S, long lds, T *D, long ldd) {
  const int r = blockIdx.x * 16 + (threadIdx.x & 15), c = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (r < n && c < n && (lower ? r >= c : r <= c)) D[(long)c * ldd + r] = S[(long)c * lds + r];
}

// Block b inverts the nb x nb diagonal block of op(A) starting at b * nb.
template <typename T, typename ST, int NMAX>
__global__ void __launch_bounds__(DIAG_THREADS)
trsm_inv_kernel(int lowerlike, int op, int unit, long m, int nb, const ST *A, long lda, ST *W) {
  constexpr int LS = Lds<T, NMAX>::LS;
  __shared__ T S[NMAX * (NMAX + 1)];
  __shared__ T vec[2 * NMAX];
  __shared__ T dinv[NMAX];
  const long off = (long)blockIdx.x * nb;
  const int n = (int)((m - off) < nb ? (m - off) : nb);
  // lower form: S(r, c) = op(A)(r, c) when op(A) is lower-like, else
  // conj(op(A)(c, r)); op(A)(i, j) = A(i, j), A(j, i) or conj(A(j, i)).
  gather_block<NMAX, DIAG_THREADS, 16>(n, A + off * lda + off, lda, [&](int a, int b, ST x) {
    // memory element A(off + a, off + b)
    const T v = convert<T>(x);
    // as an element of op(A): (a, b) for N, (b, a) for T/C
    int r = (op == OP_N) ? a : b, c = (op == OP_N) ? b : a;
    T e = (op == OP_C) ? conjx(v) : v;
    if (!lowerlike) { const int t = r; r = c; c = t; e = conjx(e); }
    if (r >= c) S[c * LS + r] = e;
  });
  __syncthreads();
  if (!unit) diag_reciprocals<T, NMAX>(S, dinv, n);
  __syncthreads();
  RegTile<T, NMAX> x;
  reg_trtri<T, NMAX>(S, x, vec, dinv, n, unit != 0);
  __syncthreads();
  reg_to_lds<T, NMAX>(x, S, n);
  __syncthreads();
  store_dense<T, NMAX, DIAG_THREADS, ST>(S, !lowerlike, n, unit != 0, W + (long)blockIdx.x * nb * nb, nb);
}

// One workgroup scans the diagonal; info = first (upper) or last (lower)
// 1-based index of an exact zero, 0 when none.
template <typename T>
__global__ void __launch_bounds__(1024) zero_scan_kernel(int lower, long n, const T *A, long lda, int *info) {
  __shared__ int best[1024];
  int b = lower ? 0 : INT_MAX;
  for (long i = threadIdx.x; i < n; i += 1024) {
    if (is_zero(A[i * lda + i])) {
      const int v = (int)(i + 1);
      b = lower ? (v > b ? v : b) : (v < b ? v : b);
    }
  }
  best[threadIdx.x] = b;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const int o = best[threadIdx.x + s], c = best[threadIdx.x];
      best[threadIdx.x] = lower ? (o > c ? o : c) : (o < c ? o : c);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *info = (best[0] == INT_MAX) ? 0 : best[0];
}

// ---- compensated (double-double) summation of logs ----
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd dd_add(dd a, dd b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b.hi - bb) + a.lo + b.lo;
  dd r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}
__device__ __forceinline__ dd dd_shfl_xor(dd v, int m) {
  dd r;
  r.hi = __shfl_xor(v.hi, m, 64);
  r.lo = __shfl_xor(v.lo, m, 64);
  return r;
}
__device__ dd block_sum(dd v) {
  __shared__ double part[2 * 16];
  for (int off = 32; off > 0; off >>= 1) v = dd_add(v, dd_shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) { part[2 * wave] = v.hi; part[2 * wave + 1] = v.lo; }
  __syncthreads();
  dd t = {0.0, 0.0};
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; w++) t = dd_add(t, dd{part[2 * w], part[2 * w + 1]});
  return t;
}

template <typename T>
__global__ void __launch_bounds__(256) logdet_partial_kernel(const T *x, long incx, long n, double *partials) {
  dd acc = {0.0, 0.0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc = dd_add(acc, dd{log(realx(x[i * incx])), 0.0});
  const dd t = block_sum(acc);
  if (threadIdx.x == 0) { partials[2 * blockIdx.x] = t.hi; partials[2 * blockIdx.x + 1] = t.lo; }
}

__global__ void __launch_bounds__(256) logdet_final_kernel(const double *partials, int nblocks, double *out) {
  dd acc = {0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    acc = dd_add(acc, dd{partials[2 * b], partials[2 * b + 1]});
  const dd t = block_sum(acc);
  if (threadIdx.x == 0) out[0] = 2.0 * (t.hi + t.lo);
}

}  // namespace

// Compute type of the diagonal-block kernels for a storage type.
template <typename ST> struct Compute { using type = ST; };
template <> struct Compute<float> { using type = double; };
template <> struct Compute<zf> { using type = zd; };

template <typename ST>
hipError_t diag_potrf(hipStream_t s, char uplo, int n, ST *A, long lda, ST *W, long ldw, int *info,
                      long info_off) {
  using T = typename Compute<ST>::type;
  constexpr int NMAX = diag_max<T>();
  if (n <= 0) return hipSuccess;
  if (n > NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((diag_potrf_kernel<T, ST, NMAX>), dim3(1), dim3(BLK_THREADS), 0, s,
                     uplo == 'L' ? 1 : 0, n, A, lda, W, ldw, info, info_off);
  return hipGetLastError();
}

template <typename ST>
hipError_t diag_trtri(hipStream_t s, char uplo, char diag, int n, const ST *A, long lda, ST *B,
                      long ldb, ST *W, long ldw, const int *abort_flag) {
  using T = typename Compute<ST>::type;
  constexpr int NMAX = diag_max<T>();
  if (n <= 0) return hipSuccess;
  if (n > NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((diag_trtri_kernel<T, ST, NMAX>), dim3(1), dim3(DIAG_THREADS), 0, s,
                     uplo == 'L' ? 1 : 0, diag == 'U' ? 1 : 0, n, n, A, lda, B, ldb, W, ldw,
                     abort_flag);
  return hipGetLastError();
}

template <typename ST>
hipError_t diag_trtri_blocks(hipStream_t s, char uplo, char diag, long n, int nb, ST *A, long lda) {
  using T = typename Compute<ST>::type;
  constexpr int NMAX = diag_max<T>();
  if (n <= 0) return hipSuccess;
  if (nb <= 0 || nb > NMAX || n > (long)INT32_MAX) return hipErrorInvalidValue;
  const long blocks = (n + nb - 1) / nb;
  hipLaunchKernelGGL((diag_trtri_kernel<T, ST, NMAX>), dim3((unsigned)blocks), dim3(DIAG_THREADS), 0, s,
                     uplo == 'L' ? 1 : 0, diag == 'U' ? 1 : 0, (int)n, nb, A, lda, A, lda,
                     static_cast<ST *>(nullptr), 0L, static_cast<const int *>(nullptr));
  return hipGetLastError();
}

template <typename ST>
hipError_t diag_lauum(hipStream_t s, char uplo, int n, ST *A, long lda) {
  using T = typename Compute<ST>::type;
  constexpr int NMAX = diag_max<T>();
  if (n <= 0) return hipSuccess;
  if (n > NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((diag_lauum_kernel<T, ST, NMAX>), dim3(1), dim3(LAUUM_THREADS), 0, s,
                     uplo == 'L' ? 1 : 0, n, n, static_cast<const ST *>(A), lda, A, lda, 0L);
  return hipGetLastError();
}

template <typename ST>
hipError_t diag_lauum_blocks(hipStream_t s, char uplo, long n, int nb, const ST *A, long lda, ST *W) {
  using T = typename Compute<ST>::type;
  constexpr int NMAX = diag_max<T>();
  if (n <= 0) return hipSuccess;
  if (nb <= 0 || nb > NMAX || n > (long)INT32_MAX) return hipErrorInvalidValue;
  const long blocks = (n + nb - 1) / nb;
  hipLaunchKernelGGL((diag_lauum_kernel<T, ST, NMAX>), dim3((unsigned)blocks), dim3(LAUUM_THREADS), 0, s,
                     uplo == 'L' ? 1 : 0, (int)n, nb, A, lda, W, (long)nb, (long)nb * nb);
  return hipGetLastError();
}

template <typename T>
hipError_t tri_copy(hipStream_t s, char uplo, int n, const T *S, long lds, T *D, long ldd) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 15) / 16), (unsigned)((n + 15) / 16));
  hipLaunchKernelGGL((tri_copy_kernel<T>), grid, dim3(256), 0, s, uplo == 'L' ? 1 : 0, n, S, lds, D, ldd);
  return hipGetLastError();
}

template <typename ST>
hipError_t trsm_diag_inverses(hipStream_t s, char uplo, char trans, char diag, long m, int nb,
                              const ST *A, long lda, ST *W) {
  using T = typename Compute<ST>::type;
  constexpr int NMAX = diag_max<T>();
  if (m <= 0) return hipSuccess;
  if (nb > NMAX) return hipErrorInvalidValue;
  const int op = is_complex<T>::value ? op_code(trans) : (trans == 'N' ? OP_N : OP_T);
  const int lowerlike = ((uplo == 'L') == (trans == 'N')) ? 1 : 0;
  const unsigned nblk = (unsigned)((m + nb - 1) / nb);
  hipLaunchKernelGGL((trsm_inv_kernel<T, ST, NMAX>), dim3(nblk), dim3(DIAG_THREADS), 0, s, lowerlike,
                     op, diag == 'U' ? 1 : 0, m, nb, A, lda, W);
  return hipGetLastError();
}

template <typename T>
hipError_t diag_zero_scan(hipStream_t s, char uplo, long n, const T *A, long lda, int *info) {
  hipLaunchKernelGGL((zero_scan_kernel<T>), dim3(1), dim3(1024), 0, s, uplo == 'L' ? 1 : 0, n, A,
                     lda, info);
  return hipGetLastError();
}

template <typename T>
hipError_t logdet(hipStream_t s, const T *x, long incx, long n, double *partials, double *out) {
  long blocks = (n + 255) / 256;
  if (blocks > kLogdetMaxBlocks) blocks = kLogdetMaxBlocks;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((logdet_partial_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, x, incx,
                     n, partials);
  HC_TRY(hipGetLastError());
  hipLaunchKernelGGL(logdet_final_kernel, dim3(1), dim3(256), 0, s, partials, (int)blocks, out);
  return hipGetLastError();
}

template <typename ST>
hipError_t panel_potrf(hipStream_t s, long m, long jb, ST *A, long lda, ST *ws, int *flags, int epoch, int *info,
                       long info_off) {
  if (jb <= 0) return hipSuccess;
  if (jb > kPanelMaxCols || m < jb || m > (long)INT32_MAX) return hipErrorInvalidValue;
  const long C = (jb + PK_B - 1) / PK_B, grid = C + (m - jb + PK_B - 1) / PK_B;
  hipLaunchKernelGGL((panel_potrf_kernel<ST>), dim3((unsigned)grid), dim3(PK_THREADS), 0, s, (int)m, (int)jb, A,
                     lda, ws, flags, epoch, info, info_off);
  return hipGetLastError();
}
template hipError_t panel_potrf<float>(hipStream_t, long, long, float *, long, float *, int *, int, int *, long);
template hipError_t panel_potrf<double>(hipStream_t, long, long, double *, long, double *, int *, int, int *,
                                        long);

#define HC_INST(T)                                                                            \
  template hipError_t diag_potrf<T>(hipStream_t, char, int, T *, long, T *, long, int *, long); \
  template hipError_t diag_trtri<T>(hipStream_t, char, char, int, const T *, long, T *, long,   \
                                    T *, long, const int *);                                    \
  template hipError_t diag_trtri_blocks<T>(hipStream_t, char, char, long, int, T *, long);      \
  template hipError_t diag_lauum<T>(hipStream_t, char, int, T *, long);                         \
  template hipError_t diag_lauum_blocks<T>(hipStream_t, char, long, int, const T *, long, T *); \
  template hipError_t tri_copy<T>(hipStream_t, char, int, const T *, long, T *, long);          \
  template hipError_t trsm_diag_inverses<T>(hipStream_t, char, char, char, long, int,           \
                                            const T *, long, T *);                              \
  template hipError_t diag_zero_scan<T>(hipStream_t, char, long, const T *, long, int *);       \
  template hipError_t logdet<T>(hipStream_t, const T *, long, long, double *, double *);
// single / single-complex storage: computed in double / double complex
// (the mixed-precision factorisation and the s / c routines); the
// log-determinants take their logs in FP64
HC_INST(double)
HC_INST(zd)
HC_INST(float)
HC_INST(zf)
#undef HC_INST

}  // namespace hc
