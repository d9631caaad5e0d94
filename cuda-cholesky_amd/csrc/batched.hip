// Batched Cholesky factorisation + log-determinant of many small SPD
// matrices of uneven sizes, packed contiguously (matrix i: n_i x n_i,
// column-major with leading dimension n_i, at element offset off_i).
//
// One persistent kernel: workgroup w factors matrices order[w], order[w + G],
// ... (G = grid size; `order` lists the matrices by decreasing n, so the
// round-robin deal is a longest-first schedule).  Each matrix is factored in
// place, right-looking over 16-column leaves:
//   * wave 0 factors leaf p in registers (DPP row broadcasts), writes L_pp,
//     accumulates log of the pivots and inverts L_pp into LDS;
//   * all waves form the panel L_Ip = A_Ip inv(L_pp)^T on the matrix cores
//     (written to the matrix and kept in LDS);
//   * all waves update the trailing lower tiles A_IJ -= L_Ip L_Jp^T with
//     operands from LDS.
// With `inverse`, the same workgroup then overwrites the factor with A^-1:
//   * inv(L) in place, leaves right to left: X_pp = inv(L_pp) (wave 0,
//     registers), X_Ip = -(sum_{p<K<=I} X_IK L_Kp) X_pp for the tiles below
//     (all Y_I = sum X_IK L_Kp formed into LDS before column p is
//     overwritten);
//   * A^-1 = inv(L)^T inv(L), tile rows top-down: row i of the result is
//     sum_{K>=i} X_Ki^T X_KC (C <= i), formed into LDS, then written; later
//     rows never read row i again.
// Storage is FP64 or FP32 (ST); arithmetic is FP64 throughout.  Only the
// referenced triangle (uplo) is read or written; tiles reaching past n read
// the identity and are not written back.
#include <algorithm>

#include "internal.h"
#include "tile16.h"

namespace hc {
namespace {

using t16::LEAF;
using t16::TileAcc;
using t16::kNegA;

constexpr int BT_THREADS = 256;                // 4 waves per matrix
constexpr int BT_NW = BT_THREADS / 64;
constexpr int BT_NPMAX = kBatchedMaxN / LEAF;  // leaves per matrix at most
constexpr int LL = LEAF + 1;                   // LDS tile leading dimension
constexpr int TG = 4;                          // trailing tiles loaded together per wave

// Lower-form view of one packed matrix: element (r, c), r >= c, is A(r, c)
// for lower storage and A(c, r) for upper storage.
template <typename ST, bool LOWER>
struct View {
  ST *M;
  int n;
  __device__ long off(int r, int c) const { return LOWER ? (long)c * n + r : (long)r * n + c; }
  // value of lower-form (r, c): zero above the diagonal, identity beyond n.
  // The load is unconditional, from a clamped (always valid) address, and the
  // value is selected afterwards: no branch per element, so loads stay in
  // flight together and register arrays filled from them stay in registers.
  __device__ double get(int r, int c) const { return fix(raw(r, c), r, c); }
  // the two halves of get: the load from the clamped address, and the select
  __device__ double raw(int r, int c) const { return (double)M[off(min(r, n - 1), min(c, n - 1))]; }
  __device__ double fix(double x, int r, int c) const {
    return (r < n && c < n) ? (r >= c ? x : 0.0) : (r == c ? 1.0 : 0.0);
  }
  __device__ void put(int r, int c, double v) const {
    if (r < n && c < n && r >= c) M[off(r, c)] = (ST)v;
  }
};

// Wave 0: factor the leaf at g0 whose rows are in a[] (lane r + 16 q holds
// row r), write L_pp to the matrix, accumulate log L_rr^2 of the valid rows
// into logpart and put inv(L_pp) into Li; returns the failing pivot's column
// (g0 + k) or -1.
template <typename ST, bool LOWER>
__device__ __forceinline__ int leaf_step(double (&a)[LEAF], const View<ST, LOWER> &v, int g0, double *Li,
                                         double &logpart) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  double x[4], rpiv;
  const int b = t16::leaf_factor_invert<double>(a, x, rpiv);
  if (b >= 0) return g0 + b;
  // log of this lane's own pivot L_rr^2 = -2 log(1 / L_rr), all rows at once
  // (off the pivot chain); summed over the 16 rows once per matrix
  if (q == 0 && g0 + r < v.n) logpart -= 2.0 * log(rpiv);
#pragma unroll
  for (int j = 0; j < 4; j++) {  // lane group q writes columns q + 4 j of its row
    // a value select (a conditional over the array elements themselves
    // would select their addresses and keep the array in memory)
    const double f = t16::sel(q == 0, a[4 * j], t16::sel(q == 1, a[4 * j + 1], t16::sel(q == 2, a[4 * j + 2], a[4 * j + 3])));
    v.put(g0 + r, g0 + q + 4 * j, f);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) Li[(q + 4 * j) * LL + r] = x[j];
  return -1;
}

// The factor-only kernel is held to three waves per SIMD (168 registers):
// more matrices in flight per CU hide the per-matrix pivot chain.
template <bool INVERSE>
constexpr int batched_waves_per_eu() { return INVERSE ? 2 : 3; }

template <typename ST, bool LOWER, bool INVERSE>
__global__ void __launch_bounds__(BT_THREADS)
__attribute__((amdgpu_waves_per_eu(batched_waves_per_eu<INVERSE>(), 8)))
potrf_batched_kernel(int count, const int *nvec, const long *offs, const int *order, ST *A,
                     double *logdet, int *info) {
  __shared__ double Li[LEAF * LL];                 // inv(L_pp)
  __shared__ double P[BT_NPMAX * LEAF * LL];       // panel tiles L_Ip
  __shared__ double Dg[LEAF * LL];                 // wave 0: the next diagonal tile, layout change
  __shared__ int flag;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  for (int t = blockIdx.x; t < count; t += gridDim.x) {
    const int idx = order[t];
    const int n = nvec[idx];
    const View<ST, LOWER> v{A + offs[idx], n};
    const int np = (n + LEAF - 1) / LEAF;
    double logpart = 0.0;  // wave 0, lanes 0..15: sum of log(L_rr^2) of row r of every leaf
    int bad = -1;
    // Leaf p + 1 is factored by wave 0 during step p's trailing update
    // (look-ahead), from the tile (p + 1, p + 1) it updates itself; step
    // p = -1 only factors leaf 0 (one call site keeps the leaf inlined once).
    for (int p = -1; p < np; p++) {
      const int g0 = p * LEAF;
      if (p >= 0) {
        bad = flag;
        if (bad >= 0) break;
      }
      // panel: L_Ip = A_Ip inv(L_pp)^T, to the matrix and to LDS
      for (int I = p + 1 + wave; p >= 0 && I < np; I += BT_NW) {
        TileAcc<double> acc;
        t16::acc_zero(acc);
#pragma unroll
        for (int k0 = 0; k0 < LEAF; k0 += 4) {
          const double x = v.get(I * LEAF + li, g0 + k0 + lk);
          const double y = Li[(k0 + lk) * LL + li];
          acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc.re, 0, 0, 0);
        }
        double *Pt = P + I * LEAF * LL;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = lk + 4 * r, col = li;
          Pt[col * LL + row] = acc.re[r];
          v.put(I * LEAF + row, g0 + col, acc.re[r]);
        }
      }
      __syncthreads();
      const int m = np - 1 - p;  // trailing tile rows / columns
      if (m > 0 && wave == 0) {
        // look-ahead: tile (p + 1, p + 1) updated in registers, then leaf p + 1
        const int D = p + 1;
        TileAcc<double> acc;
#pragma unroll
        for (int r = 0; r < 4; r++) acc.re[r] = v.get(D * LEAF + lk + 4 * r, D * LEAF + li);
        if (p >= 0) {
          const double *PD = P + D * LEAF * LL;
#pragma unroll
          for (int k0 = 0; k0 < LEAF; k0 += 4) {
            const double x = -PD[(k0 + lk) * LL + li];
            const double y = PD[(k0 + lk) * LL + li];
            acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc.re, 0, 0, 0);
          }
        }
        t16::acc_store(acc, Dg, LL);  // element (row, col) at col LL + row
        t16::wave_lds_fence();
        double a[LEAF];
        t16::static_for<0, LEAF>([&](auto C) { a[decltype(C)::value] = Dg[decltype(C)::value * LL + li]; });
        const int b = leaf_step<ST, LOWER>(a, v, D * LEAF, Li, logpart);
        if (lane == 0) flag = b;
      } else if (m > 0 && p >= 0) {
        // the other trailing tiles (I, J), p < J <= I, without (p+1, p+1),
        // dealt round-robin to waves 1..3; each wave loads TG of its tiles
        // before computing any, so their global-memory latencies overlap
        const int ntile = m * (m + 1) / 2 - 1;
        constexpr int NWT = BT_NW - 1;
        for (int t0 = (wave - 1) * TG; t0 < ntile; t0 += NWT * TG) {
          int ti[TG], tj[TG];
          TileAcc<double> acc[TG];
#pragma unroll
          for (int g = 0; g < TG; g++) {
            // tile t0 + g + 1 of the column-major lower triangle of the m x m tile block
            const bool live = t0 + g < ntile;
            int tt = live ? t0 + g + 1 : 1, J = 0;  // tt < m (m + 1) / 2: the walk ends
            while (tt >= m - J) {
              tt -= m - J;
              J++;
            }
            ti[g] = p + 1 + J + tt;
            tj[g] = p + 1 + J;
#pragma unroll
            for (int r = 0; r < 4; r++)
              acc[g].re[r] = live ? v.get(ti[g] * LEAF + lk + 4 * r, tj[g] * LEAF + li) : 0.0;
          }
#pragma unroll
          for (int g = 0; g < TG; g++) {
            if (t0 + g >= ntile) break;
            const double *PI = P + ti[g] * LEAF * LL, *PJ = P + tj[g] * LEAF * LL;
#pragma unroll
            for (int k0 = 0; k0 < LEAF; k0 += 4) {
              const double x = -PI[(k0 + lk) * LL + li];
              const double y = PJ[(k0 + lk) * LL + li];
              acc[g].re = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[g].re, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) v.put(ti[g] * LEAF + lk + 4 * r, tj[g] * LEAF + li, acc[g].re[r]);
          }
        }
      }
      __syncthreads();
    }
    if (INVERSE && bad < 0) {
      // ---- inv(L) in place, leaves right to left ----
      for (int p = np - 1; p >= 0; p--) {
        const int g0 = p * LEAF;
        if (wave == 0) {
          double a[LEAF], x[4];
          t16::static_for<0, LEAF>([&](auto C) { a[decltype(C)::value] = v.get(g0 + li, g0 + decltype(C)::value); });
          t16::leaf_inverse_cols<double>(a, x);
#pragma unroll
          for (int j = 0; j < 4; j++) Li[(lk + 4 * j) * LL + li] = x[j];
        }
        // Y_I = sum_{K=p+1}^{I} X_IK L_Kp into LDS (L_Kp still the factor)
        for (int I = p + 1 + wave; I < np; I += BT_NW) {
          TileAcc<double> acc;
          t16::acc_zero(acc);
          for (int K = p + 1; K <= I; K++) {
#pragma unroll
            for (int k0 = 0; k0 < LEAF; k0 += 4) {
              const double x = v.get(I * LEAF + li, K * LEAF + k0 + lk);
              const double y = v.get(K * LEAF + k0 + lk, g0 + li);
              acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc.re, 0, 0, 0);
            }
          }
          double *Pt = P + I * LEAF * LL;
#pragma unroll
          for (int r = 0; r < 4; r++) Pt[li * LL + lk + 4 * r] = acc.re[r];
        }
        __syncthreads();
        // X_Ip = -Y_I X_pp, then X_pp itself
        for (int I = p + 1 + wave; I < np; I += BT_NW) {
          TileAcc<double> acc;
          t16::acc_zero(acc);
          const double *Pt = P + I * LEAF * LL;
#pragma unroll
          for (int k0 = 0; k0 < LEAF; k0 += 4) {
            const double x = -Pt[(k0 + lk) * LL + li];
            const double y = Li[li * LL + k0 + lk];
            acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc.re, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 4; r++) v.put(I * LEAF + lk + 4 * r, g0 + li, acc.re[r]);
        }
        if (wave == 0) {
#pragma unroll
          for (int j = 0; j < 4; j++) v.put(g0 + li, g0 + lk + 4 * j, Li[(lk + 4 * j) * LL + li]);
        }
        __syncthreads();
      }
      // ---- A^-1 = X^T X, tile rows top-down ----
      for (int i = 0; i < np; i++) {
        for (int C = wave; C <= i; C += BT_NW) {
          TileAcc<double> acc;
          t16::acc_zero(acc);
          for (int K = i; K < np; K++) {
#pragma unroll
            for (int k0 = 0; k0 < LEAF; k0 += 4) {
              const double x = v.get(K * LEAF + k0 + lk, i * LEAF + li);  // X_Ki^T
              const double y = v.get(K * LEAF + k0 + lk, C * LEAF + li);  // X_KC
              acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc.re, 0, 0, 0);
            }
          }
          double *Pt = P + C * LEAF * LL;
#pragma unroll
          for (int r = 0; r < 4; r++) Pt[li * LL + lk + 4 * r] = acc.re[r];
        }
        __syncthreads();
        for (int C = wave; C <= i; C += BT_NW) {
          const double *Pt = P + C * LEAF * LL;
#pragma unroll
          for (int r = 0; r < 4; r++) v.put(i * LEAF + lk + 4 * r, C * LEAF + li, Pt[li * LL + lk + 4 * r]);
        }
        __syncthreads();
      }
    }
    if (wave == 0) {
      double logsum = logpart;  // lanes 16..63 hold 0
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) logsum += __shfl_xor(logsum, off, 64);
      if (lane == 0) {
        info[idx] = bad >= 0 ? bad + 1 : 0;
        if (logdet != nullptr) logdet[idx] = bad >= 0 ? __builtin_nan("") : logsum;
      }
    }
    __syncthreads();  // flag and LDS are reused by the next matrix
  }
}

// ---------------------------------------------------------------------------
// Register-resident factorisation (factor + log-determinant, no inverse).
//
// One workgroup factors one matrix (n <= 16 NPMAX) at a time and holds the
// matrix's whole lower tile triangle in registers, so that its HBM traffic is
// one read of the triangle and one write of L.  The critical path of a
// matrix is its chain of 16-column leaves: leaf p + 1 needs tile (p + 1, p + 1)
// after its update with panel tile (p + 1, p), which needs inv(L_pp) from
// leaf p.  The roles are cut so that nothing else lies on that chain:
//   * the leaf wave (the last) holds no tiles: it waits (LDS flag `dready`)
//     for tile (D, D), factors and inverts leaf D, records its pivots (their
//     logs are taken once per matrix, off the chain), publishes inv(L_DD)
//     (flag `lready`) and waits for the next tile.  It takes part in no
//     barrier inside a matrix;
//   * the NW - 1 tile waves hold every tile.  The sub-diagonal tile
//     (J + 1, J) and the diagonal tile (J + 1, J + 1) are in the same wave,
//     (J + 1) % (NW - 1): at step p that wave forms its panel tile (p + 1, p)
//     first, updates tile (p + 1, p + 1) with it straight from registers and
//     hands it to the leaf wave before it forms its other panel tiles, so leaf
//     p + 1 starts one panel tile and one update after leaf p has finished.
//     The other off-diagonal tiles follow round-robin: in the sequence g =
//     0, 1, ... of the np - 1 sub-diagonal tiles (g: tile (g + 1, g)) and then
//     the tiles I >= J + 2 column by column, tile g is in wave
//     (g + 1) % (NW - 1), slot g / (NW - 1) (np = ceil(n / 16)).  Diagonal tile
//     (J, J) is in wave J % (NW - 1), diagonal slot J / (NW - 1).
// A tile is held as T = A_IJ^T in the MFMA result layout (lane l, register
// r: T(l / 16 + 4 r, l % 16) = A_IJ(l % 16, l / 16 + 4 r)), which is also the
// operand fragment of A_IJ for k-step r: tiles enter the matrix cores with no
// layout change.  Per step p, in the tile waves:
//   wait for inv(L_pp) (flag `lready`);
//   panel:    T_Ip <- inv(L_pp) T_Ip = L_Ip^T (I > p), to LDS buffer p % 2
//             (element (row, k) at P_I[k LL + row]) and to the matrix;
//   tile-wave barrier (an LDS counter: the leaf wave is not held);
//   trailing: T_IJ -= L_Jp L_Ip^T (p < J <= I, (p + 1, p + 1) excepted).
// The panel buffers, inv(L_pp) and the hand-off tile alternate or are
// consumed before they are written again: P and inv(L) by step parity, the
// hand-off tile (D, D) is read by the leaf wave before leaf D publishes, and
// the next hand-off waits for that.  Every wait is bounded (4 s of the
// 100 MHz real-time counter); an abandoned one makes the matrix report
// info = -1.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

// The slot code (I | J << 8, or -1) of slot s, read from lane s of `codes`
// into a scalar register.  The empty asm makes it opaque, so that nothing
// derived from it is hoisted out of the step loop and kept live per slot.
__device__ __forceinline__ int slot_code(int codes, int s) {
  int c = __builtin_amdgcn_readlane(codes, s);
  asm volatile("" : "+s"(c));
  return c;
}

// T -= L_J L_I^T with the panel tiles L_I, L_J in LDS (P_I, P_J); lo is the
// lane's fragment offset at k-step 0.  The minus sign is the instruction's
// negate-A modifier, not a VALU negation of each operand.
__device__ __forceinline__ void tile_update(v4d &T, const double *PI, const double *PJ, int lo) {
#pragma unroll
  for (int r = 0; r < 4; r++)
    T = __builtin_amdgcn_mfma_f64_16x16x4f64(PJ[r * 4 * LL + lo], PI[r * 4 * LL + lo], T, 0, 0, kNegA);
}

// The leaf wave's step: factor the leaf at g0 (rows in a[], lane r + 16 q
// holds row r), write L_pp to the matrix, 1 / L_rr (rows below n) to
// rpiv[g0 + r] and inv(L_pp) to Li; returns the failing pivot's column
// (g0 + k) or -1.
template <typename ST, bool LOWER>
__device__ __forceinline__ int leaf_step_piv(double (&a)[LEAF], const View<ST, LOWER> &v, int g0, double *Li,
                                             double *rpiv) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  double x[4], rp;
  const int b = t16::leaf_factor_invert<double>(a, x, rp);
  if (b >= 0) return g0 + b;
  if (q == 0 && g0 + r < v.n) rpiv[g0 + r] = rp;
#pragma unroll
  for (int j = 0; j < 4; j++) {  // lane group q writes columns q + 4 j of its row
    const double f = t16::sel(q == 0, a[4 * j], t16::sel(q == 1, a[4 * j + 1], t16::sel(q == 2, a[4 * j + 2], a[4 * j + 3])));
    v.put(g0 + r, g0 + q + 4 * j, f);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) Li[(q + 4 * j) * LL + r] = x[j];
  return -1;
}

// Size classes of the register-resident kernel, by tile rows np (order n):
//   0: np <= 4  (n <= 64):  2 waves, 3 waves per SIMD (6 workgroups per CU);
//   1: np <= 8  (n <= 128): 4 waves, 3 waves per SIMD (3 workgroups per CU);
//   2: np <= 16 (n <= 256): 8 waves, 2 waves per SIMD (1 workgroup per CU).
// More workgroups per CU for the small classes overlap more leaf chains.
template <int CLS>
struct RegClass;
template <>
struct RegClass<0> {
  static constexpr int NPMAX = 4, NW = 2, OCC = 3;
};
template <>
struct RegClass<1> {
  static constexpr int NPMAX = 8, NW = 4, OCC = 3;
};
template <>
struct RegClass<2> {
  static constexpr int NPMAX = 16, NW = 8, OCC = 2;
};

// Bounded wait of a whole wave until `pred()` holds (4 s of the 100 MHz
// real-time counter); false when abandoned.
template <typename F>
__device__ __forceinline__ bool wave_wait(F pred) {
  if (pred()) return true;
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while (!pred()) {
    __builtin_amdgcn_s_sleep(1);
    if (__builtin_amdgcn_s_memrealtime() - t0 > 400000000ull) return false;
  }
  return true;
}

__device__ __forceinline__ int lds_acquire(const int *f) {
  return __builtin_amdgcn_readfirstlane(__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}

// cnt[c]: matrices of class c; `order` lists class 2, then 1, then 0.
template <typename ST, bool LOWER, int CLS, int NPMAX = RegClass<CLS>::NPMAX, int NW = RegClass<CLS>::NW,
          int OCC = RegClass<CLS>::OCC>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
potrf_batched_reg_kernel(const int *cnt, const int *nvec, const long *offs, const int *order, ST *A,
                         double *logdet, int *info) {
  const int count = cnt[CLS];
  order += CLS == 2 ? 0 : (CLS == 1 ? cnt[2] : cnt[2] + cnt[1]);
  constexpr int NWT = NW - 1;                                            // tile waves
  constexpr int NO = NPMAX * (NPMAX - 1) / 2, S = (NO + NWT - 1) / NWT;  // off-diagonal tiles, slots
  constexpr int SD = (NPMAX + NWT - 1) / NWT;                           // diagonal slots
  constexpr int PB = NPMAX * LEAF * LL;                                 // one panel buffer
  static_assert(S <= 64, "slot codes live in one lane each");
  __shared__ double P[2 * PB];           // panel tiles L_Ip, by step parity
  __shared__ double Li[2 * LEAF * LL];   // inv(L_pp), by step parity
  __shared__ double Dg[LEAF * LL];       // tile (D, D) on its way to the leaf wave
  __shared__ double rpiv[NPMAX * LEAF];  // 1 / L_rr
  __shared__ double red[NW];
  // flag: failing column, -1 (none) or -2 (the leaf wave abandoned a wait);
  // fleaf: the leaf that failed (set once, before lready; a large value
  // while none has), so that every tile wave decides at the same step;
  // dready: D once tile (D, D) is in Dg; lready: D once leaf D is done;
  // tbar: tile-wave barrier arrivals; werr: a tile wave abandoned a wait
  __shared__ int flag, fleaf, dready, lready, tbar, werr;
  const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int lo = lk * LL + li;  // this lane's operand fragment in an LDS tile (k-step 0)
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (tid == 0) {
    flag = -1;
    fleaf = 1 << 30;
    dready = -1;
    lready = -1;
    tbar = 0;
    werr = 0;
  }
  __syncthreads();
  for (int t = blockIdx.x; t < count; t += gridDim.x) {
    const int idx = order[t];
    const int n = nvec[idx];
    const View<ST, LOWER> v{A + offs[idx], n};
    if (n == 0) {
      // nothing to load: the clamped addresses below assume n >= 1
      if (tid == 0) {
        info[idx] = 0;
        if (logdet != nullptr) logdet[idx] = 0.0;
      }
      continue;
    }
    const int np = (n + LEAF - 1) / LEAF;
    if (wave == NWT) {
      // ---- leaf wave ----
      for (int D = 0; D < np; D++) {
        int b = -2;
        if (wave_wait([&] { return lds_acquire(&dready) == D; })) {
          double a[LEAF];
          t16::static_for<0, LEAF>([&](auto C) { a[decltype(C)::value] = Dg[decltype(C)::value * LL + li]; });
          b = leaf_step_piv<ST, LOWER>(a, v, D * LEAF, Li + (D & 1) * LEAF * LL, rpiv);
        }
        if (lane == 0) {
          if (b != -1) {
            flag = b;
            fleaf = D;
          }
          __hip_atomic_store(&lready, D, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (b != -1) break;
      }
    } else {
      // ---- tile waves ----
      const int no = np * (np - 1) / 2;
      int codes = -1;  // lane s < S: the code of this wave's off-diagonal slot s
      {
        const int g = (wave + NWT - 1) % NWT + NWT * lane;
        if (lane < S && g < no) {
          if (g < np - 1) {
            codes = (g + 1) | (g << 8);  // sub-diagonal tile (g + 1, g)
          } else {
            int k = g - (np - 1), J = 0;  // k-th tile with I >= J + 2, column by column
            while (k >= np - 2 - J) {
              k -= np - 2 - J;
              J++;
            }
            codes = (J + 2 + k) | (J << 8);
          }
        }
      }
      v4d T[S], Td[SD];
      // every load is in flight before the first select, and each select
      // overwrites its own load (no second register set); a slot without a
      // tile loads tile (0, 0) and never uses it (no branch)
      t16::static_for<0, S>([&](auto Sc) {
        constexpr int s = decltype(Sc)::value;
        const int c = max(slot_code(codes, s), 0);
        const int I = c & 255, J = c >> 8;
#pragma unroll
        for (int r = 0; r < 4; r++) T[s][r] = v.raw(I * LEAF + li, J * LEAF + lk + 4 * r);
      });
#pragma unroll
      for (int q = 0; q < SD; q++) {
        const int J = min(q * NWT + wave, np - 1);
#pragma unroll
        for (int r = 0; r < 4; r++) Td[q][r] = v.raw(J * LEAF + li, J * LEAF + lk + 4 * r);
      }
      t16::static_for<0, S>([&](auto Sc) {
        constexpr int s = decltype(Sc)::value;
        const int c = max(slot_code(codes, s), 0);
        const int I = c & 255, J = c >> 8;
#pragma unroll
        for (int r = 0; r < 4; r++) T[s][r] = v.fix(T[s][r], I * LEAF + li, J * LEAF + lk + 4 * r);
      });
#pragma unroll
      for (int q = 0; q < SD; q++) {
        const int J = min(q * NWT + wave, np - 1);
#pragma unroll
        for (int r = 0; r < 4; r++) Td[q][r] = v.fix(Td[q][r], J * LEAF + li, J * LEAF + lk + 4 * r);
      }
      if (wave == 0) {  // tile (0, 0) to the leaf wave
#pragma unroll
        for (int r = 0; r < 4; r++) Dg[r * 4 * LL + lo] = Td[0][r];  // A_00(li, lk + 4 r)
        if (lane == 0) __hip_atomic_store(&dready, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      // the last column has no panel tiles and no trailing update
      for (int p = 0; p + 1 < np; p++) {
        const int D = p + 1;
        if (!wave_wait([&] { return lds_acquire(&lready) >= p; })) {
          if (lane == 0) werr = 1;
          break;
        }
        // leaf p failed (a later leaf may already have failed too: the
        // decision is made on the step, the same in every tile wave)
        if (fleaf <= p) break;
        const double *Lp = Li + (p & 1) * LEAF * LL;
        double *Pp = P + (p & 1) * PB;
        // panel of column p: T_Ip <- inv(L_pp) T_Ip, to LDS and to the
        // matrix; the sub-diagonal tile (D, p) comes first (lowest slot), and
        // its owner updates tile (D, D) with it and hands that over at once
        // this lane's fragments of inv(L_pp), read once for the wave's panel
        // tiles in the largest class (the smaller ones have no registers
        // to spare and re-read them)
        double lf[4];
#pragma unroll
        for (int r = 0; r < 4; r++) lf[r] = CLS == 2 ? Lp[r * 4 * LL + lo] : 0.0;
        t16::static_for<0, S>([&](auto Sc) {
          constexpr int s = decltype(Sc)::value;
          const int c = slot_code(codes, s);
          if (c >= 0 && (c >> 8) == p) {
            v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 4; r++)
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(CLS == 2 ? lf[r] : Lp[r * 4 * LL + lo], T[s][r], acc, 0, 0, 0);
            T[s] = acc;
            const int I = c & 255;
            if (I == D) {
              // tile (D, D) -= L_Dp L_Dp^T: acc[r] is the operand fragment of
              // L_Dp for k-step r, on both sides
#pragma unroll
              for (int q = 0; q < SD; q++)
                if (q == D / NWT) {
#pragma unroll
                  for (int r = 0; r < 4; r++)
                    Td[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[r], acc[r], Td[q], 0, 0, kNegA);
#pragma unroll
                  for (int r = 0; r < 4; r++) Dg[r * 4 * LL + lo] = Td[q][r];  // A_DD(li, lk + 4 r)
                }
              if (lane == 0) __hip_atomic_store(&dready, D, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            double *Pt = Pp + I * LEAF * LL;
#pragma unroll
            for (int r = 0; r < 4; r++) {
              Pt[r * 4 * LL + lo] = acc[r];
              v.put(I * LEAF + li, p * LEAF + lk + 4 * r, acc[r]);  // L_Ip(li, lk + 4 r)
            }
          }
        });
        // tile-wave barrier: every panel tile of column p is in LDS
        if (lane == 0) __hip_atomic_fetch_add(&tbar, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (!wave_wait([&] { return lds_acquire(&tbar) >= NWT * D; })) {
          if (lane == 0) werr = 1;
          break;
        }
        // the other trailing tiles
#pragma unroll
        for (int q = 0; q < SD; q++) {
          const int J = q * NWT + wave;
          if (J > D && J < np) tile_update(Td[q], Pp + J * LEAF * LL, Pp + J * LEAF * LL, lo);
        }
        t16::static_for<0, S>([&](auto Sc) {
          constexpr int s = decltype(Sc)::value;
          const int c = slot_code(codes, s);
          if (c >= 0 && (c >> 8) > p) tile_update(T[s], Pp + (c & 255) * LEAF * LL, Pp + (c >> 8) * LEAF * LL, lo);
        });
      }
    }
    __syncthreads();  // the leaf wave is done: flag, rpiv and werr are final
    const int fl = flag;
    // bad: failing column, -1 (none) or -2 (a wait was abandoned)
    const int bad = fl != -1 ? fl : (werr != 0 ? -2 : -1);
    // log-determinant: 2 sum log L_rr = -2 sum log(1 / L_rr), one pivot per thread
    double lg = 0.0;
    if (bad == -1)
      for (int r = tid; r < n; r += NW * 64) lg -= log(rpiv[r]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lg += __shfl_xor(lg, off, 64);
    if (lane == 0) red[wave] = lg;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < NW; w++) sum += red[w];
      info[idx] = bad >= 0 ? bad + 1 : (bad == -2 ? -1 : 0);
      if (logdet != nullptr) logdet[idx] = bad != -1 ? __builtin_nan("") : 2.0 * sum;
      // every wave is past its last use of the flags
      flag = -1;
      fleaf = 1 << 30;
      dready = -1;
      lready = -1;
      tbar = 0;
      werr = 0;
    }
    __syncthreads();  // flags, rpiv and LDS tiles are reused by the next matrix
  }
}

template <typename ST, bool LOWER, int CLS>
constexpr auto reg_kernel_ptr() {
  return potrf_batched_reg_kernel<ST, LOWER, CLS>;
}

template <typename ST, bool LOWER, int CLS>
void reg_launch(hipStream_t s, int grid, const int *cnt, const int *nvec, const long *offs, const int *order, ST *A,
                double *logdet, int *info) {
  hipLaunchKernelGGL((reg_kernel_ptr<ST, LOWER, CLS>()), dim3((unsigned)std::max(1, grid)),
                     dim3(RegClass<CLS>::NW * 64), 0, s, cnt, nvec, offs, order, A, logdet, info);
}

template <typename ST, bool LOWER, int CLS>
int reg_occupancy() {
  int nb = 0;
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reg_kernel_ptr<ST, LOWER, CLS>(), RegClass<CLS>::NW * 64, 0);
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

}  // namespace

template <typename ST>
hipError_t potrf_batched_reg(hipStream_t s, char uplo, const int *d_cnt, const int *h_cnt, const int grid[3],
                             const int *nvec, const long *offs, const int *order, ST *A, double *logdet, int *info) {
  for (int c = 2; c >= 0; c--) {
    if (h_cnt != nullptr && h_cnt[c] == 0) continue;
    const int g = h_cnt != nullptr ? std::min(grid[c], h_cnt[c]) : grid[c];
    if (uplo == 'L') {
      if (c == 2) reg_launch<ST, true, 2>(s, g, d_cnt, nvec, offs, order, A, logdet, info);
      else if (c == 1) reg_launch<ST, true, 1>(s, g, d_cnt, nvec, offs, order, A, logdet, info);
      else reg_launch<ST, true, 0>(s, g, d_cnt, nvec, offs, order, A, logdet, info);
    } else {
      if (c == 2) reg_launch<ST, false, 2>(s, g, d_cnt, nvec, offs, order, A, logdet, info);
      else if (c == 1) reg_launch<ST, false, 1>(s, g, d_cnt, nvec, offs, order, A, logdet, info);
      else reg_launch<ST, false, 0>(s, g, d_cnt, nvec, offs, order, A, logdet, info);
    }
    HC_TRY(hipGetLastError());
  }
  return hipSuccess;
}

template <typename ST>
int potrf_batched_reg_blocks_per_cu(char uplo, int cls) {
  if (uplo == 'L') return cls == 2 ? reg_occupancy<ST, true, 2>() : (cls == 1 ? reg_occupancy<ST, true, 1>() : reg_occupancy<ST, true, 0>());
  return cls == 2 ? reg_occupancy<ST, false, 2>() : (cls == 1 ? reg_occupancy<ST, false, 1>() : reg_occupancy<ST, false, 0>());
}

template int potrf_batched_reg_blocks_per_cu<double>(char, int);
template int potrf_batched_reg_blocks_per_cu<float>(char, int);
template hipError_t potrf_batched_reg<double>(hipStream_t, char, const int *, const int *, const int *,
                                              const int *, const long *, const int *, double *, double *, int *);
template hipError_t potrf_batched_reg<float>(hipStream_t, char, const int *, const int *, const int *,
                                             const int *, const long *, const int *, float *, double *, int *);

template <typename ST>
hipError_t potrf_batched(hipStream_t s, char uplo, int count, const int *nvec, const long *offs,
                         const int *order, ST *A, double *logdet, int *info, int grid, bool inverse) {
  if (count <= 0) return hipSuccess;
  if (grid < 1) grid = 1;
  const dim3 g((unsigned)grid), b(BT_THREADS);
  if (uplo == 'L') {
    if (inverse)
      hipLaunchKernelGGL((potrf_batched_kernel<ST, true, true>), g, b, 0, s, count, nvec, offs, order, A,
                         logdet, info);
    else
      hipLaunchKernelGGL((potrf_batched_kernel<ST, true, false>), g, b, 0, s, count, nvec, offs, order, A,
                         logdet, info);
  } else {
    if (inverse)
      hipLaunchKernelGGL((potrf_batched_kernel<ST, false, true>), g, b, 0, s, count, nvec, offs, order, A,
                         logdet, info);
    else
      hipLaunchKernelGGL((potrf_batched_kernel<ST, false, false>), g, b, 0, s, count, nvec, offs, order,
                         A, logdet, info);
  }
  return hipGetLastError();
}

template <typename ST>
int potrf_batched_blocks_per_cu(char uplo, bool inverse) {
  int nb = 0;
  hipError_t e;
  if (uplo == 'L')
    e = inverse ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, potrf_batched_kernel<ST, true, true>,
                                                               BT_THREADS, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, potrf_batched_kernel<ST, true, false>,
                                                               BT_THREADS, 0);
  else
    e = inverse ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, potrf_batched_kernel<ST, false, true>,
                                                               BT_THREADS, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, potrf_batched_kernel<ST, false, false>,
                                                               BT_THREADS, 0);
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

template int potrf_batched_blocks_per_cu<double>(char, bool);
template int potrf_batched_blocks_per_cu<float>(char, bool);
template hipError_t potrf_batched<double>(hipStream_t, char, int, const int *, const long *,
                                          const int *, double *, double *, int *, int, bool);
template hipError_t potrf_batched<float>(hipStream_t, char, int, const int *, const long *,
                                         const int *, float *, double *, int *, int, bool);

}  // namespace hc
