// Tile GEMM on CDNA4 matrix cores for FP32 storage (the mixed-precision
// factorisation and its triangular solves).
//
// D <- alpha * op(A) * op(B) + beta * C, column-major FP32, op in {N, T}, FP32
// accumulation.  The products run either in FP32 (v_mfma_f32_16x16x4_f32) or,
// with the operands rounded to bfloat16 (round to nearest even) as they are
// staged, in BF16 (v_mfma_f32_16x16x32_bf16).
//
// A 256-thread workgroup (4 waves, 2 x 2) owns an MT x NT tile of D and walks K
// in steps of 32.  Both operands are staged in LDS row-major with K
// contiguous (op(A) as MT rows, op(B) as NT columns), so every lane fetches
// its MFMA fragment with one 16-byte LDS read: 4 consecutive k (FP32) or 8
// (BF16).  For the FP32 instruction, lane group g supplies k = 4 g + t to the
// t-th of four 4-deep products, a permutation of the 16 k values that A and B
// share, so the sum is unchanged.  The op(B) fragment is the first operand, so
// the 16 x 16 result comes out with D's row index across lanes 0..15 (lane
// l, register r: row l % 16, column 4 (l / 16) + r) and stores write 64-byte
// column segments.  Tiles are walked XCD-contiguously in groups of 8 tile rows
// (see gemm.hip).
#include "internal.h"

#include <cmath>

namespace hc {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int KT = 32;
constexpr int NTHREADS = 256;
constexpr int NXCD = 8;
constexpr int GROUP = 8;

enum { MODE_F32 = 0, MODE_BF16 = 1 };

template <int MODE> struct Stage;
template <> struct Stage<MODE_F32> {
  using type = float;
  static constexpr int LK = KT + 4;  // row stride (elements): 144 bytes
};
template <> struct Stage<MODE_BF16> {
  using type = __bf16;
  static constexpr int LK = KT + 8;  // 80 bytes
};

struct LpParams {
  int m, n, k;
  float alpha, beta;
  const float *A;
  long lda;
  const float *B;
  long ldb;
  const float *C;
  long ldc;
  float *D;
  long ldd;
  int tri_out;
  const int *abort_flag;
};

__device__ __forceinline__ int xcd_contiguous(int w, int nwg) {
  const int x = w % NXCD, i = w / NXCD, base = nwg / NXCD, rem = nwg % NXCD;
  return x < rem ? x * (base + 1) + i : rem * (base + 1) + (x - rem) * base + i;
}

// Global -> register pattern for one operand tile of R x KT elements.
// R_CONTIG: the R index is contiguous in memory (op N for A, op T for B): a
// thread keeps one R index and KT * R / 256 consecutive k.  Otherwise k is
// contiguous: 8 lanes cover 32 consecutive k (4 each) of one R index.
template <int R, bool R_CONTIG>
struct Pattern {
  static constexpr int E = R * KT / NTHREADS;  // elements per thread
  __device__ static int r0(int tid) { return R_CONTIG ? tid % R : tid / 8; }
  __device__ static int k0(int tid) { return R_CONTIG ? (tid / R) * E : (tid % 8) * 4; }
  // element q: (r0 + dr(q), k0 + dk(q))
  __device__ static int dr(int q) { return R_CONTIG ? 0 : (q / 4) * (NTHREADS / 8); }
  __device__ static int dk(int q) { return R_CONTIG ? q : q % 4; }
};

template <int MODE, int TA, int TB, int MT, int NT, bool TG>
__global__ void __launch_bounds__(NTHREADS, 2) gemm_lp_kernel(LpParams p) {
  if (p.abort_flag != nullptr && *p.abort_flag != 0) return;
  using SE = typename Stage<MODE>::type;
  constexpr int LK = Stage<MODE>::LK;
  constexpr int WM = MT / 2, WN = NT / 2, BI = WM / 16, BJ = WN / 16;
  using PA = Pattern<MT, TA == OP_N>;
  using PB = Pattern<NT, TB != OP_N>;
  constexpr int EA = PA::E, EB = PB::E;
  static_assert(EA % 4 == 0 && EB % 4 == 0, "staging is in groups of 4");

  __shared__ __attribute__((aligned(16))) SE sA[2][MT * LK];
  __shared__ __attribute__((aligned(16))) SE sB[2][NT * LK];

  // ---- tile coordinates (XCD-contiguous, grouped) ----
  const int nwg = TG ? (int)gridDim.x : (int)(gridDim.x * gridDim.y);
  const int wid = TG ? (int)blockIdx.x : (int)(blockIdx.x + blockIdx.y * gridDim.x);
  const int logical = xcd_contiguous(wid, nwg);
  int bm, bn;
  if constexpr (TG) {
    const int id = logical;
    int r = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= id) r++;
    while (r * (r + 1) / 2 > id) r--;
    const int nt = (int)((sqrt(8.0 * (double)nwg + 1.0) - 1.0) * 0.5 + 0.5);
    const int first = (r / GROUP) * GROUP, gs = min(GROUP, nt - first);
    int o = id - first * (first + 1) / 2, tr, tc;
    if (o < gs * first) {
      tc = o / gs;
      tr = first + o % gs;
    } else {
      o -= gs * first;
      int cl = 0;
      while (o >= gs - cl) {
        o -= gs - cl;
        cl++;
      }
      tc = first + cl;
      tr = tc + o;
    }
    if (p.tri_out == TRI_LOWER) { bm = tr; bn = tc; } else { bm = tc; bn = tr; }
  } else {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int per = GROUP * gy, first = (logical / per) * GROUP, gs = min(GROUP, gx - first);
    const int o = logical % per;
    bm = first + o % gs;
    bn = o / gs;
  }
  const int m0 = bm * MT, n0 = bn * NT;
  const int kend = p.alpha == 0.0f ? 0 : p.k;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;

  v4f acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; i++)
#pragma unroll
    for (int j = 0; j < BJ; j++) acc[i][j] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};

  if (kend > 0) {
    const int ar = m0 + PA::r0(tid), ak = PA::k0(tid);
    const int br = n0 + PB::r0(tid), bk = PB::k0(tid);
    float ra[EA], rb[EB];
    // Interior tiles with whole K steps load through per-thread pointers
    // advanced by one K step at a time: every element offset is a
    // compile-time multiple of the (uniform) leading dimension, so the loads
    // carry no per-element address arithmetic and no predication.
    const bool interior = m0 + MT <= p.m && n0 + NT <= p.n && kend % KT == 0;
    const float *pa = (TA == OP_N) ? p.A + (long)ak * p.lda + ar : p.A + (long)ar * p.lda + ak;
    const float *pb = (TB == OP_N) ? p.B + (long)br * p.ldb + bk : p.B + (long)bk * p.ldb + br;
    const long a_kstride = (TA == OP_N) ? (long)KT * p.lda : (long)KT;
    const long b_kstride = (TB == OP_N) ? (long)KT : (long)KT * p.ldb;
    auto load = [&](int kb) {
      if (interior) {
#pragma unroll
        for (int q = 0; q < EA; q++)
          ra[q] = pa[(TA == OP_N) ? (long)PA::dk(q) * p.lda + PA::dr(q) : (long)PA::dr(q) * p.lda + PA::dk(q)];
#pragma unroll
        for (int q = 0; q < EB; q++)
          rb[q] = pb[(TB == OP_N) ? (long)PB::dr(q) * p.ldb + PB::dk(q) : (long)PB::dk(q) * p.ldb + PB::dr(q)];
        return;
      }
#pragma unroll
      for (int q = 0; q < EA; q++) {
        const int gi = ar + PA::dr(q), gk = kb + ak + PA::dk(q);
        const bool ok = gi < p.m && gk < kend;
        const long off = (TA == OP_N) ? (long)gk * p.lda + gi : (long)gi * p.lda + gk;
        const float v = p.A[ok ? off : 0];
        ra[q] = ok ? v : 0.0f;
      }
#pragma unroll
      for (int q = 0; q < EB; q++) {
        const int gj = br + PB::dr(q), gk = kb + bk + PB::dk(q);
        const bool ok = gj < p.n && gk < kend;
        const long off = (TB == OP_N) ? (long)gj * p.ldb + gk : (long)gk * p.ldb + gj;
        const float v = p.B[ok ? off : 0];
        rb[q] = ok ? v : 0.0f;
      }
    };
    auto store = [&](int buf) {
#pragma unroll
      for (int q = 0; q < EA; q += 4) {
        SE *d = &sA[buf][(PA::r0(tid) + PA::dr(q)) * LK + ak + PA::dk(q)];
#pragma unroll
        for (int e = 0; e < 4; e++) d[e] = (SE)ra[q + e];
      }
#pragma unroll
      for (int q = 0; q < EB; q += 4) {
        SE *d = &sB[buf][(PB::r0(tid) + PB::dr(q)) * LK + bk + PB::dk(q)];
#pragma unroll
        for (int e = 0; e < 4; e++) d[e] = (SE)rb[q + e];
      }
    };

    const int ntiles = (kend + KT - 1) / KT;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < ntiles; kt++) {
      const bool more = kt + 1 < ntiles;
      if (more) {
        pa += a_kstride;
        pb += b_kstride;
        load((kt + 1) * KT);
      }
      const SE *cA = sA[kt & 1], *cB = sB[kt & 1];
      if constexpr (MODE == MODE_F32) {
#pragma unroll
        for (int h = 0; h < KT; h += 16) {
          v4f fa[BI], fb[BJ];
#pragma unroll
          for (int i = 0; i < BI; i++)
            fa[i] = *reinterpret_cast<const v4f *>(&cA[(wm0 + 16 * i + lc) * LK + h + 4 * lg]);
#pragma unroll
          for (int j = 0; j < BJ; j++)
            fb[j] = *reinterpret_cast<const v4f *>(&cB[(wn0 + 16 * j + lc) * LK + h + 4 * lg]);
#pragma unroll
          for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < BI; i++)
#pragma unroll
              for (int j = 0; j < BJ; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j][t], fa[i][t], acc[i][j], 0, 0, 0);
        }
      } else {
        v8bf fa[BI], fb[BJ];
#pragma unroll
        for (int i = 0; i < BI; i++)
          fa[i] = *reinterpret_cast<const v8bf *>(&cA[(wm0 + 16 * i + lc) * LK + 8 * lg]);
#pragma unroll
        for (int j = 0; j < BJ; j++)
          fb[j] = *reinterpret_cast<const v8bf *>(&cB[(wn0 + 16 * j + lc) * LK + 8 * lg]);
#pragma unroll
        for (int i = 0; i < BI; i++)
#pragma unroll
          for (int j = 0; j < BJ; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0);
      }
      if (more) store((kt + 1) & 1);
      __syncthreads();
    }
  }

  // ---- epilogue: lane lc holds row m0 + wm0 + 16 i + lc; register r of lane
  // group lg holds column n0 + wn0 + 16 j + 4 lg + r.  With beta != 0, C is
  // read in one batch from clamped (valid) addresses, so that all its loads
  // are in flight together, then combined.  C enters here and not as the
  // accumulators' starting value: in a trailing update C and the products
  // nearly cancel, and FP32 sums that start from C round every partial sum
  // at C's magnitude (measured: the FP32 factor of a cond ~4e6 matrix then
  // no longer refines) ----
#pragma unroll
  for (int i = 0; i < BI; i++)
#pragma unroll
    for (int j = 0; j < BJ; j++) acc[i][j] *= p.alpha;
  if (p.beta != 0.0f) {
#pragma unroll
    for (int i = 0; i < BI; i++) {
      const int gm = min(m0 + wm0 + 16 * i + lc, p.m - 1);
#pragma unroll
      for (int j = 0; j < BJ; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gn = min(n0 + wn0 + 16 * j + 4 * lg + r, p.n - 1);
          acc[i][j][r] += p.beta * p.C[(long)gn * p.ldc + gm];
        }
    }
  }
#pragma unroll
  for (int i = 0; i < BI; i++) {
    const int gm = m0 + wm0 + 16 * i + lc;
    if (gm >= p.m) continue;
#pragma unroll
    for (int j = 0; j < BJ; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gn = n0 + wn0 + 16 * j + 4 * lg + r;
        if (gn >= p.n) continue;
        if (p.tri_out == TRI_LOWER && gm < gn) continue;
        if (p.tri_out == TRI_UPPER && gm > gn) continue;
        p.D[(long)gn * p.ldd + gm] = acc[i][j][r];
      }
    }
  }
}

template <int MODE, int TA, int TB, int MT, int NT>
hipError_t launch(hipStream_t s, const LpParams &p, bool tri_grid) {
  if (tri_grid) {
    if constexpr (MT == NT) {
      const long t = (p.m + MT - 1) / MT;
      hipLaunchKernelGGL((gemm_lp_kernel<MODE, TA, TB, MT, NT, true>),
                         dim3((unsigned)(t * (t + 1) / 2)), dim3(NTHREADS), 0, s, p);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  const dim3 grid((unsigned)((p.m + MT - 1) / MT), (unsigned)((p.n + NT - 1) / NT));
  hipLaunchKernelGGL((gemm_lp_kernel<MODE, TA, TB, MT, NT, false>), grid, dim3(NTHREADS), 0, s, p);
  return hipGetLastError();
}

template <int MODE, int TA, int TB>
hipError_t dispatch_tile(hipStream_t s, const LpParams &p, int mt, int nt, bool tri_grid) {
  if (mt == 128 && nt == 128) return launch<MODE, TA, TB, 128, 128>(s, p, tri_grid);
  if (mt == 64 && nt == 64) return launch<MODE, TA, TB, 64, 64>(s, p, tri_grid);
  if (mt == 64 && nt == 128) return launch<MODE, TA, TB, 64, 128>(s, p, tri_grid);
  return launch<MODE, TA, TB, 128, 64>(s, p, tri_grid);
}

template <int MODE>
hipError_t dispatch_ops(hipStream_t s, int ta, int tb, const LpParams &p, int mt, int nt,
                        bool tri_grid) {
  const bool a = ta != OP_N, b = tb != OP_N;
  if (!a && !b) return dispatch_tile<MODE, OP_N, OP_N>(s, p, mt, nt, tri_grid);
  if (!a && b) return dispatch_tile<MODE, OP_N, OP_T>(s, p, mt, nt, tri_grid);
  if (a && !b) return dispatch_tile<MODE, OP_T, OP_N>(s, p, mt, nt, tri_grid);
  return dispatch_tile<MODE, OP_T, OP_T>(s, p, mt, nt, tri_grid);
}

// ---------------------------------------------------------------------------
// BF16 products from BF16 operands stored with K contiguous (the panel of the
// mixed-precision factorisation, converted once per panel):
//   D (FP32, column-major) <- alpha * A B^T + beta * C,
//   A: m x k, row r at A + r * lda; B: n x k, row c at B + c * ldb.
// K is a multiple of BK and the operand rows are zero beyond k.  Both tiles
// go global -> LDS with 16-byte loads and stores (8 k values), and every MFMA
// fragment is one 16-byte LDS read.
// ---------------------------------------------------------------------------
constexpr int BK = 64;           // K step
constexpr int BLK = BK + 8;      // LDS row stride (bf16 elements): 144 bytes

typedef unsigned short u16;
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

struct Bf16Params {
  int m, n, k;
  float alpha, beta;
  const u16 *A;
  long lda;
  const u16 *B;
  long ldb;
  const float *C;
  long ldc;
  float *D;
  long ldd;
  int tri_out;
  const int *abort_flag;
};

typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// Tile (bm, bn) of this workgroup: XCD-contiguous, in groups of GROUP tile
// rows; on a triangular grid (TG) the lower (tri_out TRI_LOWER) or upper
// triangle of tiles, numbered row by row.
template <bool TG>
__device__ __forceinline__ void tile16_coords(int tri_out, int &bm, int &bn) {
  const int nwg = TG ? (int)gridDim.x : (int)(gridDim.x * gridDim.y);
  const int wid = TG ? (int)blockIdx.x : (int)(blockIdx.x + blockIdx.y * gridDim.x);
  const int logical = xcd_contiguous(wid, nwg);
  if constexpr (TG) {
    const int id = logical;
    int r = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= id) r++;
    while (r * (r + 1) / 2 > id) r--;
    const int nt = (int)((sqrt(8.0 * (double)nwg + 1.0) - 1.0) * 0.5 + 0.5);
    const int first = (r / GROUP) * GROUP, gs = min(GROUP, nt - first);
    int o = id - first * (first + 1) / 2, tr, tc;
    if (o < gs * first) {
      tc = o / gs;
      tr = first + o % gs;
    } else {
      o -= gs * first;
      int cl = 0;
      while (o >= gs - cl) {
        o -= gs - cl;
        cl++;
      }
      tc = first + cl;
      tr = tc + o;
    }
    if (tri_out == TRI_LOWER) { bm = tr; bn = tc; } else { bm = tc; bn = tr; }
  } else {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int per = GROUP * gy, first = (logical / per) * GROUP, gs = min(GROUP, gx - first);
    const int o = logical % per;
    bm = first + o % gs;
    bn = o / gs;
  }
}

// F16: the 16-bit copies are IEEE binary16 (v_mfma_f32_16x16x32_f16) instead
// of bfloat16; the two formats share the layout and the sign bit.
template <int MT, int NT, bool TG, bool F16>
__global__ void __launch_bounds__(NTHREADS, 2) gemm_bf16_nt_kernel(Bf16Params p) {
  if (p.abort_flag != nullptr && *p.abort_flag != 0) return;
  constexpr int WM = MT / 2, WN = NT / 2, BI = WM / 16, BJ = WN / 16;
  constexpr int CA = MT * BK / 8 / NTHREADS;  // 16-byte chunks per thread
  constexpr int CB = NT * BK / 8 / NTHREADS;
  static_assert(CA >= 1 && CB >= 1, "tile too small");
  __shared__ __attribute__((aligned(16))) u16 sA[2][MT * BLK];
  __shared__ __attribute__((aligned(16))) u16 sB[2][NT * BLK];

  int bm, bn;
  tile16_coords<TG>(p.tri_out, bm, bn);
  const int m0 = bm * MT, n0 = bn * NT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;

  // |alpha| is a power of two (the launcher checks): the accumulators start
  // from (beta / |alpha|) C (loaded at the start, so the C read overlaps the
  // whole K loop), the sign of alpha goes into the A fragments and the sums
  // are scaled by |alpha| at the end; both scalings are exact.
  const float aabs = fabsf(p.alpha);
  const bool use_beta = p.beta != 0.0f;
  v4f acc[BI][BJ];
  // C loads from clamped (valid) addresses with the condition outside the
  // loops, so that all of them are in flight together
  auto load_c = [&](bool unit, float scale) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < BI; i++) {
      const int gm = min(m0 + wm0 + 16 * i + lc, p.m - 1);
#pragma unroll
      for (int j = 0; j < BJ; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gn = min(n0 + wn0 + 16 * j + 4 * lg + r, p.n - 1);
          const float c = p.C[(long)gn * p.ldc + gm];
          acc[i][j][r] = unit ? c : scale * c;
        }
    }
  };
  // issued after the first operand loads, consumed by the first products
  auto seed_acc = [&]() __attribute__((always_inline)) {
    if (use_beta) {
      const float cs = p.beta / aabs;
      if (cs == 1.0f) load_c(true, 1.0f);
      else load_c(false, cs);
    } else {
#pragma unroll
      for (int i = 0; i < BI; i++)
#pragma unroll
        for (int j = 0; j < BJ; j++) acc[i][j] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
    }
  };
  const unsigned sgn = p.alpha < 0.0f ? 0x80008000u : 0u;  // bf16 sign bits of a fragment

  const int kend = p.k;
  if (kend > 0) {
    // chunk q of a thread: row (tid + q 256) / 8, k offset 8 ((tid + q 256) % 8).
    // Two register sets: the loads of K step kt + 2 are issued while step kt
    // is computed and step kt + 1 waits in registers for its LDS buffer.
    v4u ra[2][CA], rb[2][CB];
    auto load = [&](int set, int kb) __attribute__((always_inline)) {
#pragma unroll
      for (int q = 0; q < CA; q++) {
        const int e = tid + q * NTHREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        const int gr = m0 + row;
        const v4u *src = reinterpret_cast<const v4u *>(p.A + (long)min(gr, p.m - 1) * p.lda + kb + kc);
        const unsigned msk = gr < p.m ? 0xffffffffu : 0u;  // rows past m read as zero
        ra[set][q] = *src & (v4u){msk, msk, msk, msk};
      }
#pragma unroll
      for (int q = 0; q < CB; q++) {
        const int e = tid + q * NTHREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        const int gc = n0 + row;
        const v4u *src = reinterpret_cast<const v4u *>(p.B + (long)min(gc, p.n - 1) * p.ldb + kb + kc);
        const unsigned msk = gc < p.n ? 0xffffffffu : 0u;
        rb[set][q] = *src & (v4u){msk, msk, msk, msk};
      }
    };
    auto store = [&](int set, int buf) __attribute__((always_inline)) {
#pragma unroll
      for (int q = 0; q < CA; q++) {
        const int e = tid + q * NTHREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        *reinterpret_cast<v4u *>(&sA[buf][row * BLK + kc]) = ra[set][q];
      }
#pragma unroll
      for (int q = 0; q < CB; q++) {
        const int e = tid + q * NTHREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        *reinterpret_cast<v4u *>(&sB[buf][row * BLK + kc]) = rb[set][q];
      }
    };
    auto compute = [&](int buf) __attribute__((always_inline)) {
      const u16 *cA = sA[buf], *cB = sB[buf];
#pragma unroll
      for (int h = 0; h < BK; h += 32) {
        v8bf fa[BI], fb[BJ];
#pragma unroll
        for (int i = 0; i < BI; i++) {
          v4u x = *reinterpret_cast<const v4u *>(&cA[(wm0 + 16 * i + lc) * BLK + h + 8 * lg]);
          x ^= (v4u){sgn, sgn, sgn, sgn};
          fa[i] = __builtin_bit_cast(v8bf, x);
        }
#pragma unroll
        for (int j = 0; j < BJ; j++)
          fb[j] = *reinterpret_cast<const v8bf *>(&cB[(wn0 + 16 * j + lc) * BLK + h + 8 * lg]);
#pragma unroll
        for (int i = 0; i < BI; i++)
#pragma unroll
          for (int j = 0; j < BJ; j++) {
            if constexpr (F16)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, fb[j]),
                                                                 __builtin_bit_cast(v8h, fa[i]), acc[i][j],
                                                                 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0);
          }
      }
    };
    const int ntiles = kend / BK;
    load(0, 0);
    load(1, ntiles > 1 ? BK : 0);  // a repeat of step 0 when there is no step 1 (never stored)
    seed_acc();
    store(0, 0);
    __syncthreads();
    // step kt: compute buffer kt & 1 (set kt & 1 is free: it was stored);
    // set (kt + 1) & 1 holds step kt + 1.  Two steps per pass keep the set
    // and buffer indices static.
    for (int kt = 0; kt < ntiles; kt += 2) {
      if (kt + 2 < ntiles) load(0, (kt + 2) * BK);
      compute(0);
      if (kt + 1 < ntiles) store(1, 1);
      __syncthreads();
      if (kt + 1 >= ntiles) break;
      if (kt + 3 < ntiles) load(1, (kt + 3) * BK);
      compute(1);
      if (kt + 2 < ntiles) store(0, 0);
      __syncthreads();
    }
  } else {
    seed_acc();
  }

  if (aabs != 1.0f) {
#pragma unroll
    for (int i = 0; i < BI; i++)
#pragma unroll
      for (int j = 0; j < BJ; j++) acc[i][j] *= aabs;
  }
#pragma unroll
  for (int i = 0; i < BI; i++) {
    const int gm = m0 + wm0 + 16 * i + lc;
    if (gm >= p.m) continue;
#pragma unroll
    for (int j = 0; j < BJ; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gn = n0 + wn0 + 16 * j + 4 * lg + r;
        if (gn >= p.n) continue;
        if (p.tri_out == TRI_LOWER && gm < gn) continue;
        if (p.tri_out == TRI_UPPER && gm > gn) continue;
        p.D[(long)gn * p.ldd + gm] = acc[i][j][r];
      }
    }
  }
}

// The same product on 256 x 256 tiles for large updates: 8 waves (a 2 x 4
// grid, each wave 128 x 64 of D), so each staged operand byte feeds twice the
// products of the 128 x 128 kernel (half the L2 traffic per flop, and 12
// fragment reads per 32 products per wave instead of 8 per 16).  LDS holds
// two K steps of both operands (2 x 2 x 256 x BLK x 2 bytes = 144 KiB: one
// workgroup per CU); one register set prefetches the next K step while the
// current one is computed.
constexpr int T256_THREADS = 512;

template <bool TG, bool F16>
__global__ void __launch_bounds__(T256_THREADS, 1) gemm16_nt_256_kernel(Bf16Params p) {
  if (p.abort_flag != nullptr && *p.abort_flag != 0) return;
  constexpr int MT = 256, NT = 256, WGM = 2;
  constexpr int WM = MT / WGM, WN = NT / (T256_THREADS / 64 / WGM), BI = WM / 16, BJ = WN / 16;
  constexpr int CA = MT * BK / 8 / T256_THREADS;  // 16-byte chunks per thread
  constexpr int CB = NT * BK / 8 / T256_THREADS;
  static_assert(BI == 8 && BJ == 4 && CA == 4 && CB == 4, "256 x 256 tile, 8 waves of 128 x 64");
  __shared__ __attribute__((aligned(16))) u16 sA[2][MT * BLK];
  __shared__ __attribute__((aligned(16))) u16 sB[2][NT * BLK];

  int bm, bn;
  tile16_coords<TG>(p.tri_out, bm, bn);
  const int m0 = bm * MT, n0 = bn * NT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int wm0 = (wave % WGM) * WM, wn0 = (wave / WGM) * WN;

  // accumulators seeded with (beta / |alpha|) C, sign of alpha in the A
  // fragments, |alpha| applied at the end (a power of two: exact), as in
  // gemm_bf16_nt_kernel
  const float aabs = fabsf(p.alpha);
  const bool use_beta = p.beta != 0.0f;
  v4f acc[BI][BJ];
  auto seed_acc = [&]() __attribute__((always_inline)) {
    if (use_beta) {
      const float cs = p.beta / aabs;
#pragma unroll
      for (int i = 0; i < BI; i++) {
        const int gm = min(m0 + wm0 + 16 * i + lc, p.m - 1);
#pragma unroll
        for (int j = 0; j < BJ; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int gn = min(n0 + wn0 + 16 * j + 4 * lg + r, p.n - 1);
            acc[i][j][r] = cs * p.C[(long)gn * p.ldc + gm];
          }
      }
    } else {
#pragma unroll
      for (int i = 0; i < BI; i++)
#pragma unroll
        for (int j = 0; j < BJ; j++) acc[i][j] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
    }
  };
  const unsigned sgn = p.alpha < 0.0f ? 0x80008000u : 0u;

  const int kend = p.k;
  if (kend > 0) {
    v4u ra[CA], rb[CB];
    auto load = [&](int kb) __attribute__((always_inline)) {
#pragma unroll
      for (int q = 0; q < CA; q++) {
        const int e = tid + q * T256_THREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        const int gr = m0 + row;
        const v4u *src = reinterpret_cast<const v4u *>(p.A + (long)min(gr, p.m - 1) * p.lda + kb + kc);
        const unsigned msk = gr < p.m ? 0xffffffffu : 0u;  // rows past m read as zero
        ra[q] = *src & (v4u){msk, msk, msk, msk};
      }
#pragma unroll
      for (int q = 0; q < CB; q++) {
        const int e = tid + q * T256_THREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        const int gc = n0 + row;
        const v4u *src = reinterpret_cast<const v4u *>(p.B + (long)min(gc, p.n - 1) * p.ldb + kb + kc);
        const unsigned msk = gc < p.n ? 0xffffffffu : 0u;
        rb[q] = *src & (v4u){msk, msk, msk, msk};
      }
    };
    auto store = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
      for (int q = 0; q < CA; q++) {
        const int e = tid + q * T256_THREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        *reinterpret_cast<v4u *>(&sA[buf][row * BLK + kc]) = ra[q];
      }
#pragma unroll
      for (int q = 0; q < CB; q++) {
        const int e = tid + q * T256_THREADS, row = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        *reinterpret_cast<v4u *>(&sB[buf][row * BLK + kc]) = rb[q];
      }
    };
    auto compute = [&](int buf) __attribute__((always_inline)) {
      const u16 *cA = sA[buf], *cB = sB[buf];
#pragma unroll
      for (int h = 0; h < BK; h += 32) {
        v8bf fb[BJ];
#pragma unroll
        for (int j = 0; j < BJ; j++)
          fb[j] = *reinterpret_cast<const v8bf *>(&cB[(wn0 + 16 * j + lc) * BLK + h + 8 * lg]);
#pragma unroll
        for (int i = 0; i < BI; i++) {
          v4u x = *reinterpret_cast<const v4u *>(&cA[(wm0 + 16 * i + lc) * BLK + h + 8 * lg]);
          x ^= (v4u){sgn, sgn, sgn, sgn};
          const v8bf fa = __builtin_bit_cast(v8bf, x);
#pragma unroll
          for (int j = 0; j < BJ; j++) {
            if constexpr (F16)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, fb[j]),
                                                                 __builtin_bit_cast(v8h, fa), acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa, acc[i][j], 0, 0, 0);
          }
        }
      }
    };
    const int ntiles = kend / BK;
    load(0);
    seed_acc();
    store(0);
    __syncthreads();
    for (int kt = 0; kt < ntiles; kt++) {
      // buffer (kt + 1) & 1 was last read in step kt - 1, before the barrier
      const bool more = kt + 1 < ntiles;
      if (more) load((kt + 1) * BK);
      compute(kt & 1);
      if (more) store((kt + 1) & 1);
      __syncthreads();
    }
  } else {
    seed_acc();
  }

  if (aabs != 1.0f) {
#pragma unroll
    for (int i = 0; i < BI; i++)
#pragma unroll
      for (int j = 0; j < BJ; j++) acc[i][j] *= aabs;
  }
#pragma unroll
  for (int i = 0; i < BI; i++) {
    const int gm = m0 + wm0 + 16 * i + lc;
    if (gm >= p.m) continue;
#pragma unroll
    for (int j = 0; j < BJ; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gn = n0 + wn0 + 16 * j + 4 * lg + r;
        if (gn >= p.n) continue;
        if (p.tri_out == TRI_LOWER && gm < gn) continue;
        if (p.tri_out == TRI_UPPER && gm > gn) continue;
        p.D[(long)gn * p.ldd + gm] = acc[i][j][r];
      }
    }
  }
}

// T (bf16, row r at T + r * ldt, k contiguous) <- P (FP32, column-major, ld
// ldp) for rows 0..m-1 and k < kp, zeros for kp <= k < ldt.
__global__ void panel_to_bf16_kernel(int m, int kp, const float *P, long ldp, u16 *T, long ldt) {
  __shared__ float t[32][33];
  const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + tx, k = k0 + j;
    t[j][tx] = (r < m && k < kp) ? P[(long)k * ldp + r] : 0.0f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, k = k0 + tx;
    if (r < m && k < ldt) {
      const __bf16 b = (__bf16)t[tx][j];
      T[(long)r * ldt + k] = __builtin_bit_cast(u16, b);
    }
  }
}

// T (binary16, row r at T + r * ldt, k contiguous) <- scale P for rows 0..m-1
// and k < kp, zeros for kp <= k < ldt.  P starts at a diagonal element: its
// strictly upper part (r < k) is not referenced and is stored as zero, so
// that whatever that triangle holds cannot be taken for an overflow.  scale
// is a power of two chosen so that every entry of an SPD factor is below
// 2^15 in magnitude; an entry that is not (or not finite) records column
// col0 + k + 1 in *info, which stops the factorisation as a failed pivot would.
__global__ void panel_to_f16_kernel(int m, int kp, const float *P, long ldp, u16 *T, long ldt, float scale,
                                    int *info, long col0) {
  __shared__ float t[32][33];
  const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + tx, k = k0 + j;
    t[j][tx] = (r < m && k < kp) ? scale * P[(long)k * ldp + r] : 0.0f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, k = k0 + tx;
    if (r < m && k < ldt) {
      const float x = r >= k ? t[tx][j] : 0.0f;
      if (!(fabsf(x) <= 65504.0f)) atomicCAS(info, 0, (int)(col0 + k + 1));
      const _Float16 h = (_Float16)x;
      T[(long)r * ldt + k] = __builtin_bit_cast(u16, h);
    }
  }
}

// Split copy for three-product BF16 updates: x = hi + lo (hi = bf16(x),
// lo = bf16(x - hi)), stored as rows of 3 ldt values
//   TA row r: [hi | hi | lo],  TB row r: [hi | lo | hi]
// so that TA TB^T = hi hi^T + hi lo^T + lo hi^T (the lo lo^T term, 2^-16
// relative, is dropped); zeros for kp <= k < ldt.
__global__ void panel_to_bf16_split_kernel(int m, int kp, const float *P, long ldp, u16 *TA, u16 *TB,
                                           long ldt) {
  __shared__ float t[32][33];
  const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + tx, k = k0 + j;
    t[j][tx] = (r < m && k < kp) ? P[(long)k * ldp + r] : 0.0f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, k = k0 + tx;
    if (r < m && k < ldt) {
      const float x = t[tx][j];
      const __bf16 hi = (__bf16)x;
      const __bf16 lo = (__bf16)(x - (float)hi);
      const u16 h = __builtin_bit_cast(u16, hi), l = __builtin_bit_cast(u16, lo);
      u16 *a = TA + (long)r * 3 * ldt + k, *b = TB + (long)r * 3 * ldt + k;
      a[0] = h;
      a[ldt] = h;
      a[2 * ldt] = l;
      b[0] = h;
      b[ldt] = l;
      b[2 * ldt] = h;
    }
  }
}

}  // namespace

hipError_t panel_to_bf16_split(hipStream_t s, long m, long kp, const float *P, long ldp, void *TA, void *TB,
                               long ldt) {
  if (m <= 0) return hipSuccess;
  const dim3 grid((unsigned)((m + 31) / 32), (unsigned)((ldt + 31) / 32));
  hipLaunchKernelGGL(panel_to_bf16_split_kernel, grid, dim3(256), 0, s, (int)m, (int)kp, P, ldp,
                     static_cast<u16 *>(TA), static_cast<u16 *>(TB), ldt);
  return hipGetLastError();
}

hipError_t panel_to_bf16(hipStream_t s, long m, long kp, const float *P, long ldp, void *T, long ldt) {
  if (m <= 0) return hipSuccess;
  const dim3 grid((unsigned)((m + 31) / 32), (unsigned)((ldt + 31) / 32));
  hipLaunchKernelGGL(panel_to_bf16_kernel, grid, dim3(256), 0, s, (int)m, (int)kp, P, ldp,
                     static_cast<u16 *>(T), ldt);
  return hipGetLastError();
}

hipError_t panel_to_f16(hipStream_t s, long m, long kp, const float *P, long ldp, void *T, long ldt,
                        float scale, int *info, long col0) {
  if (m <= 0) return hipSuccess;
  const dim3 grid((unsigned)((m + 31) / 32), (unsigned)((ldt + 31) / 32));
  hipLaunchKernelGGL(panel_to_f16_kernel, grid, dim3(256), 0, s, (int)m, (int)kp, P, ldp,
                     static_cast<u16 *>(T), ldt, scale, info, col0);
  return hipGetLastError();
}

// Tile choice: 256 x 256 (one 8-wave workgroup per CU) when the grid has at
// least the measured number of such tiles (family W; by default two per CU),
// 128 x 128 when it has enough of those (family H; about one per CU), else
// 64 x 64.
template <bool F16>
hipError_t launch_nt16(hipStream_t s, const Bf16Params &p, long m, long n, bool tri_grid) {
  const long tm = (m + 127) / 128, tn = (n + 127) / 128;
  const long tiles = tri_grid ? tm * (tm + 1) / 2 : tm * tn;
  const long um = (m + 255) / 256, un = (n + 255) / 256;
  const long tiles256 = tri_grid ? um * (um + 1) / 2 : um * un;
  if (tiles256 >= tuned_tile_min('W', 512)) {
    if (tri_grid) {
      if (m != n) return hipErrorInvalidValue;
      hipLaunchKernelGGL((gemm16_nt_256_kernel<true, F16>), dim3((unsigned)tiles256), dim3(T256_THREADS), 0, s, p);
    } else {
      hipLaunchKernelGGL((gemm16_nt_256_kernel<false, F16>), dim3((unsigned)um, (unsigned)un), dim3(T256_THREADS),
                         0, s, p);
    }
    return hipGetLastError();
  }
  const bool big = tiles >= tuned_tile_min('H', 240);
  if (tri_grid) {
    if (m != n) return hipErrorInvalidValue;
    if (big) {
      hipLaunchKernelGGL((gemm_bf16_nt_kernel<128, 128, true, F16>), dim3((unsigned)tiles),
                         dim3(NTHREADS), 0, s, p);
    } else {
      const long t = (m + 63) / 64;
      hipLaunchKernelGGL((gemm_bf16_nt_kernel<64, 64, true, F16>), dim3((unsigned)(t * (t + 1) / 2)),
                         dim3(NTHREADS), 0, s, p);
    }
    return hipGetLastError();
  }
  if (big)
    hipLaunchKernelGGL((gemm_bf16_nt_kernel<128, 128, false, F16>), dim3((unsigned)tm, (unsigned)tn),
                       dim3(NTHREADS), 0, s, p);
  else
    hipLaunchKernelGGL((gemm_bf16_nt_kernel<64, 64, false, F16>),
                       dim3((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64)), dim3(NTHREADS), 0, s, p);
  return hipGetLastError();
}

hipError_t gemm_bf16_nt(hipStream_t s, long m, long n, long k, float alpha, const void *A, long lda,
                        const void *B, long ldb, float beta, const float *C, long ldc, float *D,
                        long ldd, int tri_out, bool tri_grid, const int *abort_flag, bool f16) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k % BK != 0 || lda % 8 != 0 || ldb % 8 != 0) return hipErrorInvalidValue;
  // |alpha| must be a normal power of two (the scaling of the seeded
  // accumulators is then exact); every caller passes -2^-2e.
  int e2 = 0;
  const float mant = std::frexp(std::fabs(alpha), &e2);
  if (!(mant == 0.5f) || !std::isnormal(alpha)) return hipErrorInvalidValue;
  Bf16Params p;
  p.m = (int)m; p.n = (int)n; p.k = (int)k;
  p.alpha = alpha; p.beta = beta;
  p.A = static_cast<const u16 *>(A); p.lda = lda;
  p.B = static_cast<const u16 *>(B); p.ldb = ldb;
  p.C = C; p.ldc = ldc; p.D = D; p.ldd = ldd;
  p.tri_out = tri_out;
  p.abort_flag = abort_flag;
  return f16 ? launch_nt16<true>(s, p, m, n, tri_grid) : launch_nt16<false>(s, p, m, n, tri_grid);
}

template <>
hipError_t gemm<float>(hipStream_t s, int ta, int tb, long m, long n, long k, float alpha,
                       const float *A, long lda, const float *B, long ldb, float beta,
                       const float *C, long ldc, float *D, long ldd, const GemmOpts &o) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (o.tri_a != TRI_NONE || o.tri_b != TRI_NONE)
    return gemm_tile<float>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, D, ldd, o);
  LpParams p;
  p.m = (int)m; p.n = (int)n; p.k = (int)(k > 0 ? k : 0);
  p.alpha = alpha; p.beta = beta;
  p.A = A; p.lda = lda; p.B = B; p.ldb = ldb; p.C = C; p.ldc = ldc; p.D = D; p.ldd = ldd;
  p.tri_out = o.tri_out;
  p.abort_flag = o.abort_flag;
  const long tb_m = (m + 127) / 128, tb_n = (n + 127) / 128;
  const long tiles_big = o.tri_grid ? tb_m * (tb_m + 1) / 2 : tb_m * tb_n;
  int mt = tiles_big >= tuned_tile_min('S', 240) ? 128 : 64, nt = mt;
  if (o.min_mt > mt) mt = 128;
  if (o.min_nt > nt) nt = 128;
  if (o.min_mt > 128 || o.min_nt > 128) return hipErrorInvalidValue;
  if (o.tri_grid) {
    if (m != n) return hipErrorInvalidValue;
    mt = nt = (mt > nt ? mt : nt);
  }
  if (o.bf16) return dispatch_ops<MODE_BF16>(s, ta, tb, p, mt, nt, o.tri_grid);
  return dispatch_ops<MODE_F32>(s, ta, tb, p, mt, nt, o.tri_grid);
}

}  // namespace hc
