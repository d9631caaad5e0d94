// Tile GEMM on CDNA4 matrix cores: double and double complex on the FP64
// matrix cores (v_mfma_f64_16x16x4_f64), single and single complex on the
// FP32 ones (v_mfma_f32_16x16x4_f32, twice the rate).
//
// D <- alpha * op(A) * op(B) + beta * C, column-major, op in {N, T, C}.  One
// kernel serves every BLAS-3 shape on the Cholesky path:
//   * GEMM                (reference blas/dgemm.cu:198-324, zgemm.cu:53-330)
//   * SYRK / HERK         tri_out + tri_grid (reference blas/dsyrk.cu:261-455)
//   * TRMM                tri_a / tri_b masks (reference blas/dtrmm.cu:941-1727)
//   * TRSM block steps    with inverted diagonal blocks (blas_ops.cpp)
//
// Work decomposition: a 256-thread workgroup (4 waves in a 2 x 2 grid) owns an
// MT x NT tile of D and walks K in steps of 16.  Each step stages op(A)
// (MT x 16) and op(B) (16 x NT) into LDS, transposed so that both are stored
// k-major with the M / N index contiguous; the next step's global loads are
// issued into registers before the current step's matrix-core work so the
// HBM/L2 latency overlaps with compute.  Each wave multiplies its
// (MT/2) x (NT/2) sub-tile with v_mfma_f64_16x16x4_f64; the op(B) fragment is
// fed as the instruction's first operand and the op(A) fragment as its second,
// so the 16 x 16 result block comes out with the row index of D across lanes
// 0..15 (lane group g = lane / 16 and register r hold column g + 4 r) and the
// epilogue writes 128-byte column segments (coalesced).
// Complex products use four real MFMAs (re*re - im*im, re*im + im*re) on
// separate real / imaginary LDS planes.
#include "internal.h"

namespace hc {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Matrix-core arithmetic per element type: the real type R (LDS planes and
// fragments), the accumulator V and the 16 x 16 x 4 product.  The FP64 and
// FP32 instructions take the same operand fragments (lane l: row l % 16,
// k = l / 16) but return the 16 x 16 result in different lane orders:
// register r of lane group g holds result row g + 4 r (FP64) or 4 g + r
// (FP32), column l % 16.
template <typename T> struct Prec;
template <> struct Prec<double> {
  using R = double;
  using V = v4d;
  static constexpr bool F32 = false;
};
template <> struct Prec<zd> : Prec<double> {};
template <> struct Prec<float> {
  using R = float;
  using V = v4f;
  static constexpr bool F32 = true;
};
template <> struct Prec<zf> : Prec<float> {};

__device__ __forceinline__ v4d mma(double b, double a, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mma(float b, float a, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c, 0, 0, 0);
}
// result row (the D column here) of register r in lane group g
template <typename T>
__device__ __forceinline__ int res_row(int g, int r) { return Prec<T>::F32 ? 4 * g + r : g + 4 * r; }

template <typename T> __device__ __forceinline__ T cmake(typename Prec<T>::R re, typename Prec<T>::R im);
template <> __device__ __forceinline__ zd cmake<zd>(double re, double im) { return make_zd(re, im); }
template <> __device__ __forceinline__ zf cmake<zf>(float re, float im) { return make_zf(re, im); }

constexpr int KT = 16;         // K step
constexpr int NTHREADS = 256;  // 4 waves
constexpr int PAD = 2;         // doubles of padding per LDS row

template <typename T>
struct KParams {
  int m, n, k;
  T alpha, beta;
  const T *A;
  long lda;
  const T *B;
  long ldb;
  const T *C;
  long ldc;
  T *D;
  long ldd;
  int tri_out, tri_grid, real_diag, tri_a, tri_b, unit;
  const int *abort_flag;
  // split K: blockIdx.z takes K range [z kchunk, (z + 1) kchunk) and writes
  // its raw sum to part + z * m * n (m x n column-major, ld m)
  int kchunk;
  T *part;
  // tri-grid launches: the number of tiles when the grid is smaller (each
  // workgroup then takes tiles w, w + gridDim.x, ...); 0: one per tile
  int vtiles;
};

// Element (r, c) of op(X).
template <typename T, int OP>
__device__ __forceinline__ T fetch(const T *X, long ld, int r, int c) {
  if (OP == OP_N) return X[(long)c * ld + r];
  const T v = X[(long)r * ld + c];
  return OP == OP_C ? conjx(v) : v;
}

// Element (r, c) of a triangular operand: zero outside the triangle (selected,
// not multiplied, so garbage there never propagates), one on a unit diagonal
// (unit == 1) or zero on an excluded one (unit == 2, a strict triangle).
template <typename T>
__device__ __forceinline__ T tri_mask(T v, int tri, int unit, int r, int c) {
  if ((tri == TRI_LOWER && c > r) || (tri == TRI_UPPER && c < r)) return from_real<T>(0.0);
  if (unit && r == c) return from_real<T>(unit == 1 ? 1.0 : 0.0);
  return v;
}

__device__ __forceinline__ void lds_put(double *s, int plane_stride, int idx, double v) {
  (void)plane_stride;
  s[idx] = v;
}
__device__ __forceinline__ void lds_put(double *s, int plane_stride, int idx, zd v) {
  s[idx] = v.x;
  s[plane_stride + idx] = v.y;
}
__device__ __forceinline__ void lds_put(float *s, int plane_stride, int idx, float v) {
  (void)plane_stride;
  s[idx] = v;
}
__device__ __forceinline__ void lds_put(float *s, int plane_stride, int idx, zf v) {
  s[idx] = v.x;
  s[plane_stride + idx] = v.y;
}

// Per-thread global-load pattern for one operand tile of R x KT elements
// (R = MT for op(A), NT for op(B)).  With the R index contiguous in memory
// (op N for A, op T/C for B) a thread keeps one R index and walks K; with K
// contiguous (op T/C for A, op N for B) it keeps one K index and walks R.
template <int R, bool R_CONTIG>
struct LoadPattern {
  static constexpr int E = R * KT / NTHREADS;           // elements per thread
  static constexpr int RSTEP = R_CONTIG ? 0 : NTHREADS / KT;
  static constexpr int KSTEP = R_CONTIG ? NTHREADS / R : 0;
  __device__ static int r0(int tid) { return R_CONTIG ? tid % R : tid / KT; }
  __device__ static int k0(int tid) { return R_CONTIG ? tid / R : tid % KT; }
};

// TG: the grid enumerates only the tiles meeting the tri_out triangle (the
// SYRK / HERK launch shape); otherwise it is the full (M / MT) x (N / NT) grid.
constexpr int NXCD = 8;   // MI355X: 8 accelerator complex dies of 32 CUs
constexpr int GROUP = 8;  // tile rows per traversal group

// Logical index of workgroup w of nwg when workgroup w runs on XCD w % 8:
// XCD x gets the contiguous logical range of its own workgroups.
__device__ __forceinline__ int xcd_contiguous(int w, int nwg) {
  const int x = w % NXCD, i = w / NXCD, base = nwg / NXCD, rem = nwg % NXCD;
  return x < rem ? x * (base + 1) + i : rem * (base + 1) + (x - rem) * base + i;
}

// Position i of the order n - 1, 0, n - 2, 1, ... (high_first) or
// 0, n - 1, 1, n - 2, ...: a permutation of [0, n).
__device__ __forceinline__ int serpentine(int i, int n, bool high_first) {
  const int j = i >> 1;
  return ((i & 1) == 0) == high_first ? n - 1 - j : j;
}

template <typename T, int TA, int TB, int MT, int NT, bool TG>
__global__ void __launch_bounds__(NTHREADS, 2) gemm_kernel(KParams<T> p) {
  if (p.abort_flag != nullptr && *p.abort_flag != 0) return;

  constexpr bool CPLX = is_complex<T>::value;
  using R = typename Prec<T>::R;
  using V = typename Prec<T>::V;
  constexpr int P = CPLX ? 2 : 1;
  constexpr int LSA = MT + PAD, LSB = NT + PAD;
  constexpr int PSA = KT * LSA, PSB = KT * LSB;  // plane strides (elements of R)
  constexpr int BUFA = P * PSA, BUFB = P * PSB;  // one LDS buffer
  // 2 x 2 waves; a 16-wide tile (skinny right-hand sides) uses 4 x 1
  constexpr int WAVES_N = NT >= 32 ? 2 : 1, WAVES_M = 4 / WAVES_N;
  constexpr int WM = MT / WAVES_M, WN = NT / WAVES_N, BI = WM / 16, BJ = WN / 16;
  using PA = LoadPattern<MT, TA == OP_N>;
  using PB = LoadPattern<NT, TB != OP_N>;
  constexpr int EA = PA::E, EB = PB::E;
  static_assert(EA >= 1 && EB >= 1 && BI >= 1 && BJ >= 1, "tile too small");

  __shared__ R sA[2 * BUFA];
  __shared__ R sB[2 * BUFB];

  // ---- tile coordinates ----
  // Workgroups are dealt round-robin to the 8 XCDs (each with its own L2):
  // renumber so that every XCD owns a contiguous range of logical tiles, and
  // walk tiles in groups of GROUP tile rows, column by column inside a group,
  // so the tiles in flight on one XCD share their op(A) and op(B) panels.
  auto tile_body = [&](const int wid, const int nwg) __attribute__((always_inline)) {
  const int logical = xcd_contiguous(wid, nwg);
  int bm, bn;
  if constexpr (TG) {
    // row-major position of the logical tile in the lower triangle of tiles
    const int id = logical;
    int r = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= id) r++;
    while (r * (r + 1) / 2 > id) r--;
    // rows first .. first + gs - 1 form one group (same tile set as row-major)
    const int nt = (int)((sqrt(8.0 * (double)nwg + 1.0) - 1.0) * 0.5 + 0.5);  // tile rows
    const int first = (r / GROUP) * GROUP, gs = min(GROUP, nt - first);
    int o = id - first * (first + 1) / 2, tr, tc;
    if (o < gs * first) {  // the rectangular part left of the group's diagonal
      tc = o / gs;
      tr = first + o % gs;
    } else {                // the group's own triangle, column by column
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
    // A triangular operand makes a tile's K range grow along its row (op(A))
    // or column (op(B)) index, and each XCD runs a contiguous band of the
    // logical order (one group of row blocks with all columns).  op(A):
    // the row blocks are taken in serpentine order (every band holds heavy
    // and light rows alike), and inside a band row by row, heaviest first,
    // so that the longest tiles start first.  op(B): every band holds all
    // columns already; they are taken heaviest first.
    if (p.tri_a != TRI_NONE) {
      const int q = o / gy, ne = (gs + 1) / 2;  // q-th heaviest row of the group
      const int pos = q < ne ? 2 * q : 2 * (gs - ne - 1 - (q - ne)) + 1;
      bm = serpentine(first + pos, gx, p.tri_a == TRI_LOWER);
      bn = o % gy;
    }
    if (p.tri_b != TRI_NONE && p.tri_b == TRI_UPPER) bn = gy - 1 - bn;
  }
  const int m0 = bm * MT, n0 = bn * NT;

  // ---- K range: a triangular operand bounds it ----
  int kbeg = 0, kend = p.k;
  if (p.tri_a == TRI_LOWER) kend = min(kend, m0 + MT);
  else if (p.tri_a == TRI_UPPER) kbeg = m0;
  if (p.tri_b == TRI_LOWER) kbeg = max(kbeg, n0);
  else if (p.tri_b == TRI_UPPER) kend = min(kend, n0 + NT);
  if (p.part != nullptr) {
    kbeg = max(kbeg, (int)blockIdx.z * p.kchunk);
    kend = min(kend, ((int)blockIdx.z + 1) * p.kchunk);
  }
  kbeg &= ~(KT - 1);
  if (is_zero(p.alpha)) kend = kbeg;

  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));  // thread-derived values are formed per tile, not kept across tiles
  const int lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;

  V acc[P][BI][BJ];
#pragma unroll
  for (int q = 0; q < P; q++)
#pragma unroll
    for (int i = 0; i < BI; i++)
#pragma unroll
      for (int j = 0; j < BJ; j++) acc[q][i][j] = (V){0, 0, 0, 0};

  // Real data with alpha = +-1 and beta != 0: the accumulators start from
  // (beta / alpha) C (exact scaling), read from clamped addresses right after
  // the first operand loads, so the C read overlaps them and the first
  // products; the epilogue then only stores alpha * acc.  npre = the number
  // of accumulator blocks seeded this way (all of them, or none).  Not with
  // split K: there each split writes its raw partial sum and C enters once,
  // in the reduction.
  int npre = 0;
  R cfac = 0;
  if constexpr (!CPLX) {
    if (!is_zero(p.beta) && (p.alpha == 1.0 || p.alpha == -1.0) && p.part == nullptr) {
      cfac = p.beta / p.alpha;
      npre = BI * BJ;
    }
  }
  auto seed_c = [&]() __attribute__((always_inline)) {
    if constexpr (!CPLX) {
      if (npre > 0) {  // uniform
#pragma unroll
        for (int i = 0; i < BI; i++) {
          const int gm = min(m0 + wm0 + 16 * i + lc, p.m - 1);
#pragma unroll
          for (int j = 0; j < BJ; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int gn = min(n0 + wn0 + 16 * j + res_row<T>(lg, r), p.n - 1);
              acc[0][i][j][r] = cfac * p.C[(long)gn * p.ldc + gm];
            }
        }
      }
    }
  };

  if (kbeg < kend) {
    // Per-thread source pointers at k = kbeg, and their per-element /
    // per-K-step strides; every address is formed by additions only.
    const int ar0 = m0 + PA::r0(tid), ak0 = kbeg + PA::k0(tid);
    const int br0 = n0 + PB::r0(tid), bk0 = kbeg + PB::k0(tid);
    const T *pa = (TA == OP_N) ? p.A + (long)ak0 * p.lda + ar0 : p.A + (long)ar0 * p.lda + ak0;
    const T *pb = (TB == OP_N) ? p.B + (long)br0 * p.ldb + bk0 : p.B + (long)bk0 * p.ldb + br0;
    // element q sits PA::RSTEP rows / PA::KSTEP k further on
    const long a_qstride = (TA == OP_N) ? (long)PA::KSTEP * p.lda : (long)PA::RSTEP * p.lda;
    const long b_qstride = (TB == OP_N) ? (long)PB::RSTEP * p.ldb : (long)PB::KSTEP * p.ldb;
    const long a_kstride = (TA == OP_N) ? (long)KT * p.lda : (long)KT;
    const long b_kstride = (TB == OP_N) ? (long)KT : (long)KT * p.ldb;
    // LDS offsets of the thread's elements (k-major, R contiguous)
    const int sa0 = PA::k0(tid) * LSA + PA::r0(tid);
    const int sb0 = PB::k0(tid) * LSB + PB::r0(tid);
    constexpr int SA_Q = PA::KSTEP * LSA + PA::RSTEP;
    constexpr int SB_Q = PB::KSTEP * LSB + PB::RSTEP;

    T ra[EA], rb[EB];
    // Out-of-range elements read a valid address (the operand's origin) and
    // are replaced by zero, so the loads carry no branches.
    // Interior tiles with whole K steps take unpredicated loads on every K
    // step that does not meet a triangular operand's diagonal block: the K
    // range is already cut to the triangle, so such a step lies strictly
    // inside it.  Edge tiles and diagonal steps check every element.
    const bool interior = m0 + MT <= p.m && n0 + NT <= p.n && (kend - kbeg) % KT == 0;
    const bool tri_ops = p.tri_a != TRI_NONE || p.tri_b != TRI_NONE;
    auto load = [&](int kt_off, const T *qa, const T *qb) {
      bool fast = interior;
      if (tri_ops) {
        const int k0 = kbeg + kt_off;
        fast = fast && (p.tri_a == TRI_NONE || k0 + KT <= m0 || k0 >= m0 + MT) &&
               (p.tri_b == TRI_NONE || k0 + KT <= n0 || k0 >= n0 + NT);
      }
      if (fast) {
#pragma unroll
        for (int q = 0; q < EA; q++) {
          T v = qa[q * a_qstride];
          if (TA == OP_C) v = conjx(v);
          ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < EB; q++) {
          T v = qb[q * b_qstride];
          if (TB == OP_C) v = conjx(v);
          rb[q] = v;
        }
        return;
      }
#pragma unroll
      for (int q = 0; q < EA; q++) {
        const int gi = ar0 + q * PA::RSTEP, gk = ak0 + kt_off + q * PA::KSTEP;
        const bool ok = gi < p.m && gk < kend;
        T v = *(ok ? qa + q * a_qstride : p.A);
        v = ok ? v : from_real<T>(0.0);
        if (TA == OP_C) v = conjx(v);
        if (p.tri_a) v = tri_mask(v, p.tri_a, p.unit, gi, gk);
        ra[q] = v;
      }
#pragma unroll
      for (int q = 0; q < EB; q++) {
        const int gj = br0 + q * PB::RSTEP, gk = bk0 + kt_off + q * PB::KSTEP;
        const bool ok = gj < p.n && gk < kend;
        T v = *(ok ? qb + q * b_qstride : p.B);
        v = ok ? v : from_real<T>(0.0);
        if (TB == OP_C) v = conjx(v);
        if (p.tri_b) v = tri_mask(v, p.tri_b, p.unit, gk, gj);
        rb[q] = v;
      }
    };
    auto store = [&](int buf) {
#pragma unroll
      for (int q = 0; q < EA; q++) lds_put(sA + buf * BUFA, PSA, sa0 + q * SA_Q, ra[q]);
#pragma unroll
      for (int q = 0; q < EB; q++) lds_put(sB + buf * BUFB, PSB, sb0 + q * SB_Q, rb[q]);
    };

    const int ntiles = (kend - kbeg + KT - 1) / KT;
    load(0, pa, pb);
    seed_c();
    store(0);
    __syncthreads();
    for (int kt = 0; kt < ntiles; kt++) {
      const bool more = kt + 1 < ntiles;
      if (more) {
        pa += a_kstride;
        pb += b_kstride;
        load((kt + 1) * KT, pa, pb);
      }
      const R *cA = sA + (kt & 1) * BUFA, *cB = sB + (kt & 1) * BUFB;
#pragma unroll
      for (int ks = 0; ks < KT; ks += 4) {
        const int kr = ks + lg;
        R ar[BI], br[BJ];
#pragma unroll
        for (int i = 0; i < BI; i++) ar[i] = cA[kr * LSA + wm0 + 16 * i + lc];
#pragma unroll
        for (int j = 0; j < BJ; j++) br[j] = cB[kr * LSB + wn0 + 16 * j + lc];
        if constexpr (!CPLX) {
#pragma unroll
          for (int i = 0; i < BI; i++)
#pragma unroll
            for (int j = 0; j < BJ; j++)
              acc[0][i][j] = mma(br[j], ar[i], acc[0][i][j]);
        } else {
          R ai[BI], bi[BJ];
#pragma unroll
          for (int i = 0; i < BI; i++) ai[i] = cA[PSA + kr * LSA + wm0 + 16 * i + lc];
#pragma unroll
          for (int j = 0; j < BJ; j++) bi[j] = cB[PSB + kr * LSB + wn0 + 16 * j + lc];
#pragma unroll
          for (int i = 0; i < BI; i++)
#pragma unroll
            for (int j = 0; j < BJ; j++) {
              acc[0][i][j] = mma(br[j], ar[i], acc[0][i][j]);
              acc[0][i][j] = mma(-bi[j], ai[i], acc[0][i][j]);
              acc[1][i][j] = mma(bi[j], ar[i], acc[1][i][j]);
              acc[1][i][j] = mma(br[j], ai[i], acc[1][i][j]);
            }
        }
      }
      // The other buffer was last read in the previous step, which every
      // thread finished before the previous barrier.
      if (more) store((kt + 1) & 1);
      __syncthreads();
    }
  } else {
    seed_c();
  }

  // ---- epilogue: lane lc holds row m; register r of lane group lg holds
  // column res_row(lg, r) (the result rows of the 16 x 16 x 4 product index
  // N here, its columns M) ----
  if (p.part != nullptr) {  // split K: raw partial sums
    T *part = p.part + (long)blockIdx.z * p.m * p.n;
#pragma unroll
    for (int i = 0; i < BI; i++) {
      const int gm = m0 + wm0 + 16 * i + lc;
      if (gm >= p.m) continue;
#pragma unroll
      for (int j = 0; j < BJ; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gn = n0 + wn0 + 16 * j + res_row<T>(lg, r);
          if (gn >= p.n) continue;
          T v;
          if constexpr (CPLX) v = cmake<T>(acc[0][i][j][r], acc[1][i][j][r]);
          else v = acc[0][i][j][r];
          part[(long)gn * p.m + gm] = v;
        }
    }
    return;
  }
  const bool use_beta = !is_zero(p.beta);
#pragma unroll
  for (int i = 0; i < BI; i++) {
    // C of the blocks not pre-added, one batch per block row (clamped, always
    // valid addresses) so that its loads are in flight together
    T cv[BJ][4];
    const bool need_c = use_beta && (i + 1) * BJ > npre;  // uniform
    if (need_c) {
      const int gmc = min(m0 + wm0 + 16 * i + lc, p.m - 1);
#pragma unroll
      for (int j = 0; j < BJ; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int gnc = min(n0 + wn0 + 16 * j + res_row<T>(lg, r), p.n - 1);
          cv[j][r] = p.C[(long)gnc * p.ldc + gmc];
        }
    }
    const int gm = m0 + wm0 + 16 * i + lc;
    if (gm >= p.m) continue;
#pragma unroll
    for (int j = 0; j < BJ; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gn = n0 + wn0 + 16 * j + res_row<T>(lg, r);
        if (gn >= p.n) continue;
        if (p.tri_out == TRI_LOWER && gm < gn) continue;
        if (p.tri_out == TRI_UPPER && gm > gn) continue;
        T v;
        if constexpr (CPLX) v = cmake<T>(acc[0][i][j][r], acc[1][i][j][r]);
        else v = acc[0][i][j][r];
        v = p.alpha * v;
        if (need_c && i * BJ + j >= npre) v = v + p.beta * cv[j][r];
        if constexpr (CPLX) {
          if (p.real_diag && gm == gn) v.y = 0.0;
        }
        p.D[(long)gn * p.ldd + gm] = v;
      }
    }
  }
  };
  if (TG && p.vtiles > 0) {
    // fewer workgroups than tiles: a workgroup's tiles keep its XCD
    // (gridDim.x is a multiple of 8) and reuse its LDS after the last
    // barrier of the previous tile
    for (int w = (int)blockIdx.x; w < p.vtiles; w += (int)gridDim.x) tile_body(w, p.vtiles);
  } else {
    tile_body(TG ? (int)blockIdx.x : (int)(blockIdx.x + blockIdx.y * gridDim.x),
              TG ? (int)gridDim.x : (int)(gridDim.x * gridDim.y));
  }
}

// D = alpha * sum_z part[z] + beta * C (split-K epilogue, fixed order).
template <typename T>
__global__ void splitk_reduce_kernel(int m, int n, int nsplit, const T *part, T alpha, T beta,
                                     const T *C, long ldc, T *D, long ldd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m || j >= n) return;
  T acc = from_real<T>(0.0);
  for (int z = 0; z < nsplit; z++) acc = acc + part[(long)z * m * n + (long)j * m + i];
  T v = alpha * acc;
  if (!is_zero(beta)) v = v + beta * C[(long)j * ldc + i];
  D[(long)j * ldd + i] = v;
}

template <typename T, int TA, int TB, int MT, int NT>
hipError_t launch(hipStream_t s, const KParams<T> &p) {
  if (p.part != nullptr) {
    const int nsplit = (p.k + p.kchunk - 1) / p.kchunk;
    const dim3 grid((unsigned)((p.m + MT - 1) / MT), (unsigned)((p.n + NT - 1) / NT), (unsigned)nsplit);
    hipLaunchKernelGGL((gemm_kernel<T, TA, TB, MT, NT, false>), grid, dim3(NTHREADS), 0, s, p);
    HC_TRY(hipGetLastError());
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3((unsigned)((p.m + 255) / 256), (unsigned)p.n),
                       dim3(256), 0, s, p.m, p.n, nsplit, (const T *)p.part, p.alpha, p.beta, p.C,
                       p.ldc, p.D, p.ldd);
    return hipGetLastError();
  }
  const dim3 grid((unsigned)((p.m + MT - 1) / MT), (unsigned)((p.n + NT - 1) / NT));
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB, MT, NT, false>), grid, dim3(NTHREADS), 0, s, p);
  return hipGetLastError();
}

template <typename T, int TA, int TB, int MT>
hipError_t launch_tri(hipStream_t s, KParams<T> p, int max_wgs) {
  const long t = (p.m + MT - 1) / MT, tiles = t * (t + 1) / 2;
  long grid = tiles;
  p.vtiles = 0;
  if (max_wgs > 0 && max_wgs < tiles) {
    grid = std::max(NXCD, max_wgs / NXCD * NXCD);
    p.vtiles = (int)tiles;
  }
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB, MT, MT, true>), dim3((unsigned)grid), dim3(NTHREADS), 0, s, p);
  return hipGetLastError();
}

template <typename T, int TA, int TB>
hipError_t dispatch_tile(hipStream_t s, const KParams<T> &p, int mt, int nt, int max_wgs) {
  constexpr bool CPLX = is_complex<T>::value;
  constexpr int BIG = CPLX ? 64 : 128, SMALL = CPLX ? 32 : 64;
  if (p.tri_grid) {
    // SYRK / HERK: op(B) = op(A)^H, so only (N, T/C) and (T/C, N) occur
    if constexpr ((TA == OP_N) != (TB == OP_N)) {
      if (mt == BIG) return launch_tri<T, TA, TB, BIG>(s, p, max_wgs);
      return launch_tri<T, TA, TB, SMALL>(s, p, max_wgs);
    }
    return hipErrorInvalidValue;
  }
  if constexpr (!CPLX && TB == OP_N) {
    if (nt == 16) return launch<T, TA, TB, 128, 16>(s, p);  // few right-hand sides
  }
  if (mt == BIG && nt == BIG) return launch<T, TA, TB, BIG, BIG>(s, p);
  if (mt == SMALL && nt == SMALL) return launch<T, TA, TB, SMALL, SMALL>(s, p);
  if (mt == SMALL && nt == BIG) return launch<T, TA, TB, SMALL, BIG>(s, p);
  return launch<T, TA, TB, BIG, SMALL>(s, p);
}

template <typename T>
hipError_t dispatch_ops(hipStream_t s, int ta, int tb, const KParams<T> &p, int mt, int nt, int max_wgs) {
  if constexpr (!is_complex<T>::value) {
    // for real data C and T coincide
    const bool a = ta != OP_N, b = tb != OP_N;
    if (!a && !b) return dispatch_tile<T, OP_N, OP_N>(s, p, mt, nt, max_wgs);
    if (!a && b) return dispatch_tile<T, OP_N, OP_T>(s, p, mt, nt, max_wgs);
    if (a && !b) return dispatch_tile<T, OP_T, OP_N>(s, p, mt, nt, max_wgs);
    return dispatch_tile<T, OP_T, OP_T>(s, p, mt, nt, max_wgs);
  } else {
#define HC_OPS_CASE(X, Y) \
    if (ta == X && tb == Y) return dispatch_tile<T, X, Y>(s, p, mt, nt, max_wgs);
    HC_OPS_CASE(OP_N, OP_N) HC_OPS_CASE(OP_N, OP_T) HC_OPS_CASE(OP_N, OP_C)
    HC_OPS_CASE(OP_T, OP_N) HC_OPS_CASE(OP_T, OP_T) HC_OPS_CASE(OP_T, OP_C)
    HC_OPS_CASE(OP_C, OP_N) HC_OPS_CASE(OP_C, OP_T) HC_OPS_CASE(OP_C, OP_C)
#undef HC_OPS_CASE
    return hipErrorInvalidValue;
  }
}

template <typename T>
__global__ void copy2d_kernel(int m, int n, const T *S, long lds, T *D, long ldd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < m && j < n) D[(long)j * ldd + i] = S[(long)j * lds + i];
}


}  // namespace


template <typename T>
hipError_t gemm_tile(hipStream_t s, int ta, int tb, long m, long n, long k, T alpha, const T *A,
                     long lda, const T *B, long ldb, T beta, const T *C, long ldc, T *D, long ldd,
                     const GemmOpts &o) {
  if (m <= 0 || n <= 0) return hipSuccess;
  KParams<T> p;
  p.m = (int)m; p.n = (int)n; p.k = (int)(k > 0 ? k : 0);
  p.alpha = alpha; p.beta = beta;
  p.A = A; p.lda = lda; p.B = B; p.ldb = ldb; p.C = C; p.ldc = ldc; p.D = D; p.ldd = ldd;
  p.tri_out = o.tri_out; p.tri_grid = o.tri_grid ? 1 : 0; p.real_diag = o.real_diag ? 1 : 0;
  p.tri_a = o.tri_a; p.tri_b = o.tri_b; p.unit = o.unit ? 1 : (o.zero_diag ? 2 : 0);
  p.abort_flag = o.abort_flag;
  p.kchunk = 0;
  p.part = nullptr;
  p.vtiles = 0;

  // Tile choice: the large tile when the launch has at least the measured
  // number of them (family D / Z of the tuning table; about one workgroup per
  // CU by default), otherwise the small one; in-place callers may demand a
  // tile that spans a whole panel in M or N.
  constexpr bool CPLX = is_complex<T>::value;
  constexpr int BIG = CPLX ? 64 : 128, SMALL = CPLX ? 32 : 64;
  const long tb_m = (m + BIG - 1) / BIG, tb_n = (n + BIG - 1) / BIG;
  const long tiles_big = o.tri_grid ? tb_m * (tb_m + 1) / 2 : tb_m * tb_n;
  int mt = tiles_big >= tuned_tile_min(CPLX ? 'Z' : 'D', 240) ? BIG : SMALL, nt = mt;
  // real data with at most 16 columns and plenty of row tiles: 128 x 16 tiles
  if (!CPLX && tb == OP_N && n <= 16 && o.tri_b == TRI_NONE && o.min_nt <= 16 && !o.tri_grid &&
      (m + 127) / 128 >= 64) {
    mt = 128;
    nt = 16;
    // long K: split it so about 2048 workgroups share the rows, partial sums
    // in the caller's workspace (splitk_ws elements), reduced in fixed order
    if (o.splitk_ws != nullptr && o.tri_out == TRI_NONE && !o.real_diag) {
      const long row_tiles = (m + 127) / 128;
      long nsplit = (2048 + row_tiles - 1) / row_tiles;
      long kchunk = ((k + nsplit - 1) / nsplit + KT - 1) / KT * KT;
      if (kchunk < 256) kchunk = 256;
      nsplit = (k + kchunk - 1) / kchunk;
      if (nsplit > 1 && nsplit * m * n <= o.splitk_elems) {
        p.kchunk = (int)kchunk;
        p.part = static_cast<T *>(o.splitk_ws);
      }
    }
  }
  if (o.min_mt > mt) mt = BIG;
  if (o.min_nt > nt) nt = BIG;
  if (o.min_mt > BIG || o.min_nt > BIG) return hipErrorInvalidValue;
  if (o.tri_grid) {
    if (m != n) return hipErrorInvalidValue;
    mt = nt = (mt > nt ? mt : nt);
  }
  return dispatch_ops<T>(s, ta, tb, p, mt, nt, o.tri_grid ? o.max_wgs : 0);
}

template <typename T>
hipError_t copy2d(hipStream_t s, long m, long n, const T *S, long lds, T *D, long ldd) {
  if (m <= 0 || n <= 0) return hipSuccess;
  dim3 grid((unsigned)((m + 255) / 256), (unsigned)n);
  hipLaunchKernelGGL((copy2d_kernel<T>), grid, dim3(256), 0, s, (int)m, (int)n, S, lds, D, ldd);
  return hipGetLastError();
}

// gemm<T> is this kernel for double, double complex and single complex; for
// single data gemm<float> (gemm_lp.hip) runs its own FP32 kernel and hands
// triangular operands to gemm_tile<float>.
template <typename T>
hipError_t gemm(hipStream_t s, int ta, int tb, long m, long n, long k, T alpha, const T *A, long lda,
                const T *B, long ldb, T beta, const T *C, long ldc, T *D, long ldd, const GemmOpts &o) {
  return gemm_tile<T>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, D, ldd, o);
}

#define HC_GEMM_INST(F, T)                                                                          \
  template hipError_t F<T>(hipStream_t, int, int, long, long, long, T, const T *, long, const T *, \
                           long, T, const T *, long, T *, long, const GemmOpts &);
HC_GEMM_INST(gemm, double)
HC_GEMM_INST(gemm, zd)
HC_GEMM_INST(gemm, zf)
HC_GEMM_INST(gemm_tile, float)
#undef HC_GEMM_INST
template hipError_t copy2d<double>(hipStream_t, long, long, const double *, long, double *, long);
template hipError_t copy2d<zd>(hipStream_t, long, long, const zd *, long, zd *, long);
template hipError_t copy2d<float>(hipStream_t, long, long, const float *, long, float *, long);
template hipError_t copy2d<zf>(hipStream_t, long, long, const zf *, long, zf *, long);

}  // namespace hc
