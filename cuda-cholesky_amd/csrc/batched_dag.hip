// Batched SPD solve + log-determinant + inverse of matrices of any size, as
// ONE persistent kernel that executes a task graph of 64 x 64 tiles of every
// matrix of the batch (no counterpart in the reference, whose routines take
// one matrix per call; the per-matrix steps follow lapack/dpotrf.c,
// dtrtri.c, dlauum.c and dpotri.c).
//
// Per matrix m (order n, T = ceil(n / 64) tile rows, padded order P = 64 T;
// every workspace is padded to P x P, column-major, ld P, and the padding
// holds the identity so that the padded matrix is diag(A, I)):
//   DIAG(j)      C = A_jj - sum_{k<j} L_jk L_jk^T  (left-looking), factored
//                and inverted in LDS: X_jj = inv(L_jj) -> Xw; sum of log L_kk^2
//                -> logpart; a non-positive pivot marks the matrix failed.
//   PANEL(i, j)  C = A_ij - sum_{k<j} L_ik L_jk^T, then L_ij = C X_jj^T -> F.
//   TRTRI(i, j)  X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj -> Xw   (X = inv(L)).
//   LAUUM(i, j)  (A^-1)_ij = sum_{k>=i} X_ki^T X_kj -> F (both triangles, L
//                is no longer read) and, when wanted, -> Ainv (triangle uplo).
//   SOLVE        x = A^-1 b, then FP64 refinement r = b - A x, x += A^-1 r
//                until every column meets ||r||_inf <= ||x||_inf ||A||_inf
//                eps sqrt(n) (LAPACK DSPOSV's test; at most 30 steps); the
//                log-determinant (partials summed in tile order), iter, info.
// DIAG and PANEL also copy the A tile into Af (A stored full and padded) for
// the residuals.
//
// Scheduling: the host lists the tasks in a topological order (a list
// schedule, critical path first).  A workgroup takes the next ticket from an
// atomic counter and waits on per-matrix dependency counters, which only
// ever name tasks earlier in the list: those were taken by workgroups that
// are already running, so the kernel cannot deadlock, whatever the
// occupancy.  Waits are bounded in time: an abandoned wait sets the error
// word and every matrix then reports info = -1.
//
// Dependency counters per matrix (at cbase): diag[T], col[T], trow[T], ttot,
// ltot, failed.  diag[j] = 1 once X_jj is written; col[j] counts finished
// tiles of column j of L (complete at T - j); trow[j] counts finished tiles
// of column j of X (they finish top-down); ttot / ltot count TRTRI / LAUUM
// tiles; failed holds the 1-based first non-positive pivot.
#include "internal.h"
#include "publish.h"
#include "tile16.h"
#include "batched_dag.h"

namespace hc {
namespace dag {
namespace {

using t16::LEAF;
using t16::TileAcc;
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TB = kDagTile;    // tile order
constexpr int NTHR = 256;       // 4 waves
constexpr int LD = TB + 2;      // LDS tile leading dimension (doubles)
constexpr int KS = 16;          // K step of the streamed products
constexpr int LL = LEAF + 1;    // leaf LDS leading dimension
constexpr int TILE_D = TB * LD;             // one LDS tile
constexpr int SMEM_D = 2 * TILE_D + 4 * LEAF * LL;
constexpr int QMAX = 4;                     // solve: rows per thread per pass
constexpr int SCH = 512;                    // solve: columns staged per chunk
constexpr unsigned long long kWaitTicks = 400000000ull;  // 4 s at the 100 MHz real-time counter

__device__ __forceinline__ int ld_acq(const int *c) {
  return __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block until *c >= target (wave 0 polls; every wave then acquires).
__device__ void wait_geq(const int *c, int target, int *err) {
  if (threadIdx.x < 64 && ld_acq(c) < target) {
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    while (ld_acq(c) < target) {
      if (ld_acq(err) != 0) break;
      __builtin_amdgcn_s_sleep(4);
      if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// *p read once for the whole workgroup (thread 0, broadcast through LDS):
// another workgroup may change it meanwhile, and threads that disagreed
// would take different sides of a branch that holds barriers.
__device__ int read_uniform(const int *p, int *s_word) {
  __syncthreads();
  if (threadIdx.x == 0) *s_word = ld_acq(p);
  __syncthreads();
  return *s_word;
}

// Publish everything this workgroup wrote, then add 1 to each counter: every
// wave waits for its own global stores to reach the L2 before the barrier
// (the barrier itself only waits for LDS traffic), then wave 0 writes the L2
// back once and thread 0 updates the counters.
__device__ void signal(int *c0, int *c1) {
  publish_wait_own();
  __syncthreads();
  if (threadIdx.x < 64) publish_release();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(c0, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c1 != nullptr) __hip_atomic_fetch_add(c1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------------------
// 64 x 64 accumulator: wave w owns rows 32 (w & 1) .. + 31 and columns
// 32 (w >> 1) .. + 31 as 2 x 2 blocks in the MFMA result layout (lane l,
// register r <-> element (l / 16 + 4 r, l % 16) of a 16 x 16 block).
// ---------------------------------------------------------------------------
struct Acc {
  TileAcc<double> b[2][2];
};

__device__ __forceinline__ int wrow() { return 32 * ((threadIdx.x >> 6) & 1); }
__device__ __forceinline__ int wcol() { return 32 * (threadIdx.x >> 7); }

__device__ __forceinline__ void acc_zero(Acc &a) {
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) t16::acc_zero(a.b[rb][cb]);
}
// P column-major (ld LD) <-> accumulator
__device__ __forceinline__ void acc_from_lds(Acc &a, const double *P) {
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
      t16::acc_load(a.b[rb][cb], P + (wcol() + 16 * cb) * LD + wrow() + 16 * rb, LD);
}
__device__ __forceinline__ void acc_to_lds(const Acc &a, double *P) {
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
      t16::acc_store(a.b[rb][cb], P + (wcol() + 16 * cb) * LD + wrow() + 16 * rb, LD);
}
// element (r, c) to P[r * LD + c] (the transpose in column-major terms)
__device__ __forceinline__ void acc_to_lds_t(const Acc &a, double *P) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        P[(wrow() + 16 * rb + lk + 4 * r) * LD + wcol() + 16 * cb + li] = a.b[rb][cb].re[r];
}

// a(r, c) += s sum_{k < K} Sa[k LD + r] Sb[k LD + c]   (K a multiple of 4)
__device__ __forceinline__ void lds_mma(Acc &a, const double *Sa, const double *Sb, int K, double s) {
  const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
  const double *pa = Sa + wrow() + li, *pb = Sb + wcol() + li;
  for (int k = 0; k < K; k += 4) {
    double xa[2], yb[2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++) xa[rb] = s * pa[(k + lk) * LD + 16 * rb];
#pragma unroll
    for (int cb = 0; cb < 2; cb++) yb[cb] = pb[(k + lk) * LD + 16 * cb];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int cb = 0; cb < 2; cb++)
        a.b[rb][cb].re = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[rb], yb[cb], a.b[rb][cb].re, 0, 0, 0);
  }
}

// Streamed product a(r, c) += s sum_{k < K} opA(r, k) opB(c, k), K a
// multiple of 16, operands in global memory: mode 0 (N): element (r, k) at
// base[k ld + r]; mode 1 (T): at base[r ld + k].  Stage: 2 x 2 x KS x LD
// doubles of LDS, double-buffered (one barrier per step).
template <int MA, int MB>
__device__ void tile_gemm(Acc &a, const double *Ab, long lda, const double *Bb, long ldb, int K, double s,
                          double *stage) {
  const int tid = threadIdx.x;
  double *Sa[2] = {stage, stage + KS * LD};
  double *Sb[2] = {stage + 2 * KS * LD, stage + 3 * KS * LD};
  double ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int idx = tid + NTHR * e;
      const int r1 = MA == 0 ? (idx & 63) : (idx >> 4), k1 = MA == 0 ? (idx >> 6) : (idx & 15);
      ra[e] = MA == 0 ? Ab[(long)(k0 + k1) * lda + r1] : Ab[(long)r1 * lda + k0 + k1];
      const int r2 = MB == 0 ? (idx & 63) : (idx >> 4), k2 = MB == 0 ? (idx >> 6) : (idx & 15);
      rb[e] = MB == 0 ? Bb[(long)(k0 + k2) * ldb + r2] : Bb[(long)r2 * ldb + k0 + k2];
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int idx = tid + NTHR * e;
      const int r1 = MA == 0 ? (idx & 63) : (idx >> 4), k1 = MA == 0 ? (idx >> 6) : (idx & 15);
      Sa[buf][k1 * LD + r1] = ra[e];
      const int r2 = MB == 0 ? (idx & 63) : (idx >> 4), k2 = MB == 0 ? (idx >> 6) : (idx & 15);
      Sb[buf][k2 * LD + r2] = rb[e];
    }
  };
  const int nk = K / KS;
  if (nk <= 0) return;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int st = 0; st < nk; st++) {
    if (st + 1 < nk) gload((st + 1) * KS);
    lds_mma(a, Sa[st & 1], Sb[st & 1], KS, s);
    if (st + 1 < nk) sstore((st + 1) & 1);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// tile transfers (64 x 64, 256 threads)
// ---------------------------------------------------------------------------

// P[c LD + r] <- lower-form A(64 i + r, 64 j + c) of one packed matrix
// (column-major, ld n; lower-form (R, C), R >= C, is A(R, C) for lower and
// A(C, R) for upper storage), the identity beyond n.  sym: a diagonal tile
// (i == j) made symmetric from its lower part.
template <bool UPPER>
__device__ void load_a_tile(double *P, const double *A, int n, int i, int j, bool sym) {
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int e = 0; e < TB * TB / NTHR; e++) {
    const int idx = tid + NTHR * e;
    // lower storage: consecutive threads walk a column (contiguous rows);
    // upper storage: they walk a row of the lower form (contiguous there)
    const int r = UPPER ? (idx >> 6) : (idx & 63), c = UPPER ? (idx & 63) : (idx >> 6);
    int gr = TB * i + r, gc = TB * j + c;
    if (sym && c > r) {
      gr = TB * i + c;
      gc = TB * j + r;
    }
    const int cr = min(gr, n - 1), cc = min(gc, n - 1);
    const double x = UPPER ? A[(long)cr * n + cc] : A[(long)cc * n + cr];
    P[c * LD + r] = (gr < n && gc < n) ? x : (gr == gc ? 1.0 : 0.0);
  }
}

// dst[c ld + r] <- P[c LD + r] (trans: dst[r ld + c] <- P[c LD + r]); sym:
// element (r, c) from the lower part P(max, min).
__device__ void store_tile(double *dst, long ld, const double *P, bool trans, bool sym) {
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int e = 0; e < TB * TB / NTHR; e++) {
    const int idx = tid + NTHR * e;
    // consecutive threads write consecutive addresses of dst
    const int a = idx & 63, b = idx >> 6;
    const int r = trans ? b : a, c = trans ? a : b;   // dst element (a, b) of its column b
    const int pr = sym ? max(r, c) : r, pc = sym ? min(r, c) : c;
    dst[(long)b * ld + a] = P[pc * LD + pr];
  }
}

// P[c LD + r] <- src[c ld + r]
__device__ void load_tile(double *P, const double *src, long ld) {
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int e = 0; e < TB * TB / NTHR; e++) {
    const int idx = tid + NTHR * e;
    const int r = idx & 63, c = idx >> 6;
    P[c * LD + r] = src[(long)c * ld + r];
  }
}

// Lower-form tile (i, j) of the inverse to the caller's packed matrix (ld n,
// triangle uplo), rows and columns below n only, the lower triangle of a
// diagonal tile only.
template <bool UPPER>
__device__ void store_ainv_tile(double *D, int n, int i, int j, const double *P) {
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int e = 0; e < TB * TB / NTHR; e++) {
    const int idx = tid + NTHR * e;
    const int r = UPPER ? (idx >> 6) : (idx & 63), c = UPPER ? (idx & 63) : (idx >> 6);
    const int gr = TB * i + r, gc = TB * j + c;
    if (gr < n && gc < n && gr >= gc) {
      if (UPPER) D[(long)gr * n + gc] = P[c * LD + r];
      else D[(long)gc * n + gr] = P[c * LD + r];
    }
  }
}

// ---------------------------------------------------------------------------
// 64 x 64 factorisation and triangular inverse in LDS (16-column leaves)
// ---------------------------------------------------------------------------

// P: lower 64 x 64 block (column-major, ld LD), factored in place.  Returns
// the first failing column (0-based, every thread) or -1; in wave 0, lane r
// (r < 16) accumulates log L_kk^2 of the rows k = r mod 16 below kvalid into
// logsum (sum the wave's lanes for the block's total).
__device__ int lds_potrf64(double *P, double *Li, int *flag, int kvalid, double &logsum) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll 1
  for (int p = 0; p < TB / LEAF; p++) {
    const int g0 = p * LEAF;
    if (wave == 0) {
      double a[LEAF];  // row li of the leaf
      t16::static_for<0, LEAF>([&](auto C) { a[decltype(C)::value] = P[(g0 + decltype(C)::value) * LD + g0 + li]; });
      double x[4], rpiv;
      const int b = t16::leaf_factor_invert<double>(a, x, rpiv);
      if (b < 0) {
        // log of this lane's own pivot L_rr^2 = -2 log(1 / L_rr), off the pivot chain
        if (lk == 0 && g0 + li < kvalid) logsum -= 2.0 * log(rpiv);
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {  // lane group lk writes columns lk + 4 jj of row li
          const double f = t16::sel(lk == 0, a[4 * jj], t16::sel(lk == 1, a[4 * jj + 1],
                                    t16::sel(lk == 2, a[4 * jj + 2], a[4 * jj + 3])));
          const int c = lk + 4 * jj;
          P[(g0 + c) * LD + g0 + li] = c <= li ? f : 0.0;
        }
#pragma unroll
        for (int jj = 0; jj < 4; jj++) Li[(lk + 4 * jj) * LL + li] = x[jj];
      }
      if (lane == 0) *flag = b < 0 ? -1 : g0 + b;
    }
    __syncthreads();
    const int bad = *flag;
    if (bad >= 0) return bad;
    {  // panel L_Ip = A_Ip inv(L_pp)^T, one tile per wave
      const int I = p + 1 + wave;
      if (I < TB / LEAF) {
        TileAcc<double> acc;
        t16::acc_zero(acc);
        t16::tile_mma<true>(acc, P + g0 * LD + LEAF * I, LD, Li, LL, 1.0);
        t16::acc_store(acc, P + g0 * LD + LEAF * I, LD);
      }
    }
    __syncthreads();
    {  // trailing tiles (I, J), p < J <= I
      const int m = TB / LEAF - 1 - p, ntile = m * (m + 1) / 2;
#pragma unroll 1
      for (int t = wave; t < ntile; t += NTHR / 64) {
        int tt = t, J = 0;
        while (tt >= m - J) {
          tt -= m - J;
          J++;
        }
        const int I = p + 1 + J + tt, Jc = p + 1 + J;
        TileAcc<double> acc;
        t16::acc_load(acc, P + LEAF * Jc * LD + LEAF * I, LD);
        t16::tile_mma<true>(acc, P + g0 * LD + LEAF * I, LD, P + g0 * LD + LEAF * Jc, LD, -1.0);
        t16::acc_store(acc, P + LEAF * Jc * LD + LEAF * I, LD);
      }
    }
    __syncthreads();
  }
  return -1;
}

// Q <- inv(L) for the factored lower block L in P (zeros above the diagonal).
// Leaves right to left: X_pp = inv(L_pp), X_Ip = -(sum_{p<K<=I} X_IK L_Kp) X_pp.
__device__ void lds_trtri64(const double *P, double *Q, double *Li, double *W) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  for (int e = tid; e < TILE_D; e += NTHR) Q[e] = 0.0;
  __syncthreads();
#pragma unroll 1
  for (int p = TB / LEAF - 1; p >= 0; p--) {
    const int g0 = p * LEAF;
    if (wave == 0) {
      double a[LEAF], x[4];
      t16::static_for<0, LEAF>([&](auto C) { a[decltype(C)::value] = P[(g0 + decltype(C)::value) * LD + g0 + li]; });
      t16::leaf_inverse_cols<double>(a, x);
#pragma unroll
      for (int jj = 0; jj < 4; jj++) Li[(lk + 4 * jj) * LL + li] = x[jj];
    }
    const int I = p + 1 + wave;
    if (I < TB / LEAF) {  // Y_I = sum_{K=p+1}^{I} X_IK L_Kp
      TileAcc<double> acc;
      t16::acc_zero(acc);
      for (int K = p + 1; K <= I; K++)
        t16::tile_mma<false>(acc, Q + LEAF * K * LD + LEAF * I, LD, P + g0 * LD + LEAF * K, LD, 1.0);
      t16::acc_store(acc, W + wave * LEAF * LL, LL);
    }
    __syncthreads();
    if (I < TB / LEAF) {  // X_Ip = -Y_I X_pp
      TileAcc<double> acc;
      t16::acc_zero(acc);
      t16::tile_mma<false>(acc, W + wave * LEAF * LL, LL, Li, LL, -1.0);
      t16::acc_store(acc, Q + g0 * LD + LEAF * I, LD);
    }
    if (wave == NTHR / 64 - 1) {  // X_pp (the last wave has no I at any p)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = lk + 4 * r;
        Q[(g0 + li) * LD + g0 + row] = Li[li * LL + row];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// the solve: y = M x for a full padded matrix M (ld ldm), one workgroup;
// thread t owns rows row0 + t + 256 q (q < QMAX) in registers.
// ---------------------------------------------------------------------------
__device__ void gemv_pass(const double *M, long ldm, int n, int row0, const double *xin, int nr,
                          double (&y)[QMAX][4], double (&absrow)[QMAX], bool want_abs, double *xs) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < QMAX; q++) {
    absrow[q] = 0.0;
#pragma unroll
    for (int g = 0; g < 4; g++) y[q][g] = 0.0;
  }
  long rowoff[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; q++) rowoff[q] = min(row0 + tid + NTHR * q, n - 1);
  for (int c0 = 0; c0 < n; c0 += SCH) {
    const int cn = min(SCH, n - c0);
    __syncthreads();
    for (int e = tid; e < SCH * 4; e += NTHR) {
      const int k = e >> 2, g = e & 3;
      xs[e] = (k < cn && g < nr) ? xin[(long)g * n + c0 + k] : 0.0;
    }
    __syncthreads();
    const double *Mc = M + (long)c0 * ldm;
    for (int k = 0; k < cn; k++) {
      const double xv[4] = {xs[4 * k], xs[4 * k + 1], xs[4 * k + 2], xs[4 * k + 3]};
#pragma unroll
      for (int q = 0; q < QMAX; q++) {
        const double m = Mc[(long)k * ldm + rowoff[q]];
#pragma unroll
        for (int g = 0; g < 4; g++) y[q][g] = __builtin_fma(m, xv[g], y[q][g]);
        if (want_abs) absrow[q] += fabs(m);
      }
    }
  }
}

// max over the workgroup of v[0..4] (every thread gets the result)
__device__ void block_max5(double (&v)[5], double *red) {
  const int tid = threadIdx.x;
  __syncthreads();
#pragma unroll
  for (int g = 0; g < 5; g++) red[g * NTHR + tid] = v[g];
  __syncthreads();
  for (int s = NTHR / 2; s > 0; s >>= 1) {
    if (tid < s)
#pragma unroll
      for (int g = 0; g < 5; g++) red[g * NTHR + tid] = fmax(red[g * NTHR + tid], red[g * NTHR + tid + s]);
    __syncthreads();
  }
#pragma unroll
  for (int g = 0; g < 5; g++) v[g] = red[g * NTHR];
}

__device__ __attribute__((noinline)) void solve_task(const DagParams &p, const DagMat &mt, int m, double *smem,
                                                      int *s_word) {
  const int tid = threadIdx.x, n = mt.n, T = mt.T;
  const long P = (long)TB * T;
  int *cnt = p.cnt + mt.cbase;
  const int failed = read_uniform(cnt + 3 * T + 2, s_word);
  int iter = 0;
  double logdet = 0.0;
  if (failed == 0 && n > 0) {
    const double *Minv = p.F + mt.offP, *Af = p.Af + mt.offP;
    const double *Bm = p.B + mt.offB;
    double *Xm = p.X + mt.offB, *Rm = p.R + mt.offB;
    double *xs = smem, *red = smem + SCH * 4;
    const double tol_scale = 2.220446049250313e-16 * sqrt((double)n);
    double anorm = 0.0;
    for (int g0 = 0; g0 < p.nrhs; g0 += 4) {
      const int nr = min(4, p.nrhs - g0);
      const double *b = Bm + (long)g0 * n;
      double *x = Xm + (long)g0 * n, *r = Rm + (long)g0 * n;
      double y[QMAX][4], ab[QMAX];
      // x = A^-1 b
      for (int row0 = 0; row0 < n; row0 += NTHR * QMAX) {
        gemv_pass(Minv, P, n, row0, b, nr, y, ab, false, xs);
#pragma unroll
        for (int q = 0; q < QMAX; q++) {
          const int row = row0 + tid + NTHR * q;
          if (row < n)
#pragma unroll
            for (int g = 0; g < 4; g++)
              if (g < nr) x[(long)g * n + row] = y[q][g];
        }
      }
      int it = 0;
      for (;; it++) {
        __syncthreads();  // x complete before it is staged
        // r = b - A x, the norms of r and x, and (once) ||A||_inf
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // max |r_g| (g < 4), max row sum
        double xn[4] = {0.0, 0.0, 0.0, 0.0};
        for (int row0 = 0; row0 < n; row0 += NTHR * QMAX) {
          gemv_pass(Af, P, n, row0, x, nr, y, ab, anorm == 0.0, xs);
#pragma unroll
          for (int q = 0; q < QMAX; q++) {
            const int row = row0 + tid + NTHR * q;
            if (row < n) {
              v[4] = fmax(v[4], ab[q]);
#pragma unroll
              for (int g = 0; g < 4; g++)
                if (g < nr) {
                  const double rr = b[(long)g * n + row] - y[q][g];
                  r[(long)g * n + row] = rr;
                  v[g] = fmax(v[g], fabs(rr));
                  xn[g] = fmax(xn[g], fabs(x[(long)g * n + row]));
                }
            }
          }
        }
        double w[5] = {xn[0], xn[1], xn[2], xn[3], 0.0};
        block_max5(v, red);
        block_max5(w, red);
        if (anorm == 0.0) anorm = v[4];
        bool ok = true;
#pragma unroll
        for (int g = 0; g < 4; g++)
          if (g < nr && !(v[g] <= w[g] * anorm * tol_scale)) ok = false;
        if (ok) break;
        if (it == 30) {
          it = -31;
          break;
        }
        __syncthreads();  // r complete
        // x += A^-1 r
        for (int row0 = 0; row0 < n; row0 += NTHR * QMAX) {
          gemv_pass(Minv, P, n, row0, r, nr, y, ab, false, xs);
#pragma unroll
          for (int q = 0; q < QMAX; q++) {
            const int row = row0 + tid + NTHR * q;
            if (row < n)
#pragma unroll
              for (int g = 0; g < 4; g++)
                if (g < nr) x[(long)g * n + row] += y[q][g];
          }
        }
      }
      if (it < 0) iter = it;
      else if (iter >= 0 && it > iter) iter = it;
    }
    for (int j = 0; j < T; j++) logdet += p.logpart[mt.cbase + j];
  }
  if (tid == 0) {
    const bool abandoned = ld_acq(p.err) != 0;
    p.iter[m] = iter;
    p.info[m] = abandoned ? -1 : failed;
    p.logdet[m] = failed != 0 ? __builtin_nan("") : logdet;
  }
}

// Everything a task body needs about its matrix.
struct Ctx {
  int T, n, i, j;
  long P;
  int *diag, *col, *trow, *ttot, *ltot, *failed;
  double *F, *Xw, *Af;
  const double *A;
};

// Each task body is a separate function so that its registers are allocated
// on their own (inlined together, the bodies' live ranges interfere).
template <bool UPPER>
__device__ __attribute__((noinline)) void task_diag(const DagParams &p, const DagMat &mt, const Ctx &c,
                                                     double *smem, int *s_flag) {
  double *Pt = smem, *Qt = smem + TILE_D, *Li = smem + 2 * TILE_D, *W = Li + LEAF * LL;
  const int tid = threadIdx.x, j = c.j, n = c.n, T = c.T;
  const long P = c.P;
  if (j > 0) wait_geq(c.col + j - 1, T - j + 1, p.err);
  if (read_uniform(c.failed, s_flag) == 0) {
    load_a_tile<UPPER>(Pt, c.A, n, j, j, true);
    __syncthreads();
    store_tile(c.Af + (long)TB * j * P + TB * j, P, Pt, false, false);
    Acc a;
    acc_from_lds(a, Pt);
    __syncthreads();
    tile_gemm<0, 0>(a, c.F + TB * j, P, c.F + TB * j, P, TB * j, -1.0, Qt);
    acc_to_lds(a, Pt);
    __syncthreads();
    double logsum = 0.0;
    const int bad = lds_potrf64(Pt, Li, s_flag, n - TB * j, logsum);
    if (bad >= 0) {
      if (tid == 0) __hip_atomic_store(c.failed, TB * j + bad + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      lds_trtri64(Pt, Qt, Li, W);
      store_tile(c.Xw + (long)TB * j * P + TB * j, P, Qt, false, false);
      if (tid < 64) {  // lanes 16..63 hold 0
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) logsum += __shfl_xor(logsum, off, 64);
        if (tid == 0) p.logpart[mt.cbase + j] = logsum;
      }
    }
  }
  signal(c.diag + j, c.col + j);
}

template <bool UPPER>
__device__ __attribute__((noinline)) void task_panel(const DagParams &p, const Ctx &c, double *smem,
                                                      int *s_word) {
  double *Pt = smem, *Qt = smem + TILE_D;
  const int i = c.i, j = c.j, T = c.T;
  const long P = c.P;
  if (j > 0) wait_geq(c.col + j - 1, T - j + 1, p.err);
  const bool skip0 = read_uniform(c.failed, s_word) != 0;
  Acc a;
  if (!skip0) {
    load_a_tile<UPPER>(Pt, c.A, c.n, i, j, false);
    __syncthreads();
    store_tile(c.Af + (long)TB * j * P + TB * i, P, Pt, false, false);
    store_tile(c.Af + (long)TB * i * P + TB * j, P, Pt, true, false);
    acc_from_lds(a, Pt);
    __syncthreads();
    tile_gemm<0, 0>(a, c.F + TB * i, P, c.F + TB * j, P, TB * j, -1.0, Qt);
  }
  wait_geq(c.diag + j, 1, p.err);
  if (!skip0 && read_uniform(c.failed, s_word) == 0) {
    acc_to_lds(a, Pt);                                       // C(r, k) at [k LD + r]
    load_tile(Qt, c.Xw + (long)TB * j * P + TB * j, P);      // X_jj(c, k) at [k LD + c]
    __syncthreads();
    acc_zero(a);
    lds_mma(a, Pt, Qt, TB, 1.0);
    __syncthreads();
    acc_to_lds(a, Pt);
    __syncthreads();
    store_tile(c.F + (long)TB * j * P + TB * i, P, Pt, false, false);
  }
  signal(c.col + j, nullptr);
}

__device__ __attribute__((noinline)) void task_trtri(const DagParams &p, const Ctx &c, double *smem,
                                                      int *s_word) {
  double *Pt = smem, *Qt = smem + TILE_D;
  const int i = c.i, j = c.j;
  const long P = c.P;
  wait_geq(c.diag + i, 1, p.err);
  if (i - j - 1 > 0) wait_geq(c.trow + j, i - j - 1, p.err);
  if (read_uniform(c.failed, s_word) == 0) {
    Acc a;
    acc_zero(a);
    tile_gemm<0, 1>(a, c.F + (long)TB * j * P + TB * i, P, c.Xw + (long)TB * j * P + TB * j, P, TB * (i - j),
                    1.0, Qt);
    acc_to_lds_t(a, Qt);                                     // S(k, c) at [k LD + c]
    load_tile(Pt, c.Xw + (long)TB * i * P + TB * i, P);      // X_ii(r, k) at [k LD + r]
    __syncthreads();
    acc_zero(a);
    lds_mma(a, Pt, Qt, TB, -1.0);
    __syncthreads();
    acc_to_lds(a, Pt);
    __syncthreads();
    store_tile(c.Xw + (long)TB * j * P + TB * i, P, Pt, false, false);
  }
  signal(c.trow + j, c.ttot);
}

template <bool UPPER>
__device__ __attribute__((noinline)) void task_lauum(const DagParams &p, const DagMat &mt, const Ctx &c,
                                                      double *smem, int *s_word) {
  double *Pt = smem, *Qt = smem + TILE_D;
  const int i = c.i, j = c.j, T = c.T;
  const long P = c.P;
  wait_geq(c.ttot, T * (T - 1) / 2, p.err);
  wait_geq(c.diag + T - 1, 1, p.err);
  if (read_uniform(c.failed, s_word) == 0) {
    Acc a;
    acc_zero(a);
    tile_gemm<1, 1>(a, c.Xw + (long)TB * i * P + TB * i, P, c.Xw + (long)TB * j * P + TB * i, P,
                    (int)(P - TB * i), 1.0, Qt);
    acc_to_lds(a, Pt);
    __syncthreads();
    const bool dg = i == j;
    store_tile(c.F + (long)TB * j * P + TB * i, P, Pt, false, dg);
    if (!dg) store_tile(c.F + (long)TB * i * P + TB * j, P, Pt, true, false);
    if (p.Ainv != nullptr) store_ainv_tile<UPPER>(p.Ainv + mt.offA, c.n, i, j, Pt);
  }
  signal(c.ltot, nullptr);
}

template <bool UPPER>
__global__ void __launch_bounds__(NTHR, 2) dag_kernel(DagParams p) {
  __shared__ double smem[SMEM_D];
  __shared__ int s_task, s_flag;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_task = __hip_atomic_fetch_add(p.tick, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int t = s_task;
    if (t >= p.ntasks) break;
    const DagTask tk = p.tasks[t];
    const DagMat mt = p.mats[tk.m];
    Ctx c;
    c.T = mt.T;
    c.n = mt.n;
    c.i = tk.i;
    c.j = tk.j;
    c.P = (long)TB * mt.T;
    int *cnt = p.cnt + mt.cbase;
    c.diag = cnt;
    c.col = cnt + mt.T;
    c.trow = cnt + 2 * mt.T;
    c.ttot = cnt + 3 * mt.T;
    c.ltot = c.ttot + 1;
    c.failed = c.ttot + 2;
    c.F = p.F + mt.offP;
    c.Xw = p.Xw + mt.offP;
    c.Af = p.Af + mt.offP;
    c.A = p.A + mt.offA;
    switch (tk.type) {
      case kTaskDiag: task_diag<UPPER>(p, mt, c, smem, &s_flag); break;
      case kTaskPanel: task_panel<UPPER>(p, c, smem, &s_flag); break;
      case kTaskTrtri: task_trtri(p, c, smem, &s_flag); break;
      case kTaskLauum: task_lauum<UPPER>(p, mt, c, smem, &s_flag); break;
      default:  // kTaskSolve
        wait_geq(c.ltot, mt.T * (mt.T + 1) / 2, p.err);
        solve_task(p, mt, tk.m, smem, &s_flag);
        __syncthreads();
        break;
    }
  }
}

__global__ void dag_reset_kernel(int *cnt, long count) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x)
    cnt[e] = 0;
}

}  // namespace

int dag_blocks_per_cu(bool upper) {
  int nb = 0;
  const hipError_t e = upper ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dag_kernel<true>, NTHR, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dag_kernel<false>, NTHR, 0);
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

hipError_t dag_run(hipStream_t s, bool upper, const DagParams &p, long ncounters, int grid) {
  // counters, ticket and error word are contiguous from p.tick
  hipLaunchKernelGGL(dag_reset_kernel, dim3(64), dim3(256), 0, s, p.tick, ncounters);
  HC_TRY(hipGetLastError());
  if (upper) hipLaunchKernelGGL(dag_kernel<true>, dim3(grid), dim3(NTHR), 0, s, p);
  else hipLaunchKernelGGL(dag_kernel<false>, dim3(grid), dim3(NTHR), 0, s, p);
  return hipGetLastError();
}

}  // namespace dag
}  // namespace hc
