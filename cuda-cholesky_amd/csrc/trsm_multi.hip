// Triangular solve with a few right-hand sides, M X = B (M = op(A)
// triangular, n x n; B, X n x nrhs), for the solvers' FP64 and FP32 paths.
//
// Blocked substitution over 128-row blocks with the inverses of M's diagonal
// blocks precomputed (one batched launch per factor, reused by every solve),
// as ONE launch per solve: a chain of workgroups, one per block row of each
// 16-column slice of the right-hand side.
//   * A workgroup takes a ticket (atomic counter of its slice); ticket s is
//     the block row solved s-th (top-down for a lower-like M, bottom-up
//     otherwise).  Tickets go out in the order workgroups start running, so
//     a workgroup only ever waits for workgroups that are already resident:
//     the chain cannot deadlock, however many workgroups fit on the device.
//   * Block row i accumulates sum_j M_ij X_j over the blocks j solved before
//     it, each as soon as block j's ready flag is set (the M_ij loads of the
//     next block are in flight while it waits; one wave polls 64 flags at a
//     time and acquires a run of ready blocks with one fence), then forms
//     X_i = inv(M_ii) (B_i - sum) with inv(M_ii) staged in LDS at the start,
//     stores X_i, and sets its own flag.
//   * Solved blocks are also stored in a compact staging copy laid out in the
//     order the consumers' lanes read them (one 16-byte vector per lane and
//     k-group): X in its own layout puts the 16 columns of a block ldx apart,
//     and gathering them cost more than the rest of a chain step.  All waves
//     store and wait for their stores, one barrier, then a single wave
//     publishes (one L2 write-back and the flag).
//   * Waits are bounded: a wait that outlasts kSpinLimit polls sets the
//     error word sync[0] and stops waiting, so every wave exits; the host
//     turns a set error word into a failed call.
// Each element of M is read once per solve; the chain costs one flag hand-off
// and two small products per block.  Products use the matrix cores
// (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32); eight waves take 16 rows
// of the block each.
#include "internal.h"
#include "publish.h"

namespace hc {
namespace {

constexpr int SB = 128;            // block order
constexpr int CH_THREADS = 512;    // 8 waves x 16 rows
constexpr int kSpinLimit = 1 << 19;  // polls before a wait is abandoned (well over a second)

// error word, tickets and ready flags, padded to 256 bytes; the staging
// copies follow
__host__ __device__ inline long chain_flags_bytes(long gy, long nblk) {
  return ((1 + gy + gy * nblk) * (long)sizeof(int) + 255) / 256 * 256;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16 x 16 accumulator; lane l, register r holds element
//   double: (l / 16 + 4 r, l % 16)     float: (4 (l / 16) + r, l % 16)
template <typename T> struct Acc16;
template <> struct Acc16<double> {
  v4d v;
  __device__ static int row(int l, int r) { return (l >> 4) + 4 * r; }
  __device__ void zero() { v = (v4d){0.0, 0.0, 0.0, 0.0}; }
  __device__ void mma(double a, double b) { v = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, v, 0, 0, 0); }
};
template <> struct Acc16<float> {
  v4f v;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ void zero() { v = (v4f){0.0f, 0.0f, 0.0f, 0.0f}; }
  __device__ void mma(float a, float b) { v = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, v, 0, 0, 0); }
};

// 16-byte vector of T
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); };

// One lane's share of a 16-row x 128-deep operand.  Lane (c = l % 16,
// g = l / 16) holds k = 4 VE q + VE g + e for q < NQ, e < VE: a 16-byte run
// of k per q, so a k-contiguous operand loads as vectors.
template <typename T>
struct Slice {
  static constexpr int VE = 16 / (int)sizeof(T);
  static constexpr int NQ = SB / (4 * VE);
  T v[NQ][VE];
  __device__ static int k(int q, int g, int e) { return 4 * VE * q + VE * g + e; }
};

// Rows 16 w .. 16 w + 15 of M_ij (rows ri.., columns rb.. of M), masked to
// the ni x nb block.  TRANS: M(r, c) = A(c, r) (k-contiguous rows of A^T);
// VEC: 16-byte loads (A 16-byte aligned, lda a multiple of VE).
template <typename T, bool TRANS, bool VEC>
__device__ __forceinline__ void load_m(Slice<T> &m, const T *A, long lda, long ri, long rb, int row,
                                       int g, int ni, int nb) {
  constexpr int VE = Slice<T>::VE, NQ = Slice<T>::NQ;
  if (TRANS && VEC && nb == SB) {
    const int rr = min(row, ni - 1);  // rows past the block load a valid row, then read as 0
    const T *p = A + (ri + rr) * lda + rb;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const typename Vec16<T>::type x = *reinterpret_cast<const typename Vec16<T>::type *>(p + Slice<T>::k(q, g, 0));
#pragma unroll
      for (int e = 0; e < VE; e++) m.v[q][e] = row < ni ? x[e] : T(0);
    }
    return;
  }
  // element loads from clamped (always valid) addresses, masked afterwards
  const int rr = min(row, ni - 1);
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int e = 0; e < VE; e++) {
      const int k = Slice<T>::k(q, g, e), kk = min(k, nb - 1);
      const T x = TRANS ? A[(ri + rr) * lda + rb + kk] : A[(rb + kk) * lda + ri + rr];
      m.v[q][e] = (row < ni && k < nb) ? x : T(0);
    }
}

// Wait (one wave) until solve position sp is ready.  Lane l polls the ready
// flag of position sp + l (agent scope); the run of set flags from sp, up to
// s, is acquired with ONE agent-scope fence.  A workgroup far from the head
// of the chain polls rarely (it catches up on a run of blocks at a time), so
// the flag lines see few requests.  Returns the first position not known
// ready, or -1 when the wait was abandoned (error word set).
__device__ __forceinline__ int wait_ready(const int *ready, bool lowerlike, int nblk, int sp, int s,
                                          int lane, int *err) {
  const int p = sp + lane;
  const int *f = ready + (lowerlike ? p : nblk - 1 - p);
  for (int polls = 0;; polls++) {
    const int v = p < s ? __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 1;
    const unsigned long long set = __ballot(v != 0);
    if (set & 1ull) {
      const int run = ~set == 0ull ? 64 : __builtin_ctzll(~set);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      return min(sp + run, s);
    }
    if (polls >= kSpinLimit) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return -1;
    }
    for (int t = min(s - sp, 64); t > 0; t--) __builtin_amdgcn_s_sleep(4);
  }
}

// Row-split accumulation (one 16-row strip per wave over all of k) or, for
// the k-split (KS) form, partial sums of the four k quarters.
template <typename T, bool KS> struct ChainLds {
  static constexpr int WLD = sizeof(T) == 4 ? SB + 16 : SB + 4;  // inv(M_ii), column stride
  T W[SB * WLD];
  T Y[16 * (SB + 1)];  // B_i - sum, 128 x 16 column-major
};
template <typename T> struct ChainLds<T, true> {
  static constexpr int WLD = SB + 16;
  T W[SB * WLD];
  T Y[16 * (SB + 1)];
  T P[4 * 16 * (SB + 1)];  // k-quarter partial sums, [quarter][column][row]
};

// staging copy of block (slice y, block j), laid out for the consumers'
// lanes: element (k, c) of the 128 x 16 block at
//   row split: ((q * 4 + g) * 16 + c) * VE + e   for k = 4 VE q + VE g + e
//   k split:   ((h * 4 + g) * 16 + c) * 8 + u    for k = 32 h + 4 u + g
template <typename T>
__device__ __forceinline__ T *stage_block(int *sync, int gy, int nblk, int y, int j) {
  T *base = reinterpret_cast<T *>(reinterpret_cast<char *>(sync) + chain_flags_bytes(gy, nblk));
  return base + ((long)y * nblk + j) * SB * 16;
}
template <typename T, bool KS>
__device__ __forceinline__ int stage_index(int k, int c) {
  if (KS) return (((k >> 5) * 4 + (k & 3)) * 16 + c) * 8 + ((k & 31) >> 2);
  constexpr int VE = Slice<T>::VE;
  const int q = k / (4 * VE), g = (k % (4 * VE)) / VE, e = k % VE;
  return ((q * 4 + g) * 16 + c) * VE + e;
}

// k-split operand of the non-transposed FP32 product: wave (h, kq) takes rows
// 64 h .. 64 h + 63 and k = 32 kq .. 32 kq + 31 of M_ij.  At step u lane
// (i = l % 16, g = l / 16) loads rows 64 h + 4 i .. + 3 of column
// k = 32 kq + 4 u + g as one 16-byte vector; element t feeds the product of
// accumulator t, whose row i stands for block row 64 h + 4 i + t.
typedef float v4f_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load_mk(v4f_t (&d)[8], const float *A, long lda, long ri, long rb, int h,
                                        int kq, int i, int g, int ni, int nb) {
  if (ni == SB && nb == SB) {
    const float *p = A + (rb + 32 * kq + g) * lda + ri + 64 * h + 4 * i;
#pragma unroll
    for (int u = 0; u < 8; u++) d[u] = *reinterpret_cast<const v4f_t *>(p + 4 * u * lda);
    return;
  }
  // partial block: element loads from clamped (always valid) addresses
#pragma unroll
  for (int u = 0; u < 8; u++)
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int k = 32 * kq + 4 * u + g, r = 64 * h + 4 * i + t;
      const float x = A[(rb + min(k, nb - 1)) * lda + ri + min(r, ni - 1)];
      d[u][t] = (k < nb && r < ni) ? x : 0.0f;
    }
}

template <typename T, bool TRANS, bool VEC>
__global__ void __launch_bounds__(CH_THREADS)
trsm_chain_kernel(int lowerlike, int n, int nrhs, const T *A, long lda, const T *Winv, const T *B,
                  long ldb, T *X, long ldx, int *sync) {
  // non-transposed FP32 with 16-byte loads: k-split accumulation
  constexpr bool KS = !TRANS && VEC;
  static_assert(!KS || sizeof(T) == 4, "k-split form is FP32 only");
  __shared__ ChainLds<T, KS> sm;
  __shared__ int ticket, known_sh;
  constexpr int WLD = ChainLds<T, KS>::WLD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int nblk = (n + SB - 1) / SB;
  const int y = blockIdx.y, c0 = 16 * y;
  int *err = sync;
  int *tick = sync + 1 + y;
  int *ready = sync + 1 + gridDim.y + (long)y * nblk;
  if (tid == 0) ticket = atomicAdd(tick, 1);
  __syncthreads();
  const int s = ticket;
  const int i = lowerlike ? s : nblk - 1 - s;
  const long ri = (long)i * SB;
  const int ni = (int)min((long)SB, n - ri);
  const int row = 16 * wave + lc;

  // inv(M_ii) to LDS, zero outside the ni x ni block
  {
    const T *Wg = Winv + (long)i * SB * SB;
    for (int e = tid; e < SB * SB; e += CH_THREADS) {
      const int r = e % SB, k = e / SB;
      sm.W[k * WLD + r] = (r < ni && k < ni) ? Wg[e] : T(0);
    }
  }
  // B_i in accumulator layout
  T bv[4];
  if constexpr (!KS)
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int rr = 16 * wave + Acc16<T>::row(lane, r), col = c0 + lc;
    bv[r] = (rr < ni && col < nrhs) ? B[(long)col * ldb + ri + rr] : T(0);
  }

  Acc16<T> acc;
  acc.zero();
  int known = 0;  // positions below are ready and acquired; -1: waits abandoned
  if constexpr (KS) {
    const int h = wave >> 2, kq = wave & 3;
    // B_i in the layout of the final reduction (element tid + 512 m), loaded
    // now so that no global load sits between the last wait and the solve
    T bk[16 * SB / CH_THREADS];
#pragma unroll
    for (int m = 0; m < 16 * SB / CH_THREADS; m++) {
      const int e = tid + m * CH_THREADS, c = e / SB, rr = e % SB, col = c0 + c;
      bk[m] = (rr < ni && col < nrhs) ? B[(long)col * ldb + ri + rr] : T(0);
    }
    Acc16<float> a4[4];
#pragma unroll
    for (int t = 0; t < 4; t++) a4[t].zero();
    if (s > 0) {
      v4f_t mv[8], mvn[8];
      auto blk = [&](int sp) { return lowerlike ? sp : nblk - 1 - sp; };
      auto rows_of = [&](int j) { return (int)min((long)SB, n - (long)j * SB); };
      load_mk(mv, reinterpret_cast<const float *>(A), lda, ri, (long)blk(0) * SB, h, kq, lc, lg, ni,
              rows_of(blk(0)));
      for (int sp = 0; sp < s; sp++) {
        const int j = blk(sp);
        if (known >= 0 && sp >= known) {
          if (wave == 0) {
            const int kn = wait_ready(ready, lowerlike, nblk, sp, s, lane, err);
            if (lane == 0) known_sh = kn;
          }
          __syncthreads();
          known = known_sh;
        }
        // X_j rows 32 kq + 4 u + g, column lc: 8 values, two 16-byte vectors
        v4f_t x0, x1;
        {
          const v4f_t *xp = reinterpret_cast<const v4f_t *>(
              reinterpret_cast<const float *>(stage_block<T>(sync, gridDim.y, nblk, y, j)) +
              ((kq * 4 + lg) * 16 + lc) * 8);
          x0 = xp[0];
          x1 = xp[1];
        }
        if (sp + 1 < s) {
          const int jn = blk(sp + 1);
          load_mk(mvn, reinterpret_cast<const float *>(A), lda, ri, (long)jn * SB, h, kq, lc, lg, ni,
                  rows_of(jn));
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const float xb = u < 4 ? x0[u & 3] : x1[u & 3];
#pragma unroll
          for (int t = 0; t < 4; t++) a4[t].mma(mv[u][t], xb);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) mv[u] = mvn[u];
      }
    }
    // partial sums: accumulator t, lane l, register r is block row
    // 64 h + 16 (l / 16) + 4 r + t, column l % 16
    float *P = reinterpret_cast<float *>(sm.P) + kq * 16 * (SB + 1);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) P[lc * (SB + 1) + 64 * h + 16 * lg + 4 * r + t] = a4[t].v[r];
    __syncthreads();
    // Y = B_i - (sum of the four quarters, in a fixed order)
#pragma unroll
    for (int m = 0; m < 16 * SB / CH_THREADS; m++) {
      const int e = tid + m * CH_THREADS, c = e / SB, rr = e % SB;
      const float *Q = reinterpret_cast<const float *>(sm.P) + c * (SB + 1) + rr;
      const float sum = ((Q[0] + Q[16 * (SB + 1)]) + Q[32 * (SB + 1)]) + Q[48 * (SB + 1)];
      sm.Y[c * (SB + 1) + rr] = bk[m] - sum;
    }
  } else if (s > 0) {
    Slice<T> m, mn, xs;
    auto blk = [&](int sp) { return lowerlike ? sp : nblk - 1 - sp; };
    auto rows_of = [&](int j) { return (int)min((long)SB, n - (long)j * SB); };
    load_m<T, TRANS, VEC>(m, A, lda, ri, (long)blk(0) * SB, row, lg, ni, rows_of(blk(0)));
    for (int sp = 0; sp < s; sp++) {
      const int j = blk(sp);
      if (known >= 0 && sp >= known) {  // uniform: wave 0 polls, the others wait at the barrier
        if (wave == 0) {
          const int kn = wait_ready(ready, lowerlike, nblk, sp, s, lane, err);
          if (lane == 0) known_sh = kn;
        }
        __syncthreads();
        known = known_sh;
      }
      // X_j from its staging copy (rows past the block and columns past
      // nrhs hold exact zeros), then the next block's M slice, then the product
      {
        typedef typename Vec16<T>::type V;
        const V *xp = reinterpret_cast<const V *>(stage_block<T>(sync, gridDim.y, nblk, y, j)) + lane;
#pragma unroll
        for (int q = 0; q < Slice<T>::NQ; q++) {
          const V x = xp[q * 64];
#pragma unroll
          for (int e = 0; e < Slice<T>::VE; e++) xs.v[q][e] = x[e];
        }
      }
      if (sp + 1 < s) {
        const int jn = blk(sp + 1);
        load_m<T, TRANS, VEC>(mn, A, lda, ri, (long)jn * SB, row, lg, ni, rows_of(jn));
      }
#pragma unroll
      for (int q = 0; q < Slice<T>::NQ; q++)
#pragma unroll
        for (int e = 0; e < Slice<T>::VE; e++) acc.mma(m.v[q][e], xs.v[q][e]);
      m = mn;
    }
  }
  // Y = B_i - sum  ->  X_i = inv(M_ii) Y
  if constexpr (!KS) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int rr = 16 * wave + Acc16<T>::row(lane, r);
      sm.Y[lc * (SB + 1) + rr] = bv[r] - acc.v[r];
    }
  }
  __syncthreads();
  // four independent accumulators (k = 16 u' + 4 t + g, t = u mod 4): the
  // product is on the chain, and one accumulator would serialise its 32
  // matrix-core steps
  {
    Acc16<T> a4[4];
#pragma unroll
    for (int t = 0; t < 4; t++) a4[t].zero();
    if constexpr (sizeof(T) == 4) {
      // every operand read first (64 LDS loads in flight), then the products:
      // the solve is on the chain (FP32 forward solve 1.46 -> 1.36 ms at
      // n = 32768; FP64 has no registers to spare for it)
      T wv[SB / 4], yv[SB / 4];
#pragma unroll
      for (int u = 0; u < SB / 4; u++) {
        const int k = 4 * u + lg;
        wv[u] = sm.W[k * WLD + row];
        yv[u] = sm.Y[lc * (SB + 1) + k];
      }
#pragma unroll
      for (int u = 0; u < SB / 4; u++) a4[u & 3].mma(wv[u], yv[u]);
    } else {
#pragma unroll 8
      for (int u = 0; u < SB / 4; u++) {
        const int k = 4 * u + lg;
        a4[u & 3].mma(sm.W[k * WLD + row], sm.Y[lc * (SB + 1) + k]);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) acc.v[r] = (a4[0].v[r] + a4[1].v[r]) + (a4[2].v[r] + a4[3].v[r]);
  }
  // X_i to the staging copy (rows >= ni and columns >= nrhs are exact zeros:
  // B, the sum and inv(M_ii) are zero there)
  {
    T *xt = stage_block<T>(sync, gridDim.y, nblk, y, i);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = 16 * wave + Acc16<T>::row(lane, r);
      xt[stage_index<T, KS>(k, lc)] = acc.v[r];
    }
  }
  // publish: every wave waits for its own global stores to reach the L2 (the
  // barrier alone waits only for LDS traffic), then wave 0 writes the L2 back
  // once and sets the flag
  publish_wait_own();
  __syncthreads();
  if (wave == 0) {
    publish_release();
    if (lane == 0) __hip_atomic_store(ready + i, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // X_i in the caller's layout after the flag: no workgroup reads it, so its
  // stores stay off the chain
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int k = 16 * wave + Acc16<T>::row(lane, r), col = c0 + lc;
    if (k < ni && col < nrhs) X[(long)col * ldx + ri + k] = acc.v[r];
  }
}

// Clears the tickets and ready flags (sync[1 ..]); the error word is left.
__global__ void chain_reset_kernel(int *sync, long count) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x)
    sync[1 + e] = 0;
}

}  // namespace

long trsm_multi_ws_elems(long n) { return ((n + SB - 1) / SB) * SB * SB; }

size_t trsm_multi_sync_bytes(long n, long nrhs) {
  const long nblk = (n + SB - 1) / SB, gy = (nrhs + 15) / 16;
  return (size_t)chain_flags_bytes(gy, nblk) + (size_t)(gy * nblk) * SB * 16 * sizeof(double);
}

template <typename T>
hipError_t trsm_multi_inverses(hipStream_t s, char uplo, char trans, long n, const T *A, long lda,
                               T *Winv) {
  return trsm_diag_inverses<T>(s, uplo, trans, 'N', n, SB, A, lda, Winv);
}

template <typename T>
hipError_t trsm_multi(hipStream_t s, char uplo, char trans, long n, long nrhs, const T *A, long lda,
                      const T *Winv, const T *B, long ldb, T *X, long ldx, int *sync) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (n > (long)INT32_MAX - SB || nrhs > 16L * 65535) return hipErrorInvalidValue;
  const bool lowerlike = (uplo == 'L') == (trans == 'N');
  const bool trans_m = trans != 'N';  // M(r, c) = A(c, r)
  const int nblk = (int)((n + SB - 1) / SB);
  const unsigned gy = (unsigned)((nrhs + 15) / 16);
  const long cnt = (1 + gy + (long)gy * nblk) - 1;
  hipLaunchKernelGGL(chain_reset_kernel, dim3((unsigned)std::min<long>((cnt + 255) / 256, 1024)),
                     dim3(256), 0, s, sync, cnt);
  HC_TRY(hipGetLastError());
  // 16-byte loads of k-contiguous rows (FP32 only: the FP64 variant would spill)
  const bool vec = sizeof(T) == 4 && (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (lda % 4 == 0);
  const dim3 grid((unsigned)nblk, gy), block(CH_THREADS);
  const int ll = lowerlike ? 1 : 0, nn = (int)n, nr = (int)nrhs;
  if (trans_m) {
    if (vec)
      hipLaunchKernelGGL((trsm_chain_kernel<T, true, sizeof(T) == 4>), grid, block, 0, s, ll, nn, nr, A,
                         lda, Winv, B, ldb, X, ldx, sync);
    else
      hipLaunchKernelGGL((trsm_chain_kernel<T, true, false>), grid, block, 0, s, ll, nn, nr, A, lda,
                         Winv, B, ldb, X, ldx, sync);
  } else {
    if (vec)
      hipLaunchKernelGGL((trsm_chain_kernel<T, false, sizeof(T) == 4>), grid, block, 0, s, ll, nn, nr, A,
                         lda, Winv, B, ldb, X, ldx, sync);
    else
      hipLaunchKernelGGL((trsm_chain_kernel<T, false, false>), grid, block, 0, s, ll, nn, nr, A, lda,
                         Winv, B, ldb, X, ldx, sync);
  }
  return hipGetLastError();
}

template hipError_t trsm_multi_inverses<double>(hipStream_t, char, char, long, const double *, long,
                                                double *);
template hipError_t trsm_multi_inverses<float>(hipStream_t, char, char, long, const float *, long,
                                               float *);
template hipError_t trsm_multi<double>(hipStream_t, char, char, long, long, const double *, long,
                                       const double *, const double *, long, double *, long, int *);
template hipError_t trsm_multi<float>(hipStream_t, char, char, long, long, const float *, long,
                                      const float *, const float *, long, float *, long, int *);

}  // namespace hc
