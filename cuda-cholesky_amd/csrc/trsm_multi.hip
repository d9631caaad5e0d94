// Triangular solve with a few right-hand sides, M X = B (M = op(A)
// triangular, n x n; B, X n x nrhs), for the solvers' FP64 and FP32 paths.
//
// Blocked substitution over 128-row blocks with the inverses of M's
// diagonal blocks precomputed (one batched launch per factor, reused by
// every solve): launch b (blocks in solve order) is one workgroup per block
// that still needs it.  Every workgroup forms X_b = inv(M_bb) B_b itself
// (a 128 x 128 by 128 x 16 product, the inputs L2-resident), the
// workgroup of block b stores it, and the others apply it to their own block
// of the right-hand side, B_i -= M_ib X_b.  B_b is only read during launch
// b and X goes to its own array, so no workgroup waits for another; each
// element of M is read once per solve.  Products use the matrix cores
// (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32).
#include "internal.h"

namespace hc {
namespace {

constexpr int SB = 128;        // block order
constexpr int ST_THREADS = 256;
constexpr int KU = 8;          // k-steps whose loads are issued together

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

// M(r, c) = A(r, c), or A(c, r) when trans.
template <typename T>
__device__ __forceinline__ T mget(const T *A, long lda, bool trans, long r, long c) {
  return trans ? A[r * lda + c] : A[c * lda + r];
}

template <typename T>
__global__ void __launch_bounds__(ST_THREADS)
trsm_step_kernel(int lowerlike, int trans, int n, int nrhs, const T *A, long lda, const T *Winv,
                 T *B, long ldb, T *X, long ldx, int b) {
  __shared__ T Xs[16 * (SB + 1)];  // X_b, 128 x 16 column-major, ld SB + 1
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int i = lowerlike ? b + (int)blockIdx.x : b - (int)blockIdx.x;  // target block
  const int c0 = 16 * (int)blockIdx.y;
  const long rb = (long)b * SB, ri = (long)i * SB;
  const int nb_rows = (int)min((long)SB, n - rb), ni_rows = (int)min((long)SB, n - ri);
  const T *W = Winv + (long)b * SB * SB;
  // X_b rows 32 wave + 16 t .. : sum_k Winv(row, k) B_b(k, col)
  Acc16<T> acc[2];
#pragma unroll
  for (int t = 0; t < 2; t++) acc[t].zero();
  // A-operand lane l: (row l % 16, k l / 16); B-operand: (k l / 16, col l % 16).
  // Loads for KU k-steps are issued together so their latencies overlap.
  for (int kc = 0; kc < nb_rows; kc += 4 * KU) {
    T bv[KU], av[2][KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int k = kc + 4 * u + lg, col = c0 + lc;
      bv[u] = (k < nb_rows && col < nrhs) ? B[(long)col * ldb + rb + k] : T(0);
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int row = 32 * wave + 16 * t + lc;
        av[t][u] = (row < nb_rows && k < nb_rows) ? W[(long)k * SB + row] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < KU; u++)
#pragma unroll
      for (int t = 0; t < 2; t++) acc[t].mma(av[t][u], bv[u]);
  }
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 32 * wave + 16 * t + Acc16<T>::row(lane, r), col = lc;
      Xs[col * (SB + 1) + row] = acc[t].v[r];
    }
  __syncthreads();
  if (i == b) {
    for (int e = threadIdx.x; e < SB * 16; e += ST_THREADS) {
      const int row = e % SB, col = e / SB;
      if (row < nb_rows && c0 + col < nrhs) X[(long)(c0 + col) * ldx + rb + row] = Xs[col * (SB + 1) + row];
    }
    return;
  }
  // B_i -= M_ib X_b
#pragma unroll
  for (int t = 0; t < 2; t++) acc[t].zero();
  for (int kc = 0; kc < nb_rows; kc += 4 * KU) {
    T mv[2][KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int k = kc + 4 * u + lg;
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int row = 32 * wave + 16 * t + lc;
        mv[t][u] = (row < ni_rows && k < nb_rows) ? mget(A, lda, trans != 0, ri + row, rb + k) : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int k = kc + 4 * u + lg;
      const T xv = k < nb_rows ? Xs[lc * (SB + 1) + k] : T(0);
#pragma unroll
      for (int t = 0; t < 2; t++) acc[t].mma(mv[t][u], xv);
    }
  }
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 32 * wave + 16 * t + Acc16<T>::row(lane, r), col = c0 + lc;
      if (row < ni_rows && col < nrhs) {
        T *p = B + (long)col * ldb + ri + row;
        *p = *p - acc[t].v[r];
      }
    }
}

}  // namespace

long trsm_multi_ws_elems(long n) { return ((n + SB - 1) / SB) * SB * SB; }

template <typename T>
hipError_t trsm_multi_inverses(hipStream_t s, char uplo, char trans, long n, const T *A, long lda,
                               T *Winv) {
  return trsm_diag_inverses<T>(s, uplo, trans, 'N', n, SB, A, lda, Winv);
}

template <typename T>
hipError_t trsm_multi(hipStream_t s, char uplo, char trans, long n, long nrhs, const T *A, long lda,
                      const T *Winv, T *B, long ldb, T *X, long ldx) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  const bool lowerlike = (uplo == 'L') == (trans == 'N');
  const int nblk = (int)((n + SB - 1) / SB);
  const unsigned gy = (unsigned)((nrhs + 15) / 16);
  for (int step = 0; step < nblk; step++) {
    const int b = lowerlike ? step : nblk - 1 - step;
    const unsigned gx = (unsigned)(lowerlike ? nblk - b : b + 1);
    hipLaunchKernelGGL((trsm_step_kernel<T>), dim3(gx, gy), dim3(ST_THREADS), 0, s, lowerlike ? 1 : 0,
                       trans == 'N' ? 0 : 1, (int)n, (int)nrhs, A, lda, Winv, B, ldb, X, ldx, b);
    HC_TRY(hipGetLastError());
  }
  return hipSuccess;
}

template hipError_t trsm_multi_inverses<double>(hipStream_t, char, char, long, const double *, long,
                                                double *);
template hipError_t trsm_multi_inverses<float>(hipStream_t, char, char, long, const float *, long,
                                               float *);
template hipError_t trsm_multi<double>(hipStream_t, char, char, long, long, const double *, long,
                                       const double *, double *, long, double *, long);
template hipError_t trsm_multi<float>(hipStream_t, char, char, long, long, const float *, long,
                                      const float *, float *, long, float *, long);

}  // namespace hc
