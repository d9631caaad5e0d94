// Element-wise and reduction kernels of the mixed-precision solver: FP64 ->
// FP32 conversion of the matrix (lower form) and of right-hand sides, the
// FP32 correction added back into the FP64 solution, column max-norms, and
// the infinity norm of a symmetric matrix stored as one triangle.  All
// reductions are deterministic (fixed trees, no atomics).
#include "internal.h"

#include <cfloat>

namespace hc {
namespace {

constexpr int TT = 32;  // tile of the transposing conversion

// S (FP32, lower form, ld lds) <- A (FP64) for the lower-form tiles
// (bi >= bj); *overflow <- 1 when a value exceeds the FP32 range.
__global__ void __launch_bounds__(256)
to_float_lower_kernel(int lower, int n, const double *A, long lda, float *S, long lds, int *overflow) {
  __shared__ double t[TT][TT + 1];
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bj > bi) return;
  const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;  // 32 x 8
  const int r0 = bi * TT, c0 = bj * TT;
  bool ovf = false;
  if (lower) {
    for (int k = ty; k < TT; k += 8) {
      const int r = r0 + tx, c = c0 + k;
      if (r < n && c < n && r >= c) {
        const double v = A[(long)c * lda + r];
        ovf |= !(fabs(v) <= (double)FLT_MAX);
        S[(long)c * lds + r] = (float)v;
      }
    }
  } else {
    // lower-form (r, c) = A(c, r): read A's rows c0.. of columns r0.. along c
    for (int k = ty; k < TT; k += 8) {
      const int c = c0 + tx, r = r0 + k;
      if (r < n && c < n) t[k][tx] = A[(long)r * lda + c];
    }
    __syncthreads();
    for (int k = ty; k < TT; k += 8) {
      const int r = r0 + tx, c = c0 + k;
      if (r < n && c < n && r >= c) {
        const double v = t[tx][k];
        ovf |= !(fabs(v) <= (double)FLT_MAX);
        S[(long)c * lds + r] = (float)v;
      }
    }
  }
  if (ovf) *overflow = 1;
}

// Y (FP32) <- X (FP64), m x n.
__global__ void d2s_kernel(int m, int n, const double *X, long ldx, float *Y, long ldy, int *overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m || j >= n) return;
  const double v = X[(long)j * ldx + i];
  if (!(fabs(v) <= (double)FLT_MAX)) *overflow = 1;
  Y[(long)j * ldy + i] = (float)v;
}

// X (FP64) += Y (FP32), m x n.
__global__ void add_s2d_kernel(int m, int n, const float *Y, long ldy, double *X, long ldx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m || j >= n) return;
  X[(long)j * ldx + i] += (double)Y[(long)j * ldy + i];
}

// X (FP64) <- Y (FP32), m x n.
__global__ void s2d_kernel(int m, int n, const float *Y, long ldy, double *X, long ldx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m || j >= n) return;
  X[(long)j * ldx + i] = (double)Y[(long)j * ldy + i];
}

template <int NT>
__device__ __forceinline__ double block_max(double v) {
  __shared__ double red[NT / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < NT / 64; w++) r = fmax(r, red[w]);
  return r;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[NT / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < NT / 64; w++) r += red[w];
  return r;
}

// out[j] = max_i |X(i, j)| (a NaN propagates as NaN).
__global__ void __launch_bounds__(256) colmax_kernel(int m, const double *X, long ldx, double *out) {
  const int j = blockIdx.x;
  double v = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) {
    const double a = fabs(X[(long)j * ldx + i]);
    v = (a > v || a != a) ? a : v;
  }
  const double r = block_max<256>(v);
  if (threadIdx.x == 0) out[j] = r;
}

// Stored part of column j, diagonal included: rows j.. (lower) or ..j (upper).
__global__ void __launch_bounds__(256) abs_colsum_kernel(int lower, int n, const double *A, long lda,
                                                         double *colpart) {
  const int j = blockIdx.x;
  const int i0 = lower ? j : 0, i1 = lower ? n : j + 1;
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) s += fabs(A[(long)j * lda + i]);
  const double r = block_sum<256>(s);
  if (threadIdx.x == 0) colpart[j] = r;
}

// Stored part of row i without the diagonal, restricted to one chunk of
// columns: columns < i (lower) or > i (upper).  Threads walk rows, so every
// column step is one coalesced read.
__global__ void __launch_bounds__(256) abs_rowsum_kernel(int lower, int n, const double *A, long lda,
                                                         int chunk, double *part) {
  const int i = blockIdx.x * 256 + threadIdx.x, cb = blockIdx.y;
  if (i >= n) return;
  const int c0 = cb * chunk, c1 = min(n, c0 + chunk);
  const int lo = lower ? c0 : max(c0, i + 1), hi = lower ? min(c1, i) : c1;
  double s = 0.0;
  for (int c = lo; c < hi; c++) s += fabs(A[(long)c * lda + i]);
  part[(long)cb * n + i] = s;
}

// ||A||_inf = max_i (colpart[i] + sum_cb part[cb][i]).
__global__ void __launch_bounds__(1024) anorm_final_kernel(int n, int nchunk, const double *colpart,
                                                           const double *part, double *out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    double s = colpart[i];
    for (int cb = 0; cb < nchunk; cb++) s += part[(long)cb * n + i];
    v = (s > v || s != s) ? s : v;
  }
  const double r = block_max<1024>(v);
  if (threadIdx.x == 0) out[0] = r;
}

}  // namespace

hipError_t to_float_lower(hipStream_t s, char uplo, long n, const double *A, long lda, float *S,
                          long lds, int *overflow) {
  if (n <= 0) return hipSuccess;
  const unsigned nt = (unsigned)((n + TT - 1) / TT);
  hipLaunchKernelGGL(to_float_lower_kernel, dim3(nt, nt), dim3(256), 0, s, uplo == 'L' ? 1 : 0,
                     (int)n, A, lda, S, lds, overflow);
  return hipGetLastError();
}

hipError_t d2s(hipStream_t s, long m, long n, const double *X, long ldx, float *Y, long ldy,
               int *overflow) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(d2s_kernel, dim3((unsigned)((m + 255) / 256), (unsigned)n), dim3(256), 0, s,
                     (int)m, (int)n, X, ldx, Y, ldy, overflow);
  return hipGetLastError();
}

hipError_t s2d(hipStream_t s, long m, long n, const float *Y, long ldy, double *X, long ldx) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(s2d_kernel, dim3((unsigned)((m + 255) / 256), (unsigned)n), dim3(256), 0, s,
                     (int)m, (int)n, Y, ldy, X, ldx);
  return hipGetLastError();
}

hipError_t add_s2d(hipStream_t s, long m, long n, const float *Y, long ldy, double *X, long ldx) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_s2d_kernel, dim3((unsigned)((m + 255) / 256), (unsigned)n), dim3(256), 0,
                     s, (int)m, (int)n, Y, ldy, X, ldx);
  return hipGetLastError();
}

hipError_t colmax(hipStream_t s, long m, long n, const double *X, long ldx, double *out) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(colmax_kernel, dim3((unsigned)n), dim3(256), 0, s, (int)m, X, ldx, out);
  return hipGetLastError();
}

long sym_norm_ws_doubles(long n) {
  const long nchunk = (n + kNormChunk - 1) / kNormChunk;
  return n + nchunk * n + 1;
}

hipError_t sym_norm_inf(hipStream_t s, char uplo, long n, const double *A, long lda, double *ws,
                        double *out) {
  if (n <= 0) return hipSuccess;
  const int lower = uplo == 'L' ? 1 : 0;
  const long nchunk = (n + kNormChunk - 1) / kNormChunk;
  double *colpart = ws, *part = ws + n;
  hipLaunchKernelGGL(abs_colsum_kernel, dim3((unsigned)n), dim3(256), 0, s, lower, (int)n, A, lda,
                     colpart);
  HC_TRY(hipGetLastError());
  hipLaunchKernelGGL(abs_rowsum_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nchunk),
                     dim3(256), 0, s, lower, (int)n, A, lda, (int)kNormChunk, part);
  HC_TRY(hipGetLastError());
  hipLaunchKernelGGL(anorm_final_kernel, dim3(1), dim3(1024), 0, s, (int)n, (int)nchunk, colpart,
                     part, out);
  return hipGetLastError();
}

}  // namespace hc

// ---------------------------------------------------------------------------
// Krylov-basis helpers of GMRES-based refinement.  Right-hand side k owns the
// basis vectors V_k[:, i] = V + (k * nv + i) * n (each of length n).
// ---------------------------------------------------------------------------
namespace hc {
namespace {

// out[k * nv + i] = V_k[:, i] . w_k  for first <= i < first + nvec (w_k = W + k * ldw)
__global__ void __launch_bounds__(256) basis_dots_kernel(int n, int nv, int first, const double *V,
                                                          const double *W, long ldw, double *out) {
  const int i = first + (int)blockIdx.x, k = blockIdx.y;
  const double *v = V + ((long)k * nv + i) * n, *w = W + (long)k * ldw;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) s += v[r] * w[r];
  const double t = block_sum<256>(s);
  if (threadIdx.x == 0) out[(long)k * nv + i] = t;
}

// w_k -= sum_{i < nvec} h[k * nv + i] V_k[:, i]
__global__ void basis_sub_kernel(int n, int nv, int nvec, const double *V, const double *h, double *W,
                                 long ldw) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (r >= n) return;
  double s = 0.0;
  for (int i = 0; i < nvec; i++) s += h[(long)k * nv + i] * V[((long)k * nv + i) * n + r];
  W[(long)k * ldw + r] -= s;
}

// V_k[:, slot] = w_k * scale[k]
__global__ void basis_put_kernel(int n, int nv, int slot, const double *W, long ldw,
                                 const double *scale, double *V) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (r >= n) return;
  V[((long)k * nv + slot) * n + r] = W[(long)k * ldw + r] * scale[k];
}

// x_k += sum_{i < nvec} y[k * nv + i] V_k[:, i]
__global__ void basis_combine_kernel(int n, int nv, int nvec, const double *V, const double *y,
                                     double *X, long ldx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (r >= n) return;
  double s = 0.0;
  for (int i = 0; i < nvec; i++) s += y[(long)k * nv + i] * V[((long)k * nv + i) * n + r];
  X[(long)k * ldx + r] += s;
}

}  // namespace

hipError_t basis_dots(hipStream_t s, long n, long nrhs, int nv, int first, int nvec, const double *V,
                      const double *W, long ldw, double *out) {
  if (nvec <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(basis_dots_kernel, dim3((unsigned)nvec, (unsigned)nrhs), dim3(256), 0, s, (int)n,
                     nv, first, V, W, ldw, out);
  return hipGetLastError();
}

hipError_t basis_sub(hipStream_t s, long n, long nrhs, int nv, int nvec, const double *V,
                     const double *h, double *W, long ldw) {
  if (nvec <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(basis_sub_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nrhs), dim3(256), 0,
                     s, (int)n, nv, nvec, V, h, W, ldw);
  return hipGetLastError();
}

hipError_t basis_put(hipStream_t s, long n, long nrhs, int nv, int slot, const double *W, long ldw,
                     const double *scale, double *V) {
  hipLaunchKernelGGL(basis_put_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nrhs), dim3(256), 0,
                     s, (int)n, nv, slot, W, ldw, scale, V);
  return hipGetLastError();
}

hipError_t basis_combine(hipStream_t s, long n, long nrhs, int nv, int nvec, const double *V,
                         const double *y, double *X, long ldx) {
  if (nvec <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(basis_combine_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nrhs), dim3(256),
                     0, s, (int)n, nv, nvec, V, y, X, ldx);
  return hipGetLastError();
}

}  // namespace hc
