// Element-wise and reduction kernels of the mixed-precision solver: FP64 ->
// FP32 conversion of the matrix (lower form) and of right-hand sides, the
// FP32 correction added back into the FP64 solution, column max-norms, and
// the infinity norm of a symmetric matrix stored as one triangle.  All
// reductions are deterministic (fixed trees, no atomics).
#include "internal.h"

#include <algorithm>
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

// One 64 x 64 lower-form tile per workgroup (a 1-D grid over the lower
// tiles, column-major: tile columns bj = 0, 1, ... each from the diagonal
// down, so that a range of tile columns is a range of tiles): the FP32 copy
// as in to_float_lower_kernel, and the tile's share of the infinity norm of
// the symmetric matrix -- row sums of |a| over the tile (diagonal included)
// to rowpart[bj][row], column sums of |a| strictly below the diagonal to
// colpart[bi][column] -- so the matrix is read once for both.
constexpr int NT = 64;
__global__ void __launch_bounds__(256)
to_float_norm_kernel(int lower, int n, const double *A, long lda, float *S, long lds, int *overflow,
                     double *rowpart, double *colpart, long t0, long count) {
  __shared__ double t[NT][NT + 1];  // t[c][r]: lower-form element (r0 + r, c0 + c), 0 outside
  bool ovf = false;
  // tiles t0 .. t0 + count - 1, strided over the grid
  for (long tt = blockIdx.x; tt < count; tt += gridDim.x) {
  // counted from the last tile, the triangle is row-major: tile column
  // nt - 1 - cr holds cr + 1 tiles
  const long nt = (n + NT - 1) / NT, total = nt * (nt + 1) / 2, rev = total - 1 - (t0 + tt);
  int cr = (int)((sqrt(8.0 * (double)rev + 1.0) - 1.0) * 0.5);
  while ((long)cr * (cr + 1) / 2 > rev) cr--;
  while ((long)(cr + 1) * (cr + 2) / 2 <= rev) cr++;
  const int bj = (int)(nt - 1 - cr), bi = (int)(nt - 1 - (rev - (long)cr * (cr + 1) / 2));
  const int r0 = bi * NT, c0 = bj * NT;
  const int li = threadIdx.x & (NT - 1), g = threadIdx.x >> 6;
  // lower: lanes over rows of A (lower-form rows); upper: lanes over rows of A,
  // which are lower-form columns.  Loads from clamped addresses, all in flight.
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int k = g + 4 * u;
    const int R = lower ? r0 + li : r0 + k, C = lower ? c0 + k : c0 + li;
    const int Rc = min(R, n - 1), Cc = min(C, n - 1);
    v[u] = lower ? A[(long)Cc * lda + Rc] : A[(long)Rc * lda + Cc];
  }
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int k = g + 4 * u;
    const int r = lower ? li : k, c = lower ? k : li;
    const int R = r0 + r, C = c0 + c;
    t[c][r] = (R < n && C < n && R >= C) ? v[u] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 16; u++) {  // S column-major, lanes over rows
    const int c = g + 4 * u, R = r0 + li, C = c0 + c;
    if (R < n && C < n && R >= C) {
      const double x = t[c][li];
      ovf |= !(fabs(x) <= (double)FLT_MAX);
      S[(long)C * lds + R] = (float)x;
    }
  }
  if (threadIdx.x < NT) {  // row sums of the tile
    const int R = r0 + threadIdx.x;
    double sum = 0.0;
    for (int c = 0; c < NT; c++) sum += fabs(t[c][threadIdx.x]);
    if (R < n) rowpart[(long)bj * n + R] = sum;
  } else if (threadIdx.x < 2 * NT) {  // column sums strictly below the diagonal
    const int c = threadIdx.x - NT, C = c0 + c;
    double sum = 0.0;
    for (int r = 0; r < NT; r++)
      if (r0 + r > C) sum += fabs(t[c][r]);
    if (C < n) colpart[(long)bi * n + C] = sum;
  }
  __syncthreads();  // t is refilled by the next tile
  }
  if (ovf) *overflow = 1;
}

// Row i of the symmetric matrix: sum_{bj <= i / 64} rowpart[bj][i] +
// sum_{bi >= i / 64} colpart[bi][i]; the block's maximum to bmax.
__global__ void __launch_bounds__(256)
tile_norm_rows_kernel(int n, int nt, const double *rowpart, const double *colpart, double *bmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    // four independent sums (more loads in flight), added in a fixed order
    const int b = i / NT;
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int j = 0; j <= b; j++) a4[j & 3] += rowpart[(long)j * n + i];
#pragma unroll 4
    for (int j = b; j < nt; j++) a4[j & 3] += colpart[(long)j * n + i];
    v = (a4[0] + a4[1]) + (a4[2] + a4[3]);
  }
  const double r = block_max<256>(v);
  if (threadIdx.x == 0) bmax[blockIdx.x] = r;
}

__global__ void __launch_bounds__(256) max_final_kernel(int count, const double *bmax, double *out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < count; i += 256) v = (bmax[i] > v || bmax[i] != bmax[i]) ? bmax[i] : v;
  const double r = block_max<256>(v);
  if (threadIdx.x == 0) out[0] = r;
}

// out[0] = max_i |A(i, i)|, one workgroup
__global__ void __launch_bounds__(256) diag_absmax_kernel(int n, const double *A, long lda, double *out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double a = fabs(A[(long)i * lda + i]);
    v = (a > v || a != a) ? a : v;
  }
  const double r = block_max<256>(v);
  if (threadIdx.x == 0) out[0] = r;
}

}  // namespace

hipError_t diag_absmax(hipStream_t s, long n, const double *A, long lda, double *out) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_absmax_kernel, dim3(1), dim3(256), 0, s, (int)n, A, lda, out);
  return hipGetLastError();
}

long to_float_norm_ws_doubles(long n) {
  const long nt = (n + NT - 1) / NT;
  return 2 * nt * n + (n + 255) / 256 + 1;
}

hipError_t to_float_norm_cols(hipStream_t s, char uplo, long n, const double *A, long lda, float *S, long lds,
                              int *overflow, double *ws, long col0, long col1, long max_wgs) {
  if (n <= 0 || col1 <= col0) return hipSuccess;
  if (col0 % NT != 0 || col0 < 0 || col1 > n || (col1 % NT != 0 && col1 != n)) return hipErrorInvalidValue;
  const long nt = (n + NT - 1) / NT, b0 = col0 / NT, b1 = (col1 + NT - 1) / NT;
  // tiles before tile column b: sum_{k<b} (nt - k)
  auto before = [&](long b) { return b * nt - b * (b - 1) / 2; };
  const long t0 = before(b0), count = before(b1) - t0;
  double *rowpart = ws, *colpart = ws + nt * n;
  const long grid = max_wgs > 0 ? std::min(count, max_wgs) : count;
  hipLaunchKernelGGL(to_float_norm_kernel, dim3((unsigned)grid), dim3(256), 0, s, uplo == 'L' ? 1 : 0,
                     (int)n, A, lda, S, lds, overflow, rowpart, colpart, t0, count);
  return hipGetLastError();
}

hipError_t to_float_norm_finish(hipStream_t s, long n, double *ws, double *anorm) {
  if (n <= 0) return hipSuccess;
  const long nt = (n + NT - 1) / NT;
  double *rowpart = ws, *colpart = ws + nt * n, *bmax = ws + 2 * nt * n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(tile_norm_rows_kernel, dim3(nb), dim3(256), 0, s, (int)n, (int)nt, rowpart, colpart,
                     bmax);
  HC_TRY(hipGetLastError());
  hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, s, (int)nb, bmax, anorm);
  return hipGetLastError();
}

hipError_t to_float_norm(hipStream_t s, char uplo, long n, const double *A, long lda, float *S, long lds,
                         int *overflow, double *ws, double *anorm) {
  HC_TRY(to_float_norm_cols(s, uplo, n, A, lda, S, lds, overflow, ws, 0, n, 0));
  return to_float_norm_finish(s, n, ws, anorm);
}

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

// GMRES bookkeeping on the device: right-hand side k is thread k.  H_k is
// stored by columns (column j of H_k at H + (k m + j) nv), with the Givens
// rotations (cs, sn: k m + j) and the rotated right-hand side g_k (k nv + i).

// Start of a cycle: nrm2[k nv] = ||w_k||^2 of the preconditioned residual.
__global__ void gmres_start_kernel(int K, int nv, const double *nrm2, double *beta, double *g,
                                   double *scale) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const double b = sqrt(nrm2[(long)k * nv]);
  beta[k] = b;
  scale[k] = b > 0.0 ? 1.0 / b : 0.0;
  for (int i = 0; i < nv; i++) g[(long)k * nv + i] = i == 0 ? b : 0.0;
}

// Column j of H_k: the two Gram-Schmidt passes' coefficients (hd, hd + K nv)
// and the new vector's norm (nrm2[k nv + j + 1]); the previous rotations, a
// new one that zeroes H(j + 1, j), and g_k rotated with it.  scale[k] =
// 1 / H(j + 1, j) normalises the new basis vector; active[j] counts the
// right-hand sides whose estimated residual |g(j + 1)| is still above
// tol * beta (and whose Krylov space has not broken down).
__global__ void gmres_rotate_kernel(int K, int nv, int m, int j, const double *hd, const double *nrm2,
                                    double *H, double *cs, double *sn, double *g, const double *beta,
                                    double *scale, double tol, int *active) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double *col = H + ((long)k * m + j) * nv;
  for (int i = 0; i <= j; i++) col[i] = hd[(long)k * nv + i] + hd[(long)K * nv + (long)k * nv + i];
  const double hn = sqrt(nrm2[(long)k * nv + j + 1]);
  col[j + 1] = hn;
  scale[k] = hn > 0.0 ? 1.0 / hn : 0.0;
  double *c = cs + (long)k * m, *sv = sn + (long)k * m, *gk = g + (long)k * nv;
  for (int i = 0; i < j; i++) {  // previous rotations
    const double a = col[i], b = col[i + 1];
    col[i] = c[i] * a + sv[i] * b;
    col[i + 1] = -sv[i] * a + c[i] * b;
  }
  const double a = col[j], b = col[j + 1], rr = hypot(a, b);
  c[j] = rr > 0.0 ? a / rr : 1.0;
  sv[j] = rr > 0.0 ? b / rr : 0.0;
  col[j] = rr;
  col[j + 1] = 0.0;
  gk[j + 1] = -sv[j] * gk[j];
  gk[j] = c[j] * gk[j];
  if (!(fabs(gk[j + 1]) <= tol * beta[k]) && hn > 0.0) atomicAdd(active + j, 1);
}

// y_k = H_k^-1 g_k over the first `used` columns (upper triangular after
// the rotations); a zero pivot (breakdown) gives a zero coefficient.
__global__ void gmres_solve_y_kernel(int K, int nv, int m, int used, const double *H, const double *g,
                                     double *y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const double *Hk = H + (long)k * m * nv;
  double *yk = y + (long)k * nv;
  for (int i = nv - 1; i >= used; i--) yk[i] = 0.0;
  for (int i = used - 1; i >= 0; i--) {
    double t = g[(long)k * nv + i];
    for (int c = i + 1; c < used; c++) t -= Hk[(long)c * nv + i] * yk[c];
    const double d = Hk[(long)i * nv + i];
    yk[i] = d != 0.0 ? t / d : 0.0;
  }
}

}  // namespace

hipError_t gmres_start(hipStream_t s, long nrhs, int nv, const double *nrm2, double *beta, double *g,
                       double *scale) {
  if (nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(gmres_start_kernel, dim3((unsigned)((nrhs + 63) / 64)), dim3(64), 0, s, (int)nrhs, nv,
                     nrm2, beta, g, scale);
  return hipGetLastError();
}

hipError_t gmres_rotate(hipStream_t s, long nrhs, int nv, int m, int j, const double *hd, const double *nrm2,
                        double *H, double *cs, double *sn, double *g, const double *beta, double *scale,
                        double tol, int *active) {
  if (nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(gmres_rotate_kernel, dim3((unsigned)((nrhs + 63) / 64)), dim3(64), 0, s, (int)nrhs, nv,
                     m, j, hd, nrm2, H, cs, sn, g, beta, scale, tol, active);
  return hipGetLastError();
}

hipError_t gmres_solve_y(hipStream_t s, long nrhs, int nv, int m, int used, const double *H, const double *g,
                         double *y) {
  if (nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(gmres_solve_y_kernel, dim3((unsigned)((nrhs + 63) / 64)), dim3(64), 0, s, (int)nrhs, nv,
                     m, used, H, g, y);
  return hipGetLastError();
}

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
