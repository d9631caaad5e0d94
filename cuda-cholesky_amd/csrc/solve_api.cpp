// Solvers on top of the factorisation: xPOTRS (FP64 solve with a Cholesky
// factor) and the mixed-precision xSPOSV (low-precision factor + FP64
// iterative refinement).
//
// Neither exists in the reference; their interfaces follow LAPACK's DPOTRS
// and DSPOSV (argument order, INFO codes, ITER codes, the refinement stopping
// test and the fall-back to a full FP64 factorisation) with the reference's
// conventions added: a handle first, device byte addresses (CUdeviceptr) for
// matrices, CUresult returns and XERBLA on bad arguments.
#include "api_util.h"

#include <cmath>
#include <vector>

using hc::bad_arg;
using hc::dptr;

namespace {

#define CU_TRY(call)                       \
  do {                                     \
    const CUresult cu_r_ = (call);         \
    if (cu_r_ != CUDA_SUCCESS) return cu_r_; \
  } while (0)

constexpr long kIterMax = 30;  // refinement steps before falling back (LAPACK ITERMAX)

// B <- inv(A) B for the factor in triangle uplo (L L^T or U^T U), n x nrhs.
template <typename T>
hipError_t potrs_device(hipStream_t s, char uplo, long n, long nrhs, const T *F, long ldf, T *B,
                        long ldb, T *ws) {
  const T one = hc::from_real<T>(1.0);
  if (uplo == 'L') {
    HC_TRY(hc::trsm<T>(s, 'L', 'L', 'N', 'N', n, nrhs, one, F, ldf, B, ldb, ws));
    return hc::trsm<T>(s, 'L', 'L', 'T', 'N', n, nrhs, one, F, ldf, B, ldb, ws);
  }
  HC_TRY(hc::trsm<T>(s, 'L', 'U', 'T', 'N', n, nrhs, one, F, ldf, B, ldb, ws));
  return hc::trsm<T>(s, 'L', 'U', 'N', 'N', n, nrhs, one, F, ldf, B, ldb, ws);
}

// R <- B - A X with A symmetric in triangle uplo: the stored triangle times X
// plus the transpose of its strict part times X (two masked GEMMs).
hipError_t residual(hipStream_t s, char uplo, long n, long nrhs, const double *A, long lda,
                    const double *B, long ldb, const double *X, long ldx, double *R, long ldr) {
  HC_TRY(hc::copy2d<double>(s, n, nrhs, B, ldb, R, ldr));
  hc::GemmOpts o1;
  o1.tri_a = (uplo == 'L') ? hc::TRI_LOWER : hc::TRI_UPPER;
  HC_TRY(hc::gemm<double>(s, hc::OP_N, hc::OP_N, n, nrhs, n, -1.0, A, lda, X, ldx, 1.0, R, ldr, R,
                          ldr, o1));
  hc::GemmOpts o2;  // op(A) = A^T is upper-like (lower storage) / lower-like (upper)
  o2.tri_a = (uplo == 'L') ? hc::TRI_UPPER : hc::TRI_LOWER;
  o2.zero_diag = true;
  return hc::gemm<double>(s, hc::OP_T, hc::OP_N, n, nrhs, n, -1.0, A, lda, X, ldx, 1.0, R, ldr, R,
                          ldr, o2);
}

// Workspace layout of the mixed solver (byte offsets, 256-aligned).
struct SposvLayout {
  size_t S, W, T, R, RS, NORM, DEV, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

SposvLayout sposv_layout(long n, long nrhs) {
  SposvLayout L;
  size_t off = 0;
  L.S = off;    off = align256(off + (size_t)n * n * sizeof(float));              // factor
  L.W = off;    off = align256(off + 128 * 128 * sizeof(double));                 // diag inverse
  L.T = off;    off = align256(off + (size_t)hc::trsm_ws_elems('L', n, nrhs) * sizeof(double));
  L.R = off;    off = align256(off + (size_t)n * nrhs * sizeof(double));          // residual
  L.RS = off;   off = align256(off + (size_t)n * nrhs * sizeof(float));           // FP32 rhs
  L.NORM = off; off = align256(off + (size_t)hc::sym_norm_ws_doubles(n) * sizeof(double));
  L.DEV = off;  off = align256(off + (3 * (size_t)nrhs + 8) * sizeof(double));    // norms, flags
  L.total = off;
  return L;
}

struct Norms {
  std::vector<double> r, x;
};

}  // namespace

extern "C" {

CUresult cuDpotrs(CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, long *info) {
  const char *name = "cuDpotrs";
  if (info == nullptr) return bad_arg(name, 8);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }  // handle
  if (!hc::valid_uplo(uplo)) { *info = -1; return bad_arg(name, 1); }
  if (lda < (n > 1 ? n : 1)) { *info = -5; return bad_arg(name, 5); }
  if (ldb < (n > 1 ? n : 1)) { *info = -7; return bad_arg(name, 7); }
  if (n == 0 || nrhs == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  void *ws = nullptr;
  HC_CHECK(hc::lapack_workspace(h, (size_t)hc::trsm_ws_elems('L', (long)n, (long)nrhs) * sizeof(double),
                                &ws));
  HC_CHECK(potrs_device<double>(h->stream, (char)uplo, (long)n, (long)nrhs, dptr<double>(A), (long)lda,
                                dptr<double>(B), (long)ldb, static_cast<double *>(ws)));
  HC_CHECK(hipStreamSynchronize(h->stream));
  return CUDA_SUCCESS;
}

}  // extern "C"

namespace {

CUresult sposv(const char *name, CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs,
               CUdeviceptr A, size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx,
               char factor, long *iter, char *used, long *info) {
  if (info == nullptr) return bad_arg(name, 14);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }
  if (!hc::valid_uplo(uplo)) { *info = -1; return bad_arg(name, 1); }
  if (lda < (n > 1 ? n : 1)) { *info = -5; return bad_arg(name, 5); }
  if (ldb < (n > 1 ? n : 1)) { *info = -7; return bad_arg(name, 7); }
  if (ldx < (n > 1 ? n : 1)) { *info = -9; return bad_arg(name, 9); }
  if (factor != 'S' && factor != 'B') { *info = -10; return bad_arg(name, 10); }
  if (iter == nullptr) { *info = -11; return bad_arg(name, 11); }
  *iter = 0;
  if (used != nullptr) *used = factor;
  if (n == 0 || nrhs == 0) return CUDA_SUCCESS;

  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  const long N = (long)n, K = (long)nrhs;
  const char ul = (char)uplo;
  const SposvLayout L = sposv_layout(N, K);
  void *base = nullptr;
  HC_CHECK(hc::lapack_workspace(h, L.total, &base));
  char *wsb = static_cast<char *>(base);
  float *S = reinterpret_cast<float *>(wsb + L.S);
  float *W = reinterpret_cast<float *>(wsb + L.W);
  void *T = wsb + L.T;
  double *R = reinterpret_cast<double *>(wsb + L.R);
  float *RS = reinterpret_cast<float *>(wsb + L.RS);
  double *NW = reinterpret_cast<double *>(wsb + L.NORM);
  double *DEV = reinterpret_cast<double *>(wsb + L.DEV);
  double *dnorm = DEV;                       // ||A||_inf
  double *rn = DEV + 1, *xn = DEV + 1 + K;   // column max-norms
  int *ovf = reinterpret_cast<int *>(DEV + 1 + 2 * K);
  const double *Ad = dptr<double>(A);
  const double *Bd = dptr<double>(B);
  double *Xd = dptr<double>(X);
  hipStream_t s = h->stream;

  // ||A||_inf; refinement stops when every column has
  // ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n)   (LAPACK dsposv, eps = 2^-53)
  HC_CHECK(hc::sym_norm_inf(s, ul, N, Ad, (long)lda, NW, dnorm));
  double anorm = 0.0;
  HC_CHECK(hipMemcpyAsync(&anorm, dnorm, sizeof(double), hipMemcpyDeviceToHost, s));
  HC_CHECK(hipStreamSynchronize(s));
  const double cte = anorm * std::ldexp(1.0, -53) * std::sqrt((double)N);

  auto read_norms = [&](Norms &nm) -> CUresult {
    HC_CHECK(hc::colmax(s, N, K, R, N, rn));
    HC_CHECK(hc::colmax(s, N, K, Xd, (long)ldx, xn));
    nm.r.resize(K);
    nm.x.resize(K);
    HC_CHECK(hipMemcpyAsync(nm.r.data(), rn, K * sizeof(double), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipMemcpyAsync(nm.x.data(), xn, K * sizeof(double), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
    return CUDA_SUCCESS;
  };
  auto converged = [&](const Norms &nm) {
    for (long j = 0; j < K; j++)
      if (!(nm.r[j] <= nm.x[j] * cte)) return false;
    return true;
  };
  auto read_flag = [&](int *flag) -> CUresult {
    int v = 0;
    HC_CHECK(hipMemcpyAsync(&v, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
    *flag = v;
    return CUDA_SUCCESS;
  };

  // One attempt with a low-precision factor; returns the LAPACK ITER code.
  auto attempt = [&](bool bf16, long *code) -> CUresult {
    HC_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
    HC_CHECK(hc::to_float_lower(s, ul, N, Ad, (long)lda, S, N, ovf));
    int flag = 0;
    CU_TRY(read_flag(&flag));
    if (flag) { *code = -2; return CUDA_SUCCESS; }       // narrowing overflowed
    HC_CHECK(hipMemsetAsync(h->dinfo, 0, sizeof(int), s));
    HC_CHECK(hc::potrf_on_device<float>(h, 'L', N, S, N, W, bf16));
    long finfo = 0;
    HC_CHECK(hipMemcpyAsync(h->hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
    finfo = h->hinfo[0];
    if (finfo != 0) { *code = -3; return CUDA_SUCCESS; }  // low-precision factorisation failed
    // x0 = inv(S S^T) b in FP32
    HC_CHECK(hc::d2s(s, N, K, Bd, (long)ldb, RS, N, ovf));
    CU_TRY(read_flag(&flag));
    if (flag) { *code = -2; return CUDA_SUCCESS; }
    HC_CHECK(potrs_device<float>(s, 'L', N, K, S, N, RS, N, static_cast<float *>(T)));
    HC_CHECK(hc::s2d(s, N, K, RS, N, Xd, (long)ldx));
    Norms nm;
    for (long it = 0;; it++) {
      HC_CHECK(residual(s, ul, N, K, Ad, (long)lda, Bd, (long)ldb, Xd, (long)ldx, R, N));
      CU_TRY(read_norms(nm));
      if (converged(nm)) { *code = it; return CUDA_SUCCESS; }
      if (it == kIterMax) { *code = -(kIterMax + 1); return CUDA_SUCCESS; }
      HC_CHECK(hc::d2s(s, N, K, R, N, RS, N, ovf));
      CU_TRY(read_flag(&flag));
      if (flag) { *code = -2; return CUDA_SUCCESS; }
      HC_CHECK(potrs_device<float>(s, 'L', N, K, S, N, RS, N, static_cast<float *>(T)));
      HC_CHECK(hc::add_s2d(s, N, K, RS, N, Xd, (long)ldx));
    }
  };

  long code = 0;
  CUresult r = attempt(factor == 'B', &code);
  if (r != CUDA_SUCCESS) return r;
  if (code < 0 && factor == 'B') {  // BF16 products too coarse: FP32 next
    r = attempt(false, &code);
    if (r != CUDA_SUCCESS) return r;
    if (used != nullptr) *used = 'S';
  }
  *iter = code;
  if (code >= 0) return CUDA_SUCCESS;

  // Full FP64 factorisation and solve; A is overwritten by its factor.
  if (used != nullptr) *used = 'D';
  double *Aw = dptr<double>(A);
  HC_CHECK(hc::lapack_workspace(h, L.total, &base));
  HC_CHECK(hipMemsetAsync(h->dinfo, 0, sizeof(int), s));
  HC_CHECK(hc::potrf_on_device<double>(h, ul, N, Aw, (long)lda,
                                       reinterpret_cast<double *>(wsb + L.W), false));
  HC_CHECK(hipMemcpyAsync(h->hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HC_CHECK(hipStreamSynchronize(s));
  *info = h->hinfo[0];
  if (*info != 0) return CUDA_SUCCESS;
  HC_CHECK(hc::copy2d<double>(s, N, K, Bd, (long)ldb, Xd, (long)ldx));
  HC_CHECK(potrs_device<double>(s, ul, N, K, Aw, (long)lda, Xd, (long)ldx,
                                static_cast<double *>(T)));
  HC_CHECK(hipStreamSynchronize(s));
  return CUDA_SUCCESS;
}

}  // namespace

extern "C" {

CUresult cuDsposvx(CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                   size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, char factor,
                   long *iter, char *used, long *info) {
  return sposv("cuDsposvx", h, uplo, n, nrhs, A, lda, B, ldb, X, ldx, factor, iter, used, info);
}

CUresult cuDsposv(CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, long *iter,
                  long *info) {
  return sposv("cuDsposv", h, uplo, n, nrhs, A, lda, B, ldb, X, ldx, 'S', iter, nullptr, info);
}

}  // extern "C"
