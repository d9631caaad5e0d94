// Host-memory BLAS and LAPACK entry points of the reference headers
// (include/lapack.h:17-60,161-164 and include/blas.h:45-150): the routines
// without a prefix, which in the reference run on the CPU.  Here they run on
// the GPU: each call moves its operands to the device through pinned staging
// buffers, runs this library's device routine and moves the result back
// (the host-pointer machinery of multigpu_api.cpp), so a caller of the
// reference's dpotrf / dgemm / ... links against this library unchanged.
//
// Conventions kept from the reference:
//   * argument checks and their xerbla indices are those of the CPU
//     routines (dgemm lda / ldb / ldc -> 8 / 10 / 13, dsyrk lda / ldc -> 7 /
//     10 and zherk trans = T -> 2, dtrmm / dtrsm lda / ldb -> 9 / 11 and
//     dtrmm2 ldx -> 13; dpotrf, dlauum, dpotri lda -> info = -4, dtrtri lda
//     -> -5, dtrtri2 lda / ldb -> -5 / -7), with enum values checked first
//     (index = the argument's position);
//   * quick returns (m = 0, n = 0, alpha = 0 or k = 0 with beta = 1) touch
//     nothing, and an operand the routine does not read is not read here;
//   * a numerical failure is info > 0 (1-based column), with no other
//     report, as the CPU routines do.
// The routines are synchronous.  Each calling thread has its own staging
// buffers and device operands per device (created on its first call there),
// so that threads may call concurrently.  A device failure (no usable GPU,
// out of device memory) is reported through errorHandler; a LAPACK routine
// then sets *info = -1000 - the CUresult code.
#include "api_util.h"

#include <cmath>
#include <memory>
#include <vector>

#include "cucholesky.h"
#include "cumultigpu.h"
#include "lapack.h"

using hc::bad_arg;

namespace {

using zc = cuDoubleComplex_t;
using cc = cuComplex_t;

// Per-thread, per-device context: a one-device CUmultiGPU context with its
// BLAS and LAPACK host-pointer handles, a LAPACK handle and a device buffer
// for the log-determinant.
struct HostCtx {
  int device = -1;
  CUmultiGPU mg = nullptr;
  CUmultiGPUBLAShandle blas = nullptr;
  CUmultiGPULAPACKhandle lapack = nullptr;
  CULAPACKhandle lh = nullptr;
  void *buf = nullptr;
  size_t cap = 0;
  ~HostCtx() {
    hc::DeviceGuard g(device);
    if (buf != nullptr) (void)hipFree(buf);
    if (lh != nullptr) (void)cuLAPACKDestroy(lh);
    if (lapack != nullptr) (void)cuMultiGPULAPACKDestroy(lapack);
    if (blas != nullptr) (void)cuMultiGPUBLASDestroy(blas);
    if (mg != nullptr) (void)cuMultiGPUDestroy(mg);
  }
};

thread_local std::vector<std::unique_ptr<HostCtx>> tl_ctx;

// The calling thread's context on its current device, created on first use.
CUresult context(HostCtx **out) {
  int dev = 0;
  HC_CHECK(hipGetDevice(&dev));
  if (dev < 0) return CUDA_ERROR_INVALID_DEVICE;
  if ((size_t)dev >= tl_ctx.size()) tl_ctx.resize(dev + 1);
  if (!tl_ctx[dev]) {
    auto c = std::make_unique<HostCtx>();
    c->device = dev;
    CUdevice d = dev;
    CUresult r = cuMultiGPUCreate(&c->mg, &d, 1);
    if (r == CUDA_SUCCESS) r = cuMultiGPUBLASCreate(&c->blas, c->mg);
    if (r == CUDA_SUCCESS) r = cuMultiGPULAPACKCreate(&c->lapack, c->mg);
    if (r == CUDA_SUCCESS) r = cuLAPACKCreate(&c->lh);
    if (r != CUDA_SUCCESS) return r;  // c releases what was made
    tl_ctx[dev] = std::move(c);
  }
  *out = tl_ctx[dev].get();
  return CUDA_SUCCESS;
}

void device_failure(const char *name, CUresult r) {
  hc::report("device execution", name, __FILE__, __LINE__, r);
}

template <typename T> bool is_zero(T v) { return v == T(0); }
template <typename T> bool is_one(T v) { return v == T(1); }

// ---------------------------------------------------------------- BLAS

template <typename T>
void host_gemm(const char *name, CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n, size_t k, T alpha,
               const T *A, size_t lda, const T *B, size_t ldb, T beta, T *C, size_t ldc) {
  // reference blas/dgemm.c:34-55
  if (!hc::valid_trans(ta)) return (void)bad_arg(name, 1);
  if (!hc::valid_trans(tb)) return (void)bad_arg(name, 2);
  const size_t nRowA = (ta == CBlasNoTrans) ? m : k, nRowB = (tb == CBlasNoTrans) ? k : n;
  if (lda < nRowA) return (void)bad_arg(name, 8);
  if (ldb < nRowB) return (void)bad_arg(name, 10);
  if (ldc < m) return (void)bad_arg(name, 13);
  if (m == 0 || n == 0 || ((is_zero(alpha) || k == 0) && is_one(beta))) return;
  HostCtx *c = nullptr;
  CUresult r = context(&c);
  if (r == CUDA_SUCCESS) {
    if constexpr (std::is_same<T, double>::value)
      r = cuMultiGPUDgemm(c->blas, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if constexpr (std::is_same<T, float>::value)
      r = cuMultiGPUSgemm(c->blas, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if constexpr (std::is_same<T, zc>::value)
      r = cuMultiGPUZgemm(c->blas, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      r = cuMultiGPUCgemm(c->blas, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  if (r != CUDA_SUCCESS) device_failure(name, r);
}

// SYRK (real) and HERK (complex, real alpha and beta)
template <typename T, typename R>
void host_herk(const char *name, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k, R alpha, const T *A,
               size_t lda, R beta, T *C, size_t ldc) {
  constexpr bool cplx = std::is_same<T, zc>::value || std::is_same<T, cc>::value;
  // reference blas/dsyrk.c:34-49, blas/zherk.c:35-52
  if (!hc::valid_uplo(uplo)) return (void)bad_arg(name, 1);
  if (!hc::valid_trans(trans) || (cplx && trans == CBlasTrans)) return (void)bad_arg(name, 2);
  const size_t nRowA = (trans == CBlasNoTrans) ? n : k;
  if (lda < nRowA) return (void)bad_arg(name, 7);
  if (ldc < n) return (void)bad_arg(name, 10);
  if (n == 0 || ((alpha == R(0) || k == 0) && beta == R(1))) return;
  HostCtx *c = nullptr;
  CUresult r = context(&c);
  if (r == CUDA_SUCCESS) {
    if constexpr (std::is_same<T, double>::value)
      r = cuMultiGPUDsyrk(c->blas, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
    else if constexpr (std::is_same<T, float>::value)
      r = cuMultiGPUSsyrk(c->blas, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
    else if constexpr (std::is_same<T, zc>::value)
      r = cuMultiGPUZherk(c->blas, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
    else
      r = cuMultiGPUCherk(c->blas, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
  }
  if (r != CUDA_SUCCESS) device_failure(name, r);
}

template <typename T>
bool tri_args(const char *name, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m,
              size_t n, size_t lda, size_t ldb) {
  // reference blas/dtrmm.c:44-58, blas/dtrsm.c:34-48
  if (!hc::valid_side(side)) return bad_arg(name, 1), false;
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2), false;
  if (!hc::valid_trans(trans)) return bad_arg(name, 3), false;
  if (!hc::valid_diag(diag)) return bad_arg(name, 4), false;
  const size_t nRowA = (side == CBlasLeft) ? m : n;
  if (lda < nRowA) return bad_arg(name, 9), false;
  if (ldb < m) return bad_arg(name, 11), false;
  return true;
}

// In-place TRMM (TRSM = false) or TRSM (TRSM = true) on B.
template <typename T, bool TRSM>
void host_tri(const char *name, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m,
              size_t n, T alpha, const T *A, size_t lda, T *B, size_t ldb) {
  if (!tri_args<T>(name, side, uplo, trans, diag, m, n, lda, ldb)) return;
  if (m == 0 || n == 0) return;
  HostCtx *c = nullptr;
  CUresult r = context(&c);
  if (r == CUDA_SUCCESS) {
    if constexpr (TRSM) {
      if constexpr (std::is_same<T, double>::value)
        r = cuMultiGPUDtrsm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
      else if constexpr (std::is_same<T, float>::value)
        r = cuMultiGPUStrsm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
      else if constexpr (std::is_same<T, zc>::value)
        r = cuMultiGPUZtrsm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
      else
        r = cuMultiGPUCtrsm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
    } else {
      if constexpr (std::is_same<T, double>::value)
        r = cuMultiGPUDtrmm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
      else if constexpr (std::is_same<T, float>::value)
        r = cuMultiGPUStrmm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
      else if constexpr (std::is_same<T, zc>::value)
        r = cuMultiGPUZtrmm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
      else
        r = cuMultiGPUCtrmm(c->blas, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
    }
  }
  if (r != CUDA_SUCCESS) device_failure(name, r);
}

// Out-of-place TRMM: X <- alpha op(A) B or alpha B op(A), B unchanged
// (reference blas/dtrmm.c:206-372): B is copied into X, then multiplied in place.
template <typename T>
void host_trmm2(const char *name, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m,
                size_t n, T alpha, const T *A, size_t lda, const T *B, size_t ldb, T *X, size_t ldx) {
  if (!tri_args<T>(name, side, uplo, trans, diag, m, n, lda, ldb)) return;
  if (ldx < m) return (void)bad_arg(name, 13);
  if (m == 0 || n == 0) return;
  if (!is_zero(alpha))  // alpha = 0: X <- 0 without reading B
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < m; i++) X[j * ldx + i] = B[j * ldb + i];
  host_tri<T, false>(name, side, uplo, trans, diag, m, n, alpha, A, lda, X, ldx);
}

// ---------------------------------------------------------------- LAPACK

enum class Lop { potrf, trtri, lauum, potri };

template <typename T>
void host_lapack(const char *name, Lop op, CBlasUplo uplo, CBlasDiag diag, size_t n, T *A, size_t lda,
                 long *info) {
  if (info == nullptr) return (void)bad_arg(name, op == Lop::trtri ? 6 : 5);
  *info = 0;
  // reference lapack/dpotrf.c:99-107, lapack/dtrtri.c:98-107, dlauum.c:75-80, dpotri.c:8-13
  const long ia = op == Lop::trtri ? 5 : 4;
  if (!hc::valid_uplo(uplo)) *info = -1;
  else if (op == Lop::trtri && !hc::valid_diag(diag)) *info = -2;
  else if (lda < n) *info = -ia;
  if (*info != 0) return (void)bad_arg(name, -*info);
  if (n == 0) return;
  HostCtx *c = nullptr;
  CUresult r = context(&c);
  if (r == CUDA_SUCCESS) {
    CUmultiGPULAPACKhandle L = c->lapack;
    constexpr int P = std::is_same<T, double>::value ? 0 : std::is_same<T, float>::value ? 1
                      : std::is_same<T, zc>::value ? 2 : 3;
    auto call = [&]() -> CUresult {
      if constexpr (P == 0) {
        switch (op) {
          case Lop::potrf: return cuMultiGPUDpotrf(L, uplo, n, A, lda, info);
          case Lop::trtri: return cuMultiGPUDtrtri(L, uplo, diag, n, A, lda, info);
          case Lop::lauum: return cuMultiGPUDlauum(L, uplo, n, A, lda, info);
          default: return cuMultiGPUDpotri(L, uplo, n, A, lda, info);
        }
      } else if constexpr (P == 1) {
        switch (op) {
          case Lop::potrf: return cuMultiGPUSpotrf(L, uplo, n, A, lda, info);
          case Lop::trtri: return cuMultiGPUStrtri(L, uplo, diag, n, A, lda, info);
          case Lop::lauum: return cuMultiGPUSlauum(L, uplo, n, A, lda, info);
          default: return cuMultiGPUSpotri(L, uplo, n, A, lda, info);
        }
      } else if constexpr (P == 2) {
        switch (op) {
          case Lop::potrf: return cuMultiGPUZpotrf(L, uplo, n, A, lda, info);
          case Lop::trtri: return cuMultiGPUZtrtri(L, uplo, diag, n, A, lda, info);
          case Lop::lauum: return cuMultiGPUZlauum(L, uplo, n, A, lda, info);
          default: return cuMultiGPUZpotri(L, uplo, n, A, lda, info);
        }
      } else {
        switch (op) {
          case Lop::potrf: return cuMultiGPUCpotrf(L, uplo, n, A, lda, info);
          case Lop::trtri: return cuMultiGPUCtrtri(L, uplo, diag, n, A, lda, info);
          case Lop::lauum: return cuMultiGPUClauum(L, uplo, n, A, lda, info);
          default: return cuMultiGPUCpotri(L, uplo, n, A, lda, info);
        }
      }
    };
    r = call();
    if (r == CUDA_ERROR_INVALID_VALUE && *info > 0) r = CUDA_SUCCESS;  // numerical failure: info says it
  }
  if (r != CUDA_SUCCESS) {
    device_failure(name, r);
    *info = -1000 - (long)r;
  }
}

// Out-of-place triangular inverse B <- inv(A) (reference
// lapack/dtrtri.c:226-299): the triangle uplo of A (with its diagonal) is
// copied into B, which is then inverted in place; B's other triangle is not
// touched.
template <typename T>
void host_trtri2(const char *name, CBlasUplo uplo, CBlasDiag diag, size_t n, const T *A, size_t lda, T *B,
                 size_t ldb, long *info) {
  if (info == nullptr) return (void)bad_arg(name, 8);
  *info = 0;
  if (!hc::valid_uplo(uplo)) *info = -1;
  else if (!hc::valid_diag(diag)) *info = -2;
  else {
    if (lda < n) *info = -5;
    if (ldb < n) *info = -7;
  }
  if (*info != 0) return (void)bad_arg(name, -*info);
  if (n == 0) return;
  for (size_t j = 0; j < n; j++) {
    const size_t i0 = uplo == CBlasLower ? j : 0, i1 = uplo == CBlasLower ? n : j + 1;
    for (size_t i = i0; i < i1; i++) B[j * ldb + i] = A[j * lda + i];
  }
  host_lapack<T>(name, Lop::trtri, uplo, diag, n, B, ldb, info);
}

// 2 sum log(Re x[i incx]): the values are gathered on the host and reduced
// on the device by this library's log-determinant kernel.
template <typename T, typename R>
R host_logdet(const char *name, const T *x, size_t incx, size_t n) {
  if (n == 0) return R(0);
  HostCtx *c = nullptr;
  CUresult r = context(&c);
  double out = std::nan("");
  if (r == CUDA_SUCCESS) {
    std::vector<T> packed(n);
    for (size_t i = 0; i < n; i++) packed[i] = x[i * incx];
    const size_t bytes = n * sizeof(T);
    hc::DeviceGuard g(c->device);
    hipError_t e = hipSuccess;
    if (bytes > c->cap) {
      if (c->buf != nullptr) e = hipFree(c->buf);
      c->buf = nullptr;
      c->cap = 0;
      if (e == hipSuccess) e = hipMalloc(&c->buf, bytes);
      if (e == hipSuccess) c->cap = bytes;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c->buf, packed.data(), bytes, hipMemcpyHostToDevice, c->lh->stream);
    r = hc::to_cu(e);
    if (r == CUDA_SUCCESS) {
      const CUdeviceptr d = reinterpret_cast<CUdeviceptr>(c->buf);
      if constexpr (std::is_same<T, double>::value) r = cuDlogdet(c->lh, d, 1, n, &out, c->lh->stream);
      else if constexpr (std::is_same<T, zc>::value) r = cuZlogdet(c->lh, d, 1, n, &out, c->lh->stream);
      else {
        float f = 0.0f;
        if constexpr (std::is_same<T, float>::value) r = cuSlogdet(c->lh, d, 1, n, &f, c->lh->stream);
        else r = cuClogdet(c->lh, d, 1, n, &f, c->lh->stream);
        out = f;
      }
    }
  }
  if (r != CUDA_SUCCESS) device_failure(name, r);
  return (R)out;
}

}  // namespace

extern "C" {

// ---- LAPACK (reference include/lapack.h:17-60)
void spotrf(CBlasUplo uplo, size_t n, float *A, size_t lda, long *info) { host_lapack<float>("spotrf", Lop::potrf, uplo, CBlasNonUnit, n, A, lda, info); }
void dpotrf(CBlasUplo uplo, size_t n, double *A, size_t lda, long *info) { host_lapack<double>("dpotrf", Lop::potrf, uplo, CBlasNonUnit, n, A, lda, info); }
void cpotrf(CBlasUplo uplo, size_t n, cc *A, size_t lda, long *info) { host_lapack<cc>("cpotrf", Lop::potrf, uplo, CBlasNonUnit, n, A, lda, info); }
void zpotrf(CBlasUplo uplo, size_t n, zc *A, size_t lda, long *info) { host_lapack<zc>("zpotrf", Lop::potrf, uplo, CBlasNonUnit, n, A, lda, info); }

void strtri(CBlasUplo uplo, CBlasDiag diag, size_t n, float *A, size_t lda, long *info) { host_lapack<float>("strtri", Lop::trtri, uplo, diag, n, A, lda, info); }
void dtrtri(CBlasUplo uplo, CBlasDiag diag, size_t n, double *A, size_t lda, long *info) { host_lapack<double>("dtrtri", Lop::trtri, uplo, diag, n, A, lda, info); }
void ctrtri(CBlasUplo uplo, CBlasDiag diag, size_t n, cc *A, size_t lda, long *info) { host_lapack<cc>("ctrtri", Lop::trtri, uplo, diag, n, A, lda, info); }
void ztrtri(CBlasUplo uplo, CBlasDiag diag, size_t n, zc *A, size_t lda, long *info) { host_lapack<zc>("ztrtri", Lop::trtri, uplo, diag, n, A, lda, info); }

void strtri2(CBlasUplo uplo, CBlasDiag diag, size_t n, const float *A, size_t lda, float *B, size_t ldb, long *info) { host_trtri2<float>("strtri2", uplo, diag, n, A, lda, B, ldb, info); }
void dtrtri2(CBlasUplo uplo, CBlasDiag diag, size_t n, const double *A, size_t lda, double *B, size_t ldb, long *info) { host_trtri2<double>("dtrtri2", uplo, diag, n, A, lda, B, ldb, info); }
void ctrtri2(CBlasUplo uplo, CBlasDiag diag, size_t n, const cc *A, size_t lda, cc *B, size_t ldb, long *info) { host_trtri2<cc>("ctrtri2", uplo, diag, n, A, lda, B, ldb, info); }
void ztrtri2(CBlasUplo uplo, CBlasDiag diag, size_t n, const zc *A, size_t lda, zc *B, size_t ldb, long *info) { host_trtri2<zc>("ztrtri2", uplo, diag, n, A, lda, B, ldb, info); }

void slauum(CBlasUplo uplo, size_t n, float *A, size_t lda, long *info) { host_lapack<float>("slauum", Lop::lauum, uplo, CBlasNonUnit, n, A, lda, info); }
void dlauum(CBlasUplo uplo, size_t n, double *A, size_t lda, long *info) { host_lapack<double>("dlauum", Lop::lauum, uplo, CBlasNonUnit, n, A, lda, info); }
void clauum(CBlasUplo uplo, size_t n, cc *A, size_t lda, long *info) { host_lapack<cc>("clauum", Lop::lauum, uplo, CBlasNonUnit, n, A, lda, info); }
void zlauum(CBlasUplo uplo, size_t n, zc *A, size_t lda, long *info) { host_lapack<zc>("zlauum", Lop::lauum, uplo, CBlasNonUnit, n, A, lda, info); }

void spotri(CBlasUplo uplo, size_t n, float *A, size_t lda, long *info) { host_lapack<float>("spotri", Lop::potri, uplo, CBlasNonUnit, n, A, lda, info); }
void dpotri(CBlasUplo uplo, size_t n, double *A, size_t lda, long *info) { host_lapack<double>("dpotri", Lop::potri, uplo, CBlasNonUnit, n, A, lda, info); }
void cpotri(CBlasUplo uplo, size_t n, cc *A, size_t lda, long *info) { host_lapack<cc>("cpotri", Lop::potri, uplo, CBlasNonUnit, n, A, lda, info); }
void zpotri(CBlasUplo uplo, size_t n, zc *A, size_t lda, long *info) { host_lapack<zc>("zpotri", Lop::potri, uplo, CBlasNonUnit, n, A, lda, info); }

// ---- log-determinants (reference include/lapack.h:161-164)
float slogdet(const float *x, size_t incx, size_t n) { return host_logdet<float, float>("slogdet", x, incx, n); }
double dlogdet(const double *x, size_t incx, size_t n) { return host_logdet<double, double>("dlogdet", x, incx, n); }
float clogdet(const cc *x, size_t incx, size_t n) { return host_logdet<cc, float>("clogdet", x, incx, n); }
double zlogdet(const zc *x, size_t incx, size_t n) { return host_logdet<zc, double>("zlogdet", x, incx, n); }

// ---- BLAS (reference include/blas.h:45-150)
void ssyrk(CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k, float alpha, const float *A, size_t lda, float beta, float *C, size_t ldc) { host_herk<float, float>("ssyrk", uplo, trans, n, k, alpha, A, lda, beta, C, ldc); }
void dsyrk(CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k, double alpha, const double *A, size_t lda, double beta, double *C, size_t ldc) { host_herk<double, double>("dsyrk", uplo, trans, n, k, alpha, A, lda, beta, C, ldc); }
void cherk(CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k, float alpha, const cc *A, size_t lda, float beta, cc *C, size_t ldc) { host_herk<cc, float>("cherk", uplo, trans, n, k, alpha, A, lda, beta, C, ldc); }
void zherk(CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k, double alpha, const zc *A, size_t lda, double beta, zc *C, size_t ldc) { host_herk<zc, double>("zherk", uplo, trans, n, k, alpha, A, lda, beta, C, ldc); }

void sgemm(CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n, size_t k, float alpha, const float *A, size_t lda, const float *B, size_t ldb, float beta, float *C, size_t ldc) { host_gemm<float>("sgemm", ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc); }
void dgemm(CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n, size_t k, double alpha, const double *A, size_t lda, const double *B, size_t ldb, double beta, double *C, size_t ldc) { host_gemm<double>("dgemm", ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc); }
void cgemm(CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n, size_t k, cc alpha, const cc *A, size_t lda, const cc *B, size_t ldb, cc beta, cc *C, size_t ldc) { host_gemm<cc>("cgemm", ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc); }
void zgemm(CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n, size_t k, zc alpha, const zc *A, size_t lda, const zc *B, size_t ldb, zc beta, zc *C, size_t ldc) { host_gemm<zc>("zgemm", ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc); }

void strmm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, float alpha, const float *A, size_t lda, float *B, size_t ldb) { host_tri<float, false>("strmm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }
void dtrmm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, double alpha, const double *A, size_t lda, double *B, size_t ldb) { host_tri<double, false>("dtrmm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }
void ctrmm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, cc alpha, const cc *A, size_t lda, cc *B, size_t ldb) { host_tri<cc, false>("ctrmm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }
void ztrmm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, zc alpha, const zc *A, size_t lda, zc *B, size_t ldb) { host_tri<zc, false>("ztrmm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }

void strmm2(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, float alpha, const float *A, size_t lda, const float *B, size_t ldb, float *X, size_t ldx) { host_trmm2<float>("strmm2", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx); }
void dtrmm2(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, double alpha, const double *A, size_t lda, const double *B, size_t ldb, double *X, size_t ldx) { host_trmm2<double>("dtrmm2", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx); }
void ctrmm2(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, cc alpha, const cc *A, size_t lda, const cc *B, size_t ldb, cc *X, size_t ldx) { host_trmm2<cc>("ctrmm2", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx); }
void ztrmm2(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, zc alpha, const zc *A, size_t lda, const zc *B, size_t ldb, zc *X, size_t ldx) { host_trmm2<zc>("ztrmm2", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx); }

void strsm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, float alpha, const float *A, size_t lda, float *B, size_t ldb) { host_tri<float, true>("strsm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }
void dtrsm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, double alpha, const double *A, size_t lda, double *B, size_t ldb) { host_tri<double, true>("dtrsm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }
void ctrsm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, cc alpha, const cc *A, size_t lda, cc *B, size_t ldb) { host_tri<cc, true>("ctrsm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }
void ztrsm(CBlasSide side, CBlasUplo uplo, CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, zc alpha, const zc *A, size_t lda, zc *B, size_t ldb) { host_tri<zc, true>("ztrsm", side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb); }

}  // extern "C"
