// Public BLAS-3 entry points (include/blas.h): argument checks with the
// reference's XERBLA indices and quick returns, then the device building
// blocks of blas_ops.cpp on the caller's stream.
#include "api_util.h"

using hc::bad_arg;
using hc::dptr;
using hc::zd;

namespace {

template <typename T>
CUresult gemm_api(const char *name, CUBLAShandle handle, CBlasTranspose transA,
                  CBlasTranspose transB, size_t m, size_t n, size_t k, T alpha, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, T beta, CUdeviceptr C, size_t ldc,
                  CUdeviceptr D, size_t ldd, CUstream stream) {
  // reference blas/dgemm.c:153-171
  if (handle == nullptr) return bad_arg(name, 1);
  if (!hc::valid_trans(transA)) return bad_arg(name, 2);
  if (!hc::valid_trans(transB)) return bad_arg(name, 3);
  const size_t nRowA = (transA == CBlasNoTrans) ? m : k;
  const size_t nRowB = (transB == CBlasNoTrans) ? k : n;
  if (lda < nRowA) return bad_arg(name, 8);
  if (ldb < nRowB) return bad_arg(name, 10);
  if (ldc < m) return bad_arg(name, 13);
  if (ldd < m) return bad_arg(name, 15);
  if (m == 0 || n == 0 || (C == D && (hc::is_zero(alpha) || k == 0) && hc::is_one(beta)))
    return CUDA_SUCCESS;
  hc::DeviceGuard g(handle->device);
  HC_CHECK(g.status());
  const int ta = hc::is_complex<T>::value ? hc::op_code(transA) : (transA == 'N' ? hc::OP_N : hc::OP_T);
  const int tb = hc::is_complex<T>::value ? hc::op_code(transB) : (transB == 'N' ? hc::OP_N : hc::OP_T);
  HC_CHECK((hc::gemm<T>(stream, ta, tb, (long)m, (long)n, (long)k, alpha, dptr<const T>(A), (long)lda,
                        dptr<const T>(B), (long)ldb, beta, dptr<const T>(C), (long)ldc, dptr<T>(D),
                        (long)ldd)));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult herk_api(const char *name, CUBLAShandle handle, CBlasUplo uplo, CBlasTranspose trans,
                  size_t n, size_t k, double alpha, CUdeviceptr A, size_t lda, double beta,
                  CUdeviceptr C, size_t ldc, CUstream stream) {
  // reference blas/dsyrk.c:167-179, blas/zherk.c:194-206
  if (handle == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (!hc::valid_trans(trans) || (hc::is_complex<T>::value && trans == CBlasTrans))
    return bad_arg(name, 3);
  const size_t nRowA = (trans == CBlasNoTrans) ? n : k;
  if (lda < nRowA) return bad_arg(name, 7);
  if (ldc < n) return bad_arg(name, 10);
  if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return CUDA_SUCCESS;
  hc::DeviceGuard g(handle->device);
  HC_CHECK(g.status());
  HC_CHECK((hc::herk<T>(stream, (char)uplo, (char)trans, (long)n, (long)k, alpha, dptr<const T>(A),
                        (long)lda, beta, dptr<T>(C), (long)ldc)));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult trmm_api(const char *name, CUBLAShandle handle, CBlasSide side, CBlasUplo uplo,
                  CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, T alpha, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx,
                  CUstream stream) {
  // reference blas/dtrmm.c:379-392
  if (handle == nullptr) return bad_arg(name, 1);
  if (!hc::valid_side(side)) return bad_arg(name, 2);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 3);
  if (!hc::valid_trans(trans)) return bad_arg(name, 4);
  if (!hc::valid_diag(diag)) return bad_arg(name, 5);
  const size_t nRowA = (side == CBlasLeft) ? m : n;
  if (lda < nRowA) return bad_arg(name, 9);
  if (ldb < m) return bad_arg(name, 11);
  if (ldx < m) return bad_arg(name, 13);
  if (m == 0 || n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(handle->device);
  HC_CHECK(g.status());
  HC_CHECK((hc::trmm<T>(stream, (char)side, (char)uplo, (char)trans, (char)diag, (long)m, (long)n,
                        alpha, dptr<const T>(A), (long)lda, dptr<const T>(B), (long)ldb, dptr<T>(X),
                        (long)ldx)));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult trsm_api(const char *name, CUBLAShandle handle, CBlasSide side, CBlasUplo uplo,
                  CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, T alpha, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, CUstream stream) {
  // reference blas/dtrsm.c:222-234
  if (handle == nullptr) return bad_arg(name, 1);
  if (!hc::valid_side(side)) return bad_arg(name, 2);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 3);
  if (!hc::valid_trans(trans)) return bad_arg(name, 4);
  if (!hc::valid_diag(diag)) return bad_arg(name, 5);
  const size_t nRowA = (side == CBlasLeft) ? m : n;
  if (lda < nRowA) return bad_arg(name, 9);
  if (ldb < m) return bad_arg(name, 11);
  if (m == 0 || n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(handle->device);
  HC_CHECK(g.status());
  T *Bp = dptr<T>(B);
  if (hc::is_zero(alpha)) {  // B <- 0 without reading A (reference blas/dtrsm.c:53-59)
    const T zero = hc::from_real<T>(0.0);
    HC_CHECK((hc::gemm<T>(stream, hc::OP_N, hc::OP_N, (long)m, (long)n, 0, zero, nullptr, 1, nullptr,
                          1, zero, nullptr, 1, Bp, (long)ldb)));
    return CUDA_SUCCESS;
  }
  // Stream-ordered scratch for the inverted diagonal blocks, so concurrent
  // calls on different streams never share it.
  const size_t bytes = (size_t)hc::trsm_ws_elems((char)side, (long)m, (long)n) * sizeof(T);
  void *ws = nullptr;
  HC_CHECK(hipMallocAsync(&ws, bytes, stream));
  const hipError_t e = hc::trsm<T>(stream, (char)side, (char)uplo, (char)trans, (char)diag, (long)m,
                                   (long)n, alpha, dptr<const T>(A), (long)lda, Bp, (long)ldb,
                                   static_cast<T *>(ws));
  HC_CHECK(hipFreeAsync(ws, stream));
  HC_CHECK(e);
  return CUDA_SUCCESS;
}

}  // namespace

extern "C" {

CUresult cuDgemm2(CUBLAShandle handle, CBlasTranspose transA, CBlasTranspose transB, size_t m,
                  size_t n, size_t k, double alpha, CUdeviceptr A, size_t lda, CUdeviceptr B,
                  size_t ldb, double beta, CUdeviceptr C, size_t ldc, CUdeviceptr D, size_t ldd,
                  CUstream stream) {
  return gemm_api<double>("cuDgemm2", handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta,
                          C, ldc, D, ldd, stream);
}

CUresult cuZgemm2(CUBLAShandle handle, CBlasTranspose transA, CBlasTranspose transB, size_t m,
                  size_t n, size_t k, cuDoubleComplex_t alpha, CUdeviceptr A, size_t lda,
                  CUdeviceptr B, size_t ldb, cuDoubleComplex_t beta, CUdeviceptr C, size_t ldc,
                  CUdeviceptr D, size_t ldd, CUstream stream) {
  return gemm_api<zd>("cuZgemm2", handle, transA, transB, m, n, k, hc::to_zd(alpha), A, lda, B,
                      ldb, hc::to_zd(beta), C, ldc, D, ldd, stream);
}

CUresult cuDsyrk(CUBLAShandle handle, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                 double alpha, CUdeviceptr A, size_t lda, double beta, CUdeviceptr C, size_t ldc,
                 CUstream stream) {
  return herk_api<double>("cuDsyrk", handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream);
}

CUresult cuZherk(CUBLAShandle handle, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                 double alpha, CUdeviceptr A, size_t lda, double beta, CUdeviceptr C, size_t ldc,
                 CUstream stream) {
  return herk_api<zd>("cuZherk", handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream);
}

CUresult cuDtrmm2(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                  CBlasDiag diag, size_t m, size_t n, double alpha, CUdeviceptr A, size_t lda,
                  CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, CUstream stream) {
  return trmm_api<double>("cuDtrmm2", handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb,
                          X, ldx, stream);
}

CUresult cuZtrmm2(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                  CBlasDiag diag, size_t m, size_t n, cuDoubleComplex_t alpha, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx,
                  CUstream stream) {
  return trmm_api<zd>("cuZtrmm2", handle, side, uplo, trans, diag, m, n, hc::to_zd(alpha), A, lda,
                      B, ldb, X, ldx, stream);
}

CUresult cuDtrsm(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                 CBlasDiag diag, size_t m, size_t n, double alpha, CUdeviceptr A, size_t lda,
                 CUdeviceptr B, size_t ldb, CUstream stream) {
  return trsm_api<double>("cuDtrsm", handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb,
                          stream);
}

// Single precision (reference include/blas.h:158,193,235): FP32 data and
// accumulation on the FP32 matrix-core path of the mixed-precision solver.
CUresult cuSgemm2(CUBLAShandle handle, CBlasTranspose transA, CBlasTranspose transB, size_t m,
                  size_t n, size_t k, float alpha, CUdeviceptr A, size_t lda, CUdeviceptr B,
                  size_t ldb, float beta, CUdeviceptr C, size_t ldc, CUdeviceptr D, size_t ldd,
                  CUstream stream) {
  return gemm_api<float>("cuSgemm2", handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta,
                         C, ldc, D, ldd, stream);
}

CUresult cuSsyrk(CUBLAShandle handle, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                 float alpha, CUdeviceptr A, size_t lda, float beta, CUdeviceptr C, size_t ldc,
                 CUstream stream) {
  return herk_api<float>("cuSsyrk", handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream);
}

CUresult cuStrsm(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                 CBlasDiag diag, size_t m, size_t n, float alpha, CUdeviceptr A, size_t lda,
                 CUdeviceptr B, size_t ldb, CUstream stream) {
  return trsm_api<float>("cuStrsm", handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb,
                         stream);
}

CUresult cuZtrsm(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                 CBlasDiag diag, size_t m, size_t n, cuDoubleComplex_t alpha, CUdeviceptr A,
                 size_t lda, CUdeviceptr B, size_t ldb, CUstream stream) {
  return trsm_api<zd>("cuZtrsm", handle, side, uplo, trans, diag, m, n, hc::to_zd(alpha), A, lda, B,
                      ldb, stream);
}

}  // extern "C"
