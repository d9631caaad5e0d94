// Public BLAS-3 entry points (include/blas.h): argument checks with the
// reference's XERBLA indices and quick returns, then the device building
// blocks of blas_ops.cpp on the caller's stream.
#include "api_util.h"

using hc::bad_arg;
using hc::dptr;
using hc::zd;

namespace {

// Argument checks (the reference's xerbla indices); CUDA_SUCCESS when valid.
CUresult gemm_check(const char *name, CUBLAShandle handle, CBlasTranspose transA,
                    CBlasTranspose transB, size_t m, size_t n, size_t k, size_t lda, size_t ldb,
                    size_t ldc, size_t ldd) {
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
  return CUDA_SUCCESS;
}

CUresult herk_check(const char *name, bool complex_type, CUBLAShandle handle, CBlasUplo uplo,
                    CBlasTranspose trans, size_t n, size_t k, size_t lda, size_t ldc) {
  // reference blas/dsyrk.c:167-179, blas/zherk.c:194-206
  if (handle == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (!hc::valid_trans(trans) || (complex_type && trans == CBlasTrans)) return bad_arg(name, 3);
  const size_t nRowA = (trans == CBlasNoTrans) ? n : k;
  if (lda < nRowA) return bad_arg(name, 7);
  if (ldc < n) return bad_arg(name, 10);
  return CUDA_SUCCESS;
}

// TRMM2 (ldx checked when X_index > 0) and TRSM: reference blas/dtrmm.c:379-392,
// blas/dtrsm.c:222-234
CUresult tri_check(const char *name, CUBLAShandle handle, CBlasSide side, CBlasUplo uplo,
                   CBlasTranspose trans, CBlasDiag diag, size_t m, size_t n, size_t lda, size_t ldb,
                   size_t ldx, bool has_x) {
  if (handle == nullptr) return bad_arg(name, 1);
  if (!hc::valid_side(side)) return bad_arg(name, 2);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 3);
  if (!hc::valid_trans(trans)) return bad_arg(name, 4);
  if (!hc::valid_diag(diag)) return bad_arg(name, 5);
  const size_t nRowA = (side == CBlasLeft) ? m : n;
  if (lda < nRowA) return bad_arg(name, 9);
  if (ldb < m) return bad_arg(name, 11);
  if (has_x && ldx < m) return bad_arg(name, 13);
  return CUDA_SUCCESS;
}

template <typename T>
CUresult gemm_api(const char *name, CUBLAShandle handle, CBlasTranspose transA,
                  CBlasTranspose transB, size_t m, size_t n, size_t k, T alpha, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, T beta, CUdeviceptr C, size_t ldc,
                  CUdeviceptr D, size_t ldd, CUstream stream) {
  if (CUresult r = gemm_check(name, handle, transA, transB, m, n, k, lda, ldb, ldc, ldd); r != CUDA_SUCCESS)
    return r;
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
  if (CUresult r = herk_check(name, hc::is_complex<T>::value, handle, uplo, trans, n, k, lda, ldc);
      r != CUDA_SUCCESS)
    return r;
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
  if (CUresult r = tri_check(name, handle, side, uplo, trans, diag, m, n, lda, ldb, ldx, true);
      r != CUDA_SUCCESS)
    return r;
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
  if (CUresult r = tri_check(name, handle, side, uplo, trans, diag, m, n, lda, ldb, 0, false);
      r != CUDA_SUCCESS)
    return r;
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

hc::zf to_zf(cuComplex_t v) { return hc::make_zf(__real__ v, __imag__ v); }

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

// Single precision triangular multiply and the single complex BLAS on the
// FP32 matrix cores (reference blas/strmm.c:374, blas/cgemm.c:226,
// blas/cherk.c:190, blas/ctrmm.c:422, blas/ctrsm.c:248).
CUresult cuStrmm2(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                  CBlasDiag diag, size_t m, size_t n, float alpha, CUdeviceptr A, size_t lda,
                  CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, CUstream stream) {
  return trmm_api<float>("cuStrmm2", handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx,
                         stream);
}

CUresult cuCgemm2(CUBLAShandle handle, CBlasTranspose transA, CBlasTranspose transB, size_t m,
                  size_t n, size_t k, cuComplex_t alpha, CUdeviceptr A, size_t lda, CUdeviceptr B,
                  size_t ldb, cuComplex_t beta, CUdeviceptr C, size_t ldc, CUdeviceptr D, size_t ldd,
                  CUstream stream) {
  return gemm_api<hc::zf>("cuCgemm2", handle, transA, transB, m, n, k, to_zf(alpha), A, lda, B, ldb,
                          to_zf(beta), C, ldc, D, ldd, stream);
}

CUresult cuCherk(CUBLAShandle handle, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                 float alpha, CUdeviceptr A, size_t lda, float beta, CUdeviceptr C, size_t ldc,
                 CUstream stream) {
  return herk_api<hc::zf>("cuCherk", handle, uplo, trans, n, k, (double)alpha, A, lda, (double)beta, C, ldc,
                          stream);
}

CUresult cuCtrmm2(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                  CBlasDiag diag, size_t m, size_t n, cuComplex_t alpha, CUdeviceptr A, size_t lda,
                  CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, CUstream stream) {
  return trmm_api<hc::zf>("cuCtrmm2", handle, side, uplo, trans, diag, m, n, to_zf(alpha), A, lda, B, ldb, X,
                          ldx, stream);
}

CUresult cuCtrsm(CUBLAShandle handle, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                 CBlasDiag diag, size_t m, size_t n, cuComplex_t alpha, CUdeviceptr A, size_t lda,
                 CUdeviceptr B, size_t ldb, CUstream stream) {
  return trsm_api<hc::zf>("cuCtrsm", handle, side, uplo, trans, diag, m, n, to_zf(alpha), A, lda, B, ldb,
                          stream);
}

}  // extern "C"
