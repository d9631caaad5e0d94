/*
 * GPU BLAS-3 entry points of the MI355X Cholesky library.
 *
 * Drop-in for the device-pointer ("cu") half of the reference's
 * include/blas.h, double and double-complex precision: same names, argument
 * order, enum values and error behaviour.  Each function replaces the
 * reference function cited beside it.  Matrices are column-major; every
 * CUdeviceptr is a device byte address.  All calls are asynchronous on the
 * given stream.
 *
 * The host-memory routines of the reference header (dgemm, dsyrk, ...,
 * which the reference runs on the CPU) are exported too and run on the GPU:
 * operands are staged through pinned host buffers, the device routine runs,
 * the result comes back (synchronous).
 */
#ifndef BLAS_H
#define BLAS_H

#include <stddef.h>
#include <stdbool.h>
#include "cu_abi.h"
#include "cumultigpu.h"

#ifdef __cplusplus
typedef _Complex double cuDoubleComplex_t;
typedef _Complex float cuComplex_t;
extern "C" {
#else
#include <complex.h>
typedef double complex cuDoubleComplex_t;
typedef float complex cuComplex_t;
#endif

/** BLAS enums, castable to char (reference include/blas.h:24-27). */
typedef enum { CBlasNoTrans = 'N', CBlasTrans = 'T', CBlasConjTrans = 'C' } CBlasTranspose;
typedef enum { CBlasLower = 'L', CBlasUpper = 'U' } CBlasUplo;
typedef enum { CBlasLeft = 'L', CBlasRight = 'R' } CBlasSide;
typedef enum { CBlasNonUnit = 'N', CBlasUnit = 'U' } CBlasDiag;

/**
 * Illegal-argument hook (reference include/blas.h:37-43, blas/xerbla.c:5-9).
 * Called with the function name and the 1-based index of the bad argument.
 * The default prints one line to stderr; NULL silences it.
 */
typedef void (*xerbla_t)(const char *, long);
extern xerbla_t xerbla;

/*
 * Host-memory BLAS-3 (replace the reference's CPU implementations,
 * include/blas.h:45-150): column-major host arrays, synchronous, run on the
 * current GPU.  Argument errors go to xerbla with the reference's indices
 * (GEMM lda / ldb / ldc: 8 / 10 / 13; SYRK / HERK lda / ldc: 7 / 10, HERK
 * trans = T: 2; TRMM / TRSM lda / ldb: 9 / 11, TRMM2 ldx: 13); enum values
 * are checked first (index = position).  A device failure is reported
 * through errorHandler.
 */
/* C <- alpha op(A) op(A)^T + beta C (replaces blas/ssyrk.c, dsyrk.c:34-161). */
void ssyrk(CBlasUplo, CBlasTranspose, size_t, size_t, float, const float *, size_t, float, float *, size_t);
void dsyrk(CBlasUplo, CBlasTranspose, size_t, size_t, double, const double *, size_t, double, double *, size_t);
/* C <- alpha op(A) op(A)^H + beta C, real alpha and beta (replaces blas/cherk.c, zherk.c:35). */
void cherk(CBlasUplo, CBlasTranspose, size_t, size_t, float, const cuComplex_t *, size_t, float, cuComplex_t *, size_t);
void zherk(CBlasUplo, CBlasTranspose, size_t, size_t, double, const cuDoubleComplex_t *, size_t, double,
           cuDoubleComplex_t *, size_t);
/* C <- alpha op(A) op(B) + beta C (replaces blas/sgemm.c, dgemm.c:34-146, cgemm.c, zgemm.c:34). */
void sgemm(CBlasTranspose, CBlasTranspose, size_t, size_t, size_t, float, const float *, size_t, const float *,
           size_t, float, float *, size_t);
void dgemm(CBlasTranspose, CBlasTranspose, size_t, size_t, size_t, double, const double *, size_t,
           const double *, size_t, double, double *, size_t);
void cgemm(CBlasTranspose, CBlasTranspose, size_t, size_t, size_t, cuComplex_t, const cuComplex_t *, size_t,
           const cuComplex_t *, size_t, cuComplex_t, cuComplex_t *, size_t);
void zgemm(CBlasTranspose, CBlasTranspose, size_t, size_t, size_t, cuDoubleComplex_t, const cuDoubleComplex_t *,
           size_t, const cuDoubleComplex_t *, size_t, cuDoubleComplex_t, cuDoubleComplex_t *, size_t);
/* B <- alpha op(A) B or alpha B op(A), in place (replaces blas/strmm.c, dtrmm.c:44-204, ctrmm.c, ztrmm.c:44). */
void strmm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, float, const float *, size_t, float *,
           size_t);
void dtrmm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, double, const double *, size_t,
           double *, size_t);
void ctrmm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, cuComplex_t, const cuComplex_t *,
           size_t, cuComplex_t *, size_t);
void ztrmm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, cuDoubleComplex_t,
           const cuDoubleComplex_t *, size_t, cuDoubleComplex_t *, size_t);
/* X <- alpha op(A) B or alpha B op(A), B unchanged (replaces blas/dtrmm.c:206-372 and twins). */
void strmm2(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, float, const float *, size_t,
            const float *, size_t, float *, size_t);
void dtrmm2(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, double, const double *, size_t,
            const double *, size_t, double *, size_t);
void ctrmm2(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, cuComplex_t, const cuComplex_t *,
            size_t, const cuComplex_t *, size_t, cuComplex_t *, size_t);
void ztrmm2(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, cuDoubleComplex_t,
            const cuDoubleComplex_t *, size_t, const cuDoubleComplex_t *, size_t, cuDoubleComplex_t *, size_t);
/* B <- alpha op(A)^-1 B or alpha B op(A)^-1 (replaces blas/strsm.c, dtrsm.c:34-216, ctrsm.c, ztrsm.c:34). */
void strsm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, float, const float *, size_t, float *,
           size_t);
void dtrsm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, double, const double *, size_t,
           double *, size_t);
void ctrsm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, cuComplex_t, const cuComplex_t *,
           size_t, cuComplex_t *, size_t);
void ztrsm(CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t, size_t, cuDoubleComplex_t,
           const cuDoubleComplex_t *, size_t, cuDoubleComplex_t *, size_t);

/** Handle: binds a device and its kernels (reference include/blas.h:153-155). */
typedef struct __cublashandle_st * CUBLAShandle;
CUresult cuBLASCreate(CUBLAShandle *);
CUresult cuBLASDestroy(CUBLAShandle);

/* C <- alpha * op(A) * op(A)^T + beta * C, one triangle
 * (replaces cuDsyrk, reference blas/dsyrk.c:163-209). */
CUresult cuDsyrk(CUBLAShandle, CBlasUplo, CBlasTranspose,
                 size_t, size_t,
                 double, CUdeviceptr, size_t,
                 double, CUdeviceptr, size_t, CUstream);

/* C <- alpha * op(A) * op(A)^H + beta * C, one triangle, real alpha and beta;
 * trans must be N or C (replaces cuZherk, reference blas/zherk.c:190). */
CUresult cuZherk(CUBLAShandle, CBlasUplo, CBlasTranspose,
                 size_t, size_t,
                 double, CUdeviceptr, size_t,
                 double, CUdeviceptr, size_t, CUstream);

/* In-place matrix multiply (reference include/blas.h:183-201). */
#define cuDgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream) \
        cuDgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc, stream)
#define cuZgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream) \
        cuZgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc, stream)

/* D <- alpha * op(A) * op(B) + beta * C (D may equal C)
 * (replaces cuDgemm2, reference blas/dgemm.c:148-200). */
CUresult cuDgemm2(CUBLAShandle, CBlasTranspose, CBlasTranspose,
                  size_t, size_t, size_t,
                  double, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  double, CUdeviceptr, size_t, CUdeviceptr, size_t, CUstream);
/* (replaces cuZgemm2, reference blas/zgemm.c:226) */
CUresult cuZgemm2(CUBLAShandle, CBlasTranspose, CBlasTranspose,
                  size_t, size_t, size_t,
                  cuDoubleComplex_t, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  cuDoubleComplex_t, CUdeviceptr, size_t, CUdeviceptr, size_t, CUstream);

/* Out-of-place triangular multiply X <- alpha * op(A) * B or alpha * B * op(A)
 * (replaces cuDtrmm2, reference blas/dtrmm.c:374-426). */
CUresult cuDtrmm2(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                  size_t, size_t,
                  double, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  CUdeviceptr, size_t, CUstream);
/* (replaces cuZtrmm2, reference blas/ztrmm.c:422) */
CUresult cuZtrmm2(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                  size_t, size_t,
                  cuDoubleComplex_t, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  CUdeviceptr, size_t, CUstream);

/* In-place triangular solve B <- alpha * op(A)^-1 * B or alpha * B * op(A)^-1
 * (replaces cuDtrsm, reference blas/dtrsm.c:218-265). */
CUresult cuDtrsm(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                 size_t, size_t,
                 double, CUdeviceptr, size_t,
                 CUdeviceptr, size_t, CUstream);
/* (replaces cuZtrsm, reference blas/ztrsm.c:248) */
CUresult cuZtrsm(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                 size_t, size_t,
                 cuDoubleComplex_t, CUdeviceptr, size_t,
                 CUdeviceptr, size_t, CUstream);

/*
 * Single precision (FP32 data, FP32 accumulation on the matrix cores):
 * replaces cuSsyrk, cuSgemm2 and cuStrsm of the reference (include/blas.h:158,
 * 193, 235; blas/ssyrk.c, blas/sgemm.c, blas/strsm.c).  op(A) in {N, T}.
 */
CUresult cuSsyrk(CUBLAShandle, CBlasUplo, CBlasTranspose,
                 size_t, size_t,
                 float, CUdeviceptr, size_t,
                 float, CUdeviceptr, size_t, CUstream);
#define cuSgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream) \
        cuSgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc, stream)
CUresult cuSgemm2(CUBLAShandle, CBlasTranspose, CBlasTranspose,
                  size_t, size_t, size_t,
                  float, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  float, CUdeviceptr, size_t, CUdeviceptr, size_t, CUstream);
CUresult cuStrsm(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                 size_t, size_t,
                 float, CUdeviceptr, size_t,
                 CUdeviceptr, size_t, CUstream);

/*
 * Single precision triangular multiply and the single complex BLAS: the
 * operands are widened on the stream into stream-ordered double / double
 * complex copies, the double / double complex routine above runs on them,
 * and the result is narrowed back (at least single-precision accuracy; the
 * unreferenced triangle and padding are not written).  Same argument order,
 * xerbla indices and quick returns as the double twins.
 */
/* (replaces cuStrmm2, reference blas/strmm.c:374) */
CUresult cuStrmm2(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                  size_t, size_t,
                  float, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  CUdeviceptr, size_t, CUstream);
/* C <- alpha op(A) op(A)^H + beta C, trans N or C (replaces cuCherk,
 * reference blas/cherk.c:190) */
CUresult cuCherk(CUBLAShandle, CBlasUplo, CBlasTranspose,
                 size_t, size_t,
                 float, CUdeviceptr, size_t,
                 float, CUdeviceptr, size_t, CUstream);
#define cuCgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream) \
        cuCgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc, stream)
/* (replaces cuCgemm2, reference blas/cgemm.c:226) */
CUresult cuCgemm2(CUBLAShandle, CBlasTranspose, CBlasTranspose,
                  size_t, size_t, size_t,
                  cuComplex_t, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  cuComplex_t, CUdeviceptr, size_t, CUdeviceptr, size_t, CUstream);
/* (replaces cuCtrmm2, reference blas/ctrmm.c:422) */
CUresult cuCtrmm2(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                  size_t, size_t,
                  cuComplex_t, CUdeviceptr, size_t, CUdeviceptr, size_t,
                  CUdeviceptr, size_t, CUstream);
/* (replaces cuCtrsm, reference blas/ctrsm.c:248) */
CUresult cuCtrsm(CUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag,
                 size_t, size_t,
                 cuComplex_t, CUdeviceptr, size_t,
                 CUdeviceptr, size_t, CUstream);

/*
 * Host-pointer BLAS of the reference's multi-GPU layer
 * (include/blas.h:257-360), collapsed onto the one device of the CUmultiGPU
 * context: operands are column-major host arrays, copied to the device
 * through double-buffered pinned staging, the device routine above runs, and
 * the result is copied back before the call returns (the reference may still
 * be running its tasks until cuMultiGPUBLASSynchronize; here that call only
 * waits for the device).  Same argument order, in-place results and xerbla
 * indices as the reference.  All four precisions (single and single complex
 * run on the device routines above of the same precision).
 */
typedef struct __cumultigpublashandle_st * CUmultiGPUBLAShandle;
CUresult cuMultiGPUBLASCreate(CUmultiGPUBLAShandle *, CUmultiGPU);
CUresult cuMultiGPUBLASDestroy(CUmultiGPUBLAShandle);
CUresult cuMultiGPUBLASSynchronize(CUmultiGPUBLAShandle);

/* (replaces cuMultiGPUDsyrk, reference blas/dsyrk.c:212) */
CUresult cuMultiGPUDsyrk(CUmultiGPUBLAShandle, CBlasUplo, CBlasTranspose, size_t, size_t,
                         double, const double *, size_t, double, double *, size_t);
/* (replaces cuMultiGPUZherk, reference blas/zherk.c:241) */
CUresult cuMultiGPUZherk(CUmultiGPUBLAShandle, CBlasUplo, CBlasTranspose, size_t, size_t,
                         double, const cuDoubleComplex_t *, size_t, double, cuDoubleComplex_t *, size_t);
/* (replaces cuMultiGPUDgemm, reference blas/dgemm.c:435) */
CUresult cuMultiGPUDgemm(CUmultiGPUBLAShandle, CBlasTranspose, CBlasTranspose, size_t, size_t, size_t,
                         double, const double *, size_t, const double *, size_t, double, double *, size_t);
/* (replaces cuMultiGPUZgemm, reference blas/zgemm.c:519) */
CUresult cuMultiGPUZgemm(CUmultiGPUBLAShandle, CBlasTranspose, CBlasTranspose, size_t, size_t, size_t,
                         cuDoubleComplex_t, const cuDoubleComplex_t *, size_t, const cuDoubleComplex_t *,
                         size_t, cuDoubleComplex_t, cuDoubleComplex_t *, size_t);
/* B <- alpha op(A) B or alpha B op(A), in place (replaces cuMultiGPUDtrmm,
 * reference blas/dtrmm.c:468) */
CUresult cuMultiGPUDtrmm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, double, const double *, size_t, double *, size_t);
/* (replaces cuMultiGPUZtrmm, reference blas/ztrmm.c:519) */
CUresult cuMultiGPUZtrmm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, cuDoubleComplex_t, const cuDoubleComplex_t *, size_t, cuDoubleComplex_t *,
                         size_t);
/* (replaces cuMultiGPUDtrsm, reference blas/dtrsm.c:268) */
CUresult cuMultiGPUDtrsm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, double, const double *, size_t, double *, size_t);
/* (replaces cuMultiGPUZtrsm, reference blas/ztrsm.c:297) */
CUresult cuMultiGPUZtrsm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, cuDoubleComplex_t, const cuDoubleComplex_t *, size_t, cuDoubleComplex_t *,
                         size_t);
/* single and single complex (replace cuMultiGPUSsyrk / Cherk / Sgemm / Cgemm /
 * Strmm / Ctrmm / Strsm / Ctrsm, reference blas/ssyrk.c:214, blas/cherk.c:240,
 * blas/sgemm.c:432, blas/cgemm.c:509, blas/strmm.c:467, blas/ctrmm.c:518,
 * blas/strsm.c:267, blas/ctrsm.c:296) */
CUresult cuMultiGPUSsyrk(CUmultiGPUBLAShandle, CBlasUplo, CBlasTranspose, size_t, size_t,
                         float, const float *, size_t, float, float *, size_t);
CUresult cuMultiGPUCherk(CUmultiGPUBLAShandle, CBlasUplo, CBlasTranspose, size_t, size_t,
                         float, const cuComplex_t *, size_t, float, cuComplex_t *, size_t);
CUresult cuMultiGPUSgemm(CUmultiGPUBLAShandle, CBlasTranspose, CBlasTranspose, size_t, size_t, size_t,
                         float, const float *, size_t, const float *, size_t, float, float *, size_t);
CUresult cuMultiGPUCgemm(CUmultiGPUBLAShandle, CBlasTranspose, CBlasTranspose, size_t, size_t, size_t,
                         cuComplex_t, const cuComplex_t *, size_t, const cuComplex_t *, size_t, cuComplex_t,
                         cuComplex_t *, size_t);
CUresult cuMultiGPUStrmm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, float, const float *, size_t, float *, size_t);
CUresult cuMultiGPUCtrmm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, cuComplex_t, const cuComplex_t *, size_t, cuComplex_t *, size_t);
CUresult cuMultiGPUStrsm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, float, const float *, size_t, float *, size_t);
CUresult cuMultiGPUCtrsm(CUmultiGPUBLAShandle, CBlasSide, CBlasUplo, CBlasTranspose, CBlasDiag, size_t,
                         size_t, cuComplex_t, const cuComplex_t *, size_t, cuComplex_t *, size_t);

#ifdef __cplusplus
}
#endif

#endif
