/*
 * GPU LAPACK entry points of the MI355X Cholesky library.
 *
 * Drop-in for the device-pointer ("cu") half of the reference's
 * include/lapack.h, double and double-complex precision: same names, argument
 * order, info conventions and return codes.  Matrices are column-major and
 * device resident (CUdeviceptr = device byte address).  Like the reference,
 * the factorisation and inverse routines are synchronous: `info` is valid on
 * return.  Positive info is a numerical failure reported with CUDA_SUCCESS;
 * negative info is an illegal argument reported through xerbla with
 * CUDA_ERROR_INVALID_VALUE.
 *
 * Unlike the reference, no diagonal block travels to the host: every step runs
 * on the GPU.  The host-memory routines of the reference header (dpotrf,
 * dtrtri, ... without a prefix, which the reference runs on the CPU) are
 * exported too and run on the GPU: operands are staged through pinned host
 * buffers to the device, the device routine runs, the result comes back.
 */
#ifndef LAPACK_H
#define LAPACK_H

#include "blas.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Host-memory routines (replace the reference's CPU implementations,
 * include/lapack.h:17-60): column-major host arrays, synchronous, run on the
 * current GPU.  Argument errors: *info = -(position) and xerbla, as the
 * reference (lda: -4; dtrtri lda: -5; dtrtri2 lda / ldb: -5 / -7).  A
 * numerical failure is info > 0.  A device failure is reported through
 * errorHandler and sets *info = -1000 - CUresult.
 */
/* Cholesky factorisation in place (replaces lapack/spotrf.c, dpotrf.c:34-146, cpotrf.c, zpotrf.c:36-148). */
void spotrf(CBlasUplo, size_t, float *, size_t, long *);
void dpotrf(CBlasUplo, size_t, double *, size_t, long *);
void cpotrf(CBlasUplo, size_t, cuComplex_t *, size_t, long *);
void zpotrf(CBlasUplo, size_t, cuDoubleComplex_t *, size_t, long *);
/* Triangular inverse in place (replaces lapack/dtrtri.c:98-164 and twins). */
void strtri(CBlasUplo, CBlasDiag, size_t, float *, size_t, long *);
void dtrtri(CBlasUplo, CBlasDiag, size_t, double *, size_t, long *);
void ctrtri(CBlasUplo, CBlasDiag, size_t, cuComplex_t *, size_t, long *);
void ztrtri(CBlasUplo, CBlasDiag, size_t, cuDoubleComplex_t *, size_t, long *);
/* Triangular inverse out of place, B <- inv(A) on triangle uplo (replaces lapack/dtrtri.c:226-299 and twins). */
void strtri2(CBlasUplo, CBlasDiag, size_t, const float *, size_t, float *, size_t, long *);
void dtrtri2(CBlasUplo, CBlasDiag, size_t, const double *, size_t, double *, size_t, long *);
void ctrtri2(CBlasUplo, CBlasDiag, size_t, const cuComplex_t *, size_t, cuComplex_t *, size_t, long *);
void ztrtri2(CBlasUplo, CBlasDiag, size_t, const cuDoubleComplex_t *, size_t, cuDoubleComplex_t *, size_t, long *);
/* U U^H / L^H L in place (replaces lapack/dlauum.c:71-129 and twins). */
void slauum(CBlasUplo, size_t, float *, size_t, long *);
void dlauum(CBlasUplo, size_t, double *, size_t, long *);
void clauum(CBlasUplo, size_t, cuComplex_t *, size_t, long *);
void zlauum(CBlasUplo, size_t, cuDoubleComplex_t *, size_t, long *);
/* Inverse from the Cholesky factor (replaces lapack/dpotri.c:4-23 and twins). */
void spotri(CBlasUplo, size_t, float *, size_t, long *);
void dpotri(CBlasUplo, size_t, double *, size_t, long *);
void cpotri(CBlasUplo, size_t, cuComplex_t *, size_t, long *);
void zpotri(CBlasUplo, size_t, cuDoubleComplex_t *, size_t, long *);
/* 2 sum log(Re x[i incx]), logs in FP64 (replaces lapack/dlogdet.c:10-25,
 * slogdet.c, clogdet.c, zlogdet.c:11-26, include/lapack.h:161-164). */
float slogdet(const float *, size_t, size_t);
double dlogdet(const double *, size_t, size_t);
float clogdet(const cuComplex_t *, size_t, size_t);
double zlogdet(const cuDoubleComplex_t *, size_t, size_t);

/** Handle (reference include/lapack.h:63-65). */
typedef struct __culapackhandle_st * CULAPACKhandle;
CUresult cuLAPACKCreate(CULAPACKhandle *);
CUresult cuLAPACKDestroy(CULAPACKhandle);

/* Cholesky factorisation A = L L^T (lower) or U^T U (upper), in place
 * (replaces cuDpotrf, reference lapack/dpotrf.c:172-296). */
CUresult cuDpotrf(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* Hermitian Cholesky A = L L^H or U^H U
 * (replaces cuZpotrf, reference lapack/zpotrf.c:174-298). */
CUresult cuZpotrf(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* Single precision Cholesky of FP32 data; the diagonal blocks are factored
 * in FP64 (replaces cuSpotrf, reference lapack/spotrf.c:261-398). */
CUresult cuSpotrf(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);

/* In-place triangular inverse; info > 0 names an exactly zero diagonal
 * (replaces cuDtrtri, reference lapack/dtrtri.c:326-468). */
CUresult cuDtrtri(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, long *);
/* (replaces cuZtrtri, reference lapack/ztrtri.c:326) */
CUresult cuZtrtri(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, long *);

/* In-place macros over the out-of-place single-block inverse
 * (reference include/lapack.h:85-88). */
#define cuDtrti2(handle, uplo, diag, n, A, lda, info, stream) cuDtrti22(handle, uplo, diag, n, A, lda, A, lda, info, stream)
#define cuZtrti2(handle, uplo, diag, n, A, lda, info, stream) cuZtrti22(handle, uplo, diag, n, A, lda, A, lda, info, stream)

/* Single-block triangular inverse B <- A^-1 of order n <= 128 (double) or
 * n <= 64 (complex), asynchronous on `stream`; the int at device address `info`
 * receives 0 or the 1-based index of a zero diagonal
 * (replaces cuDtrti22, reference lapack/dtrtri.c:301-324, which accepts
 * n <= 32). */
CUresult cuDtrti22(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, CUdeviceptr, size_t, CUdeviceptr, CUstream);
/* (replaces cuZtrti22, reference lapack/ztrtri.c:301) */
CUresult cuZtrti22(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, CUdeviceptr, size_t, CUdeviceptr, CUstream);

/* Triangular square U U^T (upper) or L^T L (lower), in place
 * (replaces cuDlauum, reference lapack/dlauum.c:154-292). */
CUresult cuDlauum(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* (replaces cuZlauum, reference lapack/zlauum.c:153) */
CUresult cuZlauum(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);

/* Inverse of an SPD matrix from its Cholesky factor: TRTRI then LAUUM
 * (replaces cuDpotri, reference lapack/dpotri.c:25-46). */
CUresult cuDpotri(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* (replaces cuZpotri, reference lapack/zpotri.c:25) */
CUresult cuZpotri(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);

/* 2 * sum(log(x[i * incx])), i < n; *result is valid on return
 * (replaces cuDlogdet, reference lapack/dlogdet.c:37-74). */
CUresult cuDlogdet(CULAPACKhandle, CUdeviceptr, size_t, size_t, double *, CUstream);
/* 2 * sum(log(Re x[i * incx])), computed in double precision
 * (replaces cuZlogdet, reference lapack/zlogdet.c:38, which takes the logs in
 * single precision). */
CUresult cuZlogdet(CULAPACKhandle, CUdeviceptr, size_t, size_t, double *, CUstream);

/*
 * Single precision and single complex: the matrix is widened on the handle's
 * stream into a stream-ordered double / double complex copy, the routine
 * above runs on it, and the result is narrowed back (the whole n x n block
 * round-trips, exactly for the unreferenced triangle).  Same argument order,
 * info conventions and xerbla indices as the double twins.  cuSpotrf is the
 * native FP32 path declared above.
 */
/* (replace cuCpotrf, reference lapack/cpotrf.c:174) */
CUresult cuCpotrf(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* (replace cuStrtri / cuCtrtri, reference lapack/strtri.c:369, lapack/ctrtri.c:326) */
CUresult cuStrtri(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, long *);
CUresult cuCtrtri(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, long *);
/* (replace cuStrti22 / cuCtrti22, reference lapack/strtri.c:301, lapack/ctrtri.c:301) */
#define cuStrti2(handle, uplo, diag, n, A, lda, info, stream) cuStrti22(handle, uplo, diag, n, A, lda, A, lda, info, stream)
#define cuCtrti2(handle, uplo, diag, n, A, lda, info, stream) cuCtrti22(handle, uplo, diag, n, A, lda, A, lda, info, stream)
CUresult cuStrti22(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, CUdeviceptr, size_t, CUdeviceptr, CUstream);
CUresult cuCtrti22(CULAPACKhandle, CBlasUplo, CBlasDiag, size_t, CUdeviceptr, size_t, CUdeviceptr, size_t, CUdeviceptr, CUstream);
/* (replace cuSlauum / cuClauum, reference lapack/slauum.c:197, lapack/clauum.c:154) */
CUresult cuSlauum(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
CUresult cuClauum(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* (replace cuSpotri / cuCpotri, reference lapack/spotri.c:25, lapack/cpotri.c:25) */
CUresult cuSpotri(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
CUresult cuCpotri(CULAPACKhandle, CBlasUplo, size_t, CUdeviceptr, size_t, long *);
/* 2 * sum(log(Re x[i * incx])) of single / single complex data, logs and sum
 * in double precision, rounded to float (replace cuSlogdet / cuClogdet,
 * reference lapack/slogdet.c:37, lapack/clogdet.c:38). */
CUresult cuSlogdet(CULAPACKhandle, CUdeviceptr, size_t, size_t, float *, CUstream);
CUresult cuClogdet(CULAPACKhandle, CUdeviceptr, size_t, size_t, float *, CUstream);

/*
 * Host-pointer LAPACK of the reference's multi-GPU layer
 * (include/lapack.h:117-158), collapsed onto the one device of the
 * CUmultiGPU context: the matrix (column-major host array, in place) is
 * copied to the device through double-buffered pinned staging, factored /
 * inverted by the device routines above and copied back.  As in the
 * reference, a numerical failure (info > 0) returns
 * CUDA_ERROR_INVALID_VALUE; an illegal argument calls xerbla with the
 * reference's index.  All four precisions.
 */
typedef struct __cumultigpulapackhandle_st * CUmultiGPULAPACKhandle;
CUresult cuMultiGPULAPACKCreate(CUmultiGPULAPACKhandle *, CUmultiGPU);
CUresult cuMultiGPULAPACKDestroy(CUmultiGPULAPACKhandle);
CUresult cuMultiGPULAPACKSynchronize(CUmultiGPULAPACKhandle);
/* (replaces cuMultiGPUDpotrf / cuMultiGPUZpotrf, reference lapack/dpotrf.c:298, lapack/zpotrf.c:300) */
CUresult cuMultiGPUDpotrf(CUmultiGPULAPACKhandle, CBlasUplo, size_t, double *, size_t, long *);
CUresult cuMultiGPUZpotrf(CUmultiGPULAPACKhandle, CBlasUplo, size_t, cuDoubleComplex_t *, size_t, long *);
/* (replaces cuMultiGPUDtrtri / cuMultiGPUZtrtri, reference lapack/dtrtri.c:470, lapack/ztrtri.c:470) */
CUresult cuMultiGPUDtrtri(CUmultiGPULAPACKhandle, CBlasUplo, CBlasDiag, size_t, double *, size_t, long *);
CUresult cuMultiGPUZtrtri(CUmultiGPULAPACKhandle, CBlasUplo, CBlasDiag, size_t, cuDoubleComplex_t *, size_t,
                          long *);
/* (replaces cuMultiGPUDlauum / cuMultiGPUZlauum, reference lapack/dlauum.c:294, lapack/zlauum.c:293) */
CUresult cuMultiGPUDlauum(CUmultiGPULAPACKhandle, CBlasUplo, size_t, double *, size_t, long *);
CUresult cuMultiGPUZlauum(CUmultiGPULAPACKhandle, CBlasUplo, size_t, cuDoubleComplex_t *, size_t, long *);
/* (replaces cuMultiGPUDpotri / cuMultiGPUZpotri, reference lapack/dpotri.c:48, lapack/zpotri.c:48) */
CUresult cuMultiGPUDpotri(CUmultiGPULAPACKhandle, CBlasUplo, size_t, double *, size_t, long *);
CUresult cuMultiGPUZpotri(CUmultiGPULAPACKhandle, CBlasUplo, size_t, cuDoubleComplex_t *, size_t, long *);
/* single and single complex (replace cuMultiGPUSpotrf ... cuMultiGPUCpotri,
 * reference lapack/spotrf.c:400, lapack/cpotrf.c:300, lapack/strtri.c:474,
 * lapack/ctrtri.c:469, lapack/slauum.c:308, lapack/clauum.c:293,
 * lapack/spotri.c:48, lapack/cpotri.c:47) */
CUresult cuMultiGPUSpotrf(CUmultiGPULAPACKhandle, CBlasUplo, size_t, float *, size_t, long *);
CUresult cuMultiGPUCpotrf(CUmultiGPULAPACKhandle, CBlasUplo, size_t, cuComplex_t *, size_t, long *);
CUresult cuMultiGPUStrtri(CUmultiGPULAPACKhandle, CBlasUplo, CBlasDiag, size_t, float *, size_t, long *);
CUresult cuMultiGPUCtrtri(CUmultiGPULAPACKhandle, CBlasUplo, CBlasDiag, size_t, cuComplex_t *, size_t, long *);
CUresult cuMultiGPUSlauum(CUmultiGPULAPACKhandle, CBlasUplo, size_t, float *, size_t, long *);
CUresult cuMultiGPUClauum(CUmultiGPULAPACKhandle, CBlasUplo, size_t, cuComplex_t *, size_t, long *);
CUresult cuMultiGPUSpotri(CUmultiGPULAPACKhandle, CBlasUplo, size_t, float *, size_t, long *);
CUresult cuMultiGPUCpotri(CUmultiGPULAPACKhandle, CBlasUplo, size_t, cuComplex_t *, size_t, long *);

#ifdef __cplusplus
}
#endif

#endif
