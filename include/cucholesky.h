/*
 * Extensions of the MI355X Cholesky library that have no counterpart in the
 * reference's include/lapack.h or include/blas.h.  They share its handles,
 * CUresult codes and CUdeviceptr (byte address) convention.
 */
#ifndef CUCHOLESKY_H
#define CUCHOLESKY_H

#include "lapack.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Event timing of the trailing-matrix updates (SYRK / HERK) that cuDpotrf and
 * cuZpotrf launch on the handle's main stream.  cuLAPACKSetTiming(h, 1)
 * clears the record and starts recording; 0 stops.  cuLAPACKGetTiming waits
 * for the handle's main stream, copies up to `capacity` launch durations (ms)
 * and algorithmic flop counts (real flops: t (t + 1) k for DSYRK, four times
 * that for ZHERK), stores the number recorded in *count and clears the
 * record.  Either output array may be NULL.
 */
CUresult cuLAPACKSetTiming(CULAPACKhandle handle, int enable);
CUresult cuLAPACKGetTiming(CULAPACKhandle handle, double * ms, double * flops, long capacity,
                           long * count);

/*
 * Outer block size of the blocked factorisations on this handle, per
 * precision: 'D' (cuDpotrf), 'Z' (cuZpotrf), 'S' (cuSpotrf and the FP32
 * factor of cuDsposvx), 'B' (its BF16, split-BF16 and FP16 factors).  nb = 0
 * restores the default:
 * $CUCHOLESKY_POTRF_NB, else the table measured by autotune.py
 * (tuning.txt beside the library or $CUCHOLESKY_TUNING), else 384 (real) /
 * 128 (complex).  nb must be 0 or in [128, 8192].
 */
CUresult cuLAPACKSetBlockSize(CULAPACKhandle handle, char precision, long nb);

/*
 * Solve A X = B with the Cholesky factor from cuDpotrf (A = L L^T or U^T U),
 * B (n x nrhs) overwritten by X.  Arguments and INFO follow LAPACK DPOTRS
 * (UPLO = 1, N = 2, NRHS = 3, A = 4, LDA = 5, B = 6, LDB = 7; the handle is
 * not counted).  Synchronous.
 */
CUresult cuDpotrs(CULAPACKhandle handle, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, long * info);

/*
 * Mixed-precision SPD solve A X = B (A FP64, n x n, triangle uplo; B, X
 * n x nrhs), as LAPACK DSPOSV: A is copied to FP32 and factored there (the
 * O(n^3) trailing updates on FP32 matrix cores, the diagonal blocks in FP64),
 * then the FP64 solution is refined with FP64 residuals r = b - A x until
 * every column has ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n), at most 30
 * steps.  ITER >= 0: refinement steps taken.  ITER < 0: the low-precision
 * path failed (-2: a value overflowed FP32, -3: the FP32 factorisation
 * failed, -31: no convergence in 30 steps); then A is overwritten by its
 * FP64 Cholesky factor and X comes from cuDpotrf + cuDpotrs, with INFO > 0 if
 * A is not positive definite.  A is not modified when ITER >= 0.
 */
CUresult cuDsposv(CULAPACKhandle handle, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, long * iter,
                  long * info);

/*
 * cuDsposv with the factorisation precision and the refinement chosen:
 * factor = 'S' (FP32 products, as cuDsposv), 'B' (trailing-update products
 * in bfloat16, FP32 accumulation) or 'X' (three bfloat16 products of a
 * two-term split hi + lo of every panel value, ~16 significant bits, FP32
 * accumulation) or 'H' (products in IEEE binary16, 11 significant bits, FP32
 * accumulation; the panel copies are scaled by a power of two 2^e with
 * sqrt(max a_ii) 2^e <= 2^15, undone exactly, and an entry that still does
 * not fit fails the attempt); refine = 'I' (classical iterative refinement, at most 30
 * steps, as cuDsposv) or 'G' (GMRES-IR: each step solves inv(M) A d =
 * inv(M) r by restarted GMRES(30) in FP64, M the low-precision factor,
 * relative tolerance 1e-10; at most 10 steps).  A 'B', 'X' or 'H' attempt
 * that fails is retried with 'S' before the FP64 fall-back; *used (may be
 * NULL) reports the precision of the returned solution: 'B', 'X', 'H', 'S'
 * or 'D'.
 * ITER counts refinement steps (GMRES cycles for 'G').
 */
CUresult cuDsposvx(CULAPACKhandle handle, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                   size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, char factor,
                   char refine, long * iter, char * used, long * info);

/*
 * Batched Cholesky factorisation + log-determinant of many SPD matrices of
 * uneven sizes (n_i <= 256) packed in one device buffer: matrix i is
 * n_i x n_i, column-major with leading dimension n_i, starting `offsets[i]`
 * elements into the buffer; only triangle uplo is read and written.
 * precision: 'D' (double data) or 'S' (float data; arithmetic in FP64).
 *
 * cuPotrfBatchedPlanCreate takes the sizes and offsets on the host, sorts the
 * matrices longest first and sizes one persistent launch.  cuPotrfBatchedRun
 * factors every matrix in place, writes info[i] (int, 0 or the 1-based
 * failing column) and logdet[i] (double, log det A_i = 2 sum log L_jj; NaN
 * on failure; logdet may be 0 to skip), asynchronously on `stream`: one HIP
 * graph launch (the graph is captured on the first run with a given
 * (A, logdet, info) and replayed afterwards).  A plan runs one batch at a
 * time: runs that overlap on different streams race.
 */
typedef struct __cubatchedplan_st * CUbatchedplan;
CUresult cuPotrfBatchedPlanCreate(CULAPACKhandle handle, CBlasUplo uplo, char precision,
                                  size_t count, const size_t * n, const size_t * offsets,
                                  CUbatchedplan * plan);
CUresult cuPotrfBatchedRun(CUbatchedplan plan, CUdeviceptr A, CUdeviceptr logdet, CUdeviceptr info,
                           CUstream stream);
/*
 * cuPotriBatchedRun: the same factorisation and logdet, then every factored
 * matrix (info[i] == 0) is overwritten by its inverse (triangle uplo of
 * A_i^-1, LAPACK xPOTRI semantics), in the same single launch.
 */
CUresult cuPotriBatchedRun(CUbatchedplan plan, CUdeviceptr A, CUdeviceptr logdet, CUdeviceptr info,
                           CUstream stream);
CUresult cuPotrfBatchedPlanDestroy(CUbatchedplan plan);

/*
 * cuPotrfBatchedNoAlloc: the batched factorisation + logdet of
 * cuPotrfBatchedRun in one call that allocates nothing.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The sizes and offsets are
 * DEVICE arrays: n (int32[count], each n_i <= 256; a matrix with n_i out of
 * range gets info = -1 and logdet = NaN and is not touched) and offsets
 * (int64[count], element offsets into A); logdet (double[count]) and info
 * (int32[count]) are required.  count <= 65536 (a larger count is rejected
 * with CUDA_ERROR_INVALID_VALUE and XERBLA parameter 4).  The call launches
 * its kernels on `stream` and returns; its only working memory is a fixed
 * array in the library's device image (the size-sorted order), shared by
 * every call on the device and left holding the last call's contents: calls
 * that overlap on one device (unordered streams) race.  The handle must have
 * been created on the device that runs the call.
 */
CUresult cuPotrfBatchedNoAlloc(CULAPACKhandle handle, CBlasUplo uplo, char precision, size_t count,
                               CUdeviceptr n, CUdeviceptr offsets, CUdeviceptr A, CUdeviceptr logdet,
                               CUdeviceptr info, CUstream stream);

/*
 * Batched SPD solve + log-determinant + inverse of `count` matrices of any
 * sizes, as one persistent kernel over 64 x 64 tiles of the whole batch.
 * Matrix i is n[i] x n[i] (column-major, ld n[i]; triangle uplo is read) at
 * element offset offA[i] of A; its right-hand sides and solution are
 * n[i] x nrhs (ld n[i]) at element offset offB[i] of B and X.  Per matrix:
 *   - the Cholesky factor and its inverse in FP64 (matrix cores), in plan
 *     workspaces; A is never written;
 *   - when Ainv != 0, triangle uplo of A_i^-1 at element offset offA[i] of
 *     Ainv (LAPACK xPOTRI semantics);
 *   - X_i = A_i^-1 B_i refined in FP64 with residuals of the original A_i
 *     (r = b - A x, x += A^-1 r) until every column meets LAPACK DSPOSV's
 *     test ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n), at most 30 steps;
 *     iter[i] = steps taken (-31: no convergence);
 *   - logdet[i] = log det A_i = sum_j log L_jj^2 (NaN on failure);
 *   - info[i] = 0, or the order of the leading minor that is not positive
 *     definite (then X_i and the inverse are not written), or -1 when the
 *     kernel abandoned a dependency wait (a fault; nothing is valid).
 *
 * cuDsposvxBatchedPlanCreate takes the sizes and offsets on the host, builds
 * and orders the task graph and allocates the workspaces (about three padded
 * copies of the batch: 3 sum (64 ceil(n_i / 64))^2 doubles, plus n_i nrhs
 * doubles per matrix).  cuDsposvxBatchedRun is asynchronous on `stream`:
 * one HIP graph launch (captured on the first run with given buffers,
 * replayed afterwards); iter and info are device int arrays, logdet a device
 * double array, of `count` elements.  A plan runs one batch at a time: runs
 * that overlap race.
 */
typedef struct __cubatchedsolveplan_st * CUbatchedsolveplan;
CUresult cuDsposvxBatchedPlanCreate(CULAPACKhandle handle, CBlasUplo uplo, size_t count, const size_t * n,
                                    const size_t * offA, size_t nrhs, const size_t * offB,
                                    CUbatchedsolveplan * plan);
CUresult cuDsposvxBatchedRun(CUbatchedsolveplan plan, CUdeviceptr A, CUdeviceptr B, CUdeviceptr X,
                             CUdeviceptr Ainv, CUdeviceptr iter, CUdeviceptr logdet, CUdeviceptr info,
                             CUstream stream);
CUresult cuDsposvxBatchedPlanDestroy(CUbatchedsolveplan plan);

/*
 * Synchronous form: a plan for the batch, one run on the handle's stream,
 * iter / logdet / info copied to the host arrays, the plan destroyed.
 * Returns CUDA_ERROR_LAUNCH_FAILED if the kernel abandoned a dependency wait.
 */
CUresult cuDsposvxBatched(CULAPACKhandle handle, CBlasUplo uplo, size_t count, const size_t * n,
                          const size_t * offA, CUdeviceptr A, size_t nrhs, const size_t * offB,
                          CUdeviceptr B, CUdeviceptr X, CUdeviceptr Ainv, long * iter,
                          double * logdet, long * info);

#ifdef __cplusplus
}
#endif

#endif
