/*
 * CPU oracle: restatement of the reference's CPU BLAS/LAPACK path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library and
 * the CPU baseline of the benchmark.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links,
 * loads or calls it.
 *
 * Every routine cites the reference file:line it restates.  Enums are the
 * reference's char values ('N','T','C' / 'L','U' / 'L','R' / 'N','U').
 * Storage is column-major with an explicit leading dimension.  Routines that
 * report failure use the reference's `long *info` convention.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <complex.h>

typedef double complex zcomplex;

/* ---- BLAS-3, double (reference blas/dgemm.c:34-146, dsyrk.c:34-161,
 *      dtrsm.c:34-216, dtrmm.c:44-372) ---- */
void oracle_dgemm(char transA, char transB, size_t m, size_t n, size_t k,
                  double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                  double beta, double *C, size_t ldc);
void oracle_dsyrk(char uplo, char trans, size_t n, size_t k,
                  double alpha, const double *A, size_t lda, double beta, double *C, size_t ldc);
void oracle_dtrsm(char side, char uplo, char transA, char diag, size_t m, size_t n,
                  double alpha, const double *A, size_t lda, double *B, size_t ldb);
void oracle_dtrmm(char side, char uplo, char transA, char diag, size_t m, size_t n,
                  double alpha, const double *A, size_t lda, double *B, size_t ldb);
void oracle_dtrmm2(char side, char uplo, char transA, char diag, size_t m, size_t n,
                   double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                   double *X, size_t ldx);

/* ---- BLAS-3, double complex (reference blas/zgemm.c:34, zherk.c:35,
 *      ztrsm.c:34, ztrmm.c:44) ---- */
void oracle_zgemm(char transA, char transB, size_t m, size_t n, size_t k,
                  zcomplex alpha, const zcomplex *A, size_t lda, const zcomplex *B, size_t ldb,
                  zcomplex beta, zcomplex *C, size_t ldc);
void oracle_zherk(char uplo, char trans, size_t n, size_t k,
                  double alpha, const zcomplex *A, size_t lda, double beta, zcomplex *C, size_t ldc);
void oracle_ztrsm(char side, char uplo, char transA, char diag, size_t m, size_t n,
                  zcomplex alpha, const zcomplex *A, size_t lda, zcomplex *B, size_t ldb);
void oracle_ztrmm2(char side, char uplo, char transA, char diag, size_t m, size_t n,
                   zcomplex alpha, const zcomplex *A, size_t lda, const zcomplex *B, size_t ldb,
                   zcomplex *X, size_t ldx);

/* ---- LAPACK, double (reference lapack/dpotrf.c:34-146, dtrtri.c:43-164,
 *      dtrtri.c:166-299, dlauum.c:43-129, dpotri.c:4-23, dlogdet.c:10-25) ---- */
void oracle_dpotrf(char uplo, size_t n, double *A, size_t lda, long *info);
void oracle_dtrtri(char uplo, char diag, size_t n, double *A, size_t lda, long *info);
void oracle_dtrtri2(char uplo, char diag, size_t n, const double *A, size_t lda,
                    double *B, size_t ldb, long *info);
void oracle_dlauum(char uplo, size_t n, double *A, size_t lda, long *info);
void oracle_dpotri(char uplo, size_t n, double *A, size_t lda, long *info);
double oracle_dlogdet(const double *x, size_t incx, size_t n);

/* ---- LAPACK, double complex (reference lapack/zpotrf.c:36-148,
 *      ztrtri.c:43-164, zlauum.c:43-129, zpotri.c:4, zlogdet.c:11-26) ---- */
void oracle_zpotrf(char uplo, size_t n, zcomplex *A, size_t lda, long *info);
void oracle_ztrtri(char uplo, char diag, size_t n, zcomplex *A, size_t lda, long *info);
void oracle_zlauum(char uplo, size_t n, zcomplex *A, size_t lda, long *info);
void oracle_zpotri(char uplo, size_t n, zcomplex *A, size_t lda, long *info);
/* zlogdet exactly as the reference computes it: logs in single precision. */
double oracle_zlogdet_ref(const zcomplex *x, size_t incx, size_t n);
/* zlogdet with double-precision logs (the value the GPU library returns). */
double oracle_zlogdet(const zcomplex *x, size_t incx, size_t n);

/* ---- Unblocked textbook forms (reference test/lapack/ref/dpotrf_ref.c:1-51,
 *      dtrtri_ref.c:1-61, dlauum_ref.c:1-32): independent checks of the
 *      blocked routines above ---- */
void oracle_dpotrf_unblocked(char uplo, size_t n, double *A, size_t lda, long *info);
void oracle_dtrtri_unblocked(char uplo, char diag, size_t n, double *A, size_t lda, long *info);
void oracle_dlauum_unblocked(char uplo, size_t n, double *A, size_t lda);
void oracle_zpotrf_unblocked(char uplo, size_t n, zcomplex *A, size_t lda, long *info);

/* ---- Test-matrix generators ---- */
/* SPD matrix with eigenvalues in [1, c], 1 and c both present
 * (reference test/lapack/util/dlatmc.c:11-70); uses the C library rand(). */
int oracle_dlatmc(size_t n, double c, double *A, size_t lda);
/* Hermitian analogue (reference test/lapack/util/zlatmc.c:11-70). */
int oracle_zlatmc(size_t n, double c, zcomplex *A, size_t lda);
/* Seeds the C library generator (reference tests call srand(0)). */
void oracle_srand(unsigned int seed);
/* Draws of the C library generator, as the reference's logdet test forms
 * its input: x[j*incx] = (rand() + 1) / RAND_MAX (test/lapack/cudlogdet.c:54). */
void oracle_logdet_input(double *x, size_t incx, size_t n);

/* Counter-based N(0,1) stream: value i of stream `seed` is a Box-Muller
 * transform of two SplitMix64 outputs.  Fully specified so any host produces
 * the same G. */
void oracle_gauss(uint64_t seed, uint64_t first, size_t count, double *out);
/* A = G G^T / n + alpha I with G n-by-n from oracle_gauss(seed, 0, n*n)
 * (column-major G), lower and upper triangles both filled. */
void oracle_spd(uint64_t seed, size_t n, double alpha, double *A, size_t lda);
/* Hermitian A = H H^H / n + alpha I, H = (N + iN)/sqrt(2) from
 * oracle_gauss(seed, 0, 2*n*n) (real parts then imaginary parts). */
void oracle_hpd(uint64_t seed, size_t n, double alpha, zcomplex *A, size_t lda);

/* Number of OpenMP threads the oracle runs with. */
int oracle_threads(void);

#endif
