/*
 * CPU oracle: unblocked textbook routines and test-matrix generators.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---- unblocked textbook forms ---- */

/* reference test/lapack/ref/dpotrf_ref.c:1-51: dot-product Cholesky */
void oracle_dpotrf_unblocked(char uplo, size_t n, double *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  for (size_t j = 0; j < n; j++) {
    if (uplo == 'U') {
      for (size_t i = 0; i <= j; i++) {
        double t = A[j * lda + i];
        for (size_t k = 0; k < i; k++) t -= A[j * lda + k] * A[i * lda + k];
        if (i == j) {
          if (t <= 0.0 || isnan(t)) { A[j * lda + j] = t; *info = (long)j + 1; return; }
          A[j * lda + j] = sqrt(t);
        } else {
          A[j * lda + i] = t / A[i * lda + i];
        }
      }
    } else {
      for (size_t i = j; i < n; i++) {
        double t = A[j * lda + i];
        for (size_t k = 0; k < j; k++) t -= A[k * lda + j] * A[k * lda + i];
        if (i == j) {
          if (t <= 0.0 || isnan(t)) { A[j * lda + j] = t; *info = (long)j + 1; return; }
          A[j * lda + j] = sqrt(t);
        } else {
          A[j * lda + i] = t / A[j * lda + j];
        }
      }
    }
  }
}

/* Hermitian form of the same dot-product Cholesky */
void oracle_zpotrf_unblocked(char uplo, size_t n, zcomplex *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  for (size_t j = 0; j < n; j++) {
    if (uplo == 'U') {
      for (size_t i = 0; i <= j; i++) {
        zcomplex t = A[j * lda + i];
        for (size_t k = 0; k < i; k++) t -= conj(A[i * lda + k]) * A[j * lda + k];
        if (i == j) {
          const double d = creal(t);
          if (d <= 0.0 || isnan(d)) { A[j * lda + j] = d; *info = (long)j + 1; return; }
          A[j * lda + j] = sqrt(d);
        } else {
          A[j * lda + i] = t / creal(A[i * lda + i]);
        }
      }
    } else {
      for (size_t i = j; i < n; i++) {
        zcomplex t = A[j * lda + i];
        for (size_t k = 0; k < j; k++) t -= A[k * lda + i] * conj(A[k * lda + j]);
        if (i == j) {
          const double d = creal(t);
          if (d <= 0.0 || isnan(d)) { A[j * lda + j] = d; *info = (long)j + 1; return; }
          A[j * lda + j] = sqrt(d);
        } else {
          A[j * lda + i] = t / creal(A[j * lda + j]);
        }
      }
    }
  }
}

/* reference test/lapack/ref/dtrtri_ref.c:1-61 */
void oracle_dtrtri_unblocked(char uplo, char diag, size_t n, double *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -5; return; }
  const int nounit = (diag == 'N');
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j++) {
      double ajj = -1.0;
      if (nounit) {
        if (A[j * lda + j] == 0.0) { *info = (long)j + 1; return; }
        A[j * lda + j] = 1.0 / A[j * lda + j];
        ajj = -A[j * lda + j];
      }
      for (size_t i = 0; i < j; i++) {
        double t = A[j * lda + i];
        if (nounit) t *= A[i * lda + i];
        for (size_t k = i + 1; k < j; k++) t += A[k * lda + i] * A[j * lda + k];
        A[j * lda + i] = t * ajj;
      }
    }
  } else {
    for (size_t j = n; j-- > 0;) {
      double ajj = -1.0;
      if (nounit) {
        if (A[j * lda + j] == 0.0) { *info = (long)j + 1; return; }
        A[j * lda + j] = 1.0 / A[j * lda + j];
        ajj = -A[j * lda + j];
      }
      for (size_t i = n - 1; i > j; i--) {
        double t = A[j * lda + i];
        if (nounit) t *= A[i * lda + i];
        for (size_t k = j + 1; k < i; k++) t += A[k * lda + i] * A[j * lda + k];
        A[j * lda + i] = t * ajj;
      }
    }
  }
}

/* reference test/lapack/ref/dlauum_ref.c:1-32 */
void oracle_dlauum_unblocked(char uplo, size_t n, double *A, size_t lda) {
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j++) {
      const double ajj = A[j * lda + j];
      for (size_t i = 0; i <= j; i++) {
        A[j * lda + i] *= ajj;
        for (size_t k = j + 1; k < n; k++) A[j * lda + i] += A[k * lda + i] * A[k * lda + j];
      }
    }
  } else {
    for (size_t i = 0; i < n; i++) {
      const double aii = A[i * lda + i];
      for (size_t j = 0; j <= i; j++) {
        A[j * lda + i] *= aii;
        for (size_t k = i + 1; k < n; k++) A[j * lda + i] += A[i * lda + k] * A[j * lda + k];
      }
    }
  }
}

/* ---- generators ---- */

void oracle_srand(unsigned int seed) { srand(seed); }

/* reference test/lapack/cudlogdet.c:54 */
void oracle_logdet_input(double *x, size_t incx, size_t n) {
  for (size_t j = 0; j < n; j++) x[j * incx] = ((double)rand() + 1.0) / (double)RAND_MAX;
}

/* reference test/lapack/util/dlatmc.c:11-70: A = D - u w^T - w u^T, which is
 * H D H for the Householder reflector H = I - t u u^T, t = 2 / u^T u, so the
 * eigenvalues of A are diag(D) (1, c and n-2 draws from [1, c]). */
int oracle_dlatmc(size_t n, double c, double *A, size_t lda) {
  if (n < 2) return -1;
  if (c < 1.0) return -2;
  if (lda < n) return -4;
  double *u = malloc(3 * n * sizeof(double));
  if (u == NULL) return 1;
  double *v = u + n, *w = v + n;
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < n; i++) A[j * lda + i] = 0.0;
  A[0] = 1.0;
  A[lda + 1] = c;
  for (size_t j = 2; j < n; j++)
    A[j * lda + j] = ((double)rand() / (double)RAND_MAX) * (c - 1.0) + 1.0;
  double t = 0.0, s = 0.0;
  for (size_t j = 0; j < n; j++) {
    u[j] = (double)rand() / (double)RAND_MAX;
    v[j] = A[j * lda + j] * u[j];
    t += u[j] * u[j];
    s += u[j] * v[j];
  }
  t = 2.0 / t;
  s = t * t * s / 2.0;
  for (size_t j = 0; j < n; j++) w[j] = t * v[j] - s * u[j];
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < n; i++) A[j * lda + i] -= u[i] * w[j] + w[i] * u[j];
  free(u);
  return 0;
}

/* reference test/lapack/util/zlatmc.c:11-70.  The reference draws the real
 * and imaginary part of u[j] inside one expression, whose evaluation order C
 * leaves unspecified; this restatement draws the real part first. */
int oracle_zlatmc(size_t n, double c, zcomplex *A, size_t lda) {
  if (n < 2) return -1;
  if (c < 1.0) return -2;
  if (lda < n) return -4;
  zcomplex *u = malloc(3 * n * sizeof(zcomplex));
  if (u == NULL) return 1;
  zcomplex *v = u + n, *w = v + n;
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < n; i++) A[j * lda + i] = 0.0;
  A[0] = 1.0;
  A[lda + 1] = c;
  for (size_t j = 2; j < n; j++)
    A[j * lda + j] = ((double)rand() / (double)RAND_MAX) * (c - 1.0) + 1.0;
  double t = 0.0;
  zcomplex s = 0.0;
  for (size_t j = 0; j < n; j++) {
    const double re = (double)rand() / (double)RAND_MAX;
    const double im = (double)rand() / (double)RAND_MAX;
    u[j] = re + im * I;
    v[j] = A[j * lda + j] * u[j];
    t += creal(conj(u[j]) * u[j]);
    s += conj(u[j]) * v[j];
  }
  t = 2.0 / t;
  s = t * t * s / 2.0;
  for (size_t j = 0; j < n; j++) w[j] = t * v[j] - s * u[j];
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < n; i++) A[j * lda + i] -= u[i] * conj(w[j]) + w[i] * conj(u[j]);
  free(u);
  return 0;
}

/* SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014). */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* Draw c of stream `seed`: the SplitMix64 output after c + 1 increments of
 * the golden-ratio constant. */
static inline uint64_t draw64(uint64_t seed, uint64_t c) {
  return mix64(seed + (c + 1) * 0x9e3779b97f4a7c15ULL);
}

/* Normal i of stream `seed`: u1 = ((d(2i) >> 11) + 1) 2^-53 in (0, 1],
 * u2 = (d(2i+1) >> 11) 2^-53 in [0, 1), g = sqrt(-2 ln u1) cos(2 pi u2). */
void oracle_gauss(uint64_t seed, uint64_t first, size_t count, double *out) {
  const double scale = 1.0 / 9007199254740992.0; /* 2^-53 */
  const double two_pi = 6.283185307179586476925286766559;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < count; i++) {
    const uint64_t idx = first + i;
    const double u1 = (double)((draw64(seed, 2 * idx) >> 11) + 1) * scale;
    const double u2 = (double)(draw64(seed, 2 * idx + 1) >> 11) * scale;
    out[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
  }
}

void oracle_spd(uint64_t seed, size_t n, double alpha, double *A, size_t lda) {
  double *G = malloc(n * n * sizeof(double));
  oracle_gauss(seed, 0, n * n, G);
  const double inv_n = 1.0 / (double)n;
  /* A(i, j) = sum_k G(i, k) G(j, k) / n + alpha delta_ij, lower triangle
   * computed and mirrored so the matrix is exactly symmetric. */
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t j = 0; j < n; j++) {
    for (size_t i = j; i < n; i++) {
      double s = 0.0;
      for (size_t k = 0; k < n; k++) s += G[k * n + i] * G[k * n + j];
      s *= inv_n;
      if (i == j) s += alpha;
      A[j * lda + i] = s;
      A[i * lda + j] = s;
    }
  }
  free(G);
}

void oracle_hpd(uint64_t seed, size_t n, double alpha, zcomplex *A, size_t lda) {
  double *G = malloc(2 * n * n * sizeof(double));
  oracle_gauss(seed, 0, 2 * n * n, G);
  const double r = 1.0 / sqrt(2.0);
  const double inv_n = 1.0 / (double)n;
  /* H(i, k) = (G[k n + i] + i G[n^2 + k n + i]) / sqrt(2) */
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t j = 0; j < n; j++) {
    for (size_t i = j; i < n; i++) {
      zcomplex s = 0.0;
      for (size_t k = 0; k < n; k++) {
        const zcomplex hi = r * (G[k * n + i] + G[n * n + k * n + i] * I);
        const zcomplex hj = r * (G[k * n + j] + G[n * n + k * n + j] * I);
        s += hi * conj(hj);
      }
      s *= inv_n;
      if (i == j) s = creal(s) + alpha;
      A[j * lda + i] = s;
      A[i * lda + j] = conj(s);
    }
  }
  free(G);
}
