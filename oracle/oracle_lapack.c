/*
 * CPU oracle, LAPACK part: blocked Cholesky, triangular inverse, triangular
 * square, SPD inverse and log-determinant, double and double complex.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Block sizes and loop structure
 * are the reference's CPU ones: POTRF and LAUUM use nb = 16 (upper) / 32
 * (lower), TRTRI uses nb = 64, and the unblocked kernels run when n < nb.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline size_t min_sz(size_t a, size_t b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ */
/* double                                                               */
/* ------------------------------------------------------------------ */

/* reference lapack/dpotrf.c:34-81.  Upper: dot-product row sweep; lower:
 * left-looking column updates.  A non-positive or NaN pivot stops with info
 * set to its 1-based index. */
static void dpotf2(char uplo, size_t n, double *A, size_t lda, long *info) {
  if (uplo == 'U') {
    for (size_t i = 0; i < n; i++) {
      double s = 0.0;
      for (size_t k = 0; k < i; k++) s += A[i * lda + k] * A[i * lda + k];
      double aii = A[i * lda + i] - s;
      if (aii <= 0.0 || isnan(aii)) {
        A[i * lda + i] = aii;
        *info = (long)i + 1;
        return;
      }
      aii = sqrt(aii);
      A[i * lda + i] = aii;
      for (size_t j = i + 1; j < n; j++) {
        double t = 0.0;
        for (size_t k = 0; k < i; k++) t += A[j * lda + k] * A[i * lda + k];
        A[j * lda + i] = (A[j * lda + i] - t) / aii;
      }
    }
  } else {
    for (size_t j = 0; j < n; j++) {
      for (size_t k = 0; k < j; k++) {
        const double t = A[k * lda + j];
        for (size_t i = j; i < n; i++) A[j * lda + i] -= t * A[k * lda + i];
      }
      double ajj = A[j * lda + j];
      if (ajj <= 0.0 || isnan(ajj)) {
        *info = (long)j + 1;
        return;
      }
      ajj = sqrt(ajj);
      A[j * lda + j] = ajj;
      for (size_t i = j + 1; i < n; i++) A[j * lda + i] /= ajj;
    }
  }
}

/* reference lapack/dpotrf.c:83-146 */
void oracle_dpotrf(char uplo, size_t n, double *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  if (n == 0) return;
  const size_t nb = (uplo == 'U') ? 16 : 32;
  if (nb > n) { dpotf2(uplo, n, A, lda, info); return; }
  for (size_t j = 0; j < n; j += nb) {
    const size_t jb = min_sz(nb, n - j);
    if (uplo == 'U') {
      oracle_dsyrk('U', 'T', jb, j, -1.0, &A[j * lda], lda, 1.0, &A[j * lda + j], lda);
      dpotf2('U', jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
      if (j + jb < n) {
        oracle_dgemm('T', 'N', jb, n - j - jb, j, -1.0, &A[j * lda], lda,
                     &A[(j + jb) * lda], lda, 1.0, &A[(j + jb) * lda + j], lda);
        oracle_dtrsm('L', 'U', 'T', 'N', jb, n - j - jb, 1.0, &A[j * lda + j], lda,
                     &A[(j + jb) * lda + j], lda);
      }
    } else {
      oracle_dsyrk('L', 'N', jb, j, -1.0, &A[j], lda, 1.0, &A[j * lda + j], lda);
      dpotf2('L', jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
      if (j + jb < n) {
        oracle_dgemm('N', 'T', n - j - jb, jb, j, -1.0, &A[j + jb], lda, &A[j], lda,
                     1.0, &A[j * lda + j + jb], lda);
        oracle_dtrsm('R', 'L', 'T', 'N', n - j - jb, jb, 1.0, &A[j * lda + j], lda,
                     &A[j * lda + j + jb], lda);
      }
    }
  }
}

/* reference lapack/dtrtri.c:43-96 */
static void dtrti2(char uplo, char diag, size_t n, double *A, size_t lda, long *info) {
  const int nounit = (diag == 'N');
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j++) {
      double ajj = -1.0;
      if (nounit) {
        if (A[j * lda + j] == 0.0) { *info = (long)j + 1; return; }
        A[j * lda + j] = 1.0 / A[j * lda + j];
        ajj = -A[j * lda + j];
      }
      /* column j above the diagonal <- inv(U(0:j,0:j)) * column, in place */
      for (size_t k = 0; k < j; k++) {
        const double t = A[j * lda + k];
        if (nounit) A[j * lda + k] *= A[k * lda + k];
        for (size_t i = 0; i < k; i++) A[j * lda + i] += t * A[k * lda + i];
      }
      for (size_t i = 0; i < j; i++) A[j * lda + i] *= ajj;
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
        const double t = A[j * lda + i];
        if (nounit) A[j * lda + i] *= A[i * lda + i];
        for (size_t k = i + 1; k < n; k++) A[j * lda + k] += t * A[i * lda + k];
      }
      for (size_t i = j + 1; i < n; i++) A[j * lda + i] *= ajj;
    }
  }
}

/* reference lapack/dtrtri.c:98-164 */
void oracle_dtrtri(char uplo, char diag, size_t n, double *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -5; return; }
  if (n == 0) return;
  const size_t nb = 64;
  if (n < nb) { dtrti2(uplo, diag, n, A, lda, info); return; }
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j += nb) {
      const size_t jb = min_sz(nb, n - j);
      oracle_dtrmm('L', 'U', 'N', diag, j, jb, 1.0, A, lda, &A[j * lda], lda);
      oracle_dtrsm('R', 'U', 'N', diag, j, jb, -1.0, &A[j * lda + j], lda, &A[j * lda], lda);
      dtrti2('U', diag, jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
    }
  } else {
    size_t j = (n + nb - 1) & ~(nb - 1);
    do {
      j -= nb;
      const size_t jb = min_sz(nb, n - j);
      if (j + jb < n) {
        oracle_dtrmm('L', 'L', 'N', diag, n - j - jb, jb, 1.0, &A[(j + jb) * lda + j + jb], lda,
                     &A[j * lda + j + jb], lda);
        oracle_dtrsm('R', 'L', 'N', diag, n - j - jb, jb, -1.0, &A[j * lda + j], lda,
                     &A[j * lda + j + jb], lda);
      }
      dtrti2('L', diag, jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
    } while (j > 0);
  }
}

/* reference lapack/dtrtri.c:226-299, out of place.  The reference passes lda
 * as the leading dimension of its own output block (dtrtri.c:257); this
 * restatement uses ldb there, the routine's evident intent. */
void oracle_dtrtri2(char uplo, char diag, size_t n, const double *A, size_t lda,
                    double *B, size_t ldb, long *info) {
  *info = 0;
  if (lda < n) { *info = -5; return; }
  if (ldb < n) { *info = -7; return; }
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < n; i++)
      if ((uplo == 'U') ? (i <= j) : (i >= j)) B[j * ldb + i] = A[j * lda + i];
  oracle_dtrtri(uplo, diag, n, B, ldb, info);
}

/* reference lapack/dlauum.c:43-69 */
static void dlauu2(char uplo, size_t n, double *A, size_t lda) {
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j++) {
      const double ajj = A[j * lda + j];
      for (size_t i = 0; i <= j; i++) A[j * lda + i] *= ajj;
      for (size_t k = j + 1; k < n; k++) {
        const double t = A[k * lda + j];
        for (size_t i = 0; i <= j; i++) A[j * lda + i] += t * A[k * lda + i];
      }
    }
  } else {
    for (size_t j = 0; j < n; j++) {
      for (size_t i = j; i < n; i++) {
        A[j * lda + i] *= A[i * lda + i];
        for (size_t k = i + 1; k < n; k++) A[j * lda + i] += A[i * lda + k] * A[j * lda + k];
      }
    }
  }
}

/* reference lapack/dlauum.c:71-129 */
void oracle_dlauum(char uplo, size_t n, double *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  if (n == 0) return;
  const size_t nb = (uplo == 'U') ? 16 : 32;
  if (nb > n) { dlauu2(uplo, n, A, lda); return; }
  for (size_t i = 0; i < n; i += nb) {
    const size_t ib = min_sz(nb, n - i);
    if (uplo == 'U') {
      oracle_dtrmm('R', 'U', 'T', 'N', i, ib, 1.0, &A[i * lda + i], lda, &A[i * lda], lda);
      dlauu2('U', ib, &A[i * lda + i], lda);
      if (i + ib < n) {
        oracle_dgemm('N', 'T', i, ib, n - i - ib, 1.0, &A[(i + ib) * lda], lda,
                     &A[(i + ib) * lda + i], lda, 1.0, &A[i * lda], lda);
        oracle_dsyrk('U', 'N', ib, n - i - ib, 1.0, &A[(i + ib) * lda + i], lda,
                     1.0, &A[i * lda + i], lda);
      }
    } else {
      oracle_dtrmm('L', 'L', 'T', 'N', ib, i, 1.0, &A[i * lda + i], lda, &A[i], lda);
      dlauu2('L', ib, &A[i * lda + i], lda);
      if (i + ib < n) {
        oracle_dgemm('T', 'N', ib, i, n - i - ib, 1.0, &A[i * lda + i + ib], lda,
                     &A[i + ib], lda, 1.0, &A[i], lda);
        oracle_dsyrk('L', 'T', ib, n - i - ib, 1.0, &A[i * lda + i + ib], lda,
                     1.0, &A[i * lda + i], lda);
      }
    }
  }
}

/* reference lapack/dpotri.c:4-23 */
void oracle_dpotri(char uplo, size_t n, double *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  if (n == 0) return;
  oracle_dtrtri(uplo, 'N', n, A, lda, info);
  if (*info != 0) return;
  oracle_dlauum(uplo, n, A, lda, info);
}

/* reference lapack/dlogdet.c:10-25 */
double oracle_dlogdet(const double *x, size_t incx, size_t n) {
  double total = 0.0;
  for (size_t i = 0; i < n; i++) total += log(x[i * incx]);
  return 2.0 * total;
}

/* ------------------------------------------------------------------ */
/* double complex                                                       */
/* ------------------------------------------------------------------ */

/* reference lapack/zpotrf.c:36-83 */
static void zpotf2(char uplo, size_t n, zcomplex *A, size_t lda, long *info) {
  if (uplo == 'U') {
    for (size_t i = 0; i < n; i++) {
      double s = 0.0;
      for (size_t k = 0; k < i; k++) s += creal(A[i * lda + k] * conj(A[i * lda + k]));
      double aii = creal(A[i * lda + i]) - s;
      if (aii <= 0.0 || isnan(aii)) {
        A[i * lda + i] = aii;
        *info = (long)i + 1;
        return;
      }
      aii = sqrt(aii);
      A[i * lda + i] = aii;
      for (size_t j = i + 1; j < n; j++) {
        zcomplex t = 0.0;
        for (size_t k = 0; k < i; k++) t += A[j * lda + k] * conj(A[i * lda + k]);
        A[j * lda + i] = (A[j * lda + i] - t) / aii;
      }
    }
  } else {
    for (size_t j = 0; j < n; j++) {
      for (size_t k = 0; k < j; k++) {
        const zcomplex t = conj(A[k * lda + j]);
        for (size_t i = j; i < n; i++) A[j * lda + i] -= t * A[k * lda + i];
      }
      double ajj = creal(A[j * lda + j]);
      if (ajj <= 0.0 || isnan(ajj)) {
        *info = (long)j + 1;
        return;
      }
      ajj = sqrt(ajj);
      A[j * lda + j] = ajj;
      for (size_t i = j + 1; i < n; i++) A[j * lda + i] /= ajj;
    }
  }
}

/* reference lapack/zpotrf.c:85-148 */
void oracle_zpotrf(char uplo, size_t n, zcomplex *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  if (n == 0) return;
  const size_t nb = (uplo == 'U') ? 16 : 32;
  if (n < nb) { zpotf2(uplo, n, A, lda, info); return; }
  for (size_t j = 0; j < n; j += nb) {
    const size_t jb = min_sz(nb, n - j);
    if (uplo == 'U') {
      oracle_zherk('U', 'C', jb, j, -1.0, &A[j * lda], lda, 1.0, &A[j * lda + j], lda);
      zpotf2('U', jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
      if (j + jb < n) {
        oracle_zgemm('C', 'N', jb, n - j - jb, j, -1.0, &A[j * lda], lda,
                     &A[(j + jb) * lda], lda, 1.0, &A[(j + jb) * lda + j], lda);
        oracle_ztrsm('L', 'U', 'C', 'N', jb, n - j - jb, 1.0, &A[j * lda + j], lda,
                     &A[(j + jb) * lda + j], lda);
      }
    } else {
      oracle_zherk('L', 'N', jb, j, -1.0, &A[j], lda, 1.0, &A[j * lda + j], lda);
      zpotf2('L', jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
      if (j + jb < n) {
        oracle_zgemm('N', 'C', n - j - jb, jb, j, -1.0, &A[j + jb], lda, &A[j], lda,
                     1.0, &A[j * lda + j + jb], lda);
        oracle_ztrsm('R', 'L', 'C', 'N', n - j - jb, jb, 1.0, &A[j * lda + j], lda,
                     &A[j * lda + j + jb], lda);
      }
    }
  }
}

/* in-place complex TRMM through the out-of-place form and a copy */
static void ztrmm_inplace(char side, char uplo, char transA, char diag, size_t m, size_t n,
                          zcomplex alpha, const zcomplex *A, size_t lda, zcomplex *B, size_t ldb) {
  if (m == 0 || n == 0) return;
  zcomplex *X = malloc(m * n * sizeof(zcomplex));
  oracle_ztrmm2(side, uplo, transA, diag, m, n, alpha, A, lda, B, ldb, X, m);
  for (size_t j = 0; j < n; j++) memcpy(&B[j * ldb], &X[j * m], m * sizeof(zcomplex));
  free(X);
}

/* reference lapack/ztrtri.c:43-96 */
static void ztrti2(char uplo, char diag, size_t n, zcomplex *A, size_t lda, long *info) {
  const int nounit = (diag == 'N');
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j++) {
      zcomplex ajj = -1.0;
      if (nounit) {
        if (A[j * lda + j] == 0.0) { *info = (long)j + 1; return; }
        A[j * lda + j] = 1.0 / A[j * lda + j];
        ajj = -A[j * lda + j];
      }
      for (size_t k = 0; k < j; k++) {
        const zcomplex t = A[j * lda + k];
        if (nounit) A[j * lda + k] *= A[k * lda + k];
        for (size_t i = 0; i < k; i++) A[j * lda + i] += t * A[k * lda + i];
      }
      for (size_t i = 0; i < j; i++) A[j * lda + i] *= ajj;
    }
  } else {
    for (size_t j = n; j-- > 0;) {
      zcomplex ajj = -1.0;
      if (nounit) {
        if (A[j * lda + j] == 0.0) { *info = (long)j + 1; return; }
        A[j * lda + j] = 1.0 / A[j * lda + j];
        ajj = -A[j * lda + j];
      }
      for (size_t i = n - 1; i > j; i--) {
        const zcomplex t = A[j * lda + i];
        if (nounit) A[j * lda + i] *= A[i * lda + i];
        for (size_t k = i + 1; k < n; k++) A[j * lda + k] += t * A[i * lda + k];
      }
      for (size_t i = j + 1; i < n; i++) A[j * lda + i] *= ajj;
    }
  }
}

/* reference lapack/ztrtri.c:98-164 */
void oracle_ztrtri(char uplo, char diag, size_t n, zcomplex *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -5; return; }
  if (n == 0) return;
  const size_t nb = 64;
  if (n < nb) { ztrti2(uplo, diag, n, A, lda, info); return; }
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j += nb) {
      const size_t jb = min_sz(nb, n - j);
      ztrmm_inplace('L', 'U', 'N', diag, j, jb, 1.0, A, lda, &A[j * lda], lda);
      oracle_ztrsm('R', 'U', 'N', diag, j, jb, -1.0, &A[j * lda + j], lda, &A[j * lda], lda);
      ztrti2('U', diag, jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
    }
  } else {
    size_t j = (n + nb - 1) & ~(nb - 1);
    do {
      j -= nb;
      const size_t jb = min_sz(nb, n - j);
      if (j + jb < n) {
        ztrmm_inplace('L', 'L', 'N', diag, n - j - jb, jb, 1.0, &A[(j + jb) * lda + j + jb], lda,
                      &A[j * lda + j + jb], lda);
        oracle_ztrsm('R', 'L', 'N', diag, n - j - jb, jb, -1.0, &A[j * lda + j], lda,
                     &A[j * lda + j + jb], lda);
      }
      ztrti2('L', diag, jb, &A[j * lda + j], lda, info);
      if (*info != 0) { *info += (long)j; return; }
    } while (j > 0);
  }
}

/* reference lapack/zlauum.c:43-69 */
static void zlauu2(char uplo, size_t n, zcomplex *A, size_t lda) {
  if (uplo == 'U') {
    for (size_t j = 0; j < n; j++) {
      const zcomplex ajj = conj(A[j * lda + j]);
      for (size_t i = 0; i <= j; i++) A[j * lda + i] *= ajj;
      for (size_t k = j + 1; k < n; k++) {
        const zcomplex t = conj(A[k * lda + j]);
        for (size_t i = 0; i <= j; i++) A[j * lda + i] += t * A[k * lda + i];
      }
    }
  } else {
    for (size_t j = 0; j < n; j++) {
      for (size_t i = j; i < n; i++) {
        A[j * lda + i] *= conj(A[i * lda + i]);
        for (size_t k = i + 1; k < n; k++) A[j * lda + i] += conj(A[i * lda + k]) * A[j * lda + k];
      }
    }
  }
}

/* reference lapack/zlauum.c:71-129 */
void oracle_zlauum(char uplo, size_t n, zcomplex *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  if (n == 0) return;
  const size_t nb = (uplo == 'U') ? 16 : 32;
  if (nb > n) { zlauu2(uplo, n, A, lda); return; }
  for (size_t i = 0; i < n; i += nb) {
    const size_t ib = min_sz(nb, n - i);
    if (uplo == 'U') {
      ztrmm_inplace('R', 'U', 'C', 'N', i, ib, 1.0, &A[i * lda + i], lda, &A[i * lda], lda);
      zlauu2('U', ib, &A[i * lda + i], lda);
      if (i + ib < n) {
        oracle_zgemm('N', 'C', i, ib, n - i - ib, 1.0, &A[(i + ib) * lda], lda,
                     &A[(i + ib) * lda + i], lda, 1.0, &A[i * lda], lda);
        oracle_zherk('U', 'N', ib, n - i - ib, 1.0, &A[(i + ib) * lda + i], lda,
                     1.0, &A[i * lda + i], lda);
      }
    } else {
      ztrmm_inplace('L', 'L', 'C', 'N', ib, i, 1.0, &A[i * lda + i], lda, &A[i], lda);
      zlauu2('L', ib, &A[i * lda + i], lda);
      if (i + ib < n) {
        oracle_zgemm('C', 'N', ib, i, n - i - ib, 1.0, &A[i * lda + i + ib], lda,
                     &A[i + ib], lda, 1.0, &A[i], lda);
        oracle_zherk('L', 'C', ib, n - i - ib, 1.0, &A[i * lda + i + ib], lda,
                     1.0, &A[i * lda + i], lda);
      }
    }
  }
}

/* reference lapack/zpotri.c:4 */
void oracle_zpotri(char uplo, size_t n, zcomplex *A, size_t lda, long *info) {
  *info = 0;
  if (lda < n) { *info = -4; return; }
  if (n == 0) return;
  oracle_ztrtri(uplo, 'N', n, A, lda, info);
  if (*info != 0) return;
  oracle_zlauum(uplo, n, A, lda, info);
}

/* reference lapack/zlogdet.c:11-26, logs taken in single precision as there */
double oracle_zlogdet_ref(const zcomplex *x, size_t incx, size_t n) {
  double total = 0.0;
  for (size_t i = 0; i < n; i++) total += logf((float)creal(x[i * incx]));
  return 2.0 * total;
}

double oracle_zlogdet(const zcomplex *x, size_t incx, size_t n) {
  double total = 0.0;
  for (size_t i = 0; i < n; i++) total += log(creal(x[i * incx]));
  return 2.0 * total;
}
