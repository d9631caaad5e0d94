/*
 * CPU oracle, BLAS-3 part: restatement of the reference's OpenMP CPU BLAS.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The loops follow the reference's
 * column-oriented algorithms, which are the netlib reference BLAS loop orders
 * with `#pragma omp parallel for` over independent columns (or rows).
 */
#include "oracle.h"

static const double zero = 0.0;
static const double one = 1.0;

/* reference blas/dgemm.c:34-146 */
void oracle_dgemm(char transA, char transB, size_t m, size_t n, size_t k,
                  double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                  double beta, double *C, size_t ldc) {
  if (m == 0 || n == 0 || ((alpha == zero || k == 0) && beta == one))
    return;
  if (alpha == zero) {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < m; i++)
        C[j * ldc + i] = (beta == zero) ? zero : beta * C[j * ldc + i];
    return;
  }
  const int nA = (transA == 'N'), nB = (transB == 'N');
  if (nA) {
    /* C(:,j) accumulates alpha * op(B)(l,j) * A(:,l), axpy form */
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      double *c = &C[j * ldc];
      if (beta == zero)
        for (size_t i = 0; i < m; i++) c[i] = zero;
      else if (beta != one)
        for (size_t i = 0; i < m; i++) c[i] *= beta;
      for (size_t l = 0; l < k; l++) {
        const double b = nB ? B[j * ldb + l] : B[l * ldb + j];
        if (b != zero) {
          const double t = alpha * b;
          const double *a = &A[l * lda];
          for (size_t i = 0; i < m; i++) c[i] += t * a[i];
        }
      }
    }
  } else {
    /* dot-product form over the contiguous columns of A */
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      for (size_t i = 0; i < m; i++) {
        double t = zero;
        const double *a = &A[i * lda];
        if (nB) {
          const double *b = &B[j * ldb];
          for (size_t l = 0; l < k; l++) t += a[l] * b[l];
        } else {
          for (size_t l = 0; l < k; l++) t += a[l] * B[l * ldb + j];
        }
        C[j * ldc + i] = (beta == zero) ? alpha * t : alpha * t + beta * C[j * ldc + i];
      }
    }
  }
}

/* reference blas/dsyrk.c:34-161 */
void oracle_dsyrk(char uplo, char trans, size_t n, size_t k,
                  double alpha, const double *A, size_t lda, double beta, double *C, size_t ldc) {
  if (n == 0 || ((alpha == zero || k == 0) && beta == one))
    return;
  const int up = (uplo == 'U');
  if (alpha == zero) {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      const size_t i0 = up ? 0 : j, i1 = up ? j + 1 : n;
      for (size_t i = i0; i < i1; i++)
        C[j * ldc + i] = (beta == zero) ? zero : beta * C[j * ldc + i];
    }
    return;
  }
  if (trans == 'N') {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      const size_t i0 = up ? 0 : j, i1 = up ? j + 1 : n;
      double *c = &C[j * ldc];
      if (beta == zero)
        for (size_t i = i0; i < i1; i++) c[i] = zero;
      else if (beta != one)
        for (size_t i = i0; i < i1; i++) c[i] *= beta;
      for (size_t l = 0; l < k; l++) {
        if (A[l * lda + j] != zero) {
          const double t = alpha * A[l * lda + j];
          for (size_t i = i0; i < i1; i++) c[i] += t * A[l * lda + i];
        }
      }
    }
  } else {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      const size_t i0 = up ? 0 : j, i1 = up ? j + 1 : n;
      for (size_t i = i0; i < i1; i++) {
        double t = zero;
        for (size_t l = 0; l < k; l++) t += A[i * lda + l] * A[j * lda + l];
        C[j * ldc + i] = (beta == zero) ? alpha * t : alpha * t + beta * C[j * ldc + i];
      }
    }
  }
}

/* reference blas/dtrsm.c:34-216 */
void oracle_dtrsm(char side, char uplo, char transA, char diag, size_t m, size_t n,
                  double alpha, const double *A, size_t lda, double *B, size_t ldb) {
  if (m == 0 || n == 0) return;
  const int nounit = (diag == 'N');
  if (alpha == zero) {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < m; i++) B[j * ldb + i] = zero;
    return;
  }
  if (side == 'L') {
    if (transA == 'N') {
      /* B := alpha inv(A) B, one column of B per iteration */
#pragma omp parallel for schedule(static)
      for (size_t j = 0; j < n; j++) {
        double *b = &B[j * ldb];
        if (alpha != one)
          for (size_t i = 0; i < m; i++) b[i] *= alpha;
        if (uplo == 'U') {
          for (size_t kk = m; kk-- > 0;) {
            if (b[kk] != zero) {
              if (nounit) b[kk] /= A[kk * lda + kk];
              const double t = b[kk];
              for (size_t i = 0; i < kk; i++) b[i] -= t * A[kk * lda + i];
            }
          }
        } else {
          for (size_t kk = 0; kk < m; kk++) {
            if (b[kk] != zero) {
              if (nounit) b[kk] /= A[kk * lda + kk];
              const double t = b[kk];
              for (size_t i = kk + 1; i < m; i++) b[i] -= t * A[kk * lda + i];
            }
          }
        }
      }
    } else {
      /* B := alpha inv(A^T) B */
#pragma omp parallel for schedule(static)
      for (size_t j = 0; j < n; j++) {
        double *b = &B[j * ldb];
        if (uplo == 'U') {
          for (size_t i = 0; i < m; i++) {
            double t = alpha * b[i];
            for (size_t kk = 0; kk < i; kk++) t -= A[i * lda + kk] * b[kk];
            if (nounit) t /= A[i * lda + i];
            b[i] = t;
          }
        } else {
          for (size_t i = m; i-- > 0;) {
            double t = alpha * b[i];
            for (size_t kk = i + 1; kk < m; kk++) t -= A[i * lda + kk] * b[kk];
            if (nounit) t /= A[i * lda + i];
            b[i] = t;
          }
        }
      }
    }
  } else {
    /* Right side: columns of B depend on each other, rows are independent;
     * the reference runs these loops serially (blas/dtrsm.c:126-212). */
    if (transA == 'N') {
      if (uplo == 'U') {
        for (size_t j = 0; j < n; j++) {
          double *bj = &B[j * ldb];
          if (alpha != one)
            for (size_t i = 0; i < m; i++) bj[i] *= alpha;
          for (size_t kk = 0; kk < j; kk++) {
            const double a = A[j * lda + kk];
            if (a != zero)
              for (size_t i = 0; i < m; i++) bj[i] -= a * B[kk * ldb + i];
          }
          if (nounit) {
            const double t = one / A[j * lda + j];
            for (size_t i = 0; i < m; i++) bj[i] *= t;
          }
        }
      } else {
        for (size_t j = n; j-- > 0;) {
          double *bj = &B[j * ldb];
          if (alpha != one)
            for (size_t i = 0; i < m; i++) bj[i] *= alpha;
          for (size_t kk = j + 1; kk < n; kk++) {
            const double a = A[j * lda + kk];
            if (a != zero)
              for (size_t i = 0; i < m; i++) bj[i] -= a * B[kk * ldb + i];
          }
          if (nounit) {
            const double t = one / A[j * lda + j];
            for (size_t i = 0; i < m; i++) bj[i] *= t;
          }
        }
      }
    } else {
      if (uplo == 'U') {
        for (size_t kk = n; kk-- > 0;) {
          double *bk = &B[kk * ldb];
          if (nounit) {
            const double t = one / A[kk * lda + kk];
            for (size_t i = 0; i < m; i++) bk[i] *= t;
          }
          for (size_t j = 0; j < kk; j++) {
            const double a = A[kk * lda + j];
            if (a != zero)
              for (size_t i = 0; i < m; i++) B[j * ldb + i] -= a * bk[i];
          }
          if (alpha != one)
            for (size_t i = 0; i < m; i++) bk[i] *= alpha;
        }
      } else {
        for (size_t kk = 0; kk < n; kk++) {
          double *bk = &B[kk * ldb];
          if (nounit) {
            const double t = one / A[kk * lda + kk];
            for (size_t i = 0; i < m; i++) bk[i] *= t;
          }
          for (size_t j = kk + 1; j < n; j++) {
            const double a = A[kk * lda + j];
            if (a != zero)
              for (size_t i = 0; i < m; i++) B[j * ldb + i] -= a * bk[i];
          }
          if (alpha != one)
            for (size_t i = 0; i < m; i++) bk[i] *= alpha;
        }
      }
    }
  }
}

/* reference blas/dtrmm.c:44-204 (in place) */
void oracle_dtrmm(char side, char uplo, char transA, char diag, size_t m, size_t n,
                  double alpha, const double *A, size_t lda, double *B, size_t ldb) {
  if (m == 0 || n == 0) return;
  const int nounit = (diag == 'N');
  if (alpha == zero) {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < m; i++) B[j * ldb + i] = zero;
    return;
  }
  if (side == 'L') {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      double *b = &B[j * ldb];
      if (transA == 'N') {
        if (uplo == 'U') {
          for (size_t kk = 0; kk < m; kk++) {
            if (b[kk] != zero) {
              double t = alpha * b[kk];
              for (size_t i = 0; i < kk; i++) b[i] += t * A[kk * lda + i];
              if (nounit) t *= A[kk * lda + kk];
              b[kk] = t;
            }
          }
        } else {
          for (size_t kk = m; kk-- > 0;) {
            if (b[kk] != zero) {
              const double t = alpha * b[kk];
              b[kk] = nounit ? t * A[kk * lda + kk] : t;
              for (size_t i = kk + 1; i < m; i++) b[i] += t * A[kk * lda + i];
            }
          }
        }
      } else {
        if (uplo == 'U') {
          for (size_t i = m; i-- > 0;) {
            double t = b[i];
            if (nounit) t *= A[i * lda + i];
            for (size_t kk = 0; kk < i; kk++) t += A[i * lda + kk] * b[kk];
            b[i] = alpha * t;
          }
        } else {
          for (size_t i = 0; i < m; i++) {
            double t = b[i];
            if (nounit) t *= A[i * lda + i];
            for (size_t kk = i + 1; kk < m; kk++) t += A[i * lda + kk] * b[kk];
            b[i] = alpha * t;
          }
        }
      }
    }
  } else {
    if (transA == 'N') {
      if (uplo == 'U') {
        for (size_t j = n; j-- > 0;) {
          double t = alpha;
          if (nounit) t *= A[j * lda + j];
          for (size_t i = 0; i < m; i++) B[j * ldb + i] *= t;
          for (size_t kk = 0; kk < j; kk++) {
            if (A[j * lda + kk] != zero) {
              const double s = alpha * A[j * lda + kk];
              for (size_t i = 0; i < m; i++) B[j * ldb + i] += s * B[kk * ldb + i];
            }
          }
        }
      } else {
        for (size_t j = 0; j < n; j++) {
          double t = alpha;
          if (nounit) t *= A[j * lda + j];
          for (size_t i = 0; i < m; i++) B[j * ldb + i] *= t;
          for (size_t kk = j + 1; kk < n; kk++) {
            if (A[j * lda + kk] != zero) {
              const double s = alpha * A[j * lda + kk];
              for (size_t i = 0; i < m; i++) B[j * ldb + i] += s * B[kk * ldb + i];
            }
          }
        }
      }
    } else {
      if (uplo == 'U') {
        for (size_t kk = 0; kk < n; kk++) {
          for (size_t j = 0; j < kk; j++) {
            if (A[kk * lda + j] != zero) {
              const double s = alpha * A[kk * lda + j];
              for (size_t i = 0; i < m; i++) B[j * ldb + i] += s * B[kk * ldb + i];
            }
          }
          double t = alpha;
          if (nounit) t *= A[kk * lda + kk];
          if (t != one)
            for (size_t i = 0; i < m; i++) B[kk * ldb + i] *= t;
        }
      } else {
        for (size_t kk = n; kk-- > 0;) {
          for (size_t j = kk + 1; j < n; j++) {
            if (A[kk * lda + j] != zero) {
              const double s = alpha * A[kk * lda + j];
              for (size_t i = 0; i < m; i++) B[j * ldb + i] += s * B[kk * ldb + i];
            }
          }
          double t = alpha;
          if (nounit) t *= A[kk * lda + kk];
          if (t != one)
            for (size_t i = 0; i < m; i++) B[kk * ldb + i] *= t;
        }
      }
    }
  }
}

/* reference blas/dtrmm.c:206-372 (out of place).  The reference's Right/Trans
 * branch (dtrmm.c:336-368) skips the copy of column k when alpha*A(k,k) == 1,
 * leaving X(:,k) unset; this restatement always forms X = alpha op(A) B, the
 * documented meaning of the routine, by copying B to X and multiplying in
 * place. */
void oracle_dtrmm2(char side, char uplo, char transA, char diag, size_t m, size_t n,
                   double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                   double *X, size_t ldx) {
  if (m == 0 || n == 0) return;
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < m; i++) X[j * ldx + i] = B[j * ldb + i];
  oracle_dtrmm(side, uplo, transA, diag, m, n, alpha, A, lda, X, ldx);
}
