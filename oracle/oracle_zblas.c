/*
 * CPU oracle, double-complex BLAS-3 part.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates the reference's complex
 * CPU BLAS (blas/zgemm.c:34, zherk.c:35-190, ztrsm.c:34-246, ztrmm.c:44-420):
 * netlib loop orders with OpenMP over independent columns, op(A) in
 * {A, A^T, A^H} selected by 'N', 'T', 'C'.
 */
#include "oracle.h"

static const zcomplex czero = 0.0;
static const zcomplex cone = 1.0;

/* element (i, j) of op(X) for X column-major with leading dimension ld */
static inline zcomplex opel(const zcomplex *X, size_t ld, char trans, size_t i, size_t j) {
  if (trans == 'N') return X[j * ld + i];
  if (trans == 'T') return X[i * ld + j];
  return conj(X[i * ld + j]);
}

/* reference blas/zgemm.c:34-224 */
void oracle_zgemm(char transA, char transB, size_t m, size_t n, size_t k,
                  zcomplex alpha, const zcomplex *A, size_t lda, const zcomplex *B, size_t ldb,
                  zcomplex beta, zcomplex *C, size_t ldc) {
  if (m == 0 || n == 0 || ((alpha == czero || k == 0) && beta == cone))
    return;
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < n; j++) {
    zcomplex *c = &C[j * ldc];
    if (alpha == czero) {
      for (size_t i = 0; i < m; i++) c[i] = (beta == czero) ? czero : beta * c[i];
      continue;
    }
    if (transA == 'N') {
      if (beta == czero)
        for (size_t i = 0; i < m; i++) c[i] = czero;
      else if (beta != cone)
        for (size_t i = 0; i < m; i++) c[i] *= beta;
      for (size_t l = 0; l < k; l++) {
        const zcomplex b = opel(B, ldb, transB, l, j);
        if (b != czero) {
          const zcomplex t = alpha * b;
          const zcomplex *a = &A[l * lda];
          for (size_t i = 0; i < m; i++) c[i] += t * a[i];
        }
      }
    } else {
      for (size_t i = 0; i < m; i++) {
        zcomplex t = czero;
        for (size_t l = 0; l < k; l++)
          t += opel(A, lda, transA, i, l) * opel(B, ldb, transB, l, j);
        c[i] = (beta == czero) ? alpha * t : alpha * t + beta * c[i];
      }
    }
  }
}

/* reference blas/zherk.c:35-188: C = alpha op(A) op(A)^H + beta C, real alpha
 * and beta, the diagonal kept real. */
void oracle_zherk(char uplo, char trans, size_t n, size_t k,
                  double alpha, const zcomplex *A, size_t lda, double beta, zcomplex *C, size_t ldc) {
  if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0))
    return;
  const int up = (uplo == 'U');
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < n; j++) {
    const size_t i0 = up ? 0 : j, i1 = up ? j + 1 : n;
    for (size_t i = i0; i < i1; i++) {
      zcomplex t = czero;
      if (alpha != 0.0) {
        if (trans == 'N')
          for (size_t l = 0; l < k; l++) t += A[l * lda + i] * conj(A[l * lda + j]);
        else
          for (size_t l = 0; l < k; l++) t += conj(A[i * lda + l]) * A[j * lda + l];
      }
      zcomplex c = (beta == 0.0) ? czero : beta * C[j * ldc + i];
      if (i == j) c = creal(c) + alpha * creal(t);
      else c += alpha * t;
      C[j * ldc + i] = c;
    }
  }
}

/* reference blas/ztrsm.c:34-246 */
void oracle_ztrsm(char side, char uplo, char transA, char diag, size_t m, size_t n,
                  zcomplex alpha, const zcomplex *A, size_t lda, zcomplex *B, size_t ldb) {
  if (m == 0 || n == 0) return;
  const int nounit = (diag == 'N');
  if (alpha == czero) {
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < m; i++) B[j * ldb + i] = czero;
    return;
  }
  /* op(A) is lower-like when (uplo == L) and no transpose, or (uplo == U)
   * and a transpose: then the solve runs forwards. */
  const int lowerlike = ((uplo == 'L') == (transA == 'N'));
  if (side == 'L') {
    /* op(A) X = alpha B: each column of B independently */
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
      zcomplex *b = &B[j * ldb];
      if (lowerlike) {
        for (size_t i = 0; i < m; i++) {
          zcomplex t = alpha * b[i];
          for (size_t kk = 0; kk < i; kk++) t -= opel(A, lda, transA, i, kk) * b[kk];
          if (nounit) t /= opel(A, lda, transA, i, i);
          b[i] = t;
        }
      } else {
        for (size_t i = m; i-- > 0;) {
          zcomplex t = alpha * b[i];
          for (size_t kk = i + 1; kk < m; kk++) t -= opel(A, lda, transA, i, kk) * b[kk];
          if (nounit) t /= opel(A, lda, transA, i, i);
          b[i] = t;
        }
      }
    }
  } else {
    /* X op(A) = alpha B: each row of B independently */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < m; i++) {
      if (lowerlike) {
        for (size_t j = n; j-- > 0;) {
          zcomplex t = alpha * B[j * ldb + i];
          for (size_t kk = j + 1; kk < n; kk++) t -= B[kk * ldb + i] * opel(A, lda, transA, kk, j);
          if (nounit) t /= opel(A, lda, transA, j, j);
          B[j * ldb + i] = t;
        }
      } else {
        for (size_t j = 0; j < n; j++) {
          zcomplex t = alpha * B[j * ldb + i];
          for (size_t kk = 0; kk < j; kk++) t -= B[kk * ldb + i] * opel(A, lda, transA, kk, j);
          if (nounit) t /= opel(A, lda, transA, j, j);
          B[j * ldb + i] = t;
        }
      }
    }
  }
}

/* reference blas/ztrmm.c:206-420 (out of place): X = alpha op(A) B or
 * alpha B op(A) with op(A) triangular; entries of A outside the triangle are
 * never read and a unit diagonal is implicit. */
void oracle_ztrmm2(char side, char uplo, char transA, char diag, size_t m, size_t n,
                   zcomplex alpha, const zcomplex *A, size_t lda, const zcomplex *B, size_t ldb,
                   zcomplex *X, size_t ldx) {
  if (m == 0 || n == 0) return;
  const int nounit = (diag == 'N');
  const int lowerlike = ((uplo == 'L') == (transA == 'N'));
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < n; j++) {
    for (size_t i = 0; i < m; i++) {
      zcomplex t = czero;
      if (side == 'L') {
        const size_t k0 = lowerlike ? 0 : i + 1, k1 = lowerlike ? i : m;
        for (size_t kk = k0; kk < k1; kk++) t += opel(A, lda, transA, i, kk) * B[j * ldb + kk];
        t += (nounit ? opel(A, lda, transA, i, i) : cone) * B[j * ldb + i];
      } else {
        const size_t k0 = lowerlike ? j + 1 : 0, k1 = lowerlike ? n : j;
        for (size_t kk = k0; kk < k1; kk++) t += B[kk * ldb + i] * opel(A, lda, transA, kk, j);
        t += B[j * ldb + i] * (nounit ? opel(A, lda, transA, j, j) : cone);
      }
      X[j * ldx + i] = alpha * t;
    }
  }
}
