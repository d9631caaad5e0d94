// Level-3 building blocks on top of the tile GEMM: TRMM, TRSM and SYRK/HERK.
//
// These run entirely on the GPU and are stream-ordered; argument checking is
// done by the public entry points (blas_api.cpp).
#include "internal.h"

namespace hc {

namespace {

// op(A) is lower-like when A is lower and untransposed or upper and transposed.
inline int lowerlike_tri(char uplo, char trans) {
  return ((uplo == 'L') == (trans == 'N')) ? TRI_LOWER : TRI_UPPER;
}

template <typename T>
inline int op_of(char trans) {
  if (!is_complex<T>::value) return trans == 'N' ? OP_N : OP_T;
  return op_code(trans);
}

// Address of element (r, c) of op(A).
template <typename T>
inline const T *op_at(const T *A, long lda, int op, long r, long c) {
  return op == OP_N ? A + c * lda + r : A + r * lda + c;
}

// Order of the diagonal blocks the TRSM inverts: the largest block one
// workgroup inverts in LDS (128 real, 64 complex), so that the leaves are
// few and the coupling updates have K >= 128.
template <typename T> constexpr int trsm_nb() { return diag_max<T>(); }

}  // namespace

// reference blas/dtrmm.c:374-426 (cuDtrmm2): X = alpha op(A) B or alpha B op(A)
// as one masked GEMM whose K range skips the zero triangle.
template <typename T>
hipError_t trmm(hipStream_t s, char side, char uplo, char trans, char diag, long m, long n,
                T alpha, const T *A, long lda, const T *B, long ldb, T *X, long ldx) {
  GemmOpts o;
  o.unit = (diag == 'U');
  const int top = op_of<T>(trans);
  const T zero = from_real<T>(0.0);
  if (side == 'L') {
    o.tri_a = lowerlike_tri(uplo, trans);
    return gemm<T>(s, top, OP_N, m, n, m, alpha, A, lda, B, ldb, zero, nullptr, 0, X, ldx, o);
  }
  o.tri_b = lowerlike_tri(uplo, trans);
  return gemm<T>(s, OP_N, top, m, n, n, alpha, B, ldb, A, lda, zero, nullptr, 0, X, ldx, o);
}

long trsm_ws_elems(char side, long m, long n) {  // enough for either block order
  const long t = (side == 'L') ? m : n;
  return ((t + 127) / 128) * 128 * 128;
}

namespace {

template <typename T>
struct TrsmCtx {
  hipStream_t s;
  bool left;
  bool lower;  // op(A) lower-like
  int op;
  const T *A;
  long lda;
  T *B;
  long ldb;
  long m, n;  // B is m x n
  const T *W; // inverted diagonal blocks of op(A)
};

// Solve on the triangle range [off, off + len) of op(A) (off a multiple of
// trsm_nb<T>()).  Recursive halving keeps the coupling updates as large GEMMs; the
// leaves multiply by the inverted diagonal block in place (the GEMM tile spans
// the whole block, so every workgroup reads its operand before writing it).
template <typename T>
hipError_t trsm_rec(const TrsmCtx<T> &c, long off, long len, T alpha) {
  const T one = from_real<T>(1.0), mone = from_real<T>(-1.0), zero = from_real<T>(0.0);
  constexpr int kTrsmNb = trsm_nb<T>();
  if (len <= kTrsmNb) {
    const T *Wb = c.W + (off / kTrsmNb) * kTrsmNb * kTrsmNb;
    GemmOpts o;
    if (c.left) {
      T *Bb = c.B + off;
      o.min_mt = (int)len;
      return gemm<T>(c.s, OP_N, OP_N, len, c.n, len, alpha, Wb, kTrsmNb, Bb, c.ldb, zero, nullptr, 0,
                     Bb, c.ldb, o);
    }
    T *Bb = c.B + off * c.ldb;
    o.min_nt = (int)len;
    return gemm<T>(c.s, OP_N, OP_N, c.m, len, len, alpha, Bb, c.ldb, Wb, kTrsmNb, zero, nullptr, 0, Bb,
                   c.ldb, o);
  }
  long len1 = ((len / 2 + kTrsmNb - 1) / kTrsmNb) * kTrsmNb;
  const long len2 = len - len1;
  const long o1 = off, o2 = off + len1;
  if (c.left) {
    T *B1 = c.B + o1, *B2 = c.B + o2;
    if (c.lower) {  // forward: X1, then B2 -= M21 X1, then X2
      HC_TRY(trsm_rec(c, o1, len1, alpha));
      HC_TRY(gemm<T>(c.s, c.op, OP_N, len2, c.n, len1, mone, op_at(c.A, c.lda, c.op, o2, o1), c.lda,
                     B1, c.ldb, alpha, B2, c.ldb, B2, c.ldb));
      return trsm_rec(c, o2, len2, one);
    }
    HC_TRY(trsm_rec(c, o2, len2, alpha));  // backward: X2, B1 -= M12 X2, X1
    HC_TRY(gemm<T>(c.s, c.op, OP_N, len1, c.n, len2, mone, op_at(c.A, c.lda, c.op, o1, o2), c.lda, B2,
                   c.ldb, alpha, B1, c.ldb, B1, c.ldb));
    return trsm_rec(c, o1, len1, one);
  }
  T *B1 = c.B + o1 * c.ldb, *B2 = c.B + o2 * c.ldb;
  if (c.lower) {  // X op(A) = B with op(A) lower: X2 first, B1 -= X2 M21, X1
    HC_TRY(trsm_rec(c, o2, len2, alpha));
    HC_TRY(gemm<T>(c.s, OP_N, c.op, c.m, len1, len2, mone, B2, c.ldb, op_at(c.A, c.lda, c.op, o2, o1),
                   c.lda, alpha, B1, c.ldb, B1, c.ldb));
    return trsm_rec(c, o1, len1, one);
  }
  HC_TRY(trsm_rec(c, o1, len1, alpha));  // upper: X1, B2 -= X1 M12, X2
  HC_TRY(gemm<T>(c.s, OP_N, c.op, c.m, len2, len1, mone, B1, c.ldb, op_at(c.A, c.lda, c.op, o1, o2),
                 c.lda, alpha, B2, c.ldb, B2, c.ldb));
  return trsm_rec(c, o2, len2, one);
}

}  // namespace

// reference blas/dtrsm.c:218-265 (cuDtrsm): B <- alpha op(A)^-1 B or
// alpha B op(A)^-1.  The diagonal blocks of op(A) are inverted in one batched
// launch, then the solve is a sequence of large GEMMs.
template <typename T>
hipError_t trsm(hipStream_t s, char side, char uplo, char trans, char diag, long m, long n,
                T alpha, const T *A, long lda, T *B, long ldb, T *ws) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const long t = (side == 'L') ? m : n;
  HC_TRY(trsm_diag_inverses<T>(s, uplo, trans, diag, t, trsm_nb<T>(), A, lda, ws));
  TrsmCtx<T> c;
  c.s = s;
  c.left = (side == 'L');
  c.lower = lowerlike_tri(uplo, trans) == TRI_LOWER;
  c.op = op_of<T>(trans);
  c.A = A; c.lda = lda; c.B = B; c.ldb = ldb; c.m = m; c.n = n; c.W = ws;
  return trsm_rec(c, 0, t, alpha);
}

// reference blas/dsyrk.c:163-209 (cuDsyrk) and blas/zherk.c:190 (cuZherk):
// only the tiles meeting the stored triangle are launched.
template <typename T>
hipError_t herk(hipStream_t s, char uplo, char trans, long n, long k, double alpha, const T *A,
                long lda, double beta, T *C, long ldc, const int *abort_flag, bool bf16, int max_wgs) {
  GemmOpts o;
  o.bf16 = bf16;
  o.max_wgs = max_wgs;
  o.tri_out = (uplo == 'L') ? TRI_LOWER : TRI_UPPER;
  o.tri_grid = true;
  o.real_diag = is_complex<T>::value;
  o.abort_flag = abort_flag;
  const int cop = is_complex<T>::value ? OP_C : OP_T;
  const T a = from_real<T>(alpha), b = from_real<T>(beta);
  if (trans == 'N')
    return gemm<T>(s, OP_N, cop, n, n, k, a, A, lda, A, lda, b, C, ldc, C, ldc, o);
  return gemm<T>(s, cop, OP_N, n, n, k, a, A, lda, A, lda, b, C, ldc, C, ldc, o);
}

template hipError_t trmm<double>(hipStream_t, char, char, char, char, long, long, double,
                                 const double *, long, const double *, long, double *, long);
template hipError_t trmm<zd>(hipStream_t, char, char, char, char, long, long, zd, const zd *, long,
                             const zd *, long, zd *, long);
template hipError_t trsm<double>(hipStream_t, char, char, char, char, long, long, double,
                                 const double *, long, double *, long, double *);
template hipError_t trsm<zd>(hipStream_t, char, char, char, char, long, long, zd, const zd *, long,
                             zd *, long, zd *);
template hipError_t herk<double>(hipStream_t, char, char, long, long, double, const double *, long,
                                 double, double *, long, const int *, bool, int);
template hipError_t herk<zd>(hipStream_t, char, char, long, long, double, const zd *, long, double,
                             zd *, long, const int *, bool, int);
template hipError_t herk<float>(hipStream_t, char, char, long, long, double, const float *, long,
                                double, float *, long, const int *, bool, int);
template hipError_t trsm<float>(hipStream_t, char, char, char, char, long, long, float,
                                const float *, long, float *, long, float *);
// single and single complex (FP32 matrix cores; diagonal blocks inverted in
// double / double complex)
template hipError_t trmm<float>(hipStream_t, char, char, char, char, long, long, float, const float *,
                                long, const float *, long, float *, long);
template hipError_t trmm<zf>(hipStream_t, char, char, char, char, long, long, zf, const zf *, long,
                             const zf *, long, zf *, long);
template hipError_t trsm<zf>(hipStream_t, char, char, char, char, long, long, zf, const zf *, long,
                             zf *, long, zf *);
template hipError_t herk<zf>(hipStream_t, char, char, long, long, double, const zf *, long, double,
                             zf *, long, const int *, bool, int);

}  // namespace hc
