// Internal declarations of the MI355X Cholesky library (not installed).
//
// Device scalar types, the tile-GEMM parameter block shared by every BLAS-3
// operation, the host launchers, the handle structures and the error helpers
// behind the public C ABI in include/blas.h and include/lapack.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

// hip_runtime.h maps CUDA_SUCCESS onto hipSuccess by macro; the public ABI
// defines CUDA_SUCCESS as its own enumerator.
#undef CUDA_SUCCESS

#include "blas.h"
#include "lapack.h"
#include "error.h"

namespace hc {

// ---------------------------------------------------------------------------
// scalar types
// ---------------------------------------------------------------------------

// Double complex, layout-compatible with C99 `double complex`.
struct __attribute__((aligned(16))) zd {
  double x, y;
};

__host__ __device__ inline zd make_zd(double re, double im) { zd r; r.x = re; r.y = im; return r; }
__host__ __device__ inline double conjx(double a) { return a; }
__host__ __device__ inline zd conjx(zd a) { return make_zd(a.x, -a.y); }
__host__ __device__ inline double realx(double a) { return a; }
__host__ __device__ inline double realx(zd a) { return a.x; }
__host__ __device__ inline double imagx(double) { return 0.0; }
__host__ __device__ inline double imagx(zd a) { return a.y; }
__host__ __device__ inline zd operator+(zd a, zd b) { return make_zd(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd operator-(zd a, zd b) { return make_zd(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd operator-(zd a) { return make_zd(-a.x, -a.y); }
__host__ __device__ inline zd operator*(zd a, zd b) {
  return make_zd(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline zd operator*(double a, zd b) { return make_zd(a * b.x, a * b.y); }
__host__ __device__ inline zd operator/(zd a, double b) { return make_zd(a.x / b, a.y / b); }
// Complex division by the scaled method (Smith, "Algorithm 116", CACM 1962).
__host__ __device__ inline zd operator/(zd a, zd b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_zd((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.y + b.x * r;
  return make_zd((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
// Single complex, layout-compatible with C99 `float complex`.  The level-3
// products compute in it (FP32 matrix cores); the diagonal-block kernels load
// it, compute in double complex and store it back.
struct __attribute__((aligned(8))) zf {
  float x, y;
};
__host__ __device__ inline zf make_zf(float re, float im) { zf r; r.x = re; r.y = im; return r; }
__host__ __device__ inline double realx(zf a) { return a.x; }
__host__ __device__ inline zf conjx(zf a) { return make_zf(a.x, -a.y); }
__host__ __device__ inline zf operator+(zf a, zf b) { return make_zf(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zf operator-(zf a, zf b) { return make_zf(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zf operator-(zf a) { return make_zf(-a.x, -a.y); }
__host__ __device__ inline zf operator*(zf a, zf b) {
  return make_zf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline zf operator*(float a, zf b) { return make_zf(a * b.x, a * b.y); }
__host__ __device__ inline bool is_zero(zf a) { return a.x == 0.0f && a.y == 0.0f; }
__host__ __device__ inline bool is_one(zf a) { return a.x == 1.0f && a.y == 0.0f; }
__host__ __device__ inline bool is_one(float a) { return a == 1.0f; }

__host__ __device__ inline float conjx(float a) { return a; }
__host__ __device__ inline double realx(float a) { return a; }
__host__ __device__ inline bool is_zero(float a) { return a == 0.0f; }
__host__ __device__ inline bool is_zero(double a) { return a == 0.0; }
__host__ __device__ inline bool is_zero(zd a) { return a.x == 0.0 && a.y == 0.0; }
__host__ __device__ inline bool is_one(double a) { return a == 1.0; }
__host__ __device__ inline bool is_one(zd a) { return a.x == 1.0 && a.y == 0.0; }

template <typename T> __host__ __device__ inline T from_real(double v);
template <> __host__ __device__ inline double from_real<double>(double v) { return v; }
template <> __host__ __device__ inline zd from_real<zd>(double v) { return make_zd(v, 0.0); }
template <> __host__ __device__ inline float from_real<float>(double v) { return (float)v; }
template <> __host__ __device__ inline zf from_real<zf>(double v) { return make_zf((float)v, 0.0f); }

// Storage <-> compute conversion (the mixed-precision path keeps FP32 data
// and computes the diagonal blocks in FP64).
template <typename D, typename S> __host__ __device__ inline D convert(S v) { return (D)v; }
template <> __host__ __device__ inline zd convert<zd, zd>(zd v) { return v; }
template <> __host__ __device__ inline zd convert<zd, zf>(zf v) { return make_zd((double)v.x, (double)v.y); }
template <> __host__ __device__ inline zf convert<zf, zd>(zd v) { return make_zf((float)v.x, (float)v.y); }
template <> __host__ __device__ inline zf convert<zf, zf>(zf v) { return v; }

template <typename T> struct is_complex { static constexpr bool value = false; };
template <> struct is_complex<zd> { static constexpr bool value = true; };
template <> struct is_complex<zf> { static constexpr bool value = true; };

// ---------------------------------------------------------------------------
// operand codes
// ---------------------------------------------------------------------------
enum Op { OP_N = 0, OP_T = 1, OP_C = 2 };
enum Tri { TRI_NONE = 0, TRI_LOWER = 1, TRI_UPPER = 2 };

inline int op_code(char t) { return t == 'N' ? OP_N : (t == 'T' ? OP_T : OP_C); }

// ---------------------------------------------------------------------------
// tile GEMM: D <- alpha * op(A) * op(B) + beta * C
// ---------------------------------------------------------------------------
// Options turn the one kernel into GEMM, SYRK/HERK (tri_out + tri_grid),
// TRMM (tri_a / tri_b masks) and the inverse-based TRSM step.
struct GemmOpts {
  int tri_out = TRI_NONE;   // write only row >= col (lower) or row <= col (upper)
  bool tri_grid = false;    // launch only tiles meeting the tri_out triangle (m == n)
  bool real_diag = false;   // HERK: force the diagonal of D to be real
  int tri_a = TRI_NONE;     // op(A) is triangular (lower- or upper-like): TRMM left
  int tri_b = TRI_NONE;     // op(B) is triangular: TRMM right
  bool unit = false;        // the triangular operand has an implicit unit diagonal
  bool zero_diag = false;   // ... or its diagonal is excluded (strict triangle)
  int min_mt = 0;           // the tile must span at least this many rows of D
  int min_nt = 0;           // ... or columns of D (in-place safety)
  const int *abort_flag = nullptr;  // skip all work when *abort_flag != 0
  bool bf16 = false;        // float data only: products in BF16 (FP32 accumulate)
  void *splitk_ws = nullptr;  // optional workspace: few-column products may split K
  long splitk_elems = 0;      // ... of this many elements
  int max_wgs = 0;            // tri_grid: at most this many workgroups (0: one per tile)
};

template <typename T>
hipError_t gemm(hipStream_t s, int ta, int tb, long m, long n, long k, T alpha, const T *A,
                long lda, const T *B, long ldb, T beta, const T *C, long ldc, T *D, long ldd,
                const GemmOpts &o = GemmOpts());
// FP32 storage (gemm_lp.hip): op in {N, T} on the FP32 / BF16 kernel;
// products with a triangular operand run on gemm_tile<float>.
template <>
hipError_t gemm<float>(hipStream_t s, int ta, int tb, long m, long n, long k, float alpha,
                       const float *A, long lda, const float *B, long ldb, float beta,
                       const float *C, long ldc, float *D, long ldd, const GemmOpts &o);
// The tile kernel of gemm.hip for any element type (double, double complex,
// float, single complex); every GemmOpts option.
template <typename T>
hipError_t gemm_tile(hipStream_t s, int ta, int tb, long m, long n, long k, T alpha, const T *A,
                     long lda, const T *B, long ldb, T beta, const T *C, long ldc, T *D, long ldd,
                     const GemmOpts &o = GemmOpts());

// 2-D copy D <- S (m x n)
template <typename T>
hipError_t copy2d(hipStream_t s, long m, long n, const T *S, long lds, T *D, long ldd);


// ---------------------------------------------------------------------------
// level-3 building blocks (blas_ops.cpp)
// ---------------------------------------------------------------------------
// Workspace supplier: a stream-ordered device allocation of `bytes`.
struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Out-of-place triangular multiply X = alpha op(A) B (side L) or alpha B op(A).
template <typename T>
hipError_t trmm(hipStream_t s, char side, char uplo, char trans, char diag, long m, long n,
                T alpha, const T *A, long lda, const T *B, long ldb, T *X, long ldx);

// In-place triangular solve.  `ws` must hold trsm_ws_elems(side, m, n) elements.
long trsm_ws_elems(char side, long m, long n);
template <typename T>
hipError_t trsm(hipStream_t s, char side, char uplo, char trans, char diag, long m, long n,
                T alpha, const T *A, long lda, T *B, long ldb, T *ws);

// SYRK (real) / HERK (complex, real alpha/beta): one triangle of C.
template <typename T>
hipError_t herk(hipStream_t s, char uplo, char trans, long n, long k, double alpha,
                const T *A, long lda, double beta, T *C, long ldc,
                const int *abort_flag = nullptr, bool bf16 = false, int max_wgs = 0);

// ---------------------------------------------------------------------------
// diagonal-block kernels (diag.hip); the block is held in LDS
// ---------------------------------------------------------------------------
template <typename T> constexpr int diag_max() { return is_complex<T>::value ? 64 : 128; }

// Factor the n x n diagonal block in place (n <= diag_max<T>()), write
// W = inv(L) densely (ld ldw) where L is the lower form of the factor
// (L = U^H for upper), and on a non-positive pivot set *info = info_off + k + 1.
template <typename T>
hipError_t diag_potrf(hipStream_t s, char uplo, int n, T *A, long lda, T *W, long ldw, int *info,
                      long info_off);
// A whole panel in one launch (lower, FP32 or FP64 storage): the m x jb
// panel at A (top jb x jb block on the diagonal, m >= jb, jb <=
// kPanelMaxCols) becomes L.  ws: panel_ws_elems(jb) elements; flags:
// panel_flag_ints ints, zero before the first launch, each launch with a
// larger epoch than the last.  A failed pivot sets *info = info_off + k + 1;
// a dependency wait abandoned after ~4 s sets kInfoAbandoned.
constexpr long kPanelMaxCols = 2048;
constexpr int kPanelFlagInts = 2 * 32 * (kPanelMaxCols / 128);
constexpr int kInfoAbandoned = -2147483647 - 1;
inline long panel_ws_elems(long jb) { return 2 * ((jb + 127) / 128) * 128L * 128L; }
template <typename T>
hipError_t panel_potrf(hipStream_t s, long m, long jb, T *A, long lda, T *ws, int *flags, int epoch, int *info,
                       long info_off);
// Invert the triangular block in place (B = inv(A), B may equal A) and, when
// W != nullptr, also write inv of its lower form densely into W.
template <typename T>
hipError_t diag_trtri(hipStream_t s, char uplo, char diag, int n, const T *A, long lda, T *B,
                      long ldb, T *W, long ldw, const int *abort_flag);
// Every nb x nb diagonal block of the n x n triangle inverted in place, one
// workgroup each (the last block may be smaller).
template <typename T>
hipError_t diag_trtri_blocks(hipStream_t s, char uplo, char diag, long n, int nb, T *A, long lda);
// U U^H / L^H L of every nb x nb diagonal block of the n x n triangle, block
// b into W + b nb^2 (ld nb, triangle uplo only), one workgroup each.
template <typename T>
hipError_t diag_lauum_blocks(hipStream_t s, char uplo, long n, int nb, const T *A, long lda, T *W);
// D <- S on the triangle uplo (diagonal included) of an n x n block.
template <typename T>
hipError_t tri_copy(hipStream_t s, char uplo, int n, const T *S, long lds, T *D, long ldd);
// In-place U U^H (upper) or L^H L (lower) of a diagonal block.
template <typename T>
hipError_t diag_lauum(hipStream_t s, char uplo, int n, T *A, long lda);
// Batched inverses of the nb x nb diagonal blocks of op(A) (order m) into W,
// block b at W + b * nb * nb with leading dimension nb.
template <typename T>
hipError_t trsm_diag_inverses(hipStream_t s, char uplo, char trans, char diag, long m, int nb,
                              const T *A, long lda, T *W);
// Scan the diagonal for exact zeros; *info = 1-based index of the first
// (upper) or last (lower) zero, 0 when none.
template <typename T>
hipError_t diag_zero_scan(hipStream_t s, char uplo, long n, const T *A, long lda, int *info);
// 2 * sum(log(Re x[i*incx])) into *out (device), compensated summation.
template <typename T>
hipError_t logdet(hipStream_t s, const T *x, long incx, long n, double *partials, double *out);
constexpr int kLogdetMaxBlocks = 1024;

// BF16 products from BF16 operands with K contiguous (gemm_lp.hip):
// D (FP32) <- alpha A B^T + beta C with A m x k (row r at A + r lda), B n x k;
// k a multiple of 64, lda and ldb multiples of 8, rows zero-padded to k.
// f16: the operands are IEEE binary16 instead (FP16 products).
hipError_t gemm_bf16_nt(hipStream_t s, long m, long n, long k, float alpha, const void *A, long lda,
                        const void *B, long ldb, float beta, const float *C, long ldc, float *D,
                        long ldd, int tri_out, bool tri_grid, const int *abort_flag, bool f16 = false);
// T (bf16, row-major, ld ldt) <- the m x kp FP32 panel P, zeros for k >= kp.
hipError_t panel_to_bf16(hipStream_t s, long m, long kp, const float *P, long ldp, void *T, long ldt);
// T (binary16, row-major, ld ldt) <- scale P (scale a power of two); an entry
// of magnitude 2^16 or more (or not finite) records column col0 + k + 1 in
// *info when *info is still 0.
hipError_t panel_to_f16(hipStream_t s, long m, long kp, const float *P, long ldp, void *T, long ldt,
                        float scale, int *info, long col0);
// Split copies (rows of 3 ldt bf16 values): TA = [hi | hi | lo], TB = [hi | lo | hi]
// with x = hi + lo, so that a BF16 product TA TB^T over 3 ldt carries ~16
// significant bits of the panel.
hipError_t panel_to_bf16_split(hipStream_t s, long m, long kp, const float *P, long ldp, void *TA, void *TB,
                               long ldt);

// ---------------------------------------------------------------------------
// mixed-precision solver kernels (mixed.hip)
// ---------------------------------------------------------------------------
// S (FP32, lower form of A, ld lds) <- A (FP64, triangle uplo); *overflow <- 1
// if a value exceeds the FP32 range.
hipError_t to_float_lower(hipStream_t s, char uplo, long n, const double *A, long lda, float *S,
                          long lds, int *overflow);
hipError_t d2s(hipStream_t s, long m, long n, const double *X, long ldx, float *Y, long ldy,
               int *overflow);
hipError_t s2d(hipStream_t s, long m, long n, const float *Y, long ldy, double *X, long ldx);
hipError_t add_s2d(hipStream_t s, long m, long n, const float *Y, long ldy, double *X, long ldx);
// out[j] = max_i |X(i, j)|
hipError_t colmax(hipStream_t s, long m, long n, const double *X, long ldx, double *out);
// The FP32 lower-form copy of to_float_lower and the infinity norm of the
// symmetric matrix from one read of A (64 x 64 tiles; deterministic partial
// sums in ws of to_float_norm_ws_doubles(n) doubles); *anorm on the device.
long to_float_norm_ws_doubles(long n);
// out[0] = max_i |A(i, i)| (device)
hipError_t diag_absmax(hipStream_t s, long n, const double *A, long lda, double *out);
hipError_t to_float_norm(hipStream_t s, char uplo, long n, const double *A, long lda, float *S, long lds,
                         int *overflow, double *ws, double *anorm);
// The same in two parts: the copy and the partial sums of the lower-form
// columns [col0, col1) (multiples of 64, or col1 = n) on at most max_wgs
// workgroups (0: one per 64 x 64 tile), then, once every column is done, the
// norm from the partial sums.
hipError_t to_float_norm_cols(hipStream_t s, char uplo, long n, const double *A, long lda, float *S, long lds,
                              int *overflow, double *ws, long col0, long col1, long max_wgs);
hipError_t to_float_norm_finish(hipStream_t s, long n, double *ws, double *anorm);
// The FP32 copy (with the norm) made by the blocked factorisation itself: its
// first block column before the first panel, the rest on the look-ahead
// stream beside that panel.
struct PreConvert {
  char uplo;            // triangle of A
  const double *A;      // FP64 source
  long lda;
  int *overflow;        // set to 1 when a value exceeds the FP32 range
  double *ws;           // to_float_norm_ws_doubles(n) doubles
  double *anorm;        // ||A||_inf (device)
};

// ---------------------------------------------------------------------------
// symmetric matrix times a few right-hand sides (symm.hip)
// ---------------------------------------------------------------------------
// R = beta C + alpha A X, A symmetric FP64 in triangle uplo (read once), X and
// R n x nrhs; ws: symm_tall_ws_bytes(n, nrhs) bytes.  C is not read when
// beta == 0; R may alias X.
size_t symm_tall_ws_bytes(long n, long nrhs);
hipError_t symm_tall(hipStream_t s, char uplo, long n, long nrhs, double alpha, const double *A,
                     long lda, const double *X, long ldx, double beta, const double *C, long ldc,
                     double *R, long ldr, void *ws);

// ---------------------------------------------------------------------------
// triangular solve with few right-hand sides (trsm_multi.hip)
// ---------------------------------------------------------------------------
// Inverses of the 128 x 128 diagonal blocks of op(A) (uplo, trans), for
// trsm_multi; trsm_multi_ws_elems(n) elements.
long trsm_multi_ws_elems(long n);
template <typename T>
hipError_t trsm_multi_inverses(hipStream_t s, char uplo, char trans, long n, const T *A, long lda,
                               T *Winv);
// X = op(A)^-1 B (left side, non-unit), one chained launch; X may be B
// (ldx == ldb).  sync: trsm_multi_sync_bytes(n, nrhs) bytes, 256-aligned;
// its first int is an error word that a bounded wait sets (clear it before,
// read it after).
size_t trsm_multi_sync_bytes(long n, long nrhs);
template <typename T>
hipError_t trsm_multi(hipStream_t s, char uplo, char trans, long n, long nrhs, const T *A, long lda,
                      const T *Winv, const T *B, long ldb, T *X, long ldx, int *sync);

// ---------------------------------------------------------------------------
// batched factorisation (batched.hip)
// ---------------------------------------------------------------------------
constexpr int kBatchedMaxN = 256;
// resident workgroups per CU of the batched kernel (its occupancy)
template <typename ST> int potrf_batched_blocks_per_cu(char uplo, bool inverse);
// Factor count packed matrices in place (matrix order[t] for t = 0..count-1,
// n = nvec[i] <= kBatchedMaxN, at element offset offs[i]); info[i] and
// logdet[i] (may be null) per matrix; `grid` persistent workgroups.  With
// `inverse`, each factored matrix is then overwritten by its inverse.
template <typename ST>
hipError_t potrf_batched(hipStream_t s, char uplo, int count, const int *nvec, const long *offs,
                         const int *order, ST *A, double *logdet, int *info, int grid, bool inverse);

// Register-resident factorisation + logdet (no inverse), in three size
// classes (batched_reg_class: orders up to 64, up to 128, up to 256).
// `order` lists class 2, then 1, then 0 (decreasing order); d_cnt[c] (device
// memory) matrices of class c run on grid[c] persistent workgroups.  With the
// counts also on the host (h_cnt), empty classes are not launched and grids
// are cut to the class size; with h_cnt null every class is launched.
__host__ __device__ constexpr int batched_reg_class(int n) { return n <= 64 ? 0 : (n <= 128 ? 1 : 2); }
template <typename ST>
hipError_t potrf_batched_reg(hipStream_t s, char uplo, const int *d_cnt, const int *h_cnt, const int grid[3],
                             const int *nvec, const long *offs, const int *order, ST *A, double *logdet, int *info);
template <typename ST> int potrf_batched_reg_blocks_per_cu(char uplo, int cls);
// Resolves a device's copy of the no-allocation entry point's scratch array
// (batched_noalloc.hip), once per device, with that device current.
hipError_t noalloc_setup(int device);

// GMRES basis helpers: right-hand side k owns V + (k nv + i) n, i < nv.
hipError_t basis_dots(hipStream_t s, long n, long nrhs, int nv, int first, int nvec, const double *V,
                      const double *W, long ldw, double *out);
hipError_t basis_sub(hipStream_t s, long n, long nrhs, int nv, int nvec, const double *V,
                     const double *h, double *W, long ldw);
hipError_t basis_put(hipStream_t s, long n, long nrhs, int nv, int slot, const double *W, long ldw,
                     const double *scale, double *V);
hipError_t basis_combine(hipStream_t s, long n, long nrhs, int nv, int nvec, const double *V,
                         const double *y, double *X, long ldx);
// GMRES bookkeeping on the device (Hessenberg columns, Givens rotations, the
// rotated right-hand side, the basis scales and the convergence count).
hipError_t gmres_start(hipStream_t s, long nrhs, int nv, const double *nrm2, double *beta, double *g,
                       double *scale);
hipError_t gmres_rotate(hipStream_t s, long nrhs, int nv, int m, int j, const double *hd, const double *nrm2,
                        double *H, double *cs, double *sn, double *g, const double *beta, double *scale,
                        double tol, int *active);
hipError_t gmres_solve_y(hipStream_t s, long nrhs, int nv, int m, int used, const double *H, const double *g,
                         double *y);

// Tile choice of a GEMM kernel family: a launch takes the family's larger
// tile when it has at least this many of them (tuning.cpp; measured by the
// autotuner, default dflt).  Families: D (double tile kernel, 128 vs 64),
// Z (complex tile kernel, 64 vs 32), S (FP32 kernel, 128 vs 64), H (16-bit
// kernel, 128 vs 64), W (16-bit kernel, 256 vs 128).
long tuned_tile_min(char family, long dflt);
// Whether each panel of the lower real factorisation of precision prec is
// factored in one launch (panel_potrf).
bool tuned_panel_launch(char prec, bool dflt);
// Workgroup cap of the factorisation's trailing update of order t in
// precision prec (tuning table entry trailing_cap), 0 for none.
int tuned_trailing_cap(char prec, long t);

// Measured outer block size for precision D / Z / S / B at order n, 0 when
// the tuning table has no entry (tuning.cpp).
long tuned_potrf_nb(char prec, long n);
// Workgroups per CU of the batched factorisation's persistent grid measured
// by the autotuner (prec D or S), or 0.
int tuned_batched_per_cu(char prec, bool inverse);

// LAPACK internals shared by the solvers (lapack_api.cpp).
hipError_t lapack_workspace(CULAPACKhandle h, size_t bytes, void **out);
// Blocked factorisation on the handle's streams; W: diag_max^2 elements.
// bf16: FP32 data only, trailing updates with BF16 products (split: three
// BF16 products of a two-term split of every panel value; f16_scale > 0:
// FP16 products of the panel values times f16_scale, a power of two).
// pre (FP32 data only): A is not yet the FP32 copy; the factorisation makes it.
template <typename T>
hipError_t potrf_on_device(CULAPACKhandle h, char uplo, long n, T *A, long lda, T *W, bool bf16,
                           void *lp_ws = nullptr, bool split = false, float f16_scale = 0.0f,
                           const PreConvert *pre = nullptr);
// Workspace of the bf16 factorisation (two bfloat16 panel copies; six with split).
size_t potrf_bf16_ws_bytes(CULAPACKhandle h, long n, bool split = false);

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
CUresult to_cu(hipError_t e);
void report(const char *call, const char *func, const char *file, int line, CUresult err);

}  // namespace hc

#define HC_CHECK(call)                                                        \
  do {                                                                        \
    hipError_t hc_e_ = (call);                                                \
    if (hc_e_ != hipSuccess) {                                                \
      CUresult hc_r_ = hc::to_cu(hc_e_);                                      \
      hc::report(#call, __func__, __FILE__, __LINE__, hc_r_);                 \
      return hc_r_;                                                           \
    }                                                                         \
  } while (0)

// Same, inside functions that return hipError_t.
#define HC_TRY(call)                                                          \
  do {                                                                        \
    hipError_t hc_e_ = (call);                                                \
    if (hc_e_ != hipSuccess) return hc_e_;                                    \
  } while (0)

// Public handle state.
struct __cublashandle_st {
  int device;
};

struct __culapackhandle_st {
  CUBLAShandle blas;
  int device;
  hipStream_t stream;     // main stream (trailing updates)
  hipStream_t side;       // look-ahead stream (diagonal block and panel)
  hipEvent_t ev_panel;    // panel ready
  hipEvent_t ev_update;   // next panel column updated
  void *ws;               // device workspace, grown on demand
  size_t ws_bytes;
  int *dinfo;             // device info word
  int *hinfo;             // pinned host copy
  int *pflags;            // one-launch panel factorisation: flags (zeroed once), then
  void *pws;              // its workspace (allocated on first use)
  int pepoch;             // epoch of the last panel launch
  long nb_override[4];    // outer block size set by cuLAPACKSetBlockSize (D, Z, S, B), 0 = tuned
  // optional timing of the factorisation's trailing updates (event pairs on
  // the main stream and the algorithmic flops of each launch)
  int timing;
  long tcount, tcap;
  hipEvent_t *tev;
  double *tflops;
};
