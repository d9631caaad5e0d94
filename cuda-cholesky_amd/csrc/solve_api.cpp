// Solvers on top of the factorisation: xPOTRS (FP64 solve with a Cholesky
// factor) and the mixed-precision xSPOSV (low-precision factor + FP64
// iterative refinement).
//
// Neither exists in the reference; their interfaces follow LAPACK's DPOTRS
// and DSPOSV (argument order, INFO codes, ITER codes, the refinement stopping
// test and the fall-back to a full FP64 factorisation) with the reference's
// conventions added: a handle first, device byte addresses (CUdeviceptr) for
// matrices, CUresult returns and XERBLA on bad arguments.
#include "api_util.h"

#include <algorithm>
#include <cmath>
#include <vector>

using hc::bad_arg;
using hc::dptr;

namespace {

#define CU_TRY(call)                       \
  do {                                     \
    const CUresult cu_r_ = (call);         \
    if (cu_r_ != CUDA_SUCCESS) return cu_r_; \
  } while (0)

constexpr long kIterMax = 30;  // refinement steps before falling back (LAPACK ITERMAX)

// The two triangular factors of A = L L^T (uplo L) or U^T U (uplo U), in
// solve order: first the lower-like one, then its transpose.
struct PotrsOps {
  char t1, t2;
};
PotrsOps potrs_ops(char uplo) { return uplo == 'L' ? PotrsOps{'N', 'T'} : PotrsOps{'T', 'N'}; }

// Inverses of the diagonal blocks of both triangular factors
// (2 trsm_multi_ws_elems(n) elements at W).
template <typename T>
hipError_t potrs_prepare(hipStream_t s, char uplo, long n, const T *F, long ldf, T *W) {
  const PotrsOps o = potrs_ops(uplo);
  HC_TRY(hc::trsm_multi_inverses<T>(s, uplo, o.t1, n, F, ldf, W));
  return hc::trsm_multi_inverses<T>(s, uplo, o.t2, n, F, ldf, W + hc::trsm_multi_ws_elems(n));
}

// B <- inv(A) B for the factor in triangle uplo, n x nrhs; Y: n x nrhs work;
// sync: trsm_multi_sync_bytes(n, nrhs) bytes.
template <typename T>
hipError_t potrs_device(hipStream_t s, char uplo, long n, long nrhs, const T *F, long ldf,
                        const T *W, T *B, long ldb, T *Y, int *sync) {
  const PotrsOps o = potrs_ops(uplo);
  HC_TRY(hc::trsm_multi<T>(s, uplo, o.t1, n, nrhs, F, ldf, W, B, ldb, Y, n, sync));
  return hc::trsm_multi<T>(s, uplo, o.t2, n, nrhs, F, ldf, W + hc::trsm_multi_ws_elems(n), Y, n, B,
                           ldb, sync);
}

// The solves' error word (set when a chained solve abandoned a wait).
CUresult solve_status(hipStream_t s, const int *sync) {
  int v = 0;
  HC_CHECK(hipMemcpyAsync(&v, sync, sizeof(int), hipMemcpyDeviceToHost, s));
  HC_CHECK(hipStreamSynchronize(s));
  return v == 0 ? CUDA_SUCCESS : CUDA_ERROR_LAUNCH_FAILED;
}

// R <- B - A X with A symmetric in triangle uplo (the triangle read once).
hipError_t residual(hipStream_t s, char uplo, long n, long nrhs, const double *A, long lda,
                    const double *B, long ldb, const double *X, long ldx, double *R, long ldr,
                    void *ws) {
  return hc::symm_tall(s, uplo, n, nrhs, -1.0, A, lda, X, ldx, 1.0, B, ldb, R, ldr, ws);
}

// W <- A X with A symmetric in triangle uplo.
hipError_t symm_apply(hipStream_t s, char uplo, long n, long nrhs, const double *A, long lda,
                      const double *X, long ldx, double *W, long ldw, void *ws) {
  return hc::symm_tall(s, uplo, n, nrhs, 1.0, A, lda, X, ldx, 0.0, nullptr, 0, W, ldw, ws);
}

// GMRES-based refinement: restart length and inner relative tolerance.
constexpr int kGmresM = 30;
constexpr double kGmresTol = 1e-10;
constexpr long kGmresOuterMax = 10;

// Workspace layout of the mixed solver (byte offsets, 256-aligned).
struct SposvLayout {
  size_t S, W, T, Y, R, RS, PART, NORM, DEV, P16, V, WV, HD, GM, SYNC, total;
};

// Device GMRES state per right-hand side (doubles): ones, beta, g (nv), y
// (nv), H (m nv), cs (m), sn (m); then m ints of convergence counts.
size_t gmres_state_bytes(long nrhs) {
  constexpr int NV = kGmresM + 1;
  return (size_t)nrhs * (2 + 2 * NV + (size_t)kGmresM * NV + 2 * kGmresM) * sizeof(double) +
         kGmresM * sizeof(int);
}


size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

SposvLayout sposv_layout(CULAPACKhandle h, long n, long nrhs) {
  SposvLayout L;
  size_t off = 0;
  L.S = off;    off = align256(off + (size_t)n * n * sizeof(float));              // factor
  L.W = off;    off = align256(off + 128 * 128 * sizeof(double));                 // diag inverse
  L.T = off;    off = align256(off + 2 * (size_t)hc::trsm_multi_ws_elems(n) * sizeof(double));
  L.Y = off;    off = align256(off + (size_t)n * nrhs * sizeof(double));          // solve work
  L.R = off;    off = align256(off + (size_t)n * nrhs * sizeof(double));          // residual
  L.RS = off;   off = align256(off + (size_t)n * nrhs * sizeof(float));           // FP32 rhs
  L.PART = off; off = align256(off + hc::symm_tall_ws_bytes(n, nrhs));  // symmetric products
  L.NORM = off; off = align256(off + (size_t)hc::to_float_norm_ws_doubles(n) * sizeof(double));
  L.DEV = off;  off = align256(off + (3 * (size_t)nrhs + 8) * sizeof(double));    // norms, flags
  L.P16 = off;  off = align256(off + hc::potrf_bf16_ws_bytes(h, n, true));          // bf16 panels
  L.V = off;    off = align256(off + (size_t)nrhs * (kGmresM + 1) * n * sizeof(double));  // Krylov bases
  L.WV = off;   off = align256(off + (size_t)n * nrhs * sizeof(double));
  L.HD = off;   off = align256(off + 3 * (size_t)nrhs * (kGmresM + 1) * sizeof(double));
  L.GM = off;   off = align256(off + gmres_state_bytes(nrhs));
  L.SYNC = off; off = align256(off + hc::trsm_multi_sync_bytes(n, nrhs));
  L.total = off;
  return L;
}

struct Norms {
  std::vector<double> r, x;
};

}  // namespace

extern "C" {

CUresult cuDpotrs(CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, long *info) {
  const char *name = "cuDpotrs";
  if (info == nullptr) return bad_arg(name, 8);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }  // handle
  if (!hc::valid_uplo(uplo)) { *info = -1; return bad_arg(name, 1); }
  if (lda < (n > 1 ? n : 1)) { *info = -5; return bad_arg(name, 5); }
  if (ldb < (n > 1 ? n : 1)) { *info = -7; return bad_arg(name, 7); }
  if (n == 0 || nrhs == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  void *ws = nullptr;
  const long N = (long)n, K = (long)nrhs, WI = 2 * hc::trsm_multi_ws_elems(N);
  const size_t sync_off = align256((size_t)(WI + N * K) * sizeof(double));
  HC_CHECK(hc::lapack_workspace(h, sync_off + hc::trsm_multi_sync_bytes(N, K), &ws));
  double *W = static_cast<double *>(ws);
  int *sync = reinterpret_cast<int *>(static_cast<char *>(ws) + sync_off);
  HC_CHECK(hipMemsetAsync(sync, 0, sizeof(int), h->stream));
  HC_CHECK(potrs_prepare<double>(h->stream, (char)uplo, N, dptr<double>(A), (long)lda, W));
  HC_CHECK(potrs_device<double>(h->stream, (char)uplo, N, K, dptr<double>(A), (long)lda, W,
                                dptr<double>(B), (long)ldb, W + WI, sync));
  return solve_status(h->stream, sync);
}

}  // extern "C"

namespace {

CUresult sposv(const char *name, CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs,
               CUdeviceptr A, size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx,
               char factor, char refine, long *iter, char *used, long *info) {
  if (info == nullptr) return bad_arg(name, 15);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }
  if (!hc::valid_uplo(uplo)) { *info = -1; return bad_arg(name, 1); }
  if (lda < (n > 1 ? n : 1)) { *info = -5; return bad_arg(name, 5); }
  if (ldb < (n > 1 ? n : 1)) { *info = -7; return bad_arg(name, 7); }
  if (ldx < (n > 1 ? n : 1)) { *info = -9; return bad_arg(name, 9); }
  if (factor != 'S' && factor != 'B' && factor != 'X' && factor != 'H') { *info = -10; return bad_arg(name, 10); }
  if (refine != 'I' && refine != 'G') { *info = -11; return bad_arg(name, 11); }
  if (iter == nullptr) { *info = -12; return bad_arg(name, 12); }
  *iter = 0;
  if (used != nullptr) *used = factor;
  if (n == 0 || nrhs == 0) return CUDA_SUCCESS;

  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  const long N = (long)n, K = (long)nrhs;
  const char ul = (char)uplo;
  const SposvLayout L = sposv_layout(h, N, K);
  void *base = nullptr;
  HC_CHECK(hc::lapack_workspace(h, L.total, &base));
  char *wsb = static_cast<char *>(base);
  float *S = reinterpret_cast<float *>(wsb + L.S);
  float *W = reinterpret_cast<float *>(wsb + L.W);
  void *T = wsb + L.T;
  void *Yw = wsb + L.Y;
  double *PART = reinterpret_cast<double *>(wsb + L.PART);
  double *V = reinterpret_cast<double *>(wsb + L.V);
  double *WV = reinterpret_cast<double *>(wsb + L.WV);
  double *HD = reinterpret_cast<double *>(wsb + L.HD);  // dots (2 sets) and scales
  constexpr int NV = kGmresM + 1;
  double *R = reinterpret_cast<double *>(wsb + L.R);
  int *sync = reinterpret_cast<int *>(wsb + L.SYNC);
  float *RS = reinterpret_cast<float *>(wsb + L.RS);
  double *NW = reinterpret_cast<double *>(wsb + L.NORM);
  double *DEV = reinterpret_cast<double *>(wsb + L.DEV);
  double *dnorm = DEV;                       // ||A||_inf
  double *rn = DEV + 1, *xn = DEV + 1 + K;   // column max-norms
  int *ovf = reinterpret_cast<int *>(DEV + 1 + 2 * K);
  double *dmax = DEV + 2 + 2 * K;            // max |a_ii|
  const double *Ad = dptr<double>(A);
  const double *Bd = dptr<double>(B);
  double *Xd = dptr<double>(X);
  hipStream_t s = h->stream;

  // ||A||_inf; refinement stops when every column has
  // ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n)   (LAPACK dsposv, eps = 2^-53).
  // One read of A gives both the norm and the FP32 copy the first attempt
  // factors, made by that factorisation itself (its first block column ahead
  // of the first panel, the rest beside it); a later attempt converts again:
  // the factor overwrites the copy.
  HC_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
  bool first = true;
  double anorm = 0.0, amax_diag = 0.0;
  if (factor == 'H') {
    HC_CHECK(hc::diag_absmax(s, N, Ad, (long)lda, dmax));
    HC_CHECK(hipMemcpyAsync(&amax_diag, dmax, sizeof(double), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
  }
  // FP16 panel copies are scaled by 2^e with every entry of the factor below
  // 2^15: |l_ij| <= sqrt(a_ii) <= sqrt(max a_ii) for an SPD matrix
  float f16_scale = 1.0f;
  if (factor == 'H' && amax_diag > 0.0 && std::isfinite(amax_diag)) {
    int e2 = 0;
    std::frexp(std::sqrt(amax_diag), &e2);   // sqrt(max a_ii) < 2^e2
    f16_scale = std::ldexp(1.0f, std::max(-60, std::min(60, 15 - e2)));
  }
  double cte = 0.0;  // anorm eps sqrt(n), once the first factorisation has formed anorm

  // flag != nullptr: the overflow word comes back with the norms (one wait)
  auto read_norms = [&](Norms &nm, int *flag) -> CUresult {
    HC_CHECK(hc::colmax(s, N, K, R, N, rn));
    HC_CHECK(hc::colmax(s, N, K, Xd, (long)ldx, xn));
    nm.r.resize(K);
    nm.x.resize(K);
    HC_CHECK(hipMemcpyAsync(nm.r.data(), rn, K * sizeof(double), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipMemcpyAsync(nm.x.data(), xn, K * sizeof(double), hipMemcpyDeviceToHost, s));
    if (flag != nullptr) HC_CHECK(hipMemcpyAsync(flag, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
    return CUDA_SUCCESS;
  };
  auto converged = [&](const Norms &nm) {
    for (long j = 0; j < K; j++)
      if (!(nm.r[j] <= nm.x[j] * cte)) return false;
    return true;
  };
  auto read_flag = [&](int *flag) -> CUresult {
    int v = 0;
    HC_CHECK(hipMemcpyAsync(&v, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
    *flag = v;
    return CUDA_SUCCESS;
  };

  // One attempt with a low-precision factor; returns the LAPACK ITER code.
  auto attempt = [&](char kind, long *code) -> CUresult {
    const bool bf16 = kind != 'S', split = kind == 'X';
    const bool pre = first;  // this factorisation makes the FP32 copy and the norm
    first = false;
    int flag = 0;
    if (!pre) {
      HC_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
      HC_CHECK(hc::to_float_lower(s, ul, N, Ad, (long)lda, S, N, ovf));
      CU_TRY(read_flag(&flag));
      if (flag) { *code = -2; return CUDA_SUCCESS; }     // narrowing overflowed
    }
    const hc::PreConvert pc{ul, Ad, (long)lda, ovf, NW, dnorm};
    HC_CHECK(hipMemsetAsync(h->dinfo, 0, sizeof(int), s));
    HC_CHECK(hc::potrf_on_device<float>(h, 'L', N, S, N, W, bf16, wsb + L.P16, split,
                                        kind == 'H' ? f16_scale : 0.0f, pre ? &pc : nullptr));
    // x0 = inv(S S^T) b in FP32; the diagonal-block inverses serve every
    // solve.  They and the FP32 right-hand sides are queued behind the
    // factorisation before its info word is read (wasted if it failed), so
    // that one wait covers info, the norm and every overflow.
    HC_CHECK(potrs_prepare<float>(s, 'L', N, S, N, static_cast<float *>(T)));
    HC_CHECK(hc::d2s(s, N, K, Bd, (long)ldb, RS, N, ovf));
    long finfo = 0;
    HC_CHECK(hipMemcpyAsync(h->hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    if (pre) HC_CHECK(hipMemcpyAsync(&anorm, dnorm, sizeof(double), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipMemcpyAsync(&flag, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    HC_CHECK(hipStreamSynchronize(s));
    if (pre) cte = anorm * std::ldexp(1.0, -53) * std::sqrt((double)N);
    if (flag) { *code = -2; return CUDA_SUCCESS; }       // narrowing overflowed (any factor is void)
    finfo = h->hinfo[0];
    if (finfo == hc::kInfoAbandoned) return CUDA_ERROR_LAUNCH_FAILED;  // a dependency wait gave up
    if (finfo != 0) { *code = -3; return CUDA_SUCCESS; }  // low-precision factorisation failed
    HC_CHECK(potrs_device<float>(s, 'L', N, K, S, N, static_cast<const float *>(T), RS, N,
                                 static_cast<float *>(Yw), sync));
    HC_CHECK(hc::s2d(s, N, K, RS, N, Xd, (long)ldx));
    // w <- inv(S S^T) w in FP32 (w: n x nrhs FP64, ld n)
    auto precondition = [&](double *w) -> CUresult {
      HC_CHECK(hc::d2s(s, N, K, w, N, RS, N, ovf));
      HC_CHECK(potrs_device<float>(s, 'L', N, K, S, N, static_cast<const float *>(T), RS, N,
                                   static_cast<float *>(Yw), sync));
      HC_CHECK(hc::s2d(s, N, K, RS, N, w, N));
      return CUDA_SUCCESS;
    };
    // GMRES state on the device (gmres_* kernels): the Hessenberg columns,
    // rotations and basis scales never leave it.  The host reads one int per
    // inner iteration, the number of right-hand sides not yet converged, to
    // decide whether to queue the next one (reading it one iteration late
    // instead cost a whole extra iteration per cycle: ~5 ms at n = 32768).
    double *Gs = reinterpret_cast<double *>(wsb + L.GM);
    double *g_scale = HD + 2 * K * NV;  // basis scales
    double *ones = Gs, *beta = ones + K, *gv = beta + K, *yv = gv + K * NV, *Hm = yv + K * NV;
    double *csv = Hm + K * kGmresM * NV, *snv = csv + K * kGmresM;
    int *act = reinterpret_cast<int *>(snv + K * kGmresM);
    bool ones_ready = false;
    hipEvent_t ev = nullptr;
    auto gmres_cycle = [&]() -> CUresult {
      if (!ones_ready) {
        const std::vector<double> one(K, 1.0);
        HC_CHECK(hipMemcpyAsync(ones, one.data(), K * sizeof(double), hipMemcpyHostToDevice, s));
        HC_CHECK(hipStreamSynchronize(s));
        HC_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        ones_ready = true;
      }
      HC_CHECK(hipMemcpyAsync(WV, R, (size_t)N * K * sizeof(double), hipMemcpyDeviceToDevice, s));
      CU_TRY(precondition(WV));
      HC_CHECK(hc::basis_put(s, N, K, NV, 0, WV, N, ones, V));
      HC_CHECK(hc::basis_dots(s, N, K, NV, 0, 1, V, WV, N, HD));
      HC_CHECK(hc::gmres_start(s, K, NV, HD, beta, gv, g_scale));
      HC_CHECK(hipMemsetAsync(act, 0, kGmresM * sizeof(int), s));
      HC_CHECK(hc::basis_put(s, N, K, NV, 0, WV, N, g_scale, V));
      int used_j = 0;
      for (int j = 0; j < kGmresM; j++) {
        HC_CHECK(symm_apply(s, ul, N, K, Ad, (long)lda, V + (long)j * N, (long)NV * N, WV, N, PART));
        CU_TRY(precondition(WV));
        for (int pass = 0; pass < 2; pass++) {
          double *hd = HD + pass * K * NV;
          HC_CHECK(hc::basis_dots(s, N, K, NV, 0, j + 1, V, WV, N, hd));
          HC_CHECK(hc::basis_sub(s, N, K, NV, j + 1, V, hd, WV, N));
        }
        HC_CHECK(hc::basis_put(s, N, K, NV, j + 1, WV, N, ones, V));
        HC_CHECK(hc::basis_dots(s, N, K, NV, j + 1, 1, V, WV, N, HD));  // ||w||^2 at slot j + 1
        HC_CHECK(hc::gmres_rotate(s, K, NV, kGmresM, j, HD, HD, Hm, csv, snv, gv, beta, g_scale, kGmresTol,
                                  act));
        HC_CHECK(hc::basis_put(s, N, K, NV, j + 1, WV, N, g_scale, V));
        HC_CHECK(hipMemcpyAsync(h->hinfo + 1, act + j, sizeof(int), hipMemcpyDeviceToHost, s));
        HC_CHECK(hipEventRecord(ev, s));
        used_j = j + 1;
        HC_CHECK(hipEventSynchronize(ev));
        if (h->hinfo[1] == 0) break;  // every right-hand side converged
      }
      // y = H^-1 g (upper triangular), then x += V y
      HC_CHECK(hc::gmres_solve_y(s, K, NV, kGmresM, used_j, Hm, gv, yv));
      HC_CHECK(hc::basis_combine(s, N, K, NV, used_j, V, yv, Xd, (long)ldx));
      return CUDA_SUCCESS;
    };
    struct EventGuard {
      hipEvent_t *e;
      ~EventGuard() {
        if (*e != nullptr) (void)hipEventDestroy(*e);
      }
    } event_guard{&ev};

    Norms nm;
    const long itmax = refine == 'G' ? kGmresOuterMax : kIterMax;
    for (long it = 0;; it++) {
      HC_CHECK(residual(s, ul, N, K, Ad, (long)lda, Bd, (long)ldb, Xd, (long)ldx, R, N, PART));
      // the FP32 residual of the next correction, formed ahead of the norms'
      // wait (unused once converged)
      if (refine != 'G') HC_CHECK(hc::d2s(s, N, K, R, N, RS, N, ovf));
      CU_TRY(read_norms(nm, refine != 'G' ? &flag : nullptr));
      if (converged(nm)) { *code = it; return CUDA_SUCCESS; }
      if (it == itmax) { *code = -(itmax + 1); return CUDA_SUCCESS; }
      if (refine == 'G') {
        CU_TRY(gmres_cycle());
        continue;
      }
      if (flag) { *code = -2; return CUDA_SUCCESS; }
      HC_CHECK(potrs_device<float>(s, 'L', N, K, S, N, static_cast<const float *>(T), RS, N,
                                   static_cast<float *>(Yw), sync));
      HC_CHECK(hc::add_s2d(s, N, K, RS, N, Xd, (long)ldx));
    }
  };

  HC_CHECK(hipMemsetAsync(sync, 0, sizeof(int), s));
  long code = 0;
  CUresult r = attempt(factor, &code);
  if (r != CUDA_SUCCESS) return r;
  if (code < 0 && factor != 'S') {  // 16-bit products too coarse: FP32 next
    r = attempt('S', &code);
    if (r != CUDA_SUCCESS) return r;
    if (used != nullptr) *used = 'S';
  }
  *iter = code;
  if (code >= 0) return solve_status(s, sync);

  // Full FP64 factorisation and solve; A is overwritten by its factor.
  if (used != nullptr) *used = 'D';
  // (the workspace of the attempts, sized at entry, still holds W, T, Y and
  // the sync words used below)
  double *Aw = dptr<double>(A);
  HC_CHECK(hipMemsetAsync(h->dinfo, 0, sizeof(int), s));
  HC_CHECK(hc::potrf_on_device<double>(h, ul, N, Aw, (long)lda,
                                       reinterpret_cast<double *>(wsb + L.W), false));
  HC_CHECK(hipMemcpyAsync(h->hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HC_CHECK(hipStreamSynchronize(s));
  *info = h->hinfo[0];
  if (*info == hc::kInfoAbandoned) {  // a dependency wait gave up
    *info = 0;
    return CUDA_ERROR_LAUNCH_FAILED;
  }
  if (*info != 0) return CUDA_SUCCESS;
  HC_CHECK(hc::copy2d<double>(s, N, K, Bd, (long)ldb, Xd, (long)ldx));
  HC_CHECK(potrs_prepare<double>(s, ul, N, Aw, (long)lda, static_cast<double *>(T)));
  HC_CHECK(potrs_device<double>(s, ul, N, K, Aw, (long)lda, static_cast<const double *>(T), Xd,
                                (long)ldx, static_cast<double *>(Yw), sync));
  return solve_status(s, sync);
}

}  // namespace

extern "C" {

CUresult cuDsposvx(CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                   size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, char factor,
                   char refine, long *iter, char *used, long *info) {
  return sposv("cuDsposvx", h, uplo, n, nrhs, A, lda, B, ldb, X, ldx, factor, refine, iter, used,
               info);
}

CUresult cuDsposv(CULAPACKhandle h, CBlasUplo uplo, size_t n, size_t nrhs, CUdeviceptr A,
                  size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr X, size_t ldx, long *iter,
                  long *info) {
  return sposv("cuDsposv", h, uplo, n, nrhs, A, lda, B, ldb, X, ldx, 'S', 'I', iter, nullptr, info);
}

}  // extern "C"
