// Host-pointer BLAS and LAPACK of the reference's multi-GPU layer
// (include/cumultigpu.h, include/blas.h:257-360, include/lapack.h:117-158),
// collapsed onto one device.
//
// The reference tiles each host-resident operand over the GPUs of a
// CUmultiGPU context and runs a task per tile.  One MI355X holds 288 GB, so
// here every call moves its operands to the device whole, runs the
// device-pointer routine of this library on them and moves the result back:
//   * host -> device and device -> host copies go through two pinned staging
//     buffers (hipHostMalloc) of kStageBytes each, so that packing one piece
//     of a column-major host matrix overlaps the DMA of the previous piece;
//   * device operands are private to the handle, packed with a leading
//     dimension rounded up to kLdAlign elements, and grow on demand;
//   * an operand a call does not read (A and B when alpha = 0, C of a GEMM
//     when beta = 0) is not copied, so NaNs in it cannot leak into the result,
//     exactly as the reference's CPU quick paths leave them unread;
//   * a symmetric / triangular C is copied whole, so the triangle the
//     routine does not reference comes back bit-identical.
// Every call returns when the result is in the caller's host array; the
// Synchronize calls therefore only wait for the device.
#include "api_util.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <type_traits>
#include <vector>

#include "cumultigpu.h"
#include "lapack.h"

using hc::bad_arg;

// ---------------------------------------------------------------- tasks

struct __cutask_st {
  CUresult (*function)(const void *);
  void *args;
  CUresult result;
};

struct __cumultigpu_st {
  int device;
};

namespace {

constexpr size_t kStageBytes = size_t(32) << 20;       // per pinned staging buffer
constexpr size_t kLdAlign = 8;                         // device leading dimensions, elements
constexpr int kPackThreads = 4;                        // host threads packing one piece
constexpr size_t kPackParallelBytes = size_t(4) << 20;  // below this one thread packs

size_t ld_of(size_t rows) { return std::max<size_t>(kLdAlign, (rows + kLdAlign - 1) / kLdAlign * kLdAlign); }

// Copies `cols` columns of `width` bytes between two column-major layouts
// (strides ldd, lds in bytes); large rectangles are shared out by columns over
// a few threads, since one core cannot keep up with the DMA engines.
void copy_rect(char *dst, size_t ldd, const char *src, size_t lds, size_t width, size_t cols) {
  auto band = [=](size_t c0, size_t c1) {
    if (ldd == width && lds == width) {
      std::memcpy(dst + c0 * width, src + c0 * width, (c1 - c0) * width);
      return;
    }
    for (size_t c = c0; c < c1; c++) std::memcpy(dst + c * ldd, src + c * lds, width);
  };
  const size_t nt = std::min<size_t>(kPackThreads, cols);
  if (width * cols < kPackParallelBytes || nt < 2) {
    band(0, cols);
    return;
  }
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; t++) pool.emplace_back(band, cols * t / nt, cols * (t + 1) / nt);
  band(0, cols / nt);
  for (auto &th : pool) th.join();
}

struct Piece {
  size_t r0, rows, c0, cols;
};

// Pieces of a rows x cols matrix that fit one staging buffer: whole columns
// when a column fits, row ranges of single columns otherwise.
std::vector<Piece> pieces(size_t rows, size_t cols, size_t elem) {
  std::vector<Piece> ps;
  const size_t pr = std::min(rows, kStageBytes / elem);
  const size_t pc = std::max<size_t>(1, std::min(cols, kStageBytes / (pr * elem)));
  for (size_t c0 = 0; c0 < cols; c0 += pc)
    for (size_t r0 = 0; r0 < rows; r0 += pr)
      ps.push_back({r0, std::min(pr, rows - r0), c0, std::min(pc, cols - c0)});
  return ps;
}

// Pinned double-buffered staging, a copy stream and the handle's device
// operands.
class Staging {
 public:
  static constexpr int kSlots = 4;  // A, B, C, X

  hipError_t init() {
    hipError_t e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
    for (int b = 0; b < 2 && e == hipSuccess; b++) {
      e = hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev_[b], stream_);  // both buffers start free
    }
    return e;
  }

  hipError_t release() {
    hipError_t first = hipSuccess;
    auto keep = [&](hipError_t e) {
      if (first == hipSuccess) first = e;
    };
    if (stream_ != nullptr) keep(hipStreamSynchronize(stream_));
    for (int b = 0; b < 2; b++) {
      if (pinned_[b] != nullptr) keep(hipHostFree(pinned_[b]));
      if (ev_[b] != nullptr) keep(hipEventDestroy(ev_[b]));
      pinned_[b] = nullptr;
      ev_[b] = nullptr;
    }
    for (int s = 0; s < kSlots; s++) {
      if (dev_[s] != nullptr) keep(hipFree(dev_[s]));
      dev_[s] = nullptr;
      cap_[s] = 0;
    }
    if (stream_ != nullptr) keep(hipStreamDestroy(stream_));
    stream_ = nullptr;
    return first;
  }

  hipStream_t stream() const { return stream_; }

  // Device operand `slot` of at least `bytes` bytes (its old contents are not
  // kept when it grows).
  hipError_t buffer(int slot, size_t bytes, void **p) {
    if (bytes > cap_[slot]) {
      if (dev_[slot] != nullptr) {
        hipError_t e = hipStreamSynchronize(stream_);
        if (e == hipSuccess) e = hipFree(dev_[slot]);
        dev_[slot] = nullptr;
        cap_[slot] = 0;
        if (e != hipSuccess) return e;
      }
      hipError_t e = hipMalloc(&dev_[slot], bytes);
      if (e != hipSuccess) {
        dev_[slot] = nullptr;
        return e;
      }
      cap_[slot] = bytes;
    }
    *p = dev_[slot];
    return hipSuccess;
  }

  // host (ldh) -> device (ldd), rows x cols elements of `elem` bytes; queued
  // on the copy stream, the host array free again on return.
  hipError_t upload(void *dev, size_t ldd, const void *host, size_t ldh, size_t rows, size_t cols,
                    size_t elem) {
    if (rows == 0 || cols == 0) return hipSuccess;
    hipError_t e = ensure_pinned();
    if (e != hipSuccess) return e;
    const std::vector<Piece> ps = pieces(rows, cols, elem);
    for (size_t i = 0; i < ps.size(); i++) {
      const Piece &p = ps[i];
      const int b = (int)(i & 1);
      if ((e = hipEventSynchronize(ev_[b])) != hipSuccess) return e;  // its last DMA is done
      copy_rect(static_cast<char *>(pinned_[b]), p.rows * elem,
                static_cast<const char *>(host) + (p.c0 * ldh + p.r0) * elem, ldh * elem, p.rows * elem,
                p.cols);
      e = hipMemcpy2DAsync(static_cast<char *>(dev) + (p.c0 * ldd + p.r0) * elem, ldd * elem, pinned_[b],
                           p.rows * elem, p.rows * elem, p.cols, hipMemcpyHostToDevice, stream_);
      if (e == hipSuccess) e = hipEventRecord(ev_[b], stream_);
      if (e != hipSuccess) return e;
    }
    // the last pieces are still in flight from pinned memory only
    return hipSuccess;
  }

  // device (ldd) -> host (ldh), after everything queued on the copy stream;
  // the data is in the host array on return.
  hipError_t download(void *host, size_t ldh, const void *dev, size_t ldd, size_t rows, size_t cols,
                      size_t elem) {
    if (rows == 0 || cols == 0) return hipSuccess;
    hipError_t e = ensure_pinned();
    if (e != hipSuccess) return e;
    const std::vector<Piece> ps = pieces(rows, cols, elem);
    auto issue = [&](size_t i) {
      const Piece &p = ps[i];
      const int b = (int)(i & 1);
      hipError_t r = hipMemcpy2DAsync(pinned_[b], p.rows * elem, static_cast<const char *>(dev) + (p.c0 * ldd + p.r0) * elem,
                                      ldd * elem, p.rows * elem, p.cols, hipMemcpyDeviceToHost, stream_);
      if (r == hipSuccess) r = hipEventRecord(ev_[b], stream_);
      return r;
    };
    if ((e = issue(0)) != hipSuccess) return e;
    for (size_t i = 0; i < ps.size(); i++) {
      // the buffer of piece i + 1 held piece i - 1, already unpacked
      if (i + 1 < ps.size() && (e = issue(i + 1)) != hipSuccess) return e;
      const Piece &p = ps[i];
      const int b = (int)(i & 1);
      if ((e = hipEventSynchronize(ev_[b])) != hipSuccess) return e;
      copy_rect(static_cast<char *>(host) + (p.c0 * ldh + p.r0) * elem, ldh * elem,
                static_cast<const char *>(pinned_[b]), p.rows * elem, p.rows * elem, p.cols);
    }
    return hipSuccess;
  }

 private:
  hipError_t ensure_pinned() {
    for (int b = 0; b < 2; b++)
      if (pinned_[b] == nullptr) {
        hipError_t e = hipHostMalloc(&pinned_[b], kStageBytes, hipHostMallocDefault);
        if (e != hipSuccess) {
          pinned_[b] = nullptr;
          return e;
        }
      }
    return hipSuccess;
  }

  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  void *pinned_[2] = {nullptr, nullptr};
  void *dev_[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap_[kSlots] = {0, 0, 0, 0};
};

}  // namespace

struct __cumultigpublashandle_st {
  CUmultiGPU mgpu;
  CUBLAShandle blas;
  Staging st;
};

struct __cumultigpulapackhandle_st {
  CUmultiGPU mgpu;
  CULAPACKhandle lapack;
  Staging st;
};

namespace {

using zc = cuDoubleComplex_t;
using cc = cuComplex_t;

bool is_zero(double v) { return v == 0.0; }
bool is_zero(float v) { return v == 0.0f; }
bool is_zero(zc v) { return __real__ v == 0.0 && __imag__ v == 0.0; }
bool is_zero(cc v) { return __real__ v == 0.0f && __imag__ v == 0.0f; }
bool is_one(double v) { return v == 1.0; }
bool is_one(float v) { return v == 1.0f; }
bool is_one(zc v) { return __real__ v == 1.0 && __imag__ v == 0.0; }
bool is_one(cc v) { return __real__ v == 1.0f && __imag__ v == 0.0f; }
template <typename T> constexpr bool complex_v() { return std::is_same<T, zc>::value || std::is_same<T, cc>::value; }
template <typename T>
T zero_of() {
  T z{};
  return z;
}

CUdeviceptr dp(void *p) { return reinterpret_cast<CUdeviceptr>(p); }

// Device operand `slot` of ld x cols elements, filled from the host when
// `host` is given.
template <typename T>
hipError_t stage_in(Staging &s, int slot, const T *host, size_t ldh, size_t rows, size_t cols, size_t ld,
                    void **dev) {
  hipError_t e = s.buffer(slot, std::max<size_t>(1, ld * cols) * sizeof(T), dev);
  if (e == hipSuccess && host != nullptr) e = s.upload(*dev, ld, host, ldh, rows, cols, sizeof(T));
  return e;
}

// ---------------------------------------------------------------- BLAS

template <typename T>
CUresult mgemm(const char *name, CUmultiGPUBLAShandle h, CBlasTranspose ta, CBlasTranspose tb, size_t m,
               size_t n, size_t k, T alpha, const T *A, size_t lda, const T *B, size_t ldb, T beta, T *C,
               size_t ldc) {
  // reference blas/dgemm.c:438-453
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_trans(ta)) return bad_arg(name, 2);
  if (!hc::valid_trans(tb)) return bad_arg(name, 3);
  const size_t nRowA = (ta == CBlasNoTrans) ? m : k;
  const size_t nRowB = (tb == CBlasNoTrans) ? k : n;
  if (lda < nRowA) return bad_arg(name, 8);
  if (ldb < nRowB) return bad_arg(name, 10);
  if (ldc < m) return bad_arg(name, 13);
  if (m == 0 || n == 0 || ((is_zero(alpha) || k == 0) && is_one(beta))) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->mgpu->device);
  HC_CHECK(g.status());
  Staging &s = h->st;
  const size_t kk = is_zero(alpha) ? 0 : k;  // A and B unread when alpha = 0
  const size_t rA = (ta == CBlasNoTrans) ? m : kk, cA = (ta == CBlasNoTrans) ? kk : m;
  const size_t rB = (tb == CBlasNoTrans) ? kk : n, cB = (tb == CBlasNoTrans) ? n : kk;
  const size_t ldA = ld_of(rA), ldB = ld_of(rB), ldC = ld_of(m);
  void *dA, *dB, *dC;
  HC_CHECK(stage_in(s, 0, A, lda, rA, cA, ldA, &dA));
  HC_CHECK(stage_in(s, 1, B, ldb, rB, cB, ldB, &dB));
  HC_CHECK(stage_in(s, 2, is_zero(beta) ? nullptr : C, ldc, m, n, ldC, &dC));  // C unread when beta = 0
  CUresult r;
  if constexpr (std::is_same<T, double>::value)
    r = cuDgemm2(h->blas, ta, tb, m, n, kk, alpha, dp(dA), ldA, dp(dB), ldB, beta, dp(dC), ldC, dp(dC), ldC,
                 s.stream());
  else if constexpr (std::is_same<T, float>::value)
    r = cuSgemm2(h->blas, ta, tb, m, n, kk, alpha, dp(dA), ldA, dp(dB), ldB, beta, dp(dC), ldC, dp(dC), ldC,
                 s.stream());
  else if constexpr (std::is_same<T, cc>::value)
    r = cuCgemm2(h->blas, ta, tb, m, n, kk, alpha, dp(dA), ldA, dp(dB), ldB, beta, dp(dC), ldC, dp(dC), ldC,
                 s.stream());
  else
    r = cuZgemm2(h->blas, ta, tb, m, n, kk, alpha, dp(dA), ldA, dp(dB), ldB, beta, dp(dC), ldC, dp(dC), ldC,
                 s.stream());
  if (r != CUDA_SUCCESS) return r;
  HC_CHECK(s.download(C, ldc, dC, ldC, m, n, sizeof(T)));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult mherk(const char *name, CUmultiGPUBLAShandle h, CBlasUplo uplo, CBlasTranspose trans, size_t n,
               size_t k, double alpha, const T *A, size_t lda, double beta, T *C, size_t ldc) {
  constexpr bool cplx = complex_v<T>();
  // reference blas/dsyrk.c:216-226, blas/zherk.c:245-257 (ZHERK rejects
  // trans = T with index 2)
  if (h == nullptr) return bad_arg(name, 1);
  if (cplx && trans == CBlasTrans) return bad_arg(name, 2);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (!hc::valid_trans(trans)) return bad_arg(name, 3);
  const size_t nRowA = (trans == CBlasNoTrans) ? n : k;
  if (lda < nRowA) return bad_arg(name, 7);
  if (ldc < n) return bad_arg(name, 10);
  if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->mgpu->device);
  HC_CHECK(g.status());
  Staging &s = h->st;
  const size_t kk = (alpha == 0.0) ? 0 : k;
  const size_t rA = (trans == CBlasNoTrans) ? n : kk, cA = (trans == CBlasNoTrans) ? kk : n;
  const size_t ldA = ld_of(rA), ldC = ld_of(n);
  void *dA, *dC;
  HC_CHECK(stage_in(s, 0, A, lda, rA, cA, ldA, &dA));
  // C whole, also when beta = 0 (unread then): its other triangle comes back as it was
  HC_CHECK(stage_in(s, 2, C, ldc, n, n, ldC, &dC));
  CUresult r;
  if constexpr (std::is_same<T, zc>::value)
    r = cuZherk(h->blas, uplo, trans, n, kk, alpha, dp(dA), ldA, beta, dp(dC), ldC, s.stream());
  else if constexpr (std::is_same<T, cc>::value)
    r = cuCherk(h->blas, uplo, trans, n, kk, (float)alpha, dp(dA), ldA, (float)beta, dp(dC), ldC, s.stream());
  else if constexpr (std::is_same<T, float>::value)
    r = cuSsyrk(h->blas, uplo, trans, n, kk, (float)alpha, dp(dA), ldA, (float)beta, dp(dC), ldC, s.stream());
  else
    r = cuDsyrk(h->blas, uplo, trans, n, kk, alpha, dp(dA), ldA, beta, dp(dC), ldC, s.stream());
  if (r != CUDA_SUCCESS) return r;
  HC_CHECK(s.download(C, ldc, dC, ldC, n, n, sizeof(T)));
  return CUDA_SUCCESS;
}

// Triangular multiply (TRSM = false) or solve (TRSM = true), in place in B.
template <typename T, bool TRSM>
CUresult mtri(const char *name, CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
              CBlasDiag diag, size_t m, size_t n, T alpha, const T *A, size_t lda, T *B, size_t ldb) {
  // reference blas/dtrmm.c:472-483, blas/dtrsm.c:272-283
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_side(side)) return bad_arg(name, 2);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 3);
  if (!hc::valid_trans(trans)) return bad_arg(name, 4);
  if (!hc::valid_diag(diag)) return bad_arg(name, 5);
  const size_t nRowA = (side == CBlasLeft) ? m : n;
  if (lda < nRowA) return bad_arg(name, 9);
  if (ldb < m) return bad_arg(name, 11);
  if (m == 0 || n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->mgpu->device);
  HC_CHECK(g.status());
  Staging &s = h->st;
  const bool zero = is_zero(alpha);  // B <- 0, neither A nor B read (reference blas/dtrmm.c:489-492)
  const size_t ldA = ld_of(nRowA), ldB = ld_of(m);
  void *dA, *dB, *dX;
  HC_CHECK(stage_in(s, 0, zero ? nullptr : A, lda, nRowA, nRowA, ldA, &dA));
  HC_CHECK(stage_in(s, 1, zero ? nullptr : B, ldb, m, n, ldB, &dB));
  CUresult r;
  void *out = dB;
  if (zero) {  // B <- 0 on the device (a GEMM with k = 0 and beta = 0 reads nothing)
    const T z = zero_of<T>();
    if constexpr (std::is_same<T, double>::value)
      r = cuDgemm2(h->blas, CBlasNoTrans, CBlasNoTrans, m, n, 0, z, dp(dB), ldB, dp(dB), ldB, z, dp(dB), ldB,
                   dp(dB), ldB, s.stream());
    else if constexpr (std::is_same<T, float>::value)
      r = cuSgemm2(h->blas, CBlasNoTrans, CBlasNoTrans, m, n, 0, z, dp(dB), ldB, dp(dB), ldB, z, dp(dB), ldB,
                   dp(dB), ldB, s.stream());
    else if constexpr (std::is_same<T, cc>::value)
      r = cuCgemm2(h->blas, CBlasNoTrans, CBlasNoTrans, m, n, 0, z, dp(dB), ldB, dp(dB), ldB, z, dp(dB), ldB,
                   dp(dB), ldB, s.stream());
    else
      r = cuZgemm2(h->blas, CBlasNoTrans, CBlasNoTrans, m, n, 0, z, dp(dB), ldB, dp(dB), ldB, z, dp(dB), ldB,
                   dp(dB), ldB, s.stream());
  } else if constexpr (TRSM) {
    if constexpr (std::is_same<T, double>::value)
      r = cuDtrsm(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, s.stream());
    else if constexpr (std::is_same<T, float>::value)
      r = cuStrsm(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, s.stream());
    else if constexpr (std::is_same<T, cc>::value)
      r = cuCtrsm(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, s.stream());
    else
      r = cuZtrsm(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, s.stream());
  } else {
    HC_CHECK(stage_in<T>(s, 3, nullptr, 0, m, n, ldB, &dX));  // the device multiply is out of place
    if constexpr (std::is_same<T, double>::value)
      r = cuDtrmm2(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, dp(dX), ldB,
                   s.stream());
    else if constexpr (std::is_same<T, float>::value)
      r = cuStrmm2(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, dp(dX), ldB,
                   s.stream());
    else if constexpr (std::is_same<T, cc>::value)
      r = cuCtrmm2(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, dp(dX), ldB,
                   s.stream());
    else
      r = cuZtrmm2(h->blas, side, uplo, trans, diag, m, n, alpha, dp(dA), ldA, dp(dB), ldB, dp(dX), ldB,
                   s.stream());
    out = dX;
  }
  if (r != CUDA_SUCCESS) return r;
  HC_CHECK(s.download(B, ldb, out, ldB, m, n, sizeof(T)));
  return CUDA_SUCCESS;
}

// ---------------------------------------------------------------- LAPACK

enum class Lop { potrf, trtri, lauum, potri };

// One LAPACK routine on a host matrix, in place.  A numerical failure
// (info > 0) returns CUDA_ERROR_INVALID_VALUE with the matrix as the device
// routine left it, as the reference's blocked paths do (lapack/dpotrf.c:326,
// lapack/dtrtri.c:495, lapack/dpotri.c:63).
template <typename T>
CUresult mlapack(const char *name, Lop op, CUmultiGPULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n,
                 T *A, size_t lda, long *info) {
  const long ia = (op == Lop::trtri) ? 5 : 4;  // position of lda
  if (info == nullptr) return bad_arg(name, ia + 1);
  *info = 0;
  if (h == nullptr) *info = -1;
  else if (!hc::valid_uplo(uplo)) *info = -2;
  else if (op == Lop::trtri && !hc::valid_diag(diag)) *info = -3;
  else if (lda < n) *info = -ia;  // reference lapack/dpotrf.c:302-308
  if (*info != 0) return bad_arg(name, -*info);
  if (n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->mgpu->device);
  HC_CHECK(g.status());
  Staging &s = h->st;
  const size_t ldA = ld_of(n);
  void *dA;
  HC_CHECK(stage_in(s, 0, A, lda, n, n, ldA, &dA));
  HC_CHECK(hipStreamSynchronize(s.stream()));  // the routines below run on the LAPACK handle's streams
  CULAPACKhandle L = h->lapack;
  const CUdeviceptr a = dp(dA);
  // the device routine of this precision: 0 double, 1 single, 2 double complex, 3 single complex
  constexpr int P = std::is_same<T, double>::value ? 0 : std::is_same<T, float>::value ? 1 : std::is_same<T, zc>::value ? 2 : 3;
  CUresult r = CUDA_SUCCESS;
  switch (op) {
    case Lop::potrf:
      r = P == 0 ? cuDpotrf(L, uplo, n, a, ldA, info) : P == 1 ? cuSpotrf(L, uplo, n, a, ldA, info)
          : P == 2 ? cuZpotrf(L, uplo, n, a, ldA, info) : cuCpotrf(L, uplo, n, a, ldA, info);
      break;
    case Lop::trtri:
      r = P == 0 ? cuDtrtri(L, uplo, diag, n, a, ldA, info) : P == 1 ? cuStrtri(L, uplo, diag, n, a, ldA, info)
          : P == 2 ? cuZtrtri(L, uplo, diag, n, a, ldA, info) : cuCtrtri(L, uplo, diag, n, a, ldA, info);
      break;
    case Lop::lauum:
      r = P == 0 ? cuDlauum(L, uplo, n, a, ldA, info) : P == 1 ? cuSlauum(L, uplo, n, a, ldA, info)
          : P == 2 ? cuZlauum(L, uplo, n, a, ldA, info) : cuClauum(L, uplo, n, a, ldA, info);
      break;
    case Lop::potri:
      r = P == 0 ? cuDpotri(L, uplo, n, a, ldA, info) : P == 1 ? cuSpotri(L, uplo, n, a, ldA, info)
          : P == 2 ? cuZpotri(L, uplo, n, a, ldA, info) : cuCpotri(L, uplo, n, a, ldA, info);
      break;
  }
  if (r != CUDA_SUCCESS) return r;
  HC_CHECK(hipStreamSynchronize(L->stream));
  HC_CHECK(hipStreamSynchronize(L->side));
  HC_CHECK(s.download(A, lda, dA, ldA, n, n, sizeof(T)));
  return *info > 0 ? CUDA_ERROR_INVALID_VALUE : CUDA_SUCCESS;
}

template <typename H>
CUresult destroy_handle(H *h) {
  hc::DeviceGuard g(h->mgpu->device);
  HC_CHECK(g.status());
  const hipError_t e = h->st.release();
  CUresult r = CUDA_SUCCESS;
  if constexpr (std::is_same<H, __cumultigpublashandle_st>::value)
    r = h->blas != nullptr ? cuBLASDestroy(h->blas) : CUDA_SUCCESS;
  else
    r = h->lapack != nullptr ? cuLAPACKDestroy(h->lapack) : CUDA_SUCCESS;
  delete h;
  HC_CHECK(e);
  return r;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- tasks, context

CUresult cuTaskCreate(CUtask *task, CUresult (*function)(const void *), const void *args, size_t size) {
  // reference multigpu/multigpu_seq.c:28-50
  if (task == nullptr || function == nullptr || (args == nullptr && size > 0)) return CUDA_ERROR_INVALID_VALUE;
  auto *t = static_cast<__cutask_st *>(std::malloc(sizeof(__cutask_st)));
  if (t == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  t->function = function;
  t->result = CUDA_SUCCESS;
  t->args = nullptr;
  if (size > 0) {
    t->args = std::malloc(size);
    if (t->args == nullptr) {
      std::free(t);
      return CUDA_ERROR_OUT_OF_MEMORY;
    }
    std::memcpy(t->args, args, size);
  }
  *task = t;
  return CUDA_SUCCESS;
}

CUresult cuTaskDestroy(CUtask task, CUresult *result) {
  if (task == nullptr) return CUDA_ERROR_INVALID_VALUE;
  if (result != nullptr) *result = task->result;
  std::free(task->args);
  std::free(task);
  return CUDA_SUCCESS;
}

CUresult cuTaskExecute(CUtask task) {
  if (task == nullptr) return CUDA_ERROR_INVALID_VALUE;
  task->result = task->function(task->args);
  return CUDA_SUCCESS;
}

CUresult cuMultiGPUCreate(CUmultiGPU *mGPU, CUdevice *devices, int n) {
  // reference multigpu/multigpu_seq.c:102-125: one context per device; here
  // the first device only
  if (mGPU == nullptr || devices == nullptr || n <= 0) return CUDA_ERROR_INVALID_VALUE;
  int count = 0;
  HC_CHECK(hipGetDeviceCount(&count));
  if (devices[0] < 0 || devices[0] >= count) return CUDA_ERROR_INVALID_DEVICE;
  auto *c = static_cast<__cumultigpu_st *>(std::malloc(sizeof(__cumultigpu_st)));
  if (c == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  c->device = devices[0];
  *mGPU = c;
  return CUDA_SUCCESS;
}

CUresult cuMultiGPUDestroy(CUmultiGPU mGPU) {
  if (mGPU == nullptr) return CUDA_ERROR_INVALID_VALUE;
  std::free(mGPU);
  return CUDA_SUCCESS;
}

int cuMultiGPUGetContextCount(CUmultiGPU mGPU) { return mGPU == nullptr ? 0 : 1; }

CUresult cuMultiGPURunTask(CUmultiGPU mGPU, int i, CUtask task) {
  // reference multigpu/multigpu_seq.c:144-159
  if (mGPU == nullptr || task == nullptr || i != 0) return CUDA_ERROR_INVALID_VALUE;
  hc::DeviceGuard g(mGPU->device);
  HC_CHECK(g.status());
  return cuTaskExecute(task);
}

CUresult cuMultiGPUSynchronize(CUmultiGPU mGPU) {
  if (mGPU == nullptr) return CUDA_ERROR_INVALID_VALUE;
  hc::DeviceGuard g(mGPU->device);
  HC_CHECK(g.status());
  HC_CHECK(hipDeviceSynchronize());
  return CUDA_SUCCESS;
}

// ---------------------------------------------------------------- BLAS handle

CUresult cuMultiGPUBLASCreate(CUmultiGPUBLAShandle *handle, CUmultiGPU mGPU) {
  // reference blas/handle.c:109-133
  if (handle == nullptr || mGPU == nullptr) return CUDA_ERROR_INVALID_VALUE;
  hc::DeviceGuard g(mGPU->device);
  HC_CHECK(g.status());
  auto *h = new (std::nothrow) __cumultigpublashandle_st();
  if (h == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  h->mgpu = mGPU;
  h->blas = nullptr;
  CUresult r = cuBLASCreate(&h->blas);
  if (r == CUDA_SUCCESS) r = hc::to_cu(h->st.init());
  if (r != CUDA_SUCCESS) {
    (void)destroy_handle(h);
    return r;
  }
  *handle = h;
  return CUDA_SUCCESS;
}

CUresult cuMultiGPUBLASDestroy(CUmultiGPUBLAShandle handle) {
  if (handle == nullptr) return CUDA_ERROR_INVALID_VALUE;
  return destroy_handle(handle);
}

CUresult cuMultiGPUBLASSynchronize(CUmultiGPUBLAShandle handle) {
  if (handle == nullptr) return CUDA_ERROR_INVALID_VALUE;
  hc::DeviceGuard g(handle->mgpu->device);
  HC_CHECK(g.status());
  HC_CHECK(hipStreamSynchronize(handle->st.stream()));
  return CUDA_SUCCESS;
}

CUresult cuMultiGPUDsyrk(CUmultiGPUBLAShandle h, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                         double alpha, const double *A, size_t lda, double beta, double *C, size_t ldc) {
  return mherk<double>("cuMultiGPUDsyrk", h, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

CUresult cuMultiGPUZherk(CUmultiGPUBLAShandle h, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                         double alpha, const zc *A, size_t lda, double beta, zc *C, size_t ldc) {
  return mherk<zc>("cuMultiGPUZherk", h, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

CUresult cuMultiGPUDgemm(CUmultiGPUBLAShandle h, CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n,
                         size_t k, double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                         double beta, double *C, size_t ldc) {
  return mgemm<double>("cuMultiGPUDgemm", h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

CUresult cuMultiGPUZgemm(CUmultiGPUBLAShandle h, CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n,
                         size_t k, zc alpha, const zc *A, size_t lda, const zc *B, size_t ldb, zc beta, zc *C,
                         size_t ldc) {
  return mgemm<zc>("cuMultiGPUZgemm", h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

CUresult cuMultiGPUDtrmm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, double alpha, const double *A, size_t lda,
                         double *B, size_t ldb) {
  return mtri<double, false>("cuMultiGPUDtrmm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

CUresult cuMultiGPUZtrmm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, zc alpha, const zc *A, size_t lda, zc *B,
                         size_t ldb) {
  return mtri<zc, false>("cuMultiGPUZtrmm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

CUresult cuMultiGPUDtrsm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, double alpha, const double *A, size_t lda,
                         double *B, size_t ldb) {
  return mtri<double, true>("cuMultiGPUDtrsm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

CUresult cuMultiGPUZtrsm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, zc alpha, const zc *A, size_t lda, zc *B,
                         size_t ldb) {
  return mtri<zc, true>("cuMultiGPUZtrsm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

// single and single complex (reference blas/ssyrk.c:214, blas/cherk.c:240,
// blas/sgemm.c:432, blas/cgemm.c:509, blas/strmm.c:467, blas/ctrmm.c:518,
// blas/strsm.c:267, blas/ctrsm.c:296)
CUresult cuMultiGPUSsyrk(CUmultiGPUBLAShandle h, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                         float alpha, const float *A, size_t lda, float beta, float *C, size_t ldc) {
  return mherk<float>("cuMultiGPUSsyrk", h, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

CUresult cuMultiGPUCherk(CUmultiGPUBLAShandle h, CBlasUplo uplo, CBlasTranspose trans, size_t n, size_t k,
                         float alpha, const cc *A, size_t lda, float beta, cc *C, size_t ldc) {
  return mherk<cc>("cuMultiGPUCherk", h, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

CUresult cuMultiGPUSgemm(CUmultiGPUBLAShandle h, CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n,
                         size_t k, float alpha, const float *A, size_t lda, const float *B, size_t ldb,
                         float beta, float *C, size_t ldc) {
  return mgemm<float>("cuMultiGPUSgemm", h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

CUresult cuMultiGPUCgemm(CUmultiGPUBLAShandle h, CBlasTranspose ta, CBlasTranspose tb, size_t m, size_t n,
                         size_t k, cc alpha, const cc *A, size_t lda, const cc *B, size_t ldb, cc beta, cc *C,
                         size_t ldc) {
  return mgemm<cc>("cuMultiGPUCgemm", h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

CUresult cuMultiGPUStrmm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, float alpha, const float *A, size_t lda, float *B,
                         size_t ldb) {
  return mtri<float, false>("cuMultiGPUStrmm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

CUresult cuMultiGPUCtrmm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, cc alpha, const cc *A, size_t lda, cc *B,
                         size_t ldb) {
  return mtri<cc, false>("cuMultiGPUCtrmm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

CUresult cuMultiGPUStrsm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, float alpha, const float *A, size_t lda, float *B,
                         size_t ldb) {
  return mtri<float, true>("cuMultiGPUStrsm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

CUresult cuMultiGPUCtrsm(CUmultiGPUBLAShandle h, CBlasSide side, CBlasUplo uplo, CBlasTranspose trans,
                         CBlasDiag diag, size_t m, size_t n, cc alpha, const cc *A, size_t lda, cc *B,
                         size_t ldb) {
  return mtri<cc, true>("cuMultiGPUCtrsm", h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

// ---------------------------------------------------------------- LAPACK handle

CUresult cuMultiGPULAPACKCreate(CUmultiGPULAPACKhandle *handle, CUmultiGPU mGPU) {
  // reference lapack/handle.c:90-97
  if (handle == nullptr || mGPU == nullptr) return CUDA_ERROR_INVALID_VALUE;
  hc::DeviceGuard g(mGPU->device);
  HC_CHECK(g.status());
  auto *h = new (std::nothrow) __cumultigpulapackhandle_st();
  if (h == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  h->mgpu = mGPU;
  h->lapack = nullptr;
  CUresult r = cuLAPACKCreate(&h->lapack);
  if (r == CUDA_SUCCESS) r = hc::to_cu(h->st.init());
  if (r != CUDA_SUCCESS) {
    (void)destroy_handle(h);
    return r;
  }
  *handle = h;
  return CUDA_SUCCESS;
}

CUresult cuMultiGPULAPACKDestroy(CUmultiGPULAPACKhandle handle) {
  if (handle == nullptr) return CUDA_ERROR_INVALID_VALUE;
  return destroy_handle(handle);
}

CUresult cuMultiGPULAPACKSynchronize(CUmultiGPULAPACKhandle handle) {
  if (handle == nullptr) return CUDA_ERROR_INVALID_VALUE;
  hc::DeviceGuard g(handle->mgpu->device);
  HC_CHECK(g.status());
  HC_CHECK(hipStreamSynchronize(handle->st.stream()));
  return CUDA_SUCCESS;
}

CUresult cuMultiGPUDpotrf(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, double *A, size_t lda, long *info) {
  return mlapack<double>("cuMultiGPUDpotrf", Lop::potrf, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUZpotrf(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, zc *A, size_t lda, long *info) {
  return mlapack<zc>("cuMultiGPUZpotrf", Lop::potrf, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUDtrtri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, double *A, size_t lda,
                          long *info) {
  return mlapack<double>("cuMultiGPUDtrtri", Lop::trtri, h, uplo, diag, n, A, lda, info);
}
CUresult cuMultiGPUZtrtri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, zc *A, size_t lda,
                          long *info) {
  return mlapack<zc>("cuMultiGPUZtrtri", Lop::trtri, h, uplo, diag, n, A, lda, info);
}
CUresult cuMultiGPUDlauum(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, double *A, size_t lda, long *info) {
  return mlapack<double>("cuMultiGPUDlauum", Lop::lauum, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUZlauum(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, zc *A, size_t lda, long *info) {
  return mlapack<zc>("cuMultiGPUZlauum", Lop::lauum, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUDpotri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, double *A, size_t lda, long *info) {
  return mlapack<double>("cuMultiGPUDpotri", Lop::potri, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUZpotri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, zc *A, size_t lda, long *info) {
  return mlapack<zc>("cuMultiGPUZpotri", Lop::potri, h, uplo, CBlasNonUnit, n, A, lda, info);
}

// single and single complex (reference lapack/spotrf.c:400, lapack/cpotrf.c:300,
// lapack/strtri.c:474, lapack/ctrtri.c:469, lapack/slauum.c:308,
// lapack/clauum.c:293, lapack/spotri.c:48, lapack/cpotri.c:47)
CUresult cuMultiGPUSpotrf(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, float *A, size_t lda, long *info) {
  return mlapack<float>("cuMultiGPUSpotrf", Lop::potrf, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUCpotrf(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, cc *A, size_t lda, long *info) {
  return mlapack<cc>("cuMultiGPUCpotrf", Lop::potrf, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUStrtri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, float *A, size_t lda,
                          long *info) {
  return mlapack<float>("cuMultiGPUStrtri", Lop::trtri, h, uplo, diag, n, A, lda, info);
}
CUresult cuMultiGPUCtrtri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, cc *A, size_t lda,
                          long *info) {
  return mlapack<cc>("cuMultiGPUCtrtri", Lop::trtri, h, uplo, diag, n, A, lda, info);
}
CUresult cuMultiGPUSlauum(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, float *A, size_t lda, long *info) {
  return mlapack<float>("cuMultiGPUSlauum", Lop::lauum, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUClauum(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, cc *A, size_t lda, long *info) {
  return mlapack<cc>("cuMultiGPUClauum", Lop::lauum, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUSpotri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, float *A, size_t lda, long *info) {
  return mlapack<float>("cuMultiGPUSpotri", Lop::potri, h, uplo, CBlasNonUnit, n, A, lda, info);
}
CUresult cuMultiGPUCpotri(CUmultiGPULAPACKhandle h, CBlasUplo uplo, size_t n, cc *A, size_t lda, long *info) {
  return mlapack<cc>("cuMultiGPUCpotri", Lop::potri, h, uplo, CBlasNonUnit, n, A, lda, info);
}

}  // extern "C"
