// Batched mixed-precision solve + log-determinant + inverse of SPD matrices
// of uneven sizes (no counterpart in the reference).  For every matrix i of
// the batch (packed like the batched factorisation: n_i x n_i, column-major,
// ld n_i, at element offset offA[i]; right-hand sides n_i x nrhs, ld n_i, at
// offB[i]):
//   * X_i = A_i^-1 B_i by an FP32 factor with FP64 iterative refinement
//     (LAPACK DSPOSV semantics, cuDsposvx with factor 'S', refine 'I'):
//     iter[i] is its ITER code, A_i is left unchanged when iter[i] >= 0;
//   * logdet[i] and, when Ainv is given, the triangle uplo of A_i^-1 at
//     Ainv + offA[i], from an FP64 factorisation of a copy of A_i;
//   * info[i]: 0, or the order of the leading minor that is not positive
//     definite.
// The matrices are shared out, longest first, to a pool of worker threads,
// each with its own handle (streams and workspace), so that several
// matrices are in flight on the GPU at once; a matrix is an ordered sequence
// of device routines on its worker's streams.
#include "api_util.h"

#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

#include "cucholesky.h"
#include "lapack.h"

using hc::bad_arg;

namespace {

constexpr int kMaxWorkers = 8;

struct Batch {
  char uplo;
  size_t nrhs;
  const size_t *n, *offA, *offB;
  CUdeviceptr A, B, X, Ainv;
  long *iter;
  double *logdet;
  long *info;
};

// One matrix on a worker's handle; returns the first failing CUresult.
CUresult one(CULAPACKhandle h, const Batch &b, size_t i, CUdeviceptr scratch) {
  const size_t n = b.n[i];
  b.iter[i] = 0;
  b.info[i] = 0;
  b.logdet[i] = 0.0;
  if (n == 0) return CUDA_SUCCESS;
  const CUdeviceptr Ai = b.A + b.offA[i] * sizeof(double);
  const CUdeviceptr Bi = b.B + b.offB[i] * sizeof(double);
  const CUdeviceptr Xi = b.X + b.offB[i] * sizeof(double);
  const CUdeviceptr Fi = b.Ainv != 0 ? b.Ainv + b.offA[i] * sizeof(double) : scratch;
  const CBlasUplo ul = (CBlasUplo)b.uplo;
  // the FP64 factor of a copy (logdet, inverse), made before the solve,
  // which may overwrite A_i with its own FP64 factor if refinement fails
  HC_CHECK(hipMemcpy2DAsync(reinterpret_cast<void *>(Fi), n * sizeof(double),
                            reinterpret_cast<const void *>(Ai), n * sizeof(double), n * sizeof(double), n,
                            hipMemcpyDeviceToDevice, h->stream));
  long it = 0, info = 0;
  char used = 'S';
  CUresult r = cuDsposvx(h, ul, n, b.nrhs, Ai, n, Bi, n, Xi, n, 'S', 'I', &it, &used, &info);
  if (r != CUDA_SUCCESS) return r;
  b.iter[i] = it;
  if (info != 0) {
    b.info[i] = info;
    b.logdet[i] = __builtin_nan("");
    return CUDA_SUCCESS;
  }
  r = cuDpotrf(h, ul, n, Fi, n, &info);
  if (r != CUDA_SUCCESS) return r;
  if (info != 0) {  // the solve succeeded on a matrix its FP64 factor rejects: report it
    b.info[i] = info;
    b.logdet[i] = __builtin_nan("");
    return CUDA_SUCCESS;
  }
  double ld = 0.0;
  r = cuDlogdet(h, Fi, n + 1, n, &ld, h->stream);  // not the null stream: it would wait for every worker
  if (r != CUDA_SUCCESS) return r;
  b.logdet[i] = ld;
  if (b.Ainv != 0) {
    r = cuDpotri(h, ul, n, Fi, n, &info);
    if (r != CUDA_SUCCESS) return r;
    b.info[i] = info;
  }
  return CUDA_SUCCESS;
}

}  // namespace

extern "C" {

CUresult cuDsposvxBatched(CULAPACKhandle h, CBlasUplo uplo, size_t count, const size_t *n,
                          const size_t *offA, CUdeviceptr A, size_t nrhs, const size_t *offB, CUdeviceptr B,
                          CUdeviceptr X, CUdeviceptr Ainv, long *iter, double *logdet, long *info) {
  const char *name = "cuDsposvxBatched";
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (count > 0 && n == nullptr) return bad_arg(name, 4);
  if (count > 0 && offA == nullptr) return bad_arg(name, 5);
  if (count > 0 && A == 0) return bad_arg(name, 6);
  if (count > 0 && nrhs > 0 && offB == nullptr) return bad_arg(name, 8);
  if (count > 0 && nrhs > 0 && (B == 0 || X == 0)) return bad_arg(name, B == 0 ? 9 : 10);
  if (count > 0 && (iter == nullptr || logdet == nullptr || info == nullptr))
    return bad_arg(name, iter == nullptr ? 12 : (logdet == nullptr ? 13 : 14));
  if (count == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  HC_CHECK(hipDeviceSynchronize());  // inputs written on any stream are complete

  // longest first, so that the large matrices start early
  std::vector<size_t> order(count);
  std::iota(order.begin(), order.end(), size_t(0));
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b2) { return n[a] > n[b2]; });
  const size_t nmax = n[order[0]];
  const Batch b{(char)uplo, nrhs, n, offA, offB, A, B, X, Ainv, iter, logdet, info};
  const int workers = (int)std::min<size_t>(kMaxWorkers, count);
  std::atomic<size_t> next{0};
  std::vector<CUresult> status(workers, CUDA_SUCCESS);
  const int device = h->device;
  auto work = [&](int w) {
    if (hipSetDevice(device) != hipSuccess) {
      status[w] = CUDA_ERROR_INVALID_DEVICE;
      return;
    }
    CULAPACKhandle hw = nullptr;
    CUresult r = cuLAPACKCreate(&hw);
    void *scratch = nullptr;  // the FP64 copy when no inverse is wanted
    if (r == CUDA_SUCCESS && Ainv == 0 && hipMalloc(&scratch, nmax * nmax * sizeof(double)) != hipSuccess)
      r = CUDA_ERROR_OUT_OF_MEMORY;
    while (r == CUDA_SUCCESS) {
      const size_t t = next.fetch_add(1);
      if (t >= count) break;
      r = one(hw, b, order[t], reinterpret_cast<CUdeviceptr>(scratch));
    }
    if (hw != nullptr) {
      (void)hipStreamSynchronize(hw->stream);
      (void)cuLAPACKDestroy(hw);
    }
    if (scratch != nullptr) (void)hipFree(scratch);
    status[w] = r;
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; w++) pool.emplace_back(work, w);
  work(0);
  for (auto &t : pool) t.join();
  for (CUresult r : status)
    if (r != CUDA_SUCCESS) return r;
  return CUDA_SUCCESS;
}

}  // extern "C"
