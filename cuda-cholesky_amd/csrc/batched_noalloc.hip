// Batched Cholesky factorisation + log-determinant that allocates nothing:
// cuPotrfBatchedNoAlloc (no counterpart in the reference, whose routines take
// one matrix per call; it is cuPotrfBatchedPlanCreate + cuPotrfBatchedRun +
// cuPotrfBatchedPlanDestroy in one call, for callers that cannot allocate:
// code under HIP graph capture, tight serving loops).
//
// The matrix sizes and offsets are device arrays.  The only memory beyond the
// caller's inputs and outputs is `scratch` below, a fixed array in the
// library's device image (each device holds its own copy): the class counts
// and the size-sorted matrix order that the batched kernels read, built by a
// one-workgroup counting sort on the caller's stream.  A call therefore
// launches four kernels (the sort, then the three size classes of the
// register-resident factorisation) and makes no allocation, no copy and no
// host synchronisation.  Every call on a device shares that device's array:
// calls that overlap on one device (different streams, no ordering between
// them) race.
#include <algorithm>
#include <mutex>

#include "api_util.h"
#include "cucholesky.h"

namespace {

constexpr int kNoAllocMax = 1 << 16;  // matrices per call
constexpr int kScratchHead = 64;      // class counts (3 ints), rest unused: keeps order 256-byte aligned

// [0, 3): matrices per size class; [kScratchHead, kScratchHead + count): the
// order (indices by decreasing size class, then decreasing n).
static __device__ __align__(256) int scratch[kScratchHead + kNoAllocMax];

constexpr int kMaxDevices = 64;
struct DeviceSetup {
  int *ptr = nullptr;     // this device's copy of `scratch`
  int grid[2][2][3] = {};  // persistent grid per precision (D, S), uplo (L, U), size class
};
std::mutex g_setup_mutex;
DeviceSetup g_setup[kMaxDevices];

// One workgroup: counting sort of the matrices by n (0..kBatchedMaxN), largest
// first; class counts to cnt, the order to ord.  A matrix whose n is out of
// range gets info = -1 and logdet = NaN and is left out.
constexpr int kSortThreads = 1024;
__global__ void __launch_bounds__(kSortThreads)
noalloc_sort_kernel(int count, const int *nvec, int *cnt, int *ord, double *logdet, int *info) {
  constexpr int NB = hc::kBatchedMaxN + 1;
  __shared__ int hist[NB], pos[NB];
  for (int v = threadIdx.x; v < NB; v += kSortThreads) hist[v] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < count; i += kSortThreads) {
    const int v = nvec[i];
    if (v >= 0 && v <= hc::kBatchedMaxN) {
      atomicAdd(&hist[v], 1);
    } else {
      info[i] = -1;
      logdet[i] = __builtin_nan("");
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0, c[3] = {0, 0, 0};
    for (int v = NB - 1; v >= 0; v--) {  // decreasing n
      pos[v] = run;
      run += hist[v];
      c[hc::batched_reg_class(v)] += hist[v];
    }
    cnt[0] = c[0];
    cnt[1] = c[1];
    cnt[2] = c[2];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < count; i += kSortThreads) {
    const int v = nvec[i];
    if (v >= 0 && v <= hc::kBatchedMaxN) ord[atomicAdd(&pos[v], 1)] = i;
  }
}

}  // namespace

namespace hc {

// Resolves this device's copy of `scratch` and the persistent grids, once per
// device, with the device current (called by cuLAPACKCreate).
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_setup_mutex);
  DeviceSetup &d = g_setup[device];
  if (d.ptr != nullptr) return hipSuccess;
  int *ptr = nullptr;
  HC_TRY(hipGetSymbolAddress(reinterpret_cast<void **>(&ptr), HIP_SYMBOL(scratch)));
  int cus = 0;
  HC_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  for (int u = 0; u < 2; u++)
    for (int c = 0; c < 3; c++) {
      const char uplo = u == 0 ? 'L' : 'U';
      d.grid[0][u][c] = cus * potrf_batched_reg_blocks_per_cu<double>(uplo, c);
      d.grid[1][u][c] = cus * potrf_batched_reg_blocks_per_cu<float>(uplo, c);
    }
  d.ptr = ptr;
  return hipSuccess;
}

}  // namespace hc

extern "C" {

CUresult cuPotrfBatchedNoAlloc(CULAPACKhandle h, CBlasUplo uplo, char precision, size_t count, CUdeviceptr n,
                               CUdeviceptr offsets, CUdeviceptr A, CUdeviceptr logdet, CUdeviceptr info,
                               CUstream stream) {
  const char *name = "cuPotrfBatchedNoAlloc";
  if (h == nullptr) return hc::bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return hc::bad_arg(name, 2);
  if (precision != 'D' && precision != 'S') return hc::bad_arg(name, 3);
  if (count > (size_t)kNoAllocMax) {
    hc::report("count above the fixed maximum of 65536 matrices per call", name, __FILE__, __LINE__,
               CUDA_ERROR_INVALID_VALUE);
    return hc::bad_arg(name, 4);
  }
  if (count == 0) return CUDA_SUCCESS;
  if (n == 0) return hc::bad_arg(name, 5);
  if (offsets == 0) return hc::bad_arg(name, 6);
  if (A == 0) return hc::bad_arg(name, 7);
  if (logdet == 0) return hc::bad_arg(name, 8);
  if (info == 0) return hc::bad_arg(name, 9);
  if (h->device < 0 || h->device >= kMaxDevices) return CUDA_ERROR_INVALID_VALUE;
  int *base;
  int grid[3];
  {
    std::lock_guard<std::mutex> lock(g_setup_mutex);
    const DeviceSetup &d = g_setup[h->device];
    base = d.ptr;
    std::copy(d.grid[precision == 'D' ? 0 : 1][uplo == CBlasLower ? 0 : 1],
              d.grid[precision == 'D' ? 0 : 1][uplo == CBlasLower ? 0 : 1] + 3, grid);
  }
  if (base == nullptr) {  // the handle's creation resolves it; nothing to run on
    hc::report("device scratch not resolved", name, __FILE__, __LINE__, CUDA_ERROR_NOT_INITIALIZED);
    return CUDA_ERROR_NOT_INITIALIZED;
  }
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  int *cnt = base, *ord = base + kScratchHead;
  const int *nv = hc::dptr<const int>(n);
  const long *off = hc::dptr<const long>(offsets);
  double *ld = hc::dptr<double>(logdet);
  int *inf = hc::dptr<int>(info);
  hipLaunchKernelGGL(noalloc_sort_kernel, dim3(1), dim3(kSortThreads), 0, stream, (int)count, nv, cnt, ord, ld,
                     inf);
  HC_CHECK(hipGetLastError());
  if (precision == 'D')
    HC_CHECK(hc::potrf_batched_reg<double>(stream, (char)uplo, cnt, nullptr, grid, nv, off, ord,
                                           hc::dptr<double>(A), ld, inf));
  else
    HC_CHECK(hc::potrf_batched_reg<float>(stream, (char)uplo, cnt, nullptr, grid, nv, off, ord,
                                          hc::dptr<float>(A), ld, inf));
  return CUDA_SUCCESS;
}

}  // extern "C"
