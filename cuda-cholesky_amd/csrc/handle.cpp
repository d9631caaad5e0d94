// Handles: cuBLASCreate / cuBLASDestroy (reference blas/handle.c:84-100) and
// cuLAPACKCreate / cuLAPACKDestroy (reference lapack/handle.c:63-80).
//
// A BLAS handle binds the device that is current at creation.  A LAPACK
// handle additionally owns two default-priority streams (main and look-ahead),
// two events, a device info word with its pinned host mirror and a device
// workspace that grows on demand, so repeated calls allocate nothing.
#include "api_util.h"

#include <cstdlib>

extern "C" {

CUresult cuBLASCreate(CUBLAShandle *handle) {
  if (handle == nullptr) return hc::bad_arg("cuBLASCreate", 1);
  auto *h = static_cast<__cublashandle_st *>(std::malloc(sizeof(__cublashandle_st)));
  if (h == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    std::free(h);
    HC_CHECK(e);
  }
  *handle = h;
  return CUDA_SUCCESS;
}

CUresult cuBLASDestroy(CUBLAShandle handle) {
  std::free(handle);
  return CUDA_SUCCESS;
}

// Releases whatever a (possibly partly built) LAPACK handle holds; every
// member is null or valid.
static CUresult lapack_release(__culapackhandle_st *h) {
  CUresult first = CUDA_SUCCESS;
  auto keep = [&](hipError_t e) {
    if (e != hipSuccess && first == CUDA_SUCCESS) first = hc::to_cu(e);
  };
  if (h->stream != nullptr) keep(hipStreamSynchronize(h->stream));
  if (h->side != nullptr) keep(hipStreamSynchronize(h->side));
  if (h->ws != nullptr) keep(hipFree(h->ws));
  for (long i = 0; i < 2 * h->tcap; i++) keep(hipEventDestroy(h->tev[i]));
  std::free(h->tev);
  std::free(h->tflops);
  if (h->dinfo != nullptr) keep(hipFree(h->dinfo));
  if (h->pflags != nullptr) keep(hipFree(h->pflags));
  if (h->hinfo != nullptr) keep(hipHostFree(h->hinfo));
  if (h->ev_panel != nullptr) keep(hipEventDestroy(h->ev_panel));
  if (h->ev_update != nullptr) keep(hipEventDestroy(h->ev_update));
  if (h->stream != nullptr) keep(hipStreamDestroy(h->stream));
  if (h->side != nullptr) keep(hipStreamDestroy(h->side));
  if (h->blas != nullptr) cuBLASDestroy(h->blas);
  std::free(h);
  return first;
}

CUresult cuLAPACKCreate(CULAPACKhandle *handle) {
  if (handle == nullptr) return hc::bad_arg("cuLAPACKCreate", 1);
  auto *h = static_cast<__culapackhandle_st *>(std::calloc(1, sizeof(__culapackhandle_st)));
  if (h == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  CUresult r = cuBLASCreate(&h->blas);
  if (r != CUDA_SUCCESS) {
    std::free(h);
    return r;
  }
  h->device = h->blas->device;
  // Streams are created blocking (ordered after work on the null stream), at
  // the default priority.  On any failure everything created so far is
  // released before the error is returned.
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->side, hipStreamDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_update, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&h->dinfo, 64);
  if (e == hipSuccess) e = hipHostMalloc(&h->hinfo, 64, hipHostMallocDefault);
  // the no-allocation batched entry point's per-device scratch address
  if (e == hipSuccess) e = hc::noalloc_setup(h->device);
  if (e != hipSuccess) {
    const CUresult err = hc::to_cu(e);
    hc::report("cuLAPACKCreate", __func__, __FILE__, __LINE__, err);
    (void)lapack_release(h);
    return err;
  }
  *handle = h;
  return CUDA_SUCCESS;
}

CUresult cuLAPACKDestroy(CULAPACKhandle handle) {
  if (handle == nullptr) return CUDA_SUCCESS;
  hc::DeviceGuard g(handle->device);
  HC_CHECK(g.status());
  const CUresult r = lapack_release(handle);
  if (r != CUDA_SUCCESS) hc::report("cuLAPACKDestroy", __func__, __FILE__, __LINE__, r);
  return r;
}

// Timing of the trailing updates inside cuDpotrf / cuZpotrf (an extension
// with no counterpart in the reference; used by the benchmark).
CUresult cuLAPACKSetTiming(CULAPACKhandle handle, int enable) {
  if (handle == nullptr) return hc::bad_arg("cuLAPACKSetTiming", 1);
  handle->timing = enable != 0;
  handle->tcount = 0;
  return CUDA_SUCCESS;
}

CUresult cuLAPACKGetTiming(CULAPACKhandle handle, double *ms, double *flops, long capacity,
                           long *count) {
  if (handle == nullptr) return hc::bad_arg("cuLAPACKGetTiming", 1);
  if (capacity < 0) return hc::bad_arg("cuLAPACKGetTiming", 4);
  if (count == nullptr) return hc::bad_arg("cuLAPACKGetTiming", 5);
  hc::DeviceGuard g(handle->device);
  HC_CHECK(hipStreamSynchronize(handle->stream));
  const long n = handle->tcount < capacity ? handle->tcount : capacity;
  for (long i = 0; i < n; i++) {
    float t = 0.0f;
    HC_CHECK(hipEventElapsedTime(&t, handle->tev[2 * i], handle->tev[2 * i + 1]));
    if (ms != nullptr) ms[i] = t;
    if (flops != nullptr) flops[i] = handle->tflops[i];
  }
  *count = handle->tcount;
  handle->tcount = 0;
  return CUDA_SUCCESS;
}

}  // extern "C"
