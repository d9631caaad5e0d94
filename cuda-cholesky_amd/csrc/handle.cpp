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
  // the default priority.
  HC_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamDefault));
  HC_CHECK(hipStreamCreateWithFlags(&h->side, hipStreamDefault));
  HC_CHECK(hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming));
  HC_CHECK(hipEventCreateWithFlags(&h->ev_update, hipEventDisableTiming));
  HC_CHECK(hipMalloc(&h->dinfo, 64));
  HC_CHECK(hipHostMalloc(&h->hinfo, 64, hipHostMallocDefault));
  *handle = h;
  return CUDA_SUCCESS;
}

CUresult cuLAPACKDestroy(CULAPACKhandle handle) {
  if (handle == nullptr) return CUDA_SUCCESS;
  hc::DeviceGuard g(handle->device);
  HC_CHECK(hipStreamSynchronize(handle->stream));
  HC_CHECK(hipStreamSynchronize(handle->side));
  if (handle->ws != nullptr) HC_CHECK(hipFree(handle->ws));
  for (long i = 0; i < 2 * handle->tcap; i++) HC_CHECK(hipEventDestroy(handle->tev[i]));
  std::free(handle->tev);
  std::free(handle->tflops);
  HC_CHECK(hipFree(handle->dinfo));
  HC_CHECK(hipHostFree(handle->hinfo));
  HC_CHECK(hipEventDestroy(handle->ev_panel));
  HC_CHECK(hipEventDestroy(handle->ev_update));
  HC_CHECK(hipStreamDestroy(handle->stream));
  HC_CHECK(hipStreamDestroy(handle->side));
  cuBLASDestroy(handle->blas);
  std::free(handle);
  return CUDA_SUCCESS;
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
