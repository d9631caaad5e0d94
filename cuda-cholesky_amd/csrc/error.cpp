// Error hooks of the public ABI: errorHandler / cuGetErrorString (reference
// include/error.h:26-40, multigpu/error.c:4-64) and xerbla (reference
// include/blas.h:37-43, blas/xerbla.c:5-9).
#include "internal.h"

#include <cstdio>

extern "C" {

static void default_error_handler(const char *call, const char *function, const char *file,
                                  int line, int error, const char *(*strerror)(int)) {
  fprintf(stderr, "Error:\n\t%s returned %d (%s) in %s (%s:%d)\n", call, error, strerror(error),
          function, file, line);
}

errorHandler_t errorHandler = &default_error_handler;

static void default_xerbla(const char *function, long info) {
  fprintf(stderr, "On entry to %s parameter %ld had an invalid value\n", function, info);
}

xerbla_t xerbla = &default_xerbla;

const char *cuGetErrorString(CUresult result) {
  switch (result) {
    case CUDA_SUCCESS: return "No errors";
    case CUDA_ERROR_INVALID_VALUE: return "Invalid value";
    case CUDA_ERROR_OUT_OF_MEMORY: return "Out of memory";
    case CUDA_ERROR_NOT_INITIALIZED: return "Driver not initialized";
    case CUDA_ERROR_DEINITIALIZED: return "Driver deinitialized";
    case CUDA_ERROR_NO_DEVICE: return "No GPU device available";
    case CUDA_ERROR_INVALID_DEVICE: return "Invalid device";
    case CUDA_ERROR_INVALID_IMAGE: return "Invalid kernel image";
    case CUDA_ERROR_INVALID_CONTEXT: return "Invalid context";
    case CUDA_ERROR_NO_BINARY_FOR_GPU: return "No binary for GPU";
    case CUDA_ERROR_INVALID_HANDLE: return "Invalid handle";
    case CUDA_ERROR_NOT_FOUND: return "Not found";
    case CUDA_ERROR_NOT_READY: return "Not ready";
    case CUDA_ERROR_ILLEGAL_ADDRESS: return "Illegal address";
    case CUDA_ERROR_LAUNCH_OUT_OF_RESOURCES: return "Launch exceeded resources";
    case CUDA_ERROR_LAUNCH_FAILED: return "Launch failed";
    case CUDA_ERROR_NOT_SUPPORTED: return "Not supported";
    case CUDA_ERROR_UNKNOWN: return "Unknown error";
    default: break;
  }
  // any other code came from the HIP runtime unchanged
  return hipGetErrorString((hipError_t)result);
}

}  // extern "C"

namespace hc {

CUresult to_cu(hipError_t e) {
  switch (e) {
    case hipSuccess: return CUDA_SUCCESS;
    case hipErrorInvalidValue: return CUDA_ERROR_INVALID_VALUE;
    case hipErrorOutOfMemory: return CUDA_ERROR_OUT_OF_MEMORY;
    default: return (CUresult)e;  // HIP numbering equals the driver numbering
  }
}

static const char *strerror_int(int e) { return cuGetErrorString((CUresult)e); }

void report(const char *call, const char *func, const char *file, int line, CUresult err) {
  if (errorHandler != nullptr) errorHandler(call, func, file, line, (int)err, strerror_int);
}

}  // namespace hc
