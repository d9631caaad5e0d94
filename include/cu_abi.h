/*
 * Value types of the reference's C entry points, re-expressed for HIP.
 *
 * The reference's include/blas.h, include/lapack.h and include/error.h take
 * their types from the CUDA driver header (cuda.h): CUresult, CUdeviceptr and
 * CUstream.  This header gives the same three names the same meaning on an
 * MI355X so that callers of those entry points compile unchanged:
 *
 *   CUresult     an int-valued enum whose numeric codes equal the driver codes
 *                the reference returns (0 success, 1 invalid value, 2 out of
 *                memory, ...).  Codes coming from the HIP runtime are passed
 *                through with their HIP numbering, which coincides with the
 *                driver numbering for every code listed here.
 *   CUdeviceptr  a device BYTE address held in an unsigned integer; callers
 *                offset it with `+ elements * sizeof(T)` exactly as the
 *                reference does (lapack/dpotrf.c:212).
 *   CUstream     a HIP stream (struct ihipStream_t *, i.e. hipStream_t); NULL
 *                is the null stream.
 *
 * Nothing here wraps or emulates a CUDA API: no driver function is declared.
 */
#ifndef CU_ABI_H
#define CU_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uintptr_t CUdeviceptr;

typedef struct ihipStream_t * CUstream;

/* Device ordinal (the HIP device id). */
typedef int CUdevice;

typedef enum {
  CUDA_SUCCESS                              = 0,
  CUDA_ERROR_INVALID_VALUE                  = 1,
  CUDA_ERROR_OUT_OF_MEMORY                  = 2,
  CUDA_ERROR_NOT_INITIALIZED                = 3,
  CUDA_ERROR_DEINITIALIZED                  = 4,
  CUDA_ERROR_NO_DEVICE                      = 100,
  CUDA_ERROR_INVALID_DEVICE                 = 101,
  CUDA_ERROR_INVALID_IMAGE                  = 200,
  CUDA_ERROR_INVALID_CONTEXT                = 201,
  CUDA_ERROR_NO_BINARY_FOR_GPU              = 209,
  CUDA_ERROR_INVALID_HANDLE                 = 400,
  CUDA_ERROR_NOT_FOUND                      = 500,
  CUDA_ERROR_NOT_READY                      = 600,
  CUDA_ERROR_ILLEGAL_ADDRESS                = 700,
  CUDA_ERROR_LAUNCH_OUT_OF_RESOURCES        = 701,
  CUDA_ERROR_LAUNCH_FAILED                  = 719,
  CUDA_ERROR_NOT_SUPPORTED                  = 801,
  CUDA_ERROR_UNKNOWN                        = 999
} CUresult;

#ifdef __cplusplus
}
#endif

#endif
