/*
 * Multi-GPU context and task API of the reference (include/cumultigpu.h),
 * collapsed onto one device: a CUmultiGPU context holds one HIP device (the
 * first one given), tasks run on the calling thread with that device current,
 * and synchronisation waits for the device.  It exists so that callers of the
 * reference's host-pointer cuMultiGPU* routines (lapack.h, blas.h) build and
 * run unchanged on one MI355X.
 */
#ifndef CUMULTIGPU_H
#define CUMULTIGPU_H

#include <stddef.h>
#include "cu_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct __cutask_st * CUtask;

/* A task: function(args) with a private copy of `size` bytes of args
 * (reference include/cumultigpu.h:11-26).  CUDA_ERROR_INVALID_VALUE when
 * function is NULL, or args is NULL with size > 0. */
CUresult cuTaskCreate(CUtask *, CUresult (*)(const void *), const void *, size_t);
/* Destroys the task, returning its result through *result (may be NULL)
 * (reference :28-36). */
CUresult cuTaskDestroy(CUtask, CUresult *);
/* Runs the task on the calling thread (reference :38-45). */
CUresult cuTaskExecute(CUtask);

typedef struct __cumultigpu_st * CUmultiGPU;

/* A context over the first of the n devices given (n >= 1; the others are
 * accepted and not used) (reference :47-61). */
CUresult cuMultiGPUCreate(CUmultiGPU *, CUdevice *, int);
CUresult cuMultiGPUDestroy(CUmultiGPU);
/* Always 1: one device (reference :71-77). */
int cuMultiGPUGetContextCount(CUmultiGPU);
/* Runs the task synchronously with the context's device current; i must be 0
 * (reference :79-88). */
CUresult cuMultiGPURunTask(CUmultiGPU, int, CUtask);
/* Waits for the device (reference :90-96). */
CUresult cuMultiGPUSynchronize(CUmultiGPU);

#ifdef __cplusplus
}
#endif

#endif
