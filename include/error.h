/*
 * Error reporting of the MI355X Cholesky library.
 *
 * Same contract as the reference's include/error.h:26-40: a replaceable
 * error-handler hook that is called whenever a HIP runtime call made inside
 * the library fails, and an error-code-to-string function.
 */
#ifndef ERROR_H
#define ERROR_H

#include "cu_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/**
 * Error handler function type (reference include/error.h:26).
 *
 * @param call      text of the runtime call that failed.
 * @param function  library function in which it failed.
 * @param file      source file.
 * @param line      source line.
 * @param error     integer error code (a CUresult value).
 * @param strerror  function turning the code into a description.
 */
typedef void (*errorHandler_t)(const char *, const char *, const char *, int,
                               int, const char * (*)(int));

/**
 * Error handler called by the library on internal runtime errors (reference
 * include/error.h:32).  May be set to NULL to silence it.  The default prints
 * one line to stderr.
 */
extern errorHandler_t errorHandler;

/**
 * Human-readable description of an error code (reference include/error.h:40).
 */
const char * cuGetErrorString(CUresult);

#ifdef __cplusplus
}
#endif

#endif
