// Helpers shared by the public entry points.
#pragma once

#include "internal.h"

namespace hc {

// Makes the handle's device current for the duration of a call and restores
// the caller's device afterwards (the reference pushes/pops the handle's
// context the same way, blas/dgemm.c:173,197).
class DeviceGuard {
 public:
  explicit DeviceGuard(int device) {
    if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
    if (prev_ != device) {
      ok_ = hipSetDevice(device);
      switched_ = true;
    }
  }
  ~DeviceGuard() {
    if (switched_ && prev_ >= 0) (void)hipSetDevice(prev_);
  }
  hipError_t status() const { return ok_; }

 private:
  int prev_ = -1;
  bool switched_ = false;
  hipError_t ok_ = hipSuccess;
};

inline bool valid_trans(int t) { return t == 'N' || t == 'T' || t == 'C'; }
inline bool valid_uplo(int u) { return u == 'L' || u == 'U'; }
inline bool valid_side(int s) { return s == 'L' || s == 'R'; }
inline bool valid_diag(int d) { return d == 'N' || d == 'U'; }

// Reports an illegal argument the way the reference's XERBLA macro does.
inline CUresult bad_arg(const char *func, long index) {
  if (xerbla != nullptr) xerbla(func, index);
  return CUDA_ERROR_INVALID_VALUE;
}

template <typename T>
inline T *dptr(CUdeviceptr p) {
  return reinterpret_cast<T *>(p);
}

inline zd to_zd(cuDoubleComplex_t v) { return make_zd(__real__ v, __imag__ v); }

}  // namespace hc
