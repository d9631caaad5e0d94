"""Python mirror of the MI355X Cholesky library's C ABI.

The product is libcucholesky.so (built from csrc/ for gfx950), whose entry
points are the reference's include/blas.h and include/lapack.h device-pointer
API.  This module binds those entry points with ctypes under the same names,
argument order and meaning, so Python callers (the tests, the benchmark) drive
exactly the code a C caller links against.  LAPACK-style functions return
``(status, info)``; BLAS functions return ``status`` (a CUresult value).

Device memory is plain torch tensors: a column-major m x n matrix with leading
dimension ld is stored as a contiguous tensor of shape (n, ld) -- element
(i, j) at offset j * ld + i -- and passed by its device byte address.

There is no fallback: if the library is missing or fails to load, importing
this module raises.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_PATH = os.path.join(HERE, "libcucholesky.so")

CUDA_SUCCESS = 0
CUDA_ERROR_INVALID_VALUE = 1
CUDA_ERROR_OUT_OF_MEMORY = 2

# char-valued enums (reference include/blas.h:24-27)
NoTrans, Trans, ConjTrans = ord("N"), ord("T"), ord("C")
Lower, Upper = ord("L"), ord("U")
Left, Right = ord("L"), ord("R")
NonUnit, Unit = ord("N"), ord("U")


def build(jobs=8):
    """Compile every HIP source for gfx950 and link libcucholesky.so."""
    subprocess.run(["make", "-C", CSRC, "-B", f"-j{jobs}", "-s"], check=True)


class _Z(ctypes.Structure):
    """C99 `double complex` passed by value."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def zscalar(v):
    v = complex(v)
    return _Z(v.real, v.imag)


class _C(ctypes.Structure):
    """C99 `float complex` passed by value."""
    _fields_ = [("re", ctypes.c_float), ("im", ctypes.c_float)]


def cscalar(v):
    v = complex(v)
    return _C(v.real, v.imag)


_sz = ctypes.c_size_t
_ptr = ctypes.c_uint64     # CUdeviceptr: device byte address
_enum = ctypes.c_int       # enums are int-sized
_res = ctypes.c_int        # CUresult
_stream = ctypes.c_void_p  # CUstream
_longp = ctypes.POINTER(ctypes.c_long)
_h = ctypes.c_void_p
_hp = ctypes.c_void_p      # host array (the cuMultiGPU* routines)
_taskfn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)   # CUresult (*)(const void *)

# name -> argument types (restype CUresult unless noted); mirrors
# include/blas.h and include/lapack.h
SIGNATURES = {
    "cuBLASCreate": [ctypes.POINTER(_h)],
    "cuBLASDestroy": [_h],
    "cuDsyrk": [_h, _enum, _enum, _sz, _sz, ctypes.c_double, _ptr, _sz, ctypes.c_double, _ptr, _sz,
                _stream],
    "cuZherk": [_h, _enum, _enum, _sz, _sz, ctypes.c_double, _ptr, _sz, ctypes.c_double, _ptr, _sz,
                _stream],
    "cuDgemm2": [_h, _enum, _enum, _sz, _sz, _sz, ctypes.c_double, _ptr, _sz, _ptr, _sz,
                 ctypes.c_double, _ptr, _sz, _ptr, _sz, _stream],
    "cuZgemm2": [_h, _enum, _enum, _sz, _sz, _sz, _Z, _ptr, _sz, _ptr, _sz, _Z, _ptr, _sz, _ptr,
                 _sz, _stream],
    "cuDtrmm2": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_double, _ptr, _sz, _ptr, _sz,
                 _ptr, _sz, _stream],
    "cuZtrmm2": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _Z, _ptr, _sz, _ptr, _sz, _ptr, _sz,
                 _stream],
    "cuDtrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_double, _ptr, _sz, _ptr, _sz,
                _stream],
    "cuZtrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _Z, _ptr, _sz, _ptr, _sz, _stream],
    "cuLAPACKCreate": [ctypes.POINTER(_h)],
    "cuLAPACKDestroy": [_h],
    "cuDpotrf": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuZpotrf": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuDtrtri": [_h, _enum, _enum, _sz, _ptr, _sz, _longp],
    "cuZtrtri": [_h, _enum, _enum, _sz, _ptr, _sz, _longp],
    "cuDtrti22": [_h, _enum, _enum, _sz, _ptr, _sz, _ptr, _sz, _ptr, _stream],
    "cuZtrti22": [_h, _enum, _enum, _sz, _ptr, _sz, _ptr, _sz, _ptr, _stream],
    "cuDlauum": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuZlauum": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuDpotri": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuZpotri": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuDlogdet": [_h, _ptr, _sz, _sz, ctypes.POINTER(ctypes.c_double), _stream],
    "cuZlogdet": [_h, _ptr, _sz, _sz, ctypes.POINTER(ctypes.c_double), _stream],
    "cuGetErrorString": [_res],
    "cuSsyrk": [_h, _enum, _enum, _sz, _sz, ctypes.c_float, _ptr, _sz, ctypes.c_float, _ptr, _sz,
                _stream],
    "cuSgemm2": [_h, _enum, _enum, _sz, _sz, _sz, ctypes.c_float, _ptr, _sz, _ptr, _sz,
                 ctypes.c_float, _ptr, _sz, _ptr, _sz, _stream],
    "cuStrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_float, _ptr, _sz, _ptr, _sz,
                _stream],
    "cuSpotrf": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuStrmm2": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_float, _ptr, _sz, _ptr, _sz,
                 _ptr, _sz, _stream],
    "cuCherk": [_h, _enum, _enum, _sz, _sz, ctypes.c_float, _ptr, _sz, ctypes.c_float, _ptr, _sz,
                _stream],
    "cuCgemm2": [_h, _enum, _enum, _sz, _sz, _sz, _C, _ptr, _sz, _ptr, _sz, _C, _ptr, _sz, _ptr,
                 _sz, _stream],
    "cuCtrmm2": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _C, _ptr, _sz, _ptr, _sz, _ptr, _sz,
                 _stream],
    "cuCtrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _C, _ptr, _sz, _ptr, _sz, _stream],
    "cuCpotrf": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuStrtri": [_h, _enum, _enum, _sz, _ptr, _sz, _longp],
    "cuCtrtri": [_h, _enum, _enum, _sz, _ptr, _sz, _longp],
    "cuStrti22": [_h, _enum, _enum, _sz, _ptr, _sz, _ptr, _sz, _ptr, _stream],
    "cuCtrti22": [_h, _enum, _enum, _sz, _ptr, _sz, _ptr, _sz, _ptr, _stream],
    "cuSlauum": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuClauum": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuSpotri": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuCpotri": [_h, _enum, _sz, _ptr, _sz, _longp],
    "cuSlogdet": [_h, _ptr, _sz, _sz, ctypes.POINTER(ctypes.c_float), _stream],
    "cuClogdet": [_h, _ptr, _sz, _sz, ctypes.POINTER(ctypes.c_float), _stream],
    # include/cucholesky.h (extensions)
    "cuDpotrs": [_h, _enum, _sz, _sz, _ptr, _sz, _ptr, _sz, _longp],
    "cuDsposv": [_h, _enum, _sz, _sz, _ptr, _sz, _ptr, _sz, _ptr, _sz, _longp, _longp],
    "cuDsposvx": [_h, _enum, _sz, _sz, _ptr, _sz, _ptr, _sz, _ptr, _sz, ctypes.c_char, ctypes.c_char,
                  _longp, ctypes.c_char_p, _longp],
    "cuDsposvxBatched": [_h, _enum, _sz, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                         _ptr, _sz, ctypes.POINTER(ctypes.c_size_t), _ptr, _ptr, _ptr, _longp,
                         ctypes.POINTER(ctypes.c_double), _longp],
    "cuDsposvxBatchedPlanCreate": [_h, _enum, _sz, ctypes.POINTER(ctypes.c_size_t),
                                   ctypes.POINTER(ctypes.c_size_t), _sz, ctypes.POINTER(ctypes.c_size_t),
                                   ctypes.POINTER(_h)],
    "cuDsposvxBatchedRun": [_h, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _stream],
    "cuDsposvxBatchedPlanDestroy": [_h],
    "cuPotrfBatchedPlanCreate": [_h, _enum, ctypes.c_char, _sz, ctypes.POINTER(ctypes.c_size_t),
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_h)],
    "cuPotrfBatchedRun": [_h, _ptr, _ptr, _ptr, _stream],
    "cuPotriBatchedRun": [_h, _ptr, _ptr, _ptr, _stream],
    "cuPotrfBatchedPlanDestroy": [_h],
    "cuPotrfBatchedNoAlloc": [_h, _enum, ctypes.c_char, _sz, _ptr, _ptr, _ptr, _ptr, _ptr, _stream],
    "cuLAPACKSetTiming": [_h, ctypes.c_int],
    "cuLAPACKSetBlockSize": [_h, ctypes.c_char, ctypes.c_long],
    "cuLAPACKGetTiming": [_h, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.c_long, _longp],
    # include/cumultigpu.h and the host-pointer halves of blas.h / lapack.h
    "cuTaskCreate": [ctypes.POINTER(_h), _taskfn, ctypes.c_void_p, _sz],
    "cuTaskDestroy": [_h, ctypes.POINTER(_res)],
    "cuTaskExecute": [_h],
    "cuMultiGPUCreate": [ctypes.POINTER(_h), ctypes.POINTER(ctypes.c_int), ctypes.c_int],
    "cuMultiGPUDestroy": [_h],
    "cuMultiGPUGetContextCount": [_h],
    "cuMultiGPURunTask": [_h, ctypes.c_int, _h],
    "cuMultiGPUSynchronize": [_h],
    "cuMultiGPUBLASCreate": [ctypes.POINTER(_h), _h],
    "cuMultiGPUBLASDestroy": [_h],
    "cuMultiGPUBLASSynchronize": [_h],
    "cuMultiGPUDsyrk": [_h, _enum, _enum, _sz, _sz, ctypes.c_double, _hp, _sz, ctypes.c_double, _hp,
                        _sz],
    "cuMultiGPUZherk": [_h, _enum, _enum, _sz, _sz, ctypes.c_double, _hp, _sz, ctypes.c_double, _hp,
                        _sz],
    "cuMultiGPUDgemm": [_h, _enum, _enum, _sz, _sz, _sz, ctypes.c_double, _hp, _sz, _hp, _sz,
                        ctypes.c_double, _hp, _sz],
    "cuMultiGPUZgemm": [_h, _enum, _enum, _sz, _sz, _sz, _Z, _hp, _sz, _hp, _sz, _Z, _hp, _sz],
    "cuMultiGPUDtrmm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_double, _hp, _sz, _hp,
                        _sz],
    "cuMultiGPUZtrmm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _Z, _hp, _sz, _hp, _sz],
    "cuMultiGPUDtrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_double, _hp, _sz, _hp,
                        _sz],
    "cuMultiGPUZtrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _Z, _hp, _sz, _hp, _sz],
    "cuMultiGPUSsyrk": [_h, _enum, _enum, _sz, _sz, ctypes.c_float, _hp, _sz, ctypes.c_float, _hp,
                        _sz],
    "cuMultiGPUCherk": [_h, _enum, _enum, _sz, _sz, ctypes.c_float, _hp, _sz, ctypes.c_float, _hp,
                        _sz],
    "cuMultiGPUSgemm": [_h, _enum, _enum, _sz, _sz, _sz, ctypes.c_float, _hp, _sz, _hp, _sz,
                        ctypes.c_float, _hp, _sz],
    "cuMultiGPUCgemm": [_h, _enum, _enum, _sz, _sz, _sz, _C, _hp, _sz, _hp, _sz, _C, _hp, _sz],
    "cuMultiGPUStrmm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_float, _hp, _sz, _hp,
                        _sz],
    "cuMultiGPUCtrmm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _C, _hp, _sz, _hp, _sz],
    "cuMultiGPUStrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, ctypes.c_float, _hp, _sz, _hp,
                        _sz],
    "cuMultiGPUCtrsm": [_h, _enum, _enum, _enum, _enum, _sz, _sz, _C, _hp, _sz, _hp, _sz],
    "cuMultiGPUSpotrf": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUCpotrf": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUStrtri": [_h, _enum, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUCtrtri": [_h, _enum, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUSlauum": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUClauum": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUSpotri": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUCpotri": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPULAPACKCreate": [ctypes.POINTER(_h), _h],
    "cuMultiGPULAPACKDestroy": [_h],
    "cuMultiGPULAPACKSynchronize": [_h],
    "cuMultiGPUDpotrf": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUZpotrf": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUDtrtri": [_h, _enum, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUZtrtri": [_h, _enum, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUDlauum": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUZlauum": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUDpotri": [_h, _enum, _sz, _hp, _sz, _longp],
    "cuMultiGPUZpotri": [_h, _enum, _sz, _hp, _sz, _longp],
}

# host-memory routines of include/lapack.h and include/blas.h (host arrays;
# void, except the log-determinants): name -> (restype, argument types)
_f, _d = ctypes.c_float, ctypes.c_double
_fp = ctypes.c_void_p        # host array
_HOST_LAPACK = {p + r: [_enum, _sz, _fp, _sz, _longp] for p in "sdcz" for r in ("potrf", "lauum", "potri")}
_HOST_LAPACK.update({p + "trtri": [_enum, _enum, _sz, _fp, _sz, _longp] for p in "sdcz"})
_HOST_LAPACK.update({p + "trtri2": [_enum, _enum, _sz, _fp, _sz, _fp, _sz, _longp] for p in "sdcz"})
_scal = {"s": _f, "d": _d, "c": _C, "z": _Z}
_real = {"s": _f, "d": _d, "c": _f, "z": _d}
HOST_SIGNATURES = {name: (None, args) for name, args in _HOST_LAPACK.items()}
HOST_SIGNATURES.update({p + "logdet": (_real[p], [_fp, _sz, _sz]) for p in "sdcz"})
for p in "sdcz":
    herk = ("syrk" if p in "sd" else "herk")
    HOST_SIGNATURES[p + herk] = (None, [_enum, _enum, _sz, _sz, _real[p], _fp, _sz, _real[p], _fp, _sz])
    HOST_SIGNATURES[p + "gemm"] = (None, [_enum, _enum, _sz, _sz, _sz, _scal[p], _fp, _sz, _fp, _sz, _scal[p],
                                          _fp, _sz])
    for r in ("trmm", "trsm"):
        HOST_SIGNATURES[p + r] = (None, [_enum, _enum, _enum, _enum, _sz, _sz, _scal[p], _fp, _sz, _fp, _sz])
    HOST_SIGNATURES[p + "trmm2"] = (None, [_enum, _enum, _enum, _enum, _sz, _sz, _scal[p], _fp, _sz, _fp, _sz,
                                           _fp, _sz])

XERBLA_T = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_long)
ERRHANDLER_T = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                ctypes.c_int, ctypes.c_int, ctypes.c_void_p)

_lib = None


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(or make -C cuda-cholesky_amd/csrc) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _res
        L.cuGetErrorString.restype = ctypes.c_char_p
        for name, (res, args) in HOST_SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def error_string(code):
    return lib().cuGetErrorString(code).decode()


class Xerbla:
    """Context manager capturing xerbla reports (function, parameter index)."""

    def __init__(self):
        self.calls = []
        self._cb = XERBLA_T(lambda f, i: self.calls.append((f.decode(), i)))
        self._slot = ctypes.c_void_p.in_dll(lib(), "xerbla")

    def __enter__(self):
        self._old = self._slot.value
        self._slot.value = ctypes.cast(self._cb, ctypes.c_void_p).value
        return self

    def __exit__(self, *exc):
        self._slot.value = self._old
        return False


# ---- handles ----

class BLASHandle:
    def __init__(self):
        self.h = _h()
        st = lib().cuBLASCreate(ctypes.byref(self.h))
        if st != CUDA_SUCCESS:
            raise RuntimeError(f"cuBLASCreate failed: {st}")

    def close(self):
        if self.h:
            lib().cuBLASDestroy(self.h)
            self.h = _h()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class LAPACKHandle:
    def __init__(self):
        self.h = _h()
        st = lib().cuLAPACKCreate(ctypes.byref(self.h))
        if st != CUDA_SUCCESS:
            raise RuntimeError(f"cuLAPACKCreate failed: {st}")

    def close(self):
        if self.h:
            lib().cuLAPACKDestroy(self.h)
            self.h = _h()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _hv(handle):
    return handle.h if hasattr(handle, "h") else handle


def cuLAPACKSetBlockSize(handle, precision, nb):
    return lib().cuLAPACKSetBlockSize(_hv(handle), precision.encode() if isinstance(precision, str)
                                      else precision, int(nb))


def cuLAPACKSetTiming(handle, enable):
    return lib().cuLAPACKSetTiming(_hv(handle), int(enable))


def cuLAPACKGetTiming(handle, capacity=4096):
    """(status, [(ms, flops), ...]) of the recorded trailing updates."""
    ms = (ctypes.c_double * capacity)()
    fl = (ctypes.c_double * capacity)()
    count = ctypes.c_long(0)
    st = lib().cuLAPACKGetTiming(_hv(handle), ms, fl, capacity, ctypes.byref(count))
    n = min(count.value, capacity)
    return st, [(ms[i], fl[i]) for i in range(n)]


# ---- LAPACK (status, info) ----

def _info_call(fn, *args):
    info = ctypes.c_long(0)
    st = fn(*args, ctypes.byref(info))
    return st, info.value


def cuDpotrf(handle, uplo, n, A, lda):
    return _info_call(lib().cuDpotrf, _hv(handle), uplo, n, A, lda)


def cuSpotrf(handle, uplo, n, A, lda):
    return _info_call(lib().cuSpotrf, _hv(handle), uplo, n, A, lda)


def cuDpotrs(handle, uplo, n, nrhs, A, lda, B, ldb):
    return _info_call(lib().cuDpotrs, _hv(handle), uplo, n, nrhs, A, lda, B, ldb)


def cuDsposv(handle, uplo, n, nrhs, A, lda, B, ldb, X, ldx):
    """(status, iter, info)"""
    it, info = ctypes.c_long(0), ctypes.c_long(0)
    st = lib().cuDsposv(_hv(handle), uplo, n, nrhs, A, lda, B, ldb, X, ldx, ctypes.byref(it),
                        ctypes.byref(info))
    return st, it.value, info.value


def cuDsposvx(handle, uplo, n, nrhs, A, lda, B, ldb, X, ldx, factor="S", refine="I"):
    """(status, iter, used precision 'B' / 'X' / 'H' / 'S' / 'D', info)"""
    it, info = ctypes.c_long(0), ctypes.c_long(0)
    used = ctypes.create_string_buffer(1)
    enc = (lambda c: c.encode() if isinstance(c, str) else c)
    st = lib().cuDsposvx(_hv(handle), uplo, n, nrhs, A, lda, B, ldb, X, ldx, enc(factor), enc(refine),
                         ctypes.byref(it), used, ctypes.byref(info))
    return st, it.value, used.raw[:1].decode(), info.value


def cuDsposvxBatched(handle, uplo, n, offA, A, nrhs, offB, B, X, Ainv=0):
    """(status, iter[], logdet[], info[]) of the batched mixed-precision solve
    + logdet (+ inverse into Ainv when nonzero); n, offA, offB host sequences."""
    count = len(n)
    arr = lambda v: (ctypes.c_size_t * max(count, 1))(*[int(x) for x in v])
    it = (ctypes.c_long * max(count, 1))()
    ld = (ctypes.c_double * max(count, 1))()
    info = (ctypes.c_long * max(count, 1))()
    st = lib().cuDsposvxBatched(_hv(handle), uplo, count, arr(n), arr(offA), A, nrhs, arr(offB), B, X,
                                Ainv, it, ld, info)
    return st, list(it)[:count], list(ld)[:count], list(info)[:count]


class BatchedSolvePlan:
    """cuDsposvxBatchedPlanCreate / Run / Destroy: the task-graph solve +
    logdet + inverse of packed matrices of orders n (host sequences n, offA,
    offB); run() takes device addresses and int / double / int device arrays
    for iter, logdet and info."""

    def __init__(self, handle, uplo, n, offA, nrhs, offB):
        count = len(n)
        arr = lambda v: (ctypes.c_size_t * max(count, 1))(*[int(x) for x in v])
        self.h = _h()
        self.status = lib().cuDsposvxBatchedPlanCreate(_hv(handle), uplo, count, arr(n), arr(offA), nrhs,
                                                       arr(offB if offB is not None else [0] * count),
                                                       ctypes.byref(self.h))

    def run(self, A, B, X, Ainv, iter_, logdet, info, stream=None):
        return lib().cuDsposvxBatchedRun(self.h, A, B, X, Ainv, iter_, logdet, info, stream)

    def close(self):
        if self.h:
            lib().cuDsposvxBatchedPlanDestroy(self.h)
            self.h = _h()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class BatchedPlan:
    """cuPotrfBatchedPlanCreate / Run / Destroy for packed matrices of sizes n
    at element offsets `offsets` (host sequences)."""

    def __init__(self, handle, uplo, precision, n, offsets):
        count = len(n)
        nn = (ctypes.c_size_t * max(count, 1))(*n)
        oo = (ctypes.c_size_t * max(count, 1))(*offsets)
        self.h = _h()
        self.status = lib().cuPotrfBatchedPlanCreate(_hv(handle), uplo,
                                                     precision.encode() if isinstance(precision, str)
                                                     else precision, count, nn, oo,
                                                     ctypes.byref(self.h))

    def run(self, A, logdet, info, stream=None):
        return lib().cuPotrfBatchedRun(self.h, A, logdet, info, stream)

    def run_inverse(self, A, logdet, info, stream=None):
        """cuPotriBatchedRun: factor, logdet, then A_i <- inv(A_i) in the same launch."""
        return lib().cuPotriBatchedRun(self.h, A, logdet, info, stream)

    def close(self):
        if self.h:
            lib().cuPotrfBatchedPlanDestroy(self.h)
            self.h = _h()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def cuZpotrf(handle, uplo, n, A, lda):
    return _info_call(lib().cuZpotrf, _hv(handle), uplo, n, A, lda)


def cuDtrtri(handle, uplo, diag, n, A, lda):
    return _info_call(lib().cuDtrtri, _hv(handle), uplo, diag, n, A, lda)


def cuZtrtri(handle, uplo, diag, n, A, lda):
    return _info_call(lib().cuZtrtri, _hv(handle), uplo, diag, n, A, lda)


def cuDlauum(handle, uplo, n, A, lda):
    return _info_call(lib().cuDlauum, _hv(handle), uplo, n, A, lda)


def cuZlauum(handle, uplo, n, A, lda):
    return _info_call(lib().cuZlauum, _hv(handle), uplo, n, A, lda)


def cuDpotri(handle, uplo, n, A, lda):
    return _info_call(lib().cuDpotri, _hv(handle), uplo, n, A, lda)


def cuZpotri(handle, uplo, n, A, lda):
    return _info_call(lib().cuZpotri, _hv(handle), uplo, n, A, lda)


def cuDtrti22(handle, uplo, diag, n, A, lda, B, ldb, info_ptr, stream=None):
    return lib().cuDtrti22(_hv(handle), uplo, diag, n, A, lda, B, ldb, info_ptr, stream)


def cuZtrti22(handle, uplo, diag, n, A, lda, B, ldb, info_ptr, stream=None):
    return lib().cuZtrti22(_hv(handle), uplo, diag, n, A, lda, B, ldb, info_ptr, stream)


def cuDlogdet(handle, x, incx, n, stream=None):
    r = ctypes.c_double(0.0)
    st = lib().cuDlogdet(_hv(handle), x, incx, n, ctypes.byref(r), stream)
    return st, r.value


def cuZlogdet(handle, x, incx, n, stream=None):
    r = ctypes.c_double(0.0)
    st = lib().cuZlogdet(_hv(handle), x, incx, n, ctypes.byref(r), stream)
    return st, r.value


# ---- BLAS (status) ----

def cuDgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, D, ldd,
             stream=None):
    return lib().cuDgemm2(_hv(handle), transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                          D, ldd, stream)


def cuZgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, D, ldd,
             stream=None):
    return lib().cuZgemm2(_hv(handle), transA, transB, m, n, k, zscalar(alpha), A, lda, B, ldb,
                          zscalar(beta), C, ldc, D, ldd, stream)


def cuDgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    return cuDgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc,
                    stream)


def cuZgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    return cuZgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc,
                    stream)


def cuDsyrk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream=None):
    return lib().cuDsyrk(_hv(handle), uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream)


def cuZherk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream=None):
    return lib().cuZherk(_hv(handle), uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream)


def cuDtrmm2(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx, stream=None):
    return lib().cuDtrmm2(_hv(handle), side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx,
                          stream)


def cuZtrmm2(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx, stream=None):
    return lib().cuZtrmm2(_hv(handle), side, uplo, trans, diag, m, n, zscalar(alpha), A, lda, B,
                          ldb, X, ldx, stream)


def cuDtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, stream=None):
    return lib().cuDtrsm(_hv(handle), side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, stream)


def cuSgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, D, ldd,
             stream=None):
    return lib().cuSgemm2(_hv(handle), transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                          D, ldd, stream)


def cuSsyrk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream=None):
    return lib().cuSsyrk(_hv(handle), uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream)


def cuStrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, stream=None):
    return lib().cuStrsm(_hv(handle), side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, stream)


def cuZtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, stream=None):
    return lib().cuZtrsm(_hv(handle), side, uplo, trans, diag, m, n, zscalar(alpha), A, lda, B, ldb,
                         stream)


# ---- single precision and single complex (widened onto the double paths) ----

def cuStrmm2(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx, stream=None):
    return lib().cuStrmm2(_hv(handle), side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx,
                          stream)


def cuCherk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream=None):
    return lib().cuCherk(_hv(handle), uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream)


def cuCgemm2(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, D, ldd,
             stream=None):
    return lib().cuCgemm2(_hv(handle), transA, transB, m, n, k, cscalar(alpha), A, lda, B, ldb,
                          cscalar(beta), C, ldc, D, ldd, stream)


def cuCtrmm2(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, X, ldx, stream=None):
    return lib().cuCtrmm2(_hv(handle), side, uplo, trans, diag, m, n, cscalar(alpha), A, lda, B,
                          ldb, X, ldx, stream)


def cuCtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, stream=None):
    return lib().cuCtrsm(_hv(handle), side, uplo, trans, diag, m, n, cscalar(alpha), A, lda, B, ldb,
                         stream)


def cuCpotrf(handle, uplo, n, A, lda):
    return _info_call(lib().cuCpotrf, _hv(handle), uplo, n, A, lda)


def cuStrtri(handle, uplo, diag, n, A, lda):
    return _info_call(lib().cuStrtri, _hv(handle), uplo, diag, n, A, lda)


def cuCtrtri(handle, uplo, diag, n, A, lda):
    return _info_call(lib().cuCtrtri, _hv(handle), uplo, diag, n, A, lda)


def cuStrti22(handle, uplo, diag, n, A, lda, B, ldb, info_ptr, stream=None):
    return lib().cuStrti22(_hv(handle), uplo, diag, n, A, lda, B, ldb, info_ptr, stream)


def cuCtrti22(handle, uplo, diag, n, A, lda, B, ldb, info_ptr, stream=None):
    return lib().cuCtrti22(_hv(handle), uplo, diag, n, A, lda, B, ldb, info_ptr, stream)


def cuSlauum(handle, uplo, n, A, lda):
    return _info_call(lib().cuSlauum, _hv(handle), uplo, n, A, lda)


def cuClauum(handle, uplo, n, A, lda):
    return _info_call(lib().cuClauum, _hv(handle), uplo, n, A, lda)


def cuSpotri(handle, uplo, n, A, lda):
    return _info_call(lib().cuSpotri, _hv(handle), uplo, n, A, lda)


def cuCpotri(handle, uplo, n, A, lda):
    return _info_call(lib().cuCpotri, _hv(handle), uplo, n, A, lda)


def cuSlogdet(handle, x, incx, n, stream=None):
    r = ctypes.c_float(0.0)
    st = lib().cuSlogdet(_hv(handle), x, incx, n, ctypes.byref(r), stream)
    return st, r.value


def cuClogdet(handle, x, incx, n, stream=None):
    r = ctypes.c_float(0.0)
    st = lib().cuClogdet(_hv(handle), x, incx, n, ctypes.byref(r), stream)
    return st, r.value


# ---- host-pointer API (include/cumultigpu.h; cuMultiGPU* of blas.h / lapack.h) ----

TaskFunction = _taskfn


class Task:
    """cuTaskCreate / cuTaskExecute / cuTaskDestroy; `function` is a
    TaskFunction (keep it alive while the task exists), `args` bytes."""

    def __init__(self, function, args=b""):
        self.h = _h()
        self._fn = function
        buf = ctypes.create_string_buffer(args, len(args)) if args else None
        st = lib().cuTaskCreate(ctypes.byref(self.h), function, buf, len(args))
        if st != CUDA_SUCCESS:
            raise RuntimeError(f"cuTaskCreate failed: {st}")

    def execute(self):
        return lib().cuTaskExecute(self.h)

    def destroy(self):
        """(status, the task's own result)."""
        r = _res(0)
        st = lib().cuTaskDestroy(self.h, ctypes.byref(r))
        self.h = _h()
        return st, r.value


class MultiGPU:
    """A CUmultiGPU context over `devices` (HIP device ids); one device is
    used."""

    def __init__(self, devices=(0,)):
        self.h = _h()
        arr = (ctypes.c_int * len(devices))(*devices)
        st = lib().cuMultiGPUCreate(ctypes.byref(self.h), arr, len(devices))
        if st != CUDA_SUCCESS:
            raise RuntimeError(f"cuMultiGPUCreate failed: {st}")

    def context_count(self):
        return lib().cuMultiGPUGetContextCount(self.h)

    def run_task(self, i, task):
        return lib().cuMultiGPURunTask(self.h, i, _hv(task))

    def synchronize(self):
        return lib().cuMultiGPUSynchronize(self.h)

    def close(self):
        if self.h:
            lib().cuMultiGPUDestroy(self.h)
            self.h = _h()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _MultiGPUHandle:
    _create = _destroy = _sync = None

    def __init__(self, mgpu):
        self.h = _h()
        st = getattr(lib(), self._create)(ctypes.byref(self.h), _hv(mgpu))
        if st != CUDA_SUCCESS:
            raise RuntimeError(f"{self._create} failed: {st}")

    def synchronize(self):
        return getattr(lib(), self._sync)(self.h)

    def close(self):
        if self.h:
            getattr(lib(), self._destroy)(self.h)
            self.h = _h()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class MultiGPUBLASHandle(_MultiGPUHandle):
    _create, _destroy, _sync = "cuMultiGPUBLASCreate", "cuMultiGPUBLASDestroy", "cuMultiGPUBLASSynchronize"


class MultiGPULAPACKHandle(_MultiGPUHandle):
    _create, _destroy, _sync = ("cuMultiGPULAPACKCreate", "cuMultiGPULAPACKDestroy",
                                "cuMultiGPULAPACKSynchronize")


def _host(M):
    """Address of a host array: a column-major numpy array (its leading
    dimension is M.shape[0]), a raw address, or None."""
    if M is None or isinstance(M, int):
        return M
    if not (M.flags.f_contiguous or M.ndim == 1):
        raise ValueError("host matrices must be Fortran-ordered")
    return M.ctypes.data


def cuMultiGPUDgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return lib().cuMultiGPUDgemm(_hv(handle), transA, transB, m, n, k, alpha, _host(A), lda, _host(B),
                                 ldb, beta, _host(C), ldc)


def cuMultiGPUZgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return lib().cuMultiGPUZgemm(_hv(handle), transA, transB, m, n, k, zscalar(alpha), _host(A), lda,
                                 _host(B), ldb, zscalar(beta), _host(C), ldc)


def cuMultiGPUDsyrk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    return lib().cuMultiGPUDsyrk(_hv(handle), uplo, trans, n, k, alpha, _host(A), lda, beta, _host(C),
                                 ldc)


def cuMultiGPUZherk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    return lib().cuMultiGPUZherk(_hv(handle), uplo, trans, n, k, alpha, _host(A), lda, beta, _host(C),
                                 ldc)


def cuMultiGPUDtrmm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUDtrmm(_hv(handle), side, uplo, trans, diag, m, n, alpha, _host(A), lda,
                                 _host(B), ldb)


def cuMultiGPUZtrmm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUZtrmm(_hv(handle), side, uplo, trans, diag, m, n, zscalar(alpha), _host(A),
                                 lda, _host(B), ldb)


def cuMultiGPUDtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUDtrsm(_hv(handle), side, uplo, trans, diag, m, n, alpha, _host(A), lda,
                                 _host(B), ldb)


def cuMultiGPUZtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUZtrsm(_hv(handle), side, uplo, trans, diag, m, n, zscalar(alpha), _host(A),
                                 lda, _host(B), ldb)


def cuMultiGPUSgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return lib().cuMultiGPUSgemm(_hv(handle), transA, transB, m, n, k, alpha, _host(A), lda, _host(B),
                                 ldb, beta, _host(C), ldc)


def cuMultiGPUCgemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return lib().cuMultiGPUCgemm(_hv(handle), transA, transB, m, n, k, cscalar(alpha), _host(A), lda,
                                 _host(B), ldb, cscalar(beta), _host(C), ldc)


def cuMultiGPUSsyrk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    return lib().cuMultiGPUSsyrk(_hv(handle), uplo, trans, n, k, alpha, _host(A), lda, beta, _host(C),
                                 ldc)


def cuMultiGPUCherk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    return lib().cuMultiGPUCherk(_hv(handle), uplo, trans, n, k, alpha, _host(A), lda, beta, _host(C),
                                 ldc)


def cuMultiGPUStrmm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUStrmm(_hv(handle), side, uplo, trans, diag, m, n, alpha, _host(A), lda,
                                 _host(B), ldb)


def cuMultiGPUCtrmm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUCtrmm(_hv(handle), side, uplo, trans, diag, m, n, cscalar(alpha), _host(A),
                                 lda, _host(B), ldb)


def cuMultiGPUStrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUStrsm(_hv(handle), side, uplo, trans, diag, m, n, alpha, _host(A), lda,
                                 _host(B), ldb)


def cuMultiGPUCtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return lib().cuMultiGPUCtrsm(_hv(handle), side, uplo, trans, diag, m, n, cscalar(alpha), _host(A),
                                 lda, _host(B), ldb)


def _mlapack(name):
    def call(handle, uplo, *args):
        *rest, A, lda = args
        return _info_call(getattr(lib(), name), _hv(handle), uplo, *rest, _host(A), lda)
    call.__name__ = name
    return call


cuMultiGPUSpotrf = _mlapack("cuMultiGPUSpotrf")
cuMultiGPUCpotrf = _mlapack("cuMultiGPUCpotrf")
cuMultiGPUStrtri = _mlapack("cuMultiGPUStrtri")   # (handle, uplo, diag, n, A, lda)
cuMultiGPUCtrtri = _mlapack("cuMultiGPUCtrtri")
cuMultiGPUSlauum = _mlapack("cuMultiGPUSlauum")
cuMultiGPUClauum = _mlapack("cuMultiGPUClauum")
cuMultiGPUSpotri = _mlapack("cuMultiGPUSpotri")
cuMultiGPUCpotri = _mlapack("cuMultiGPUCpotri")


def cuMultiGPUDpotrf(handle, uplo, n, A, lda):
    return _info_call(lib().cuMultiGPUDpotrf, _hv(handle), uplo, n, _host(A), lda)


def cuMultiGPUZpotrf(handle, uplo, n, A, lda):
    return _info_call(lib().cuMultiGPUZpotrf, _hv(handle), uplo, n, _host(A), lda)


def cuMultiGPUDtrtri(handle, uplo, diag, n, A, lda):
    return _info_call(lib().cuMultiGPUDtrtri, _hv(handle), uplo, diag, n, _host(A), lda)


def cuMultiGPUZtrtri(handle, uplo, diag, n, A, lda):
    return _info_call(lib().cuMultiGPUZtrtri, _hv(handle), uplo, diag, n, _host(A), lda)


def cuMultiGPUDlauum(handle, uplo, n, A, lda):
    return _info_call(lib().cuMultiGPUDlauum, _hv(handle), uplo, n, _host(A), lda)


def cuMultiGPUZlauum(handle, uplo, n, A, lda):
    return _info_call(lib().cuMultiGPUZlauum, _hv(handle), uplo, n, _host(A), lda)


def cuMultiGPUDpotri(handle, uplo, n, A, lda):
    return _info_call(lib().cuMultiGPUDpotri, _hv(handle), uplo, n, _host(A), lda)


def cuMultiGPUZpotri(handle, uplo, n, A, lda):
    return _info_call(lib().cuMultiGPUZpotri, _hv(handle), uplo, n, _host(A), lda)


# ---- device matrices (torch is only the allocator here) ----

def to_device(M, device="cuda"):
    """Column-major copy of the 2-D numpy array M on the GPU: a tensor of
    shape (n, ld) with M's leading dimension ld = M.shape[0]."""
    import numpy as np
    import torch
    return torch.from_numpy(np.ascontiguousarray(M.T)).to(device)


def from_device(t):
    """numpy (ld, n) Fortran-ordered view of a device matrix made by to_device."""
    return t.detach().cpu().numpy().T


def ptr(t):
    return t.data_ptr()
