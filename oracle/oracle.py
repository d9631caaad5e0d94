"""CPU oracle: ctypes bindings of liboracle.so plus numpy helpers.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP library
(tests/), the reference of __graft_entry__.smoke() and the CPU baseline leg of
bench.py.  The product library never imports it.

Matrices are numpy arrays in Fortran order of shape (ld, n): element (i, j)
sits at offset i + j * ld, the reference's column-major layout with leading
dimension ld.  Every routine returns the modified array and, where the
reference has one, the LAPACK `info` value.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_c_size = ctypes.c_size_t
_c_dbl = ctypes.c_double
_c_char = ctypes.c_char
_c_long_p = ctypes.POINTER(ctypes.c_long)
_c_ptr = ctypes.c_void_p


class _Z(ctypes.Structure):
    """C99 double complex passed by value (two doubles, SSE class)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def build():
    """Compile liboracle.so from oracle/*.c with the committed Makefile."""
    subprocess.run(["make", "-C", HERE, "-B", "-s"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "oracle_dgemm": [_c_char, _c_char, _c_size, _c_size, _c_size, _c_dbl, _c_ptr, _c_size,
                             _c_ptr, _c_size, _c_dbl, _c_ptr, _c_size],
            "oracle_dsyrk": [_c_char, _c_char, _c_size, _c_size, _c_dbl, _c_ptr, _c_size, _c_dbl,
                             _c_ptr, _c_size],
            "oracle_dtrsm": [_c_char, _c_char, _c_char, _c_char, _c_size, _c_size, _c_dbl, _c_ptr,
                             _c_size, _c_ptr, _c_size],
            "oracle_dtrmm": [_c_char, _c_char, _c_char, _c_char, _c_size, _c_size, _c_dbl, _c_ptr,
                             _c_size, _c_ptr, _c_size],
            "oracle_dtrmm2": [_c_char, _c_char, _c_char, _c_char, _c_size, _c_size, _c_dbl, _c_ptr,
                              _c_size, _c_ptr, _c_size, _c_ptr, _c_size],
            "oracle_zgemm": [_c_char, _c_char, _c_size, _c_size, _c_size, _Z, _c_ptr, _c_size,
                             _c_ptr, _c_size, _Z, _c_ptr, _c_size],
            "oracle_zherk": [_c_char, _c_char, _c_size, _c_size, _c_dbl, _c_ptr, _c_size, _c_dbl,
                             _c_ptr, _c_size],
            "oracle_ztrsm": [_c_char, _c_char, _c_char, _c_char, _c_size, _c_size, _Z, _c_ptr,
                             _c_size, _c_ptr, _c_size],
            "oracle_ztrmm2": [_c_char, _c_char, _c_char, _c_char, _c_size, _c_size, _Z, _c_ptr,
                              _c_size, _c_ptr, _c_size, _c_ptr, _c_size],
            "oracle_dpotrf": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dtrtri": [_c_char, _c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dtrtri2": [_c_char, _c_char, _c_size, _c_ptr, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dlauum": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dpotri": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_zpotrf": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_ztrtri": [_c_char, _c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_zlauum": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_zpotri": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dpotrf_unblocked": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_zpotrf_unblocked": [_c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dtrtri_unblocked": [_c_char, _c_char, _c_size, _c_ptr, _c_size, _c_long_p],
            "oracle_dlauum_unblocked": [_c_char, _c_size, _c_ptr, _c_size],
            "oracle_dlatmc": [_c_size, _c_dbl, _c_ptr, _c_size],
            "oracle_zlatmc": [_c_size, _c_dbl, _c_ptr, _c_size],
            "oracle_srand": [ctypes.c_uint],
            "oracle_logdet_input": [_c_ptr, _c_size, _c_size],
            "oracle_gauss": [ctypes.c_uint64, ctypes.c_uint64, _c_size, _c_ptr],
            "oracle_spd": [ctypes.c_uint64, _c_size, _c_dbl, _c_ptr, _c_size],
            "oracle_hpd": [ctypes.c_uint64, _c_size, _c_dbl, _c_ptr, _c_size],
            "oracle_threads": [],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None
        for name in ("oracle_dlogdet", "oracle_zlogdet", "oracle_zlogdet_ref"):
            fn = getattr(L, name)
            fn.argtypes = [_c_ptr, _c_size, _c_size]
            fn.restype = _c_dbl
        for name in ("oracle_dlatmc", "oracle_zlatmc", "oracle_threads"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _b(c):
    return c.encode() if isinstance(c, str) else c


def _p(a):
    assert a.flags["F_CONTIGUOUS"], "oracle arrays must be Fortran ordered"
    return a.ctypes.data_as(ctypes.c_void_p)


def _z(v):
    v = complex(v)
    return _Z(v.real, v.imag)


def fmat(n_rows, n_cols, ld=None, dtype=np.float64):
    """Zeroed column-major matrix with leading dimension ld (>= n_rows)."""
    ld = n_rows if ld is None else ld
    return np.zeros((max(ld, 1), n_cols), dtype=dtype, order="F")


def with_ld(M, ld):
    """Copy of the 2-D array M embedded in an (ld, n) Fortran array."""
    m, n = M.shape
    out = fmat(m, n, ld, M.dtype)
    out[:m, :n] = M
    return out


def threads():
    return lib().oracle_threads()


# ---- BLAS ----

def dgemm(transA, transB, m, n, k, alpha, A, B, beta, C):
    lib().oracle_dgemm(_b(transA), _b(transB), m, n, k, alpha, _p(A), A.shape[0], _p(B), B.shape[0],
                       beta, _p(C), C.shape[0])
    return C


def dsyrk(uplo, trans, n, k, alpha, A, beta, C):
    lib().oracle_dsyrk(_b(uplo), _b(trans), n, k, alpha, _p(A), A.shape[0], beta, _p(C), C.shape[0])
    return C


def dtrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    lib().oracle_dtrsm(_b(side), _b(uplo), _b(trans), _b(diag), m, n, alpha, _p(A), A.shape[0],
                       _p(B), B.shape[0])
    return B


def dtrmm(side, uplo, trans, diag, m, n, alpha, A, B):
    lib().oracle_dtrmm(_b(side), _b(uplo), _b(trans), _b(diag), m, n, alpha, _p(A), A.shape[0],
                       _p(B), B.shape[0])
    return B


def dtrmm2(side, uplo, trans, diag, m, n, alpha, A, B, X):
    lib().oracle_dtrmm2(_b(side), _b(uplo), _b(trans), _b(diag), m, n, alpha, _p(A), A.shape[0],
                        _p(B), B.shape[0], _p(X), X.shape[0])
    return X


def zgemm(transA, transB, m, n, k, alpha, A, B, beta, C):
    lib().oracle_zgemm(_b(transA), _b(transB), m, n, k, _z(alpha), _p(A), A.shape[0], _p(B),
                       B.shape[0], _z(beta), _p(C), C.shape[0])
    return C


def zherk(uplo, trans, n, k, alpha, A, beta, C):
    lib().oracle_zherk(_b(uplo), _b(trans), n, k, alpha, _p(A), A.shape[0], beta, _p(C), C.shape[0])
    return C


def ztrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    lib().oracle_ztrsm(_b(side), _b(uplo), _b(trans), _b(diag), m, n, _z(alpha), _p(A), A.shape[0],
                       _p(B), B.shape[0])
    return B


def ztrmm2(side, uplo, trans, diag, m, n, alpha, A, B, X):
    lib().oracle_ztrmm2(_b(side), _b(uplo), _b(trans), _b(diag), m, n, _z(alpha), _p(A),
                        A.shape[0], _p(B), B.shape[0], _p(X), X.shape[0])
    return X


# ---- LAPACK ----

def _info_call(fn, *args):
    info = ctypes.c_long(0)
    fn(*args, ctypes.byref(info))
    return info.value


def dpotrf(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_dpotrf, _b(uplo), n, _p(A), A.shape[0])


def dpotrf_unblocked(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_dpotrf_unblocked, _b(uplo), n, _p(A), A.shape[0])


def zpotrf_unblocked(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_zpotrf_unblocked, _b(uplo), n, _p(A), A.shape[0])


def dtrtri(uplo, diag, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_dtrtri, _b(uplo), _b(diag), n, _p(A), A.shape[0])


def dtrtri_unblocked(uplo, diag, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_dtrtri_unblocked, _b(uplo), _b(diag), n, _p(A), A.shape[0])


def dtrtri2(uplo, diag, A, B, n=None):
    n = A.shape[1] if n is None else n
    return B, _info_call(lib().oracle_dtrtri2, _b(uplo), _b(diag), n, _p(A), A.shape[0], _p(B),
                         B.shape[0])


def dlauum(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_dlauum, _b(uplo), n, _p(A), A.shape[0])


def dlauum_unblocked(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    lib().oracle_dlauum_unblocked(_b(uplo), n, _p(A), A.shape[0])
    return A


def dpotri(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_dpotri, _b(uplo), n, _p(A), A.shape[0])


def dlogdet(x, incx, n):
    return lib().oracle_dlogdet(_p(x), incx, n)


def zpotrf(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_zpotrf, _b(uplo), n, _p(A), A.shape[0])


def ztrtri(uplo, diag, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_ztrtri, _b(uplo), _b(diag), n, _p(A), A.shape[0])


def zlauum(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_zlauum, _b(uplo), n, _p(A), A.shape[0])


def zpotri(uplo, A, n=None):
    n = A.shape[1] if n is None else n
    return A, _info_call(lib().oracle_zpotri, _b(uplo), n, _p(A), A.shape[0])


def kahan_logdet(x, incx, n):
    """2 * sum(log x[j*incx]) by Kahan summation: the oracle of the
    reference's logdet test (test/lapack/cudlogdet.c:62-70)."""
    s, c = 0.0, 0.0
    for v in np.asarray(x)[::incx][:n]:
        y = float(np.log(v)) - c
        t = s + y
        c = (t - s) - y
        s = t
    return 2.0 * s


def zlogdet(x, incx, n):
    return lib().oracle_zlogdet(_p(x), incx, n)


def zlogdet_ref(x, incx, n):
    return lib().oracle_zlogdet_ref(_p(x), incx, n)


# ---- generators ----

def srand(seed):
    lib().oracle_srand(seed)


def dlatmc(n, c, ld=None):
    A = fmat(n, n, ld)
    rc = lib().oracle_dlatmc(n, c, _p(A), A.shape[0])
    assert rc == 0, rc
    return A


def zlatmc(n, c, ld=None):
    A = fmat(n, n, ld, np.complex128)
    rc = lib().oracle_zlatmc(n, c, _p(A), A.shape[0])
    assert rc == 0, rc
    return A


def logdet_input(n, incx=1):
    x = np.zeros(max(n * incx, 1), dtype=np.float64)
    lib().oracle_logdet_input(_p(x), incx, n)
    return x


def gauss(seed, count, first=0):
    out = np.zeros(count, dtype=np.float64)
    lib().oracle_gauss(seed, first, count, _p(out))
    return out


def spd(seed, n, alpha, ld=None):
    """A = G G^T / n + alpha I, G from gauss(seed, n*n) column-major."""
    A = fmat(n, n, ld)
    lib().oracle_spd(seed, n, alpha, _p(A), A.shape[0])
    return A


def hpd(seed, n, alpha, ld=None):
    A = fmat(n, n, ld, np.complex128)
    lib().oracle_hpd(seed, n, alpha, _p(A), A.shape[0])
    return A


# ---- numpy restatement of the generator (used to pin the C one) ----

_M64 = (1 << 64) - 1


def gauss_numpy(seed, count, first=0):
    """Same stream as oracle_gauss, computed with numpy uint64 arithmetic."""
    idx = np.arange(first, first + count, dtype=np.uint64)
    gamma = np.uint64(0x9E3779B97F4A7C15)
    seed = np.uint64(seed)

    def draw(c):
        with np.errstate(over="ignore"):
            z = seed + (c + np.uint64(1)) * gamma
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    d1 = draw(np.uint64(2) * idx)
    d2 = draw(np.uint64(2) * idx + np.uint64(1))
    u1 = ((d1 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (d2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# ---- accuracy metrics ----

def tri(M, uplo):
    return np.tril(M) if uplo in ("L", b"L") else np.triu(M)


def herm_from(M, uplo):
    """Full Hermitian matrix from the stored triangle of M (n x n)."""
    T = tri(M, uplo)
    D = np.diag(np.diag(T).real)
    return T + T.conj().T - D


def potrf_backward_error(A, F, uplo):
    """||A - L L^H||_F / ||A||_F (lower) or ||A - U^H U||_F / ||A||_F (upper);
    A is the full Hermitian input, F the factor array (n x n)."""
    T = tri(F, uplo)
    R = T @ T.conj().T if uplo in ("L", b"L") else T.conj().T @ T
    return np.linalg.norm(A - R) / np.linalg.norm(A)


def inverse_residual(A, X, uplo):
    """||I - A X||_F / (||A||_F ||X||_F) with X rebuilt from its stored
    triangle; the north star's ||I - A inv(A)|| / cond(A) in Frobenius form."""
    Xf = herm_from(X, uplo)
    n = A.shape[0]
    return np.linalg.norm(np.eye(n) - A @ Xf) / (np.linalg.norm(A) * np.linalg.norm(Xf))
