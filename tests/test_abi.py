"""C-ABI boundary checks that need no GPU.

* libcucholesky.so loads and exports every function and variable that
  include/*.h declares;
* the public headers compile as C99 and as C++ and keep the reference's enum
  values;
* argument errors are reported exactly as the reference reports them
  (CUDA_ERROR_INVALID_VALUE, negative info, xerbla with the parameter index)
  before any device work, so they run here.
"""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(f for f in os.listdir(INCLUDE) if f.endswith(".h")):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"^\s*(?:CUresult|const char \*|void|float|double)\s+(\w+)\s*\(", text,
                             flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(cc):
    L = cc.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in sorted(names):
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
        assert name in cc.SIGNATURES or name in cc.HOST_SIGNATURES, f"{name} not bound by the Python mirror"
    for var in ("xerbla", "errorHandler"):
        assert ctypes.c_void_p.in_dll(L, var).value is not None


def test_headers_compile_as_c_and_cpp(tmp_path):
    src = tmp_path / "use.c"
    src.write_text(
        '#include "lapack.h"\n#include "error.h"\n#include "cucholesky.h"\n'
        "_Static_assert(CBlasLower == 'L' && CBlasUpper == 'U', \"uplo\");\n"
        "_Static_assert(CBlasNoTrans == 'N' && CBlasTrans == 'T' && CBlasConjTrans == 'C', \"t\");\n"
        "_Static_assert(CBlasLeft == 'L' && CBlasRight == 'R', \"side\");\n"
        "_Static_assert(CBlasNonUnit == 'N' && CBlasUnit == 'U', \"diag\");\n"
        "_Static_assert(CUDA_SUCCESS == 0 && CUDA_ERROR_INVALID_VALUE == 1 && "
        "CUDA_ERROR_OUT_OF_MEMORY == 2, \"codes\");\n"
        "_Static_assert(sizeof(CUdeviceptr) == 8, \"ptr\");\n"
        "int f(CULAPACKhandle h, CUdeviceptr A) { long info;\n"
        "  CUresult r = cuDpotrf(h, CBlasLower, 4, A, 4, &info);\n"
        "  double complex one = 1.0;\n"
        "  return (int)r + (int)cuZgemm(0, CBlasNoTrans, CBlasConjTrans, 1, 1, 1, one, A, 1, A, 1,"
        " one, A, 1, (CUstream)0); }\n"
        "int h(double *A, double complex *Z) { CUdevice dev = 0; CUmultiGPU mg; long info;\n"
        "  CUmultiGPULAPACKhandle l; CUmultiGPUBLAShandle b; double complex one = 1.0;\n"
        "  CUresult r = cuMultiGPUCreate(&mg, &dev, 1);\n"
        "  r = cuMultiGPULAPACKCreate(&l, mg); r = cuMultiGPUBLASCreate(&b, mg);\n"
        "  r = cuMultiGPUDpotrf(l, CBlasLower, 4, A, 4, &info);\n"
        "  r = cuMultiGPUZgemm(b, CBlasNoTrans, CBlasConjTrans, 1, 1, 1, one, Z, 1, Z, 1, one, Z, 1);\n"
        "  return (int)r + cuMultiGPUGetContextCount(mg); }\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{INCLUDE}", str(src)],
                   check=True)
    cpp = tmp_path / "use.cpp"
    cpp.write_text('#include "lapack.h"\n#include "error.h"\n#include "cucholesky.h"\n'
                   "int g(CULAPACKhandle h) { long i; return cuDpotri(h, CBlasUpper, 0, 0, 1, &i); }\n")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-fsyntax-only", f"-I{INCLUDE}", str(cpp)],
                   check=True)


FAKE = ctypes.c_void_p(0x1000)  # never dereferenced on the argument-error paths


@pytest.mark.parametrize("fn,expect_info", [("cuDpotrf", -4), ("cuZpotrf", -4), ("cuDlauum", -4),
                                            ("cuZlauum", -4), ("cuDpotri", -4), ("cuZpotri", -4)])
def test_lapack_lda_error(cc, fn, expect_info):
    """lda < n: info = -4, xerbla(4), CUDA_ERROR_INVALID_VALUE
    (reference lapack/dpotrf.c:173-179)."""
    with cc.Xerbla() as x:
        st, info = getattr(cc, fn)(FAKE, cc.Lower, 10, 0, 9)
    assert st == cc.CUDA_ERROR_INVALID_VALUE
    assert info == expect_info
    assert x.calls == [(fn, 4)]


@pytest.mark.parametrize("fn", ["cuDtrtri", "cuZtrtri"])
def test_trtri_lda_error(cc, fn):
    """TRTRI reports lda as parameter 5 (reference lapack/dtrtri.c:331-337)."""
    with cc.Xerbla() as x:
        st, info = getattr(cc, fn)(FAKE, cc.Upper, cc.NonUnit, 10, 0, 3)
    assert st == cc.CUDA_ERROR_INVALID_VALUE and info == -5 and x.calls == [(fn, 5)]


@pytest.mark.parametrize("fn", ["cuDpotrf", "cuDpotri", "cuDlauum", "cuZpotrf"])
def test_lapack_n_zero_quick_return(cc, fn):
    st, info = getattr(cc, fn)(FAKE, cc.Lower, 0, 0, 1)
    assert st == cc.CUDA_SUCCESS and info == 0


def test_logdet_n_zero(cc):
    st, r = cc.cuDlogdet(FAKE, 0, 1, 0)
    assert st == cc.CUDA_SUCCESS and r == 0.0
    st, r = cc.cuZlogdet(FAKE, 0, 1, 0)
    assert st == cc.CUDA_SUCCESS and r == 0.0


def test_gemm_argument_indices(cc):
    """reference blas/dgemm.c:153-167: lda -> 8, ldb -> 10, ldc -> 13, ldd -> 15."""
    cases = [((5, 4, 3, 4, 3, 5, 5), 8),     # m, n, k, lda, ldb, ldc, ldd ; NN lda < m
             ((5, 4, 3, 5, 2, 5, 5), 10),
             ((5, 4, 3, 5, 3, 4, 5), 13),
             ((5, 4, 3, 5, 3, 5, 4), 15)]
    for (m, n, k, lda, ldb, ldc, ldd), idx in cases:
        with cc.Xerbla() as x:
            st = cc.cuDgemm2(FAKE, cc.NoTrans, cc.NoTrans, m, n, k, 1.0, 0, lda, 0, ldb, 1.0, 0, ldc,
                             0, ldd)
        assert st == cc.CUDA_ERROR_INVALID_VALUE and x.calls == [("cuDgemm2", idx)]
        with cc.Xerbla() as x:
            st = cc.cuZgemm2(FAKE, cc.NoTrans, cc.NoTrans, m, n, k, 1.0, 0, lda, 0, ldb, 1.0, 0, ldc,
                             0, ldd)
        assert st == cc.CUDA_ERROR_INVALID_VALUE and x.calls == [("cuZgemm2", idx)]


def test_gemm_quick_returns(cc):
    # m == 0, n == 0, and (C == D with alpha == 0 and beta == 1) need no device
    assert cc.cuDgemm2(FAKE, cc.NoTrans, cc.NoTrans, 0, 4, 3, 1.0, 0, 1, 0, 3, 1.0, 0, 1, 0, 1) == 0
    assert cc.cuDgemm2(FAKE, cc.NoTrans, cc.NoTrans, 4, 0, 3, 1.0, 0, 4, 0, 3, 1.0, 0, 4, 0, 4) == 0
    assert cc.cuDgemm2(FAKE, cc.NoTrans, cc.NoTrans, 4, 4, 3, 0.0, 0, 4, 0, 3, 1.0, 64, 4, 64, 4) == 0


def test_syrk_herk_argument_indices(cc):
    with cc.Xerbla() as x:
        assert cc.cuDsyrk(FAKE, cc.Lower, cc.NoTrans, 5, 3, 1.0, 0, 4, 1.0, 0, 5) == 1
        assert cc.cuDsyrk(FAKE, cc.Lower, cc.Trans, 5, 3, 1.0, 0, 3, 1.0, 0, 4) == 1
        # ZHERK rejects trans = T (reference blas/zherk.c:197-198)
        assert cc.cuZherk(FAKE, cc.Lower, cc.Trans, 5, 3, 1.0, 0, 5, 1.0, 0, 5) == 1
    assert x.calls == [("cuDsyrk", 7), ("cuDsyrk", 10), ("cuZherk", 3)]


def test_trsm_trmm_argument_indices(cc):
    with cc.Xerbla() as x:
        assert cc.cuDtrsm(FAKE, cc.Left, cc.Lower, cc.NoTrans, cc.NonUnit, 6, 4, 1.0, 0, 5, 0, 6) == 1
        assert cc.cuDtrsm(FAKE, cc.Right, cc.Lower, cc.NoTrans, cc.NonUnit, 6, 4, 1.0, 0, 4, 0, 5) == 1
        assert cc.cuDtrmm2(FAKE, cc.Left, cc.Upper, cc.Trans, cc.Unit, 6, 4, 1.0, 0, 6, 0, 6, 0, 5) == 1
        assert cc.cuZtrsm(FAKE, cc.Right, cc.Upper, cc.ConjTrans, cc.Unit, 6, 4, 1.0, 0, 3, 0, 6) == 1
    assert x.calls == [("cuDtrsm", 9), ("cuDtrsm", 11), ("cuDtrmm2", 13), ("cuZtrsm", 9)]


def test_invalid_enum_rejected(cc):
    with cc.Xerbla() as x:
        st, info = cc.cuDpotrf(FAKE, ord("X"), 4, 0, 4)
    assert st == 1 and info == -2 and x.calls == [("cuDpotrf", 2)]


def test_error_strings(cc):
    assert cc.error_string(0) == "No errors"
    assert cc.error_string(1) == "Invalid value"
    assert cc.error_string(2) == "Out of memory"


def test_single_precision_argument_indices(cc):
    """The FP32 surface checks arguments like its FP64 twin (same XERBLA indices)."""
    with cc.Xerbla() as x:
        assert cc.cuSgemm2(FAKE, cc.NoTrans, cc.NoTrans, 5, 4, 3, 1.0, 0, 4, 0, 3, 1.0, 0, 5, 0,
                           5) == cc.CUDA_ERROR_INVALID_VALUE
        assert cc.cuSsyrk(FAKE, cc.Lower, cc.NoTrans, 5, 3, 1.0, 0, 4, 1.0, 0, 5) == 1
        assert cc.cuStrsm(FAKE, cc.Left, cc.Lower, cc.NoTrans, cc.NonUnit, 5, 3, 1.0, 0, 4, 0,
                          5) == 1
        st, info = cc.cuSpotrf(FAKE, cc.Lower, 10, 0, 9)
    assert st == 1 and info == -4
    assert x.calls == [("cuSgemm2", 8), ("cuSsyrk", 7), ("cuStrsm", 9), ("cuSpotrf", 4)]


def test_solver_argument_indices(cc):
    """cuDpotrs / cuDsposv follow LAPACK DPOTRS / DSPOSV numbering (handle not
    counted): LDA -> 5, LDB -> 7, LDX -> 9."""
    with cc.Xerbla() as x:
        assert cc.cuDpotrs(FAKE, cc.Lower, 10, 2, 0, 9, 0, 10) == (1, -5)
        assert cc.cuDpotrs(FAKE, cc.Lower, 10, 2, 0, 10, 0, 9) == (1, -7)
        assert cc.cuDsposv(FAKE, cc.Upper, 10, 2, 0, 10, 0, 10, 0, 9) == (1, 0, -9)
        st, it, used, info = cc.cuDsposvx(FAKE, cc.Lower, 10, 2, 0, 10, 0, 10, 0, 10, "Q")
        assert (st, info) == (1, -10)
        st, it, used, info = cc.cuDsposvx(FAKE, cc.Lower, 10, 2, 0, 10, 0, 10, 0, 10, "S", "X")
        assert (st, info) == (1, -11)
    assert x.calls == [("cuDpotrs", 5), ("cuDpotrs", 7), ("cuDsposv", 9), ("cuDsposvx", 10),
                       ("cuDsposvx", 11)]


def test_solver_quick_returns(cc):
    assert cc.cuDpotrs(FAKE, cc.Lower, 0, 3, 0, 1, 0, 1) == (0, 0)
    assert cc.cuDsposv(FAKE, cc.Lower, 5, 0, 0, 5, 0, 5, 0, 5) == (0, 0, 0)


# ---- host-pointer API (include/cumultigpu.h, cuMultiGPU* of blas.h / lapack.h)

def test_task_runs_on_calling_thread_with_private_args(cc):
    """reference multigpu/multigpu_seq.c:28-95: the task keeps a copy of its
    arguments and hands its function's result back through cuTaskDestroy."""
    seen = []

    def body(args):
        v = ctypes.cast(args, ctypes.POINTER(ctypes.c_int * 2)).contents
        seen.append((v[0], v[1]))
        return v[0] + v[1]

    fn = cc.TaskFunction(body)
    args = bytearray(ctypes.c_int(40)) + bytearray(ctypes.c_int(2))
    t = cc.Task(fn, bytes(args))
    args[0] = 99                      # the caller's buffer is not the task's
    assert t.execute() == cc.CUDA_SUCCESS
    st, result = t.destroy()
    assert st == cc.CUDA_SUCCESS and result == 42 and seen == [(40, 2)]
    t2 = cc.Task(cc.TaskFunction(lambda a: 7))   # no arguments
    assert t2.execute() == 0 and t2.destroy() == (0, 7)


def test_task_and_context_argument_errors(cc):
    L = cc.lib()
    h = ctypes.c_void_p()
    fn = cc.TaskFunction(lambda a: 0)
    assert L.cuTaskCreate(ctypes.byref(h), cc.TaskFunction(), None, 0) == cc.CUDA_ERROR_INVALID_VALUE
    assert L.cuTaskCreate(ctypes.byref(h), fn, None, 8) == cc.CUDA_ERROR_INVALID_VALUE
    dev = (ctypes.c_int * 1)(0)
    assert L.cuMultiGPUCreate(ctypes.byref(h), dev, 0) == cc.CUDA_ERROR_INVALID_VALUE
    assert L.cuMultiGPURunTask(None, 0, FAKE) == cc.CUDA_ERROR_INVALID_VALUE
    assert L.cuMultiGPUGetContextCount(None) == 0
    assert L.cuMultiGPUBLASCreate(ctypes.byref(h), None) == cc.CUDA_ERROR_INVALID_VALUE
    assert L.cuMultiGPULAPACKCreate(ctypes.byref(h), None) == cc.CUDA_ERROR_INVALID_VALUE


def test_multigpu_blas_argument_indices(cc):
    """reference blas/dgemm.c:441-451 (8, 10, 13), blas/dsyrk.c:219-226 (7, 10),
    blas/zherk.c:248-255 (trans = T -> 2), blas/dtrmm.c:474-481 and
    blas/dtrsm.c:274-281 (9, 11): checked before anything is read."""
    with cc.Xerbla() as x:
        assert cc.cuMultiGPUDgemm(FAKE, cc.NoTrans, cc.NoTrans, 5, 4, 3, 1.0, 0, 4, 0, 3, 1.0, 0, 5) == 1
        assert cc.cuMultiGPUDgemm(FAKE, cc.NoTrans, cc.Trans, 5, 4, 3, 1.0, 0, 5, 0, 3, 1.0, 0, 5) == 1
        assert cc.cuMultiGPUZgemm(FAKE, cc.Trans, cc.NoTrans, 5, 4, 3, 1.0, 0, 3, 0, 3, 1.0, 0, 4) == 1
        assert cc.cuMultiGPUDsyrk(FAKE, cc.Lower, cc.Trans, 5, 3, 1.0, 0, 2, 1.0, 0, 5) == 1
        assert cc.cuMultiGPUDsyrk(FAKE, cc.Upper, cc.NoTrans, 5, 3, 1.0, 0, 5, 1.0, 0, 4) == 1
        assert cc.cuMultiGPUZherk(FAKE, cc.Lower, cc.Trans, 5, 3, 1.0, 0, 5, 1.0, 0, 5) == 1
        assert cc.cuMultiGPUDtrmm(FAKE, cc.Left, cc.Lower, cc.NoTrans, cc.NonUnit, 6, 4, 1.0, 0, 5, 0,
                                  6) == 1
        assert cc.cuMultiGPUZtrmm(FAKE, cc.Right, cc.Upper, cc.ConjTrans, cc.Unit, 6, 4, 1.0, 0, 4, 0,
                                  5) == 1
        assert cc.cuMultiGPUDtrsm(FAKE, cc.Right, cc.Lower, cc.Trans, cc.NonUnit, 6, 4, 1.0, 0, 3, 0,
                                  6) == 1
        assert cc.cuMultiGPUZtrsm(FAKE, cc.Left, cc.Upper, cc.NoTrans, cc.Unit, 6, 4, 1.0, 0, 6, 0,
                                  5) == 1
    assert x.calls == [("cuMultiGPUDgemm", 8), ("cuMultiGPUDgemm", 10), ("cuMultiGPUZgemm", 13),
                       ("cuMultiGPUDsyrk", 7), ("cuMultiGPUDsyrk", 10), ("cuMultiGPUZherk", 2),
                       ("cuMultiGPUDtrmm", 9), ("cuMultiGPUZtrmm", 11), ("cuMultiGPUDtrsm", 9),
                       ("cuMultiGPUZtrsm", 11)]


def test_multigpu_lapack_argument_indices(cc):
    """reference lapack/dpotrf.c:302-308 (lda -> -4), lapack/dtrtri.c:475-481
    (lda -> -5), lapack/dlauum.c:299-305, lapack/dpotri.c:53-59."""
    with cc.Xerbla() as x:
        for fn in ("Dpotrf", "Zpotrf", "Dlauum", "Zlauum", "Dpotri", "Zpotri"):
            assert getattr(cc, "cuMultiGPU" + fn)(FAKE, cc.Lower, 10, 0, 9) == (1, -4)
        for fn in ("Dtrtri", "Ztrtri"):
            assert getattr(cc, "cuMultiGPU" + fn)(FAKE, cc.Upper, cc.NonUnit, 10, 0, 9) == (1, -5)
    assert [i for _, i in x.calls] == [4] * 6 + [5] * 2


def test_multigpu_quick_returns(cc):
    # nothing to do: no device, no host array touched
    assert cc.cuMultiGPUDgemm(FAKE, cc.NoTrans, cc.NoTrans, 0, 4, 3, 1.0, 0, 1, 0, 3, 1.0, 0, 1) == 0
    assert cc.cuMultiGPUDgemm(FAKE, cc.NoTrans, cc.NoTrans, 4, 4, 3, 0.0, 0, 4, 0, 3, 1.0, 0, 4) == 0
    assert cc.cuMultiGPUDsyrk(FAKE, cc.Lower, cc.NoTrans, 4, 0, 1.0, 0, 4, 1.0, 0, 4) == 0
    assert cc.cuMultiGPUDtrsm(FAKE, cc.Left, cc.Lower, cc.NoTrans, cc.NonUnit, 0, 4, 1.0, 0, 1, 0, 1) == 0
    assert cc.cuMultiGPUDpotrf(FAKE, cc.Lower, 0, 0, 1) == (0, 0)
    assert cc.cuMultiGPUZtrtri(FAKE, cc.Lower, cc.Unit, 0, 0, 1) == (0, 0)


def test_single_and_single_complex_argument_indices(cc):
    """The single / single-complex entry points check their arguments like
    their double twins, before anything is widened."""
    L, U, N, T, C = cc.Lower, cc.Upper, cc.NoTrans, cc.Trans, cc.ConjTrans
    with cc.Xerbla() as x:
        assert cc.cuCgemm2(FAKE, N, N, 5, 4, 3, 1.0, 0, 4, 0, 3, 1.0, 0, 5, 0, 5) == 1
        assert cc.cuCherk(FAKE, L, T, 5, 3, 1.0, 0, 5, 1.0, 0, 5) == 1
        assert cc.cuStrmm2(FAKE, cc.Left, U, T, cc.Unit, 6, 4, 1.0, 0, 6, 0, 6, 0, 5) == 1
        assert cc.cuCtrmm2(FAKE, cc.Right, L, C, cc.NonUnit, 6, 4, 1.0, 0, 3, 0, 6, 0, 6) == 1
        assert cc.cuCtrsm(FAKE, cc.Right, U, C, cc.Unit, 6, 4, 1.0, 0, 4, 0, 5) == 1
        assert cc.cuCpotrf(FAKE, L, 10, 0, 9) == (1, -4)
        assert cc.cuStrtri(FAKE, U, ord("X"), 10, 0, 10) == (1, -3)
        assert cc.cuCtrtri(FAKE, U, cc.NonUnit, 10, 0, 9) == (1, -5)
        assert cc.cuClauum(FAKE, L, 10, 0, 9) == (1, -4)
        assert cc.cuSpotri(FAKE, ord("X"), 10, 0, 10) == (1, -2)
        assert cc.cuStrti22(FAKE, L, cc.NonUnit, 4, 0, 3, 0, 4, 0) == 1
    assert x.calls == [("cuCgemm2", 8), ("cuCherk", 3), ("cuStrmm2", 13), ("cuCtrmm2", 9),
                       ("cuCtrsm", 11), ("cuCpotrf", 4), ("cuStrtri", 3), ("cuCtrtri", 5),
                       ("cuClauum", 4), ("cuSpotri", 2), ("cuStrti22", 6)]
    # quick returns need no device
    assert cc.cuCgemm2(FAKE, N, N, 0, 4, 3, 1.0, 0, 1, 0, 3, 1.0, 0, 1, 0, 1) == 0
    assert cc.cuCpotrf(FAKE, L, 0, 0, 1) == (0, 0)
    assert cc.cuSlogdet(FAKE, 0, 1, 0) == (0, 0.0)
    assert cc.cuClogdet(FAKE, 0, 1, 0) == (0, 0.0)


# ---------------------------------------------------------------- host-memory routines

def _host(cc, name):
    return getattr(cc.lib(), name)


@pytest.mark.parametrize("p", "sdcz")
def test_host_lapack_argument_errors(cc, p):
    """The host-memory LAPACK routines report lda < n as the reference's CPU
    routines do (lapack/dpotrf.c:99-107: info = -4; dtrtri.c:101-104: -5;
    dtrtri2 lda / ldb: -5 / -7, the ldb test last) and an unknown uplo as
    argument 1, before any device work."""
    import ctypes as C
    info = C.c_long(7)
    for r in ("potrf", "lauum", "potri"):
        with cc.Xerbla() as x:
            _host(cc, p + r)(cc.Lower, 10, None, 9, C.byref(info))
        assert info.value == -4 and x.calls == [(p + r, 4)]
        with cc.Xerbla() as x:
            _host(cc, p + r)(ord("X"), 10, None, 10, C.byref(info))
        assert info.value == -1 and x.calls == [(p + r, 1)]
        info.value = 7
        _host(cc, p + r)(cc.Lower, 0, None, 1, C.byref(info))   # n = 0: quick return
        assert info.value == 0
    with cc.Xerbla() as x:
        _host(cc, p + "trtri")(cc.Upper, cc.NonUnit, 10, None, 3, C.byref(info))
    assert info.value == -5 and x.calls == [(p + "trtri", 5)]
    with cc.Xerbla() as x:
        _host(cc, p + "trtri")(cc.Upper, ord("Q"), 10, None, 10, C.byref(info))
    assert info.value == -2 and x.calls == [(p + "trtri", 2)]
    for lda, ldb, want in ((9, 10, -5), (10, 9, -7), (9, 9, -7)):
        with cc.Xerbla() as x:
            _host(cc, p + "trtri2")(cc.Lower, cc.NonUnit, 10, None, lda, None, ldb, C.byref(info))
        assert info.value == want and x.calls == [(p + "trtri2", -want)]


@pytest.mark.parametrize("p", "sdcz")
def test_host_blas_argument_errors(cc, p):
    """Host BLAS-3: the reference CPU indices (blas/dgemm.c:41-51 lda / ldb /
    ldc -> 8 / 10 / 13; blas/dsyrk.c:40-45 lda / ldc -> 7 / 10, zherk.c:42
    trans = T -> 2; blas/dtrmm.c:50-55, dtrsm.c:40-45 lda / ldb -> 9 / 11,
    dtrmm.c:213-219 ldx -> 13) and no device work on quick returns."""
    one = {"s": 1.0, "d": 1.0, "c": cc.cscalar(1.0), "z": cc.zscalar(1.0)}[p]
    zero = {"s": 0.0, "d": 0.0, "c": cc.cscalar(0.0), "z": cc.zscalar(0.0)}[p]
    gemm = _host(cc, p + "gemm")
    for (m, n, k, lda, ldb, ldc), idx in (((5, 4, 3, 4, 3, 5), 8), ((5, 4, 3, 5, 2, 5), 10),
                                          ((5, 4, 3, 5, 3, 4), 13)):
        with cc.Xerbla() as x:
            gemm(cc.NoTrans, cc.NoTrans, m, n, k, one, None, lda, None, ldb, one, None, ldc)
        assert x.calls == [(p + "gemm", idx)]
    with cc.Xerbla() as x:
        gemm(ord("X"), cc.NoTrans, 5, 4, 3, one, None, 5, None, 3, one, None, 5)
    assert x.calls == [(p + "gemm", 1)]
    with cc.Xerbla() as x:  # quick returns read nothing (null arrays)
        gemm(cc.NoTrans, cc.NoTrans, 0, 4, 3, one, None, 1, None, 3, one, None, 1)
        gemm(cc.NoTrans, cc.NoTrans, 4, 4, 3, zero, None, 4, None, 3, one, None, 4)
    assert x.calls == []
    herk = _host(cc, p + ("syrk" if p in "sd" else "herk"))
    name = p + ("syrk" if p in "sd" else "herk")
    for (n, k, lda, ldc), idx in (((5, 3, 4, 5), 7), ((5, 3, 5, 4), 10)):
        with cc.Xerbla() as x:
            herk(cc.Lower, cc.NoTrans, n, k, 1.0, None, lda, 1.0, None, ldc)
        assert x.calls == [(name, idx)]
    if p in "cz":
        with cc.Xerbla() as x:
            herk(cc.Lower, cc.Trans, 5, 3, 1.0, None, 5, 1.0, None, 5)
        assert x.calls == [(name, 2)]
    for r in ("trmm", "trsm"):
        f = _host(cc, p + r)
        for (m, n, lda, ldb), idx in (((5, 4, 4, 5), 9), ((5, 4, 5, 4), 11)):
            with cc.Xerbla() as x:
                f(cc.Left, cc.Lower, cc.NoTrans, cc.NonUnit, m, n, one, None, lda, None, ldb)
            assert x.calls == [(p + r, idx)]
        with cc.Xerbla() as x:
            f(ord("Z"), cc.Lower, cc.NoTrans, cc.NonUnit, 5, 4, one, None, 5, None, 5)
        assert x.calls == [(p + r, 1)]
    with cc.Xerbla() as x:
        _host(cc, p + "trmm2")(cc.Left, cc.Lower, cc.NoTrans, cc.NonUnit, 5, 4, one, None, 5, None, 5, None, 4)
    assert x.calls == [(p + "trmm2", 13)]


def test_host_logdet_empty(cc):
    for p in "sdcz":
        assert _host(cc, p + "logdet")(None, 1, 0) == 0.0
