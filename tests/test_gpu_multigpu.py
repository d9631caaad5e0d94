"""GPU parity of the host-pointer API (cuMultiGPU*: the reference's
include/cumultigpu.h and the multi-GPU halves of include/blas.h and
include/lapack.h) collapsed onto one device.

Operands are numpy arrays in host memory, exactly as the reference's callers
pass them (test/blas/cumultigpudgemm.c, test/lapack/cumultigpudpotrf.c); the
results are checked against the oracle with the same bounds as the
device-pointer tests, and every array is checked outside the result too:
padding rows and the unreferenced triangle come back bit-identical.  Large
cases cross the pinned staging buffers' size (32 MiB) by column groups and, for
one very tall matrix, by row ranges of single columns.
"""
import itertools

import numpy as np
import pytest

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mgpu(gpu):
    ctx = gpu.MultiGPU((0,))
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def mblas(gpu, mgpu):
    h = gpu.MultiGPUBLASHandle(mgpu)
    yield h
    h.close()


@pytest.fixture(scope="module")
def mlapack(gpu, mgpu):
    h = gpu.MultiGPULAPACKHandle(mgpu)
    yield h
    h.close()


def rng(*seed):
    return np.random.default_rng(list(seed))


def unif(r, shape, cplx=False):
    a = r.random(shape)
    if cplx:
        a = a + 1j * r.random(shape)
    return a


def opm(M, t):
    return M if t == "N" else (M.T if t == "T" else M.conj().T)


def padded(M, pad, fill):
    out = np.full((M.shape[0] + pad, M.shape[1]), fill, dtype=M.dtype, order="F")
    out[: M.shape[0]] = M
    return out


def tri(M, uplo):
    return np.tril(M) if uplo == "L" else np.triu(M)


def other_mask(n, uplo):
    return np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)


def rel(a, b):
    d = np.linalg.norm(a - b)
    return 0.0 if d == 0.0 else d / np.linalg.norm(b)


# ---------------------------------------------------------------- context

def test_context_is_one_device_and_runs_tasks(gpu, mgpu):
    """reference multigpu/multigpu_seq.c:144-176: tasks run on context i in
    [0, count); here count is 1 whatever the number of devices given."""
    assert mgpu.context_count() == 1
    ran = []
    t = gpu.Task(gpu.TaskFunction(lambda a: ran.append(1) or 0))
    assert mgpu.run_task(0, t) == 0
    assert mgpu.run_task(1, t) == gpu.CUDA_ERROR_INVALID_VALUE
    assert t.destroy() == (0, 0) and ran == [1]
    assert mgpu.synchronize() == 0
    with gpu.MultiGPU((0, 0, 0)) as two:    # extra devices are accepted, not used
        assert two.context_count() == 1


# ---------------------------------------------------------------- GEMM

def run_gemm(gpu, mblas, ta, tb, m, n, k, cplx, al=None, be=None):
    r = rng(m, n, k, ord(ta), ord(tb), int(cplx))
    A = unif(r, (m, k) if ta == "N" else (k, m), cplx)
    B = unif(r, (k, n) if tb == "N" else (n, k), cplx)
    C = unif(r, (m, n), cplx)
    if al is None:
        al, be = (0.75 - 0.5j, 1.25 + 0.25j) if cplx else (0.75, -1.25)
    Ap, Bp, Cp = padded(A, 3, 7.0), padded(B, 1, 7.0), padded(C, 2, 5.0)
    fn = gpu.cuMultiGPUZgemm if cplx else gpu.cuMultiGPUDgemm
    st = fn(mblas, ord(ta), ord(tb), m, n, k, al, Ap, Ap.shape[0], Bp, Bp.shape[0], be, Cp, Cp.shape[0])
    assert st == 0
    ref = np.asfortranarray(C.copy())
    (O.zgemm if cplx else O.dgemm)(ta, tb, m, n, k, al, np.asfortranarray(A), np.asfortranarray(B), be, ref)
    bound = 2 * (k + 2) * EPS * (abs(al) * np.abs(opm(A, ta)) @ np.abs(opm(B, tb)) + abs(be) * np.abs(C))
    err = np.abs(Cp[:m] - ref)
    assert np.all(err <= bound), f"max err {err.max():.3e}"
    assert np.all(Cp[m:] == 5.0)                      # padding rows untouched
    assert np.all(Ap[A.shape[0]:] == 7.0) and np.all(Bp[B.shape[0]:] == 7.0)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 29, 41), (300, 257, 129)])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm(gpu, mblas, ta, tb, m, n, k):
    run_gemm(gpu, mblas, ta, tb, m, n, k, False)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "C"), ("C", "N"), ("T", "C")])
def test_zgemm(gpu, mblas, ta, tb):
    run_gemm(gpu, mblas, ta, tb, 45, 70, 33, True)


def test_dgemm_alpha_or_beta_zero_reads_nothing(gpu, mblas):
    """reference blas/dgemm.c:455-478: alpha = 0 leaves A and B unread, beta = 0
    leaves C unread, so NaNs there never reach the result."""
    m, n, k = 40, 30, 20
    nanA = np.full((m, k), np.nan, order="F")
    nanB = np.full((k, n), np.nan, order="F")
    C = np.asfortranarray(rng(1).random((m, n)))
    C0 = C.copy()
    assert gpu.cuMultiGPUDgemm(mblas, gpu.NoTrans, gpu.NoTrans, m, n, k, 0.0, nanA, m, nanB, k, -2.0, C, m) == 0
    np.testing.assert_array_equal(C, -2.0 * C0)
    r = rng(2)
    A, B = np.asfortranarray(r.random((m, k))), np.asfortranarray(r.random((k, n)))
    Cn = np.full((m, n), np.nan, order="F")
    assert gpu.cuMultiGPUDgemm(mblas, gpu.NoTrans, gpu.NoTrans, m, n, k, 1.0, A, m, B, k, 0.0, Cn, m) == 0
    assert np.all(np.abs(Cn - A @ B) <= 2 * (k + 2) * EPS * (A @ B))


def test_dgemm_crosses_staging_by_columns(gpu, mblas):
    """C (2500 x 2500, 50 MB) and A (2500 x 1500, 30 MB) span several staging
    pieces of whole columns."""
    m, n, k = 2500, 2500, 1500
    r = rng(3)
    A = np.asfortranarray(r.random((m, k)))
    B = np.asfortranarray(r.random((k, n)))
    C = np.asfortranarray(r.random((m, n)))
    want = 0.5 * (A @ B) + C
    assert gpu.cuMultiGPUDgemm(mblas, gpu.NoTrans, gpu.NoTrans, m, n, k, 0.5, A, m, B, k, 1.0, C, m) == 0
    assert np.max(np.abs(C - want) / want) <= 4 * (k + 2) * EPS


def test_dgemm_crosses_staging_by_rows(gpu, mblas):
    """A 4.5M-row column (36 MB) is larger than one staging buffer: it moves in
    row ranges."""
    m, n, k = 4_500_000, 2, 1
    r = rng(4)
    A = np.asfortranarray(r.random((m, k)))
    B = np.asfortranarray(np.array([[2.0, -3.0]]))
    C = np.asfortranarray(r.random((m, n)))
    want = A @ B + C
    assert gpu.cuMultiGPUDgemm(mblas, gpu.NoTrans, gpu.NoTrans, m, n, k, 1.0, A, m, B, k, 1.0, C, m) == 0
    assert np.max(np.abs(C - want)) <= 4 * EPS * np.max(np.abs(A @ B) + np.abs(C))


# ---------------------------------------------------------------- SYRK / HERK

@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("beta", [2.0, 0.0])
def test_dsyrk(gpu, mblas, uplo, trans, beta):
    n, k = 130, 70
    r = rng(n, k, ord(uplo), ord(trans))
    A = np.asfortranarray(unif(r, (n, k) if trans == "N" else (k, n)))
    C = unif(r, (n, n))
    if beta == 0.0:
        C[tri(np.ones((n, n)), uplo).astype(bool)] = np.nan      # unread
    Cp = padded(np.asfortranarray(C), 2, 5.0)
    st = gpu.cuMultiGPUDsyrk(mblas, ord(uplo), ord(trans), n, k, -0.5, A, A.shape[0], beta, Cp, n + 2)
    assert st == 0
    C0 = np.where(np.isnan(C), 0.0, C)
    ref = np.asfortranarray(C0)
    O.dsyrk(uplo, trans, n, k, -0.5, A, beta, ref)
    M = opm(A, trans)
    bound = 2 * (k + 2) * EPS * (0.5 * np.abs(M) @ np.abs(M).T + abs(beta) * np.abs(C0))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    assert np.all(np.abs(Cp[:n] - ref)[mask] <= bound[mask])
    np.testing.assert_array_equal(Cp[:n][other_mask(n, uplo)], C[other_mask(n, uplo)])
    assert np.all(Cp[n:] == 5.0)


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("trans", "NC")
def test_zherk(gpu, mblas, uplo, trans):
    n, k = 90, 50
    r = rng(n, k, ord(uplo), ord(trans), 1)
    A = np.asfortranarray(unif(r, (n, k) if trans == "N" else (k, n), True))
    C = np.asfortranarray(unif(r, (n, n), True))
    Cg = C.copy(order="F")
    st = gpu.cuMultiGPUZherk(mblas, ord(uplo), ord(trans), n, k, 0.75, A, A.shape[0], 1.5, Cg, n)
    assert st == 0
    ref = np.asfortranarray(C.copy())
    O.zherk(uplo, trans, n, k, 0.75, A, 1.5, ref)
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    M = opm(A, trans)
    bound = 4 * (k + 2) * EPS * (0.75 * np.abs(M) @ np.abs(M).T + 1.5 * np.abs(C))
    assert np.all(np.abs(Cg - ref)[mask] <= bound[mask])
    np.testing.assert_array_equal(Cg[other_mask(n, uplo)], C[other_mask(n, uplo)])


# ---------------------------------------------------------------- TRMM / TRSM

VARIANTS_D = list(itertools.product("LR", "LU", "NT", "NU"))


def tri_rand(r, k, cplx=False):
    T = r.standard_normal((k, k))
    if cplx:
        T = T + 1j * r.standard_normal((k, k))
    return np.asfortranarray(T / np.sqrt(k) + 3 * np.eye(k))


def poison(T, uplo, diag):
    """NaN in the unreferenced triangle (and a unit diagonal)."""
    k = T.shape[0]
    T = np.asfortranarray(T.copy())
    T[np.triu_indices(k, 1) if uplo == "L" else np.tril_indices(k, -1)] = np.nan
    if diag == "U":
        np.fill_diagonal(T, np.nan)
    return T


@pytest.mark.parametrize("m,n", [(29, 23), (130, 70)])
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS_D)
def test_dtrmm(gpu, mblas, side, uplo, trans, diag, m, n):
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag))
    k = m if side == "L" else n
    T = unif(r, (k, k))
    B = unif(r, (m, n))
    Tg = poison(T, uplo, diag)
    Bp = padded(np.asfortranarray(B), 1, 3.0)
    st = gpu.cuMultiGPUDtrmm(mblas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, -1.5, Tg, k, Bp, m + 1)
    assert st == 0
    ref = np.asfortranarray(B.copy())
    O.dtrmm(side, uplo, trans, diag, m, n, -1.5, np.asfortranarray(T), ref)
    Tm = opm(np.tril(T) if uplo == "L" else np.triu(T), trans)
    if diag == "U":
        Tm = Tm - np.diag(np.diag(Tm)) + np.eye(k)
    absprod = np.abs(Tm) @ np.abs(B) if side == "L" else np.abs(B) @ np.abs(Tm)
    assert np.all(np.abs(Bp[:m] - ref) <= 2 * (k + 2) * EPS * 1.5 * absprod)
    assert np.all(Bp[m:] == 3.0)


@pytest.mark.parametrize("side,uplo,trans,diag", [("L", "L", "C", "N"), ("R", "U", "T", "U"),
                                                  ("L", "U", "N", "U"), ("R", "L", "C", "N")])
def test_ztrmm(gpu, mblas, side, uplo, trans, diag):
    m, n = 45, 70
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag), 1)
    k = m if side == "L" else n
    T = unif(r, (k, k), True)
    B = np.asfortranarray(unif(r, (m, n), True))
    Bg = B.copy(order="F")
    al = 0.5 - 1j
    st = gpu.cuMultiGPUZtrmm(mblas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, al,
                             poison(T, uplo, diag), k, Bg, m)
    assert st == 0
    ref = O.fmat(m, n, m, np.complex128)
    O.ztrmm2(side, uplo, trans, diag, m, n, al, np.asfortranarray(T), B, ref)
    assert rel(Bg, ref) <= 4 * k * EPS


@pytest.mark.parametrize("m,n", [(29, 23), (200, 150)])
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS_D)
def test_dtrsm(gpu, mblas, side, uplo, trans, diag, m, n):
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag), 3)
    k = m if side == "L" else n
    T = poison(tri_rand(r, k), uplo, diag)
    B = np.asfortranarray(unif(r, (m, n)))
    Bp = padded(B, 2, 4.0)
    st = gpu.cuMultiGPUDtrsm(mblas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, 0.75, T, k, Bp, m + 2)
    assert st == 0
    ref = np.asfortranarray(B.copy())
    O.dtrsm(side, uplo, trans, diag, m, n, 0.75, T, ref)
    assert rel(Bp[:m], ref) < 20 * k * EPS
    assert np.all(Bp[m:] == 4.0)


@pytest.mark.parametrize("side,uplo,trans,diag", [("L", "L", "C", "N"), ("R", "U", "N", "U"),
                                                  ("R", "L", "T", "N"), ("L", "U", "C", "U")])
def test_ztrsm(gpu, mblas, side, uplo, trans, diag):
    m, n = 130, 90
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag), 5)
    k = m if side == "L" else n
    T = poison(tri_rand(r, k, True), uplo, diag)
    B = np.asfortranarray(unif(r, (m, n), True))
    Bg = B.copy(order="F")
    al = 0.5 + 0.5j
    assert gpu.cuMultiGPUZtrsm(mblas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, al, T, k, Bg, m) == 0
    ref = B.copy(order="F")
    O.ztrsm(side, uplo, trans, diag, m, n, al, T, ref)
    assert rel(Bg, ref) < 20 * k * EPS


def test_trsm_trmm_alpha_zero(gpu, mblas):
    """reference blas/dtrsm.c:285-288: B <- 0, A and B unread."""
    m, n = 20, 10
    T = np.full((m, m), np.nan, order="F")
    for fn in (gpu.cuMultiGPUDtrsm, gpu.cuMultiGPUDtrmm):
        B = np.full((m, n), np.nan, order="F")
        assert fn(mblas, gpu.Left, gpu.Lower, gpu.NoTrans, gpu.NonUnit, m, n, 0.0, T, m, B, m) == 0
        assert np.all(B == 0.0)


# ---------------------------------------------------------------- LAPACK

@pytest.mark.parametrize("n", [1, 33, 300, 1030])
@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_dpotri(gpu, mlapack, uplo, n):
    A = O.spd(n + 5, n, 1.0)
    Ap = padded(A, 3, -7.0)
    st, info = gpu.cuMultiGPUDpotrf(mlapack, ord(uplo), n, Ap, n + 3)
    assert st == 0 and info == 0
    F = Ap[:n].copy(order="F")
    ref, rinfo = O.dpotrf(uplo, A.copy(order="F"))
    assert rinfo == 0
    assert O.potrf_backward_error(A, F, uplo) <= 10 * n * EPS
    assert rel(tri(F, uplo), tri(ref, uplo)) <= 10 * n * EPS * 5
    np.testing.assert_array_equal(F[other_mask(n, uplo)], A[other_mask(n, uplo)])
    assert np.all(Ap[n:] == -7.0)
    st, info = gpu.cuMultiGPUDpotri(mlapack, ord(uplo), n, Ap, n + 3)
    assert st == 0 and info == 0
    X = Ap[:n]
    refi, _ = O.dpotri(uplo, ref)
    assert O.inverse_residual(A, X, uplo) <= 10 * n * EPS
    assert rel(O.herm_from(X, uplo), O.herm_from(refi, uplo)) <= 10 * n * EPS * np.linalg.cond(A)
    assert np.all(Ap[n:] == -7.0)


@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_not_spd(gpu, mlapack, uplo):
    """reference lapack/dpotrf.c:324-327: the order of the failing minor in
    info, CUDA_ERROR_INVALID_VALUE returned."""
    n = 300
    A = O.spd(9, n, 0.5)
    L = np.linalg.cholesky(A)
    A[150, 150] = L[150, :150] @ L[150, :150] - 0.25
    _, rinfo = O.dpotrf(uplo, A.copy(order="F"))
    Ag = A.copy(order="F")
    st, info = gpu.cuMultiGPUDpotrf(mlapack, ord(uplo), n, Ag, n)
    assert st == gpu.CUDA_ERROR_INVALID_VALUE and info == rinfo == 151


def test_dpotrf_crosses_staging(gpu, mlapack):
    """n = 4500 (162 MB): the matrix moves in six staging pieces each way."""
    n = 4500
    G = rng(11).standard_normal((n, n))
    A = np.asfortranarray(G @ G.T / n + np.eye(n))
    F = A.copy(order="F")
    st, info = gpu.cuMultiGPUDpotrf(mlapack, gpu.Lower, n, F, n)
    assert st == 0 and info == 0
    assert O.potrf_backward_error(A, F, "L") <= 10 * n * EPS
    np.testing.assert_array_equal(F[other_mask(n, "L")], A[other_mask(n, "L")])


@pytest.mark.parametrize("n", [1, 65, 300])
@pytest.mark.parametrize("uplo", "LU")
def test_zpotrf_zpotri(gpu, mlapack, uplo, n):
    A = O.hpd(n, n, 0.1)
    F = A.copy(order="F")
    st, info = gpu.cuMultiGPUZpotrf(mlapack, ord(uplo), n, F, n)
    assert st == 0 and info == 0
    assert O.potrf_backward_error(A, F, uplo) <= 10 * n * EPS
    ref, _ = O.zpotrf(uplo, A.copy(order="F"))
    condA = np.linalg.cond(A)
    assert rel(tri(F, uplo), tri(ref, uplo)) <= 10 * n * EPS * condA
    np.testing.assert_array_equal(F[other_mask(n, uplo)], A[other_mask(n, uplo)])
    st, info = gpu.cuMultiGPUZpotri(mlapack, ord(uplo), n, F, n)
    assert st == 0 and info == 0
    assert O.inverse_residual(A, F, uplo) <= 10 * n * EPS
    refi, _ = O.zpotri(uplo, ref)
    assert rel(O.herm_from(F, uplo), O.herm_from(refi, uplo)) <= 10 * n * EPS * condA


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_trtri(gpu, mlapack, cplx, uplo, diag):
    n = 257
    T = tri_rand(rng(n, ord(uplo), ord(diag), int(cplx)), n, cplx)
    X = T.copy(order="F")
    fn = gpu.cuMultiGPUZtrtri if cplx else gpu.cuMultiGPUDtrtri
    st, info = fn(mlapack, ord(uplo), ord(diag), n, X, n)
    assert st == 0 and info == 0
    ref, _ = (O.ztrtri if cplx else O.dtrtri)(uplo, diag, T.copy(order="F"))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    if diag == "U":
        np.fill_diagonal(mask, False)
        assert np.all(np.diag(X) == np.diag(T))
    assert rel(X[mask], ref[mask]) <= 20 * n * EPS
    np.testing.assert_array_equal(X[other_mask(n, uplo)], T[other_mask(n, uplo)])


@pytest.mark.parametrize("uplo", "LU")
def test_dtrtri_singular(gpu, mlapack, uplo):
    n = 300
    T = tri_rand(rng(1), n)
    T[40, 40] = 0.0
    T[250, 250] = 0.0
    _, rinfo = O.dtrtri(uplo, "N", T.copy(order="F"))
    st, info = gpu.cuMultiGPUDtrtri(mlapack, ord(uplo), ord("N"), n, T, n)
    assert st == gpu.CUDA_ERROR_INVALID_VALUE and info == rinfo == (251 if uplo == "L" else 41)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
def test_lauum(gpu, mlapack, cplx, uplo):
    n = 200
    T = tri_rand(rng(n, ord(uplo), int(cplx), 7), n, cplx)
    X = T.copy(order="F")
    fn = gpu.cuMultiGPUZlauum if cplx else gpu.cuMultiGPUDlauum
    st, info = fn(mlapack, ord(uplo), n, X, n)
    assert st == 0 and info == 0
    ref, _ = (O.zlauum if cplx else O.dlauum)(uplo, T.copy(order="F"))
    assert rel(tri(X, uplo), tri(ref, uplo)) <= 10 * n * EPS
    np.testing.assert_array_equal(X[other_mask(n, uplo)], T[other_mask(n, uplo)])


def test_handles_synchronize(gpu, mblas, mlapack):
    assert mblas.synchronize() == 0 and mlapack.synchronize() == 0


# ---------------------------------------------------------------- single and single complex

EPS32 = float(np.finfo(np.float32).eps)


def f32(M):
    return np.asfortranarray(M.astype(np.complex64 if np.iscomplexobj(M) else np.float32))


def w64(M):
    return np.asfortranarray(M.astype(np.complex128 if np.iscomplexobj(M) else np.float64))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_sc_gemm(gpu, mblas, cplx, ta, tb):
    m, n, k = 70, 45, 33
    r = rng(m, n, k, ord(ta), ord(tb), int(cplx), 9)
    A = f32(unif(r, (m, k) if ta == "N" else (k, m), cplx))
    B = f32(unif(r, (k, n) if tb == "N" else (n, k), cplx))
    C = f32(unif(r, (m, n), cplx))
    al, be = (0.75 - 0.5j, 1.25 + 0.25j) if cplx else (0.75, -1.25)
    Cp = padded(C, 2, 5.0)
    fn = gpu.cuMultiGPUCgemm if cplx else gpu.cuMultiGPUSgemm
    assert fn(mblas, ord(ta), ord(tb), m, n, k, al, A, A.shape[0], B, B.shape[0], be, Cp, m + 2) == 0
    ref = w64(C)
    (O.zgemm if cplx else O.dgemm)(ta, tb, m, n, k, al, w64(A), w64(B), be, ref)
    bound = 2 * (k + 2) * EPS32 * (abs(al) * np.abs(opm(w64(A), ta)) @ np.abs(opm(w64(B), tb)) + abs(be) * np.abs(w64(C)))
    assert np.all(np.abs(Cp[:m] - ref) <= bound)
    assert np.all(Cp[m:] == 5.0)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
def test_sc_syrk_herk(gpu, mblas, cplx, uplo):
    n, k = 80, 40
    r = rng(n, k, ord(uplo), int(cplx), 10)
    A = f32(unif(r, (n, k), cplx))
    C = f32(unif(r, (n, n), cplx))
    Cg = C.copy(order="F")
    fn = gpu.cuMultiGPUCherk if cplx else gpu.cuMultiGPUSsyrk
    assert fn(mblas, ord(uplo), gpu.NoTrans, n, k, 0.5, A, n, 2.0, Cg, n) == 0
    ref = w64(C)
    (O.zherk if cplx else O.dsyrk)(uplo, "N", n, k, 0.5, w64(A), 2.0, ref)
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    bound = 2 * (k + 2) * EPS32 * (0.5 * np.abs(w64(A)) @ np.abs(w64(A)).T + 2.0 * np.abs(w64(C)))
    assert np.all(np.abs(Cg - ref)[mask] <= bound[mask])
    np.testing.assert_array_equal(Cg[other_mask(n, uplo)], C[other_mask(n, uplo)])


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("solve", [False, True])
def test_sc_trmm_trsm(gpu, mblas, cplx, solve):
    m, n = 60, 35
    r = rng(m, n, int(cplx), int(solve), 11)
    T = tri_rand(r, m, cplx)
    T = f32(T)
    B = f32(unif(r, (m, n), cplx))
    Bg = B.copy(order="F")
    al = (0.5 + 0.5j) if cplx else 0.75
    name = ("cuMultiGPUCtr" if cplx else "cuMultiGPUStr") + ("sm" if solve else "mm")
    assert getattr(gpu, name)(mblas, gpu.Left, gpu.Lower, gpu.NoTrans, gpu.NonUnit, m, n, al,
                              poison(T, "L", "N"), m, Bg, m) == 0
    ref = w64(B)
    if solve:
        (O.ztrsm if cplx else O.dtrsm)("L", "L", "N", "N", m, n, al, w64(T), ref)
        assert rel(w64(Bg), ref) <= 20 * m * EPS32
    else:
        out = O.fmat(m, n, m, np.complex128 if cplx else np.float64)
        (O.ztrmm2 if cplx else O.dtrmm2)("L", "L", "N", "N", m, n, al, w64(T), w64(B), out)
        assert rel(w64(Bg), out) <= 4 * m * EPS32


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
def test_sc_potrf_potri_trtri_lauum(gpu, mlapack, cplx, uplo):
    n = 200
    A = f32(O.hpd(n, n, 1.0) if cplx else O.spd(n, n, 1.0))
    F = A.copy(order="F")
    p = "C" if cplx else "S"
    st, info = getattr(gpu, f"cuMultiGPU{p}potrf")(mlapack, ord(uplo), n, F, n)
    assert st == 0 and info == 0
    assert O.potrf_backward_error(w64(A), w64(F), uplo) <= 20 * n * EPS32
    np.testing.assert_array_equal(F[other_mask(n, uplo)], A[other_mask(n, uplo)])
    X = F.copy(order="F")
    st, info = getattr(gpu, f"cuMultiGPU{p}potri")(mlapack, ord(uplo), n, X, n)
    assert st == 0 and info == 0
    assert O.inverse_residual(w64(A), w64(X), uplo) <= 10 * n * EPS32
    T = F.copy(order="F")
    st, info = getattr(gpu, f"cuMultiGPU{p}trtri")(mlapack, ord(uplo), gpu.NonUnit, n, T, n)
    assert st == 0 and info == 0
    ref, _ = (O.ztrtri if cplx else O.dtrtri)(uplo, "N", w64(F))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    assert rel(w64(T)[mask], ref[mask]) <= 10 * n * EPS32
    Lm = F.copy(order="F")
    st, info = getattr(gpu, f"cuMultiGPU{p}lauum")(mlapack, ord(uplo), n, Lm, n)
    assert st == 0 and info == 0
    refl, _ = (O.zlauum if cplx else O.dlauum)(uplo, w64(F))
    assert rel(tri(w64(Lm), uplo), tri(refl, uplo)) <= 2 * (n + 4) * EPS32
