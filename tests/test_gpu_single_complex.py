"""GPU parity of the single-precision and single-complex entry points
(cuStrmm2, cuStrtri, cuStrti22, cuSlauum, cuSpotri, cuSlogdet and the c
surface: cuCgemm2, cuCherk, cuCtrmm2, cuCtrsm, cuCpotrf, cuCtrtri,
cuCtrti22, cuClauum, cuCpotri, cuClogdet).

The reference is the oracle's double / double-complex routine applied to the
same FP32 inputs widened exactly.  The library computes the level-3 products
in FP32 / single complex (FP32 matrix cores) and the diagonal blocks in
double / double complex, so the bounds are the usual error bounds in units of
eps32 = 2^-23 that grow with the inner dimension k of a product
(|D - ref| <= 2 (k + 4) eps32 (|alpha| |A| |B| + |beta| |C|) per element, as
the reference's own BLAS tests bound each element) and with the order n of a
factorisation, inverse or solve (10 n eps32, relative).  A routine small
enough to stay inside one diagonal block (trti22) is still computed in double
and rounded once (4 eps32).  Padding rows and the triangle a routine does not
reference come back bit-identical, and info / argument behaviour follows the
double twins.
"""
import itertools

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def rng(*seed):
    return np.random.default_rng(list(seed))


def unif(r, shape, cplx):
    a = r.random(shape)
    if cplx:
        a = a + 1j * r.random(shape)
    return a.astype(np.complex64 if cplx else np.float32)


def opm(M, t):
    return M if t == "N" else (M.T if t == "T" else M.conj().T)


def padded(M, pad, fill):
    out = np.full((M.shape[0] + pad, M.shape[1]), fill, dtype=M.dtype, order="F")
    out[: M.shape[0]] = M
    return out


def wide(M):
    return np.asfortranarray(M.astype(np.complex128 if np.iscomplexobj(M) else np.float64))


def tri(M, uplo):
    return np.tril(M) if uplo == "L" else np.triu(M)


def other_mask(n, uplo):
    return np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)


def rel(a, b):
    d = np.linalg.norm(a - b)
    return 0.0 if d == 0.0 else d / np.linalg.norm(b)


def sync():
    torch.cuda.synchronize()


# ---------------------------------------------------------------- BLAS

@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 29, 41), (130, 70, 200)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "C"), ("T", "N"), ("C", "T")])
def test_cgemm2(gpu, blas, ta, tb, m, n, k):
    r = rng(m, n, k, ord(ta), ord(tb))
    A = unif(r, (m, k) if ta == "N" else (k, m), True)
    B = unif(r, (k, n) if tb == "N" else (n, k), True)
    C = unif(r, (m, n), True)
    al, be = 0.75 - 0.5j, 1.25 + 0.25j
    Ap, Bp, Cp = padded(A, 3, 7.0), padded(B, 1, 7.0), padded(C, 2, 5.0)
    Dp = padded(np.zeros((m, n), np.complex64), 2, 9.0)
    dA, dB, dC, dD = gpu.to_device(Ap), gpu.to_device(Bp), gpu.to_device(Cp), gpu.to_device(Dp)
    st = gpu.cuCgemm2(blas, ord(ta), ord(tb), m, n, k, al, gpu.ptr(dA), Ap.shape[0], gpu.ptr(dB),
                      Bp.shape[0], be, gpu.ptr(dC), Cp.shape[0], gpu.ptr(dD), Dp.shape[0])
    sync()
    assert st == 0
    D = gpu.from_device(dD)
    ref = wide(C)
    O.zgemm(ta, tb, m, n, k, al, wide(A), wide(B), be, ref)
    bound = 2 * (k + 4) * EPS32 * (abs(al) * np.abs(opm(wide(A), ta)) @ np.abs(opm(wide(B), tb))
                                   + abs(be) * np.abs(wide(C)))
    assert np.all(np.abs(D[:m] - ref) <= bound)
    assert np.all(D[m:] == 9.0)


def test_cgemm2_alpha_beta_zero_reads_nothing(gpu, blas):
    m, n, k = 20, 10, 5
    nan = np.full((m, k), np.nan, np.complex64, order="F")
    dA = gpu.to_device(nan)
    dB = gpu.to_device(np.full((k, n), np.nan, np.complex64, order="F"))
    dC = gpu.to_device(np.full((m, n), np.nan, np.complex64, order="F"))
    st = gpu.cuCgemm2(blas, gpu.NoTrans, gpu.NoTrans, m, n, k, 0.0, gpu.ptr(dA), m, gpu.ptr(dB), k, 0.0,
                      gpu.ptr(dC), m, gpu.ptr(dC), m)
    sync()
    assert st == 0 and np.all(gpu.from_device(dC) == 0)


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("trans", "NC")
def test_cherk(gpu, blas, uplo, trans):
    n, k = 90, 50
    r = rng(n, k, ord(uplo), ord(trans))
    A = unif(r, (n, k) if trans == "N" else (k, n), True)
    C = unif(r, (n, n), True)
    dA, dC = gpu.to_device(A), gpu.to_device(padded(C, 1, 3.0))
    st = gpu.cuCherk(blas, ord(uplo), ord(trans), n, k, 0.75, gpu.ptr(dA), A.shape[0], 1.5, gpu.ptr(dC), n + 1)
    sync()
    assert st == 0
    R = gpu.from_device(dC)
    ref = wide(C)
    O.zherk(uplo, trans, n, k, 0.75, wide(A), 1.5, ref)
    M = opm(wide(A), trans)
    bound = 2 * (k + 4) * EPS32 * (0.75 * np.abs(M) @ np.abs(M).T + 1.5 * np.abs(wide(C)))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    assert np.all(np.abs(R[:n] - ref)[mask] <= bound[mask])
    np.testing.assert_array_equal(R[:n][other_mask(n, uplo)], C[other_mask(n, uplo)])
    assert np.all(R[n:] == 3.0)


def poison(T, uplo, diag):
    k = T.shape[0]
    T = np.asfortranarray(T.copy())
    T[np.triu_indices(k, 1) if uplo == "L" else np.tril_indices(k, -1)] = np.nan
    if diag == "U":
        np.fill_diagonal(T, np.nan)
    return T


def run_trmm(gpu, blas, side, uplo, trans, diag, m, n, cplx):
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag), int(cplx))
    k = m if side == "L" else n
    T = unif(r, (k, k), cplx)
    B = unif(r, (m, n), cplx)
    al = (0.5 - 1j) if cplx else -1.5
    dT, dB = gpu.to_device(poison(T, uplo, diag)), gpu.to_device(B)
    dX = gpu.to_device(padded(np.zeros((m, n), B.dtype), 2, 4.0))
    fn = gpu.cuCtrmm2 if cplx else gpu.cuStrmm2
    st = fn(blas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, al, gpu.ptr(dT), k, gpu.ptr(dB), m,
            gpu.ptr(dX), m + 2)
    sync()
    assert st == 0
    X = gpu.from_device(dX)
    ref = O.fmat(m, n, m, np.complex128 if cplx else np.float64)
    (O.ztrmm2 if cplx else O.dtrmm2)(side, uplo, trans, diag, m, n, al, wide(T), wide(B), ref)
    Tm = opm(tri(wide(T), uplo), trans)
    if diag == "U":
        Tm = Tm - np.diag(np.diag(Tm)) + np.eye(k)
    absprod = np.abs(Tm) @ np.abs(wide(B)) if side == "L" else np.abs(wide(B)) @ np.abs(Tm)
    assert np.all(np.abs(X[:m] - ref) <= 2 * (k + 4) * EPS32 * abs(al) * absprod)
    assert np.all(X[m:] == 4.0)


@pytest.mark.parametrize("side,uplo,trans,diag", list(itertools.product("LR", "LU", "NT", "NU")))
def test_strmm2(gpu, blas, side, uplo, trans, diag):
    run_trmm(gpu, blas, side, uplo, trans, diag, 29, 23, False)


@pytest.mark.parametrize("side,uplo,trans,diag", [("L", "L", "C", "N"), ("R", "U", "T", "U"),
                                                  ("L", "U", "N", "U"), ("R", "L", "C", "N")])
def test_ctrmm2(gpu, blas, side, uplo, trans, diag):
    run_trmm(gpu, blas, side, uplo, trans, diag, 45, 70, True)


def tri_rand(r, k, cplx=False):
    T = r.standard_normal((k, k))
    if cplx:
        T = T + 1j * r.standard_normal((k, k))
    T = T / np.sqrt(k) + 3 * np.eye(k)
    return np.asfortranarray(T.astype(np.complex64 if cplx else np.float32))


@pytest.mark.parametrize("side,uplo,trans,diag", [("L", "L", "C", "N"), ("R", "U", "N", "U"),
                                                  ("R", "L", "T", "N"), ("L", "U", "C", "U")])
def test_ctrsm(gpu, blas, side, uplo, trans, diag):
    m, n = 130, 90
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag))
    k = m if side == "L" else n
    T = poison(tri_rand(r, k, True), uplo, diag)
    B = unif(r, (m, n), True)
    al = 0.5 + 0.5j
    dT, dB = gpu.to_device(T), gpu.to_device(padded(B, 1, 6.0))
    st = gpu.cuCtrsm(blas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, al, gpu.ptr(dT), k, gpu.ptr(dB),
                     m + 1)
    sync()
    assert st == 0
    X = gpu.from_device(dB)
    ref = wide(B)
    O.ztrsm(side, uplo, trans, diag, m, n, al, wide(T), ref)
    assert rel(X[:m], ref) <= 10 * k * EPS32
    assert np.all(X[m:] == 6.0)


# ---------------------------------------------------------------- LAPACK

def hpd32(seed, n, cplx):
    A = O.hpd(seed, n, 1.0) if cplx else O.spd(seed, n, 1.0)
    return np.asfortranarray(A.astype(np.complex64 if cplx else np.float32))


@pytest.mark.parametrize("n", [1, 65, 300])
@pytest.mark.parametrize("uplo", "LU")
def test_cpotrf_cpotri(gpu, lapack, uplo, n):
    A = hpd32(n, n, True)
    Ap = padded(A, 2, 8.0)
    dA = gpu.to_device(Ap)
    st, info = gpu.cuCpotrf(lapack, ord(uplo), n, gpu.ptr(dA), n + 2)
    assert st == 0 and info == 0
    F = gpu.from_device(dA)
    ref, _ = O.zpotrf(uplo, wide(A))
    assert rel(tri(F[:n], uplo), tri(ref, uplo)) <= 10 * n * EPS32
    assert O.potrf_backward_error(wide(A), wide(F[:n]), uplo) <= 10 * n * EPS32
    np.testing.assert_array_equal(F[:n][other_mask(n, uplo)], A[other_mask(n, uplo)])
    assert np.all(F[n:] == 8.0)
    st, info = gpu.cuCpotri(lapack, ord(uplo), n, gpu.ptr(dA), n + 2)
    assert st == 0 and info == 0
    X = gpu.from_device(dA)
    refi, _ = O.zpotri(uplo, wide(F[:n]))
    assert rel(O.herm_from(wide(X[:n]), uplo), O.herm_from(refi, uplo)) <= 10 * n * EPS32
    assert O.inverse_residual(wide(A), wide(X[:n]), uplo) <= 10 * n * EPS32


@pytest.mark.parametrize("uplo", "LU")
def test_cpotrf_info(gpu, lapack, uplo):
    n = 200
    A = O.hpd(2, n, 0.5)
    L = np.linalg.cholesky(A)
    A[77, 77] = np.vdot(L[77, :77], L[77, :77]).real - 0.25
    A32 = np.asfortranarray(A.astype(np.complex64))
    _, rinfo = O.zpotrf(uplo, wide(A32))
    dA = gpu.to_device(A32)
    st, info = gpu.cuCpotrf(lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == rinfo == 78


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_trtri(gpu, lapack, cplx, uplo, diag):
    n = 257
    T = tri_rand(rng(n, ord(uplo), ord(diag), int(cplx)), n, cplx)
    dT = gpu.to_device(T)
    fn = gpu.cuCtrtri if cplx else gpu.cuStrtri
    st, info = fn(lapack, ord(uplo), ord(diag), n, gpu.ptr(dT), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dT)
    ref, _ = (O.ztrtri if cplx else O.dtrtri)(uplo, diag, wide(T))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    if diag == "U":
        np.fill_diagonal(mask, False)
        assert np.all(np.diag(X) == np.diag(T))
    assert rel(wide(X)[mask], ref[mask]) <= 10 * n * EPS32
    np.testing.assert_array_equal(X[other_mask(n, uplo)], T[other_mask(n, uplo)])


def test_strtri_singular(gpu, lapack):
    n = 100
    T = tri_rand(rng(3), n)
    T[40, 40] = 0.0
    dT = gpu.to_device(T)
    st, info = gpu.cuStrtri(lapack, gpu.Lower, gpu.NonUnit, n, gpu.ptr(dT), n)
    assert st == 0 and info == 41
    np.testing.assert_array_equal(gpu.from_device(dT), T)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
def test_lauum(gpu, lapack, cplx, uplo):
    n = 200
    T = tri_rand(rng(n, ord(uplo), int(cplx), 7), n, cplx)
    dT = gpu.to_device(T)
    fn = gpu.cuClauum if cplx else gpu.cuSlauum
    st, info = fn(lapack, ord(uplo), n, gpu.ptr(dT), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dT)
    ref, _ = (O.zlauum if cplx else O.dlauum)(uplo, wide(T))
    assert rel(tri(wide(X), uplo), tri(ref, uplo)) <= 2 * (n + 4) * EPS32
    np.testing.assert_array_equal(X[other_mask(n, uplo)], T[other_mask(n, uplo)])


@pytest.mark.parametrize("uplo", "LU")
def test_spotrf_spotri(gpu, lapack, uplo):
    n = 300
    A = hpd32(5, n, False)
    dA = gpu.to_device(A)
    st, info = gpu.cuSpotrf(lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    F = gpu.from_device(dA)
    st, info = gpu.cuSpotri(lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dA)
    refi, _ = O.dpotri(uplo, wide(F))
    assert rel(O.herm_from(wide(X), uplo), O.herm_from(refi, uplo)) <= 10 * n * EPS32
    assert O.inverse_residual(wide(A), wide(X), uplo) <= 10 * n * EPS32


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", "LU")
def test_trti22(gpu, lapack, cplx, uplo):
    n = 60
    T = tri_rand(rng(n, int(cplx), ord(uplo)), n, cplx)
    dT = gpu.to_device(T)
    dB = gpu.to_device(np.full((n, n), 5.0, T.dtype, order="F"))
    dinfo = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    fn = gpu.cuCtrti22 if cplx else gpu.cuStrti22
    st = fn(lapack, ord(uplo), gpu.NonUnit, n, gpu.ptr(dT), n, gpu.ptr(dB), n, dinfo.data_ptr())
    sync()
    assert st == 0 and int(dinfo.item()) == 0
    X = gpu.from_device(dB)
    ref, _ = (O.ztrtri if cplx else O.dtrtri)(uplo, "N", wide(T))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    assert rel(wide(X)[mask], ref[mask]) <= 4 * EPS32
    assert np.all(X[other_mask(n, uplo)] == 5.0)


@pytest.mark.parametrize("cplx", [False, True])
def test_logdet(gpu, lapack, cplx):
    n, incx = 500, 3
    r = rng(n, int(cplx))
    x = (0.5 + r.random(n * incx)).astype(np.float32)
    if cplx:
        x = (x + 1j * r.random(n * incx)).astype(np.complex64)
    dx = torch.from_numpy(x).cuda()
    fn = gpu.cuClogdet if cplx else gpu.cuSlogdet
    st, ld = fn(lapack, dx.data_ptr(), incx, n)
    assert st == 0
    want = 2.0 * np.sum(np.log(np.real(x[::incx]).astype(np.float64)))
    assert abs(ld - want) <= 2 * EPS32 * max(1.0, abs(want))
