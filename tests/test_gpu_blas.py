"""GPU BLAS-3 parity: the HIP library through its C ABI against the oracle.

Inputs follow the reference's BLAS tests (entries in [0, 1), test/blas/cudgemm.c);
the pass bound is the elementwise rounding bound of a length-k inner product
in any summation order, |D - D_ref| <= 2 (k + 2) eps (|alpha| |op(A)| |op(B)| +
|beta| |C|), which contains the reference's own (2k+2) 2 eps for O(1) data.
Every test also checks that nothing outside the m x n result was written
(padding rows and the untouched triangle keep their values exactly).
"""
import itertools

import numpy as np
import pytest
import torch

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu


def rng(*seed):
    return np.random.default_rng(list(seed))


def unif(r, shape, cplx=False):
    a = r.random(shape)
    if cplx:
        a = a + 1j * r.random(shape)
    return a


def opm(M, t):
    return M if t == "N" else (M.T if t == "T" else M.conj().T)


def padded(M, pad, fill=np.nan):
    out = np.full((M.shape[0] + pad, M.shape[1]), fill, dtype=M.dtype, order="F")
    out[: M.shape[0]] = M
    return out


def tri_rand(r, k, cplx=False):
    T = r.standard_normal((k, k))
    if cplx:
        T = T + 1j * r.standard_normal((k, k))
    return np.asfortranarray(T / np.sqrt(k) + 3 * np.eye(k))


def tri_op(T, uplo, diag):
    M = np.tril(T) if uplo == "L" else np.triu(T)
    if diag == "U":
        np.fill_diagonal(M, 1.0)
    return M


GEMM_SHAPES = [(1, 1, 1), (16, 16, 4), (37, 29, 41), (64, 128, 16), (130, 70, 200), (300, 257, 129)]


def run_gemm(gpu, blas, ta, tb, m, n, k, cplx, inplace):
    r = rng(m, n, k, ord(ta), ord(tb), int(cplx))
    A = unif(r, (m, k) if ta == "N" else (k, m), cplx)
    B = unif(r, (k, n) if tb == "N" else (n, k), cplx)
    C = unif(r, (m, n), cplx)
    al, be = (0.75 - 0.5j, 1.25 + 0.25j) if cplx else (0.75, -1.25)
    Ap, Bp, Cp = padded(A, 3, 7.0), padded(B, 1, 7.0), padded(C, 2, 5.0)
    dA, dB, dC = gpu.to_device(Ap), gpu.to_device(Bp), gpu.to_device(Cp)
    if inplace:
        dD, ldd = dC, Cp.shape[0]
    else:
        Dp = padded(np.zeros((m, n), C.dtype), 5)
        dD, ldd = gpu.to_device(Dp), Dp.shape[0]
    fn = gpu.cuZgemm2 if cplx else gpu.cuDgemm2
    st = fn(blas, ord(ta), ord(tb), m, n, k, al, gpu.ptr(dA), Ap.shape[0], gpu.ptr(dB), Bp.shape[0],
            be, gpu.ptr(dC), Cp.shape[0], gpu.ptr(dD), ldd)
    torch.cuda.synchronize()
    assert st == 0
    D = gpu.from_device(dD)
    ref = np.asfortranarray(C.copy())
    (O.zgemm if cplx else O.dgemm)(ta, tb, m, n, k, al, np.asfortranarray(A), np.asfortranarray(B), be, ref)
    bound = 2 * (k + 2) * EPS * (abs(al) * np.abs(opm(A, ta)) @ np.abs(opm(B, tb)) + abs(be) * np.abs(C))
    err = np.abs(D[:m] - ref)
    assert np.all(err <= bound), f"max err {err.max():.3e}, bound at argmax {bound.flat[err.argmax()]:.3e}"
    # padding rows untouched
    if inplace:
        np.testing.assert_array_equal(D[m:], Cp[m:])
    else:
        assert np.all(np.isnan(D[m:]))


@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm(gpu, blas, ta, tb, m, n, k):
    run_gemm(gpu, blas, ta, tb, m, n, k, False, inplace=False)


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm_inplace(gpu, blas, ta, tb):
    run_gemm(gpu, blas, ta, tb, 150, 90, 70, False, inplace=True)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (37, 29, 41), (100, 70, 90)])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_zgemm(gpu, blas, ta, tb, m, n, k):
    run_gemm(gpu, blas, ta, tb, m, n, k, True, inplace=(m == 37))


def test_dgemm_alpha_zero_beta_zero_ignores_nan(gpu, blas):
    """beta = 0 must not read C and alpha = 0 must not read A or B."""
    m, n, k = 40, 30, 20
    A = np.full((m, k), np.nan, order="F")
    B = np.full((k, n), np.nan, order="F")
    C = np.full((m, n), np.nan, order="F")
    dA, dB, dC = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C)
    assert gpu.cuDgemm2(blas, gpu.NoTrans, gpu.NoTrans, m, n, k, 0.0, gpu.ptr(dA), m, gpu.ptr(dB), k,
                        0.0, gpu.ptr(dC), m, gpu.ptr(dC), m) == 0
    torch.cuda.synchronize()
    assert np.all(gpu.from_device(dC) == 0.0)


@pytest.mark.parametrize("n,k", [(1, 1), (33, 17), (64, 128), (65, 3), (200, 64), (257, 130)])
@pytest.mark.parametrize("uplo,trans", list(itertools.product("LU", "NT")))
def test_dsyrk(gpu, blas, uplo, trans, n, k):
    r = rng(n, k, ord(uplo), ord(trans))
    A = unif(r, (n, k) if trans == "N" else (k, n))
    C0 = padded(unif(r, (n, n)), 2, 9.0)
    dA, dC = gpu.to_device(np.asfortranarray(A)), gpu.to_device(C0)
    st = gpu.cuDsyrk(blas, ord(uplo), ord(trans), n, k, -0.5, gpu.ptr(dA), A.shape[0], 2.0, gpu.ptr(dC),
                     C0.shape[0])
    torch.cuda.synchronize()
    assert st == 0
    C = gpu.from_device(dC)
    ref = np.asfortranarray(C0[:n].copy())
    O.dsyrk(uplo, trans, n, k, -0.5, np.asfortranarray(A), 2.0, ref)
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    bound = 2 * (k + 2) * EPS * (0.5 * np.abs(opm(A, trans)) @ np.abs(opm(A, trans)).T + 2 * np.abs(C0[:n]))
    assert np.all(np.abs(C[:n] - ref)[mask] <= bound[mask])
    np.testing.assert_array_equal(C[:n][~mask], C0[:n][~mask])
    np.testing.assert_array_equal(C[n:], C0[n:])


@pytest.mark.parametrize("n,k", [(1, 1), (33, 17), (64, 64), (100, 130)])
@pytest.mark.parametrize("uplo,trans", list(itertools.product("LU", "NC")))
def test_zherk(gpu, blas, uplo, trans, n, k):
    r = rng(n, k, ord(uplo), ord(trans), 7)
    A = unif(r, (n, k) if trans == "N" else (k, n), True)
    C0 = unif(r, (n, n), True)
    dA, dC = gpu.to_device(np.asfortranarray(A)), gpu.to_device(np.asfortranarray(C0))
    st = gpu.cuZherk(blas, ord(uplo), ord(trans), n, k, 1.5, gpu.ptr(dA), A.shape[0], 0.5, gpu.ptr(dC), n)
    torch.cuda.synchronize()
    assert st == 0
    C = gpu.from_device(dC)
    ref = np.asfortranarray(C0.copy())
    O.zherk(uplo, trans, n, k, 1.5, np.asfortranarray(A), 0.5, ref)
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    bound = 2 * (k + 2) * EPS * (1.5 * np.abs(opm(A, trans)) @ np.abs(opm(A, trans)).T + 0.5 * np.abs(C0)) * 2
    assert np.all(np.abs(C - ref)[mask] <= bound[mask])
    assert np.all(np.diag(C).imag == 0.0)
    np.testing.assert_array_equal(C[~mask], C0[~mask])


VARIANTS_D = list(itertools.product("LR", "LU", "NT", "NU"))
VARIANTS_Z = list(itertools.product("LR", "LU", "NTC", "NU"))


def run_trmm(gpu, blas, side, uplo, trans, diag, m, n, cplx):
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag), int(cplx))
    k = m if side == "L" else n
    T = unif(r, (k, k), cplx)
    B = unif(r, (m, n), cplx)
    al = (0.5 - 1j) if cplx else -1.5
    # garbage in the unreferenced triangle must not matter
    Tg = np.asfortranarray(T.copy())
    if uplo == "L":
        Tg[np.triu_indices(k, 1)] = np.nan
    else:
        Tg[np.tril_indices(k, -1)] = np.nan
    if diag == "U":
        np.fill_diagonal(Tg, np.nan)
    dT, dB = gpu.to_device(Tg), gpu.to_device(padded(B, 1, 3.0))
    X0 = padded(np.zeros((m, n), B.dtype), 3)
    dX = gpu.to_device(X0)
    fn = gpu.cuZtrmm2 if cplx else gpu.cuDtrmm2
    st = fn(blas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, al, gpu.ptr(dT), k, gpu.ptr(dB),
            m + 1, gpu.ptr(dX), m + 3)
    torch.cuda.synchronize()
    assert st == 0
    X = gpu.from_device(dX)
    M = opm(tri_op(T, uplo, diag), trans)
    if cplx:
        ref = O.fmat(m, n, m, np.complex128)
        O.ztrmm2(side, uplo, trans, diag, m, n, al, np.asfortranarray(T), np.asfortranarray(B), ref)
    else:
        ref = O.fmat(m, n, m)
        O.dtrmm2(side, uplo, trans, diag, m, n, al, np.asfortranarray(T), np.asfortranarray(B), ref)
    absprod = np.abs(M) @ np.abs(B) if side == "L" else np.abs(B) @ np.abs(M)
    bound = 2 * (k + 2) * EPS * abs(al) * absprod * (2 if cplx else 1)
    assert np.all(np.abs(X[:m] - ref) <= bound)
    assert np.all(np.isnan(X[m:]))


@pytest.mark.parametrize("m,n", [(1, 1), (29, 23), (130, 70)])
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS_D)
def test_dtrmm2(gpu, blas, side, uplo, trans, diag, m, n):
    run_trmm(gpu, blas, side, uplo, trans, diag, m, n, False)


@pytest.mark.parametrize("side,uplo,trans", list(itertools.product("LR", "LU", "NT")))
def test_dtrmm2_many_blocks(gpu, blas, side, uplo, trans):
    """1100 rows (left) or columns (right) of the triangular operand: several
    groups of row / column blocks, the last one partial, so that the tile
    order for triangular operands (serpentine across groups, heaviest first
    inside one) is exercised, with a diagonal block cut by the edge."""
    m, n = (1100, 300) if side == "L" else (300, 1100)
    run_trmm(gpu, blas, side, uplo, trans, "N", m, n, False)


@pytest.mark.parametrize("m,n", [(1, 1), (45, 70)])
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS_Z)
def test_ztrmm2(gpu, blas, side, uplo, trans, diag, m, n):
    run_trmm(gpu, blas, side, uplo, trans, diag, m, n, True)


def run_trsm(gpu, blas, side, uplo, trans, diag, m, n, cplx):
    r = rng(m, n, ord(side), ord(uplo), ord(trans), ord(diag), int(cplx), 3)
    k = m if side == "L" else n
    T = tri_rand(r, k, cplx)
    # the unreferenced triangle (and a unit diagonal) hold NaN: never read
    if uplo == "L":
        T[np.triu_indices(k, 1)] = np.nan
    else:
        T[np.tril_indices(k, -1)] = np.nan
    if diag == "U":
        np.fill_diagonal(T, np.nan)
    B = unif(r, (m, n), cplx)
    al = (0.5 + 0.5j) if cplx else 0.75
    dT, dB = gpu.to_device(T), gpu.to_device(padded(B, 2, 4.0))
    fn = gpu.cuZtrsm if cplx else gpu.cuDtrsm
    st = fn(blas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, al, gpu.ptr(dT), k, gpu.ptr(dB), m + 2)
    torch.cuda.synchronize()
    assert st == 0
    X = gpu.from_device(dB)
    ref = np.asfortranarray(B.copy())
    (O.ztrsm if cplx else O.dtrsm)(side, uplo, trans, diag, m, n, al, T, ref)
    err = np.linalg.norm(X[:m] - ref) / np.linalg.norm(ref)
    assert err < 20 * k * EPS, err
    np.testing.assert_array_equal(X[m:], np.full((2, n), 4.0))


@pytest.mark.parametrize("m,n", [(1, 1), (29, 23), (200, 150), (70, 300)])
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS_D)
def test_dtrsm(gpu, blas, side, uplo, trans, diag, m, n):
    run_trsm(gpu, blas, side, uplo, trans, diag, m, n, False)


@pytest.mark.parametrize("m,n", [(1, 1), (130, 90)])
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS_Z)
def test_ztrsm(gpu, blas, side, uplo, trans, diag, m, n):
    run_trsm(gpu, blas, side, uplo, trans, diag, m, n, True)


def test_dtrsm_alpha_zero(gpu, blas):
    m, n = 20, 10
    T = np.full((m, m), np.nan, order="F")
    B = np.ones((m, n), order="F")
    dT, dB = gpu.to_device(T), gpu.to_device(B)
    assert gpu.cuDtrsm(blas, gpu.Left, gpu.Lower, gpu.NoTrans, gpu.NonUnit, m, n, 0.0, gpu.ptr(dT), m,
                       gpu.ptr(dB), m) == 0
    torch.cuda.synchronize()
    assert np.all(gpu.from_device(dB) == 0.0)
