"""Pins the CPU oracle (CPU only).

The oracle restates the reference's CPU BLAS/LAPACK.  Before it may judge the
GPU library it is itself checked here against what fixes the answer
independently of it:
  * library routines: numpy/scipy LAPACK (cholesky, inv, solve_triangular) and
    plain matrix products for every BLAS-3 variant;
  * closed forms: the reference's SPD generator dlatmc is H D H with a
    Householder H, so eig(A) = diag(D), logdet(A) = sum log D_jj and
    inv(A) = H D^-1 H;
  * the reference's unblocked textbook routines (test/lapack/ref/*) restated
    separately, on tiny n exhaustively, including the info of non-SPD inputs;
  * the reference's logdet known-answer test (|Kahan sum - result| < 2 n eps,
    test/lapack/cudlogdet.c:135-146).
"""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import EPS
import oracle as O

SIZES = [1, 2, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 130]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def rng(seed):
    return np.random.default_rng(seed)


def spd_np(n, seed, alpha=1.0, cplx=False):
    r = rng(seed)
    G = r.standard_normal((n, n))
    if cplx:
        G = (G + 1j * r.standard_normal((n, n))) / np.sqrt(2)
    return G @ G.conj().T / n + alpha * np.eye(n)


# ---------------------------------------------------------------- generators

def test_gauss_matches_numpy_restatement():
    a = O.gauss(12345, 4096, first=17)
    b = O.gauss_numpy(12345, 4096, first=17)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_gauss_moments():
    g = O.gauss(1, 200000)
    assert abs(g.mean()) < 0.01
    assert abs(g.std() - 1.0) < 0.01


def test_spd_generator_is_ggt():
    n, alpha, seed = 40, 0.1, 3
    A = O.spd(seed, n, alpha)
    G = O.gauss(seed, n * n).reshape(n, n, order="F")
    np.testing.assert_allclose(A, G @ G.T / n + alpha * np.eye(n), rtol=1e-13, atol=1e-13)
    H = O.hpd(seed, n, alpha)
    g = O.gauss(seed, 2 * n * n)
    Hm = (g[: n * n].reshape(n, n, order="F") + 1j * g[n * n:].reshape(n, n, order="F")) / np.sqrt(2)
    np.testing.assert_allclose(H, Hm @ Hm.conj().T / n + alpha * np.eye(n), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [2, 3, 10, 64, 101])
def test_dlatmc_closed_form(n):
    O.srand(0)
    A = O.dlatmc(n, 2.0)
    O.srand(0)
    # replay the diagonal draws exactly as dlatmc makes them
    import ctypes
    libc = ctypes.CDLL(None)
    d = [1.0, 2.0] + [libc.rand() / 2147483647.0 * (2.0 - 1.0) + 1.0 for _ in range(n - 2)]
    np.testing.assert_allclose(A, A.T, atol=1e-15)
    np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(A)), np.sort(d), atol=1e-13)
    sign, ld = np.linalg.slogdet(A)
    assert sign > 0
    assert abs(ld - np.sum(np.log(d))) < 1e-12


def test_zlatmc_hermitian_spectrum():
    n = 50
    O.srand(0)
    A = O.zlatmc(n, 2.0)
    np.testing.assert_allclose(A, A.conj().T, atol=1e-15)
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > 1.0 - 1e-12 and ev.max() < 2.0 + 1e-12
    assert abs(ev.min() - 1.0) < 1e-12 and abs(ev.max() - 2.0) < 1e-12


# ---------------------------------------------------------------- BLAS-3

def opm(M, t):
    return M if t == "N" else (M.T if t == "T" else M.conj().T)


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm(ta, tb):
    r = rng(1)
    m, n, k = 37, 29, 41
    A = np.asfortranarray(r.standard_normal((m, k) if ta == "N" else (k, m)))
    B = np.asfortranarray(r.standard_normal((k, n) if tb == "N" else (n, k)))
    C = O.with_ld(r.standard_normal((m, n)), m + 3)
    want = 0.7 * opm(A, ta) @ opm(B, tb) - 1.3 * C[:m]
    O.dgemm(ta, tb, m, n, k, 0.7, A, B, -1.3, C)
    np.testing.assert_allclose(C[:m], want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_zgemm(ta, tb):
    r = rng(2)
    m, n, k = 23, 19, 31
    mk = (m, k) if ta == "N" else (k, m)
    kn = (k, n) if tb == "N" else (n, k)
    A = np.asfortranarray(r.standard_normal(mk) + 1j * r.standard_normal(mk))
    B = np.asfortranarray(r.standard_normal(kn) + 1j * r.standard_normal(kn))
    C = np.asfortranarray(r.standard_normal((m, n)) + 1j * r.standard_normal((m, n)))
    al, be = 0.5 - 0.25j, 1.5 + 0.5j
    want = al * opm(A, ta) @ opm(B, tb) + be * C
    O.zgemm(ta, tb, m, n, k, al, A, B, be, C)
    np.testing.assert_allclose(C, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("uplo,trans", list(itertools.product("LU", "NT")))
def test_dsyrk_touches_one_triangle(uplo, trans):
    r = rng(3)
    n, k = 33, 21
    A = np.asfortranarray(r.standard_normal((n, k) if trans == "N" else (k, n)))
    C0 = np.asfortranarray(r.standard_normal((n, n)))
    C = C0.copy(order="F")
    O.dsyrk(uplo, trans, n, k, -1.0, A, 0.5, C)
    full = -opm(A, trans) @ opm(A, trans).T + 0.5 * C0
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    np.testing.assert_allclose(C[mask], full[mask], rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("uplo,trans", list(itertools.product("LU", "NC")))
def test_zherk(uplo, trans):
    r = rng(4)
    n, k = 27, 15
    sh = (n, k) if trans == "N" else (k, n)
    A = np.asfortranarray(r.standard_normal(sh) + 1j * r.standard_normal(sh))
    C0 = spd_np(n, 5, cplx=True)
    C = np.asfortranarray(C0.copy())
    O.zherk(uplo, trans, n, k, 2.0, A, 0.5, C)
    full = 2.0 * opm(A, trans) @ opm(A, trans).conj().T + 0.5 * C0
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    np.testing.assert_allclose(C[mask], full[mask], rtol=1e-13, atol=1e-13)
    assert np.all(np.diag(C).imag == 0)


def tri_rand(r, k, cplx=False):
    """Well-conditioned triangular test operand (also with a unit diagonal):
    small off-diagonal entries and a dominant diagonal."""
    T = r.standard_normal((k, k))
    if cplx:
        T = T + 1j * r.standard_normal((k, k))
    T = T / np.sqrt(k) + 3 * np.eye(k)
    return np.asfortranarray(T)


def tri_op(T, uplo, diag):
    M = np.tril(T) if uplo == "L" else np.triu(T)
    if diag == "U":
        np.fill_diagonal(M, 1.0)
    return M


VARIANTS = list(itertools.product("LR", "LU", "NT", "NU"))


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_dtrsm(side, uplo, trans, diag):
    r = rng(6)
    m, n = 29, 23
    k = m if side == "L" else n
    T = tri_rand(r, k)
    B0 = r.standard_normal((m, n))
    B = O.with_ld(B0, m + 2)
    O.dtrsm(side, uplo, trans, diag, m, n, 0.75, T, B)
    M = opm(tri_op(T, uplo, diag), trans)
    want = np.linalg.solve(M, 0.75 * B0) if side == "L" else np.linalg.solve(M.T, 0.75 * B0.T).T
    np.testing.assert_allclose(B[:m], want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_dtrmm_and_dtrmm2(side, uplo, trans, diag):
    r = rng(7)
    m, n = 26, 31
    k = m if side == "L" else n
    T = np.asfortranarray(r.standard_normal((k, k)))
    B0 = np.asfortranarray(r.standard_normal((m, n)))
    M = opm(tri_op(T, uplo, diag), trans)
    want = -1.5 * (M @ B0 if side == "L" else B0 @ M)
    B = B0.copy(order="F")
    O.dtrmm(side, uplo, trans, diag, m, n, -1.5, T, B)
    np.testing.assert_allclose(B, want, rtol=1e-13, atol=1e-13)
    X = O.fmat(m, n, m + 1)
    O.dtrmm2(side, uplo, trans, diag, m, n, -1.5, T, B0, X)
    np.testing.assert_allclose(X[:m], want, rtol=1e-13, atol=1e-13)


ZVARIANTS = list(itertools.product("LR", "LU", "NTC", "NU"))


@pytest.mark.parametrize("side,uplo,trans,diag", ZVARIANTS)
def test_ztrsm_ztrmm2(side, uplo, trans, diag):
    r = rng(8)
    m, n = 17, 13
    k = m if side == "L" else n
    T = tri_rand(r, k, True)
    B0 = np.asfortranarray(r.standard_normal((m, n)) + 1j * r.standard_normal((m, n)))
    M = opm(tri_op(T, uplo, diag), trans)
    al = 0.5 + 1j
    B = B0.copy(order="F")
    O.ztrsm(side, uplo, trans, diag, m, n, al, T, B)
    want = np.linalg.solve(M, al * B0) if side == "L" else np.linalg.solve(M.T, al * B0.T).T
    np.testing.assert_allclose(B, want, rtol=1e-12, atol=1e-12)
    X = O.fmat(m, n, m, np.complex128)
    O.ztrmm2(side, uplo, trans, diag, m, n, al, T, B0, X)
    want = al * (M @ B0 if side == "L" else B0 @ M)
    np.testing.assert_allclose(X, want, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- LAPACK

@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_vs_lapack_and_textbook(n, uplo):
    A = spd_np(n, n)
    L = np.linalg.cholesky(A)
    F, info = O.dpotrf(uplo, O.with_ld(A, n + 1), n)
    assert info == 0
    got = np.tril(F[:n]) if uplo == "L" else np.triu(F[:n]).T
    assert rel(got, L) < 1e-13
    U, info_u = O.dpotrf_unblocked(uplo, O.with_ld(A, n), n)
    assert info_u == 0
    got_u = np.tril(U) if uplo == "L" else np.triu(U).T
    assert rel(got_u, L) < 1e-13
    # the untouched triangle stays as it was
    other = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
    np.testing.assert_array_equal(F[:n][other], A[other])


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("bad", [0, 5, 31, 32, 33, 47, 99])
def test_dpotrf_info_non_spd(uplo, bad):
    """A leading minor of order bad+1 that is not positive: every form reports
    info = bad + 1 (the reference's contract, test/lapack/cudpotrf.c:100)."""
    n = 100
    A = spd_np(n, 11)
    L = np.linalg.cholesky(A)
    # make pivot `bad` exactly -1: A_bb = sum_{k<bad} L_bk^2 - 1
    A[bad, bad] = np.dot(L[bad, :bad], L[bad, :bad]) - 1.0
    for fn in (O.dpotrf, O.dpotrf_unblocked):
        _, info = fn(uplo, O.with_ld(A, n), n)
        assert info == bad + 1, (fn.__name__, info)


@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_info_nan(uplo):
    n = 70
    A = spd_np(n, 12)
    A[40, 40] = np.nan
    for fn in (O.dpotrf, O.dpotrf_unblocked):
        _, info = fn(uplo, O.with_ld(A, n), n)
        assert info == 41


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 200])
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_dtrtri(n, uplo, diag):
    T = tri_rand(rng(n), n)
    M = tri_op(T, uplo, diag)
    want = np.linalg.inv(M)
    X, info = O.dtrtri(uplo, diag, O.with_ld(T, n + 3), n)
    assert info == 0
    got = tri_op(X[:n], uplo, diag) if diag == "U" else (np.tril(X[:n]) if uplo == "L" else np.triu(X[:n]))
    assert rel(got, want) < 1e-12
    Y, info = O.dtrtri_unblocked(uplo, diag, O.with_ld(T, n), n)
    got_u = tri_op(Y, uplo, diag) if diag == "U" else (np.tril(Y) if uplo == "L" else np.triu(Y))
    assert rel(got_u, want) < 1e-12
    B, info = O.dtrtri2(uplo, diag, O.with_ld(T, n), O.fmat(n, n, n + 2), n)
    got_b = tri_op(B[:n], uplo, diag) if diag == "U" else (np.tril(B[:n]) if uplo == "L" else np.triu(B[:n]))
    assert rel(got_b, want) < 1e-12


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n,zero", [(10, 3), (100, 70), (130, 5)])
def test_dtrtri_info_zero_diagonal(uplo, n, zero):
    T = rng(3).standard_normal((n, n)) + 3 * np.eye(n)
    T[zero, zero] = 0.0
    _, info = O.dtrtri(uplo, "N", O.with_ld(T, n), n)
    assert info == zero + 1
    _, info = O.dtrtri_unblocked(uplo, "N", O.with_ld(T, n), n)
    assert info == zero + 1


@pytest.mark.parametrize("n", [1, 7, 31, 32, 33, 70, 129])
@pytest.mark.parametrize("uplo", "LU")
def test_dlauum(n, uplo):
    T = rng(n + 1).standard_normal((n, n))
    M = np.tril(T) if uplo == "L" else np.triu(T)
    want = M.T @ M if uplo == "L" else M @ M.T
    X, info = O.dlauum(uplo, O.with_ld(T, n + 1), n)
    tri = np.tril if uplo == "L" else np.triu
    assert rel(tri(X[:n]), tri(want)) < 1e-13
    Y = O.dlauum_unblocked(uplo, O.with_ld(T, n), n)
    assert rel(tri(Y), tri(want)) < 1e-13


@pytest.mark.parametrize("n", [2, 33, 100, 257])
@pytest.mark.parametrize("uplo", "LU")
def test_dpotri_closed_form(n, uplo):
    """dlatmc = H D H, so inv(A) = H D^-1 H exactly (up to rounding)."""
    O.srand(0)
    A = O.dlatmc(n, 2.0)
    want = np.linalg.inv(A)   # cond(A) = 2
    F, info = O.dpotrf(uplo, A.copy(order="F"))
    X, info2 = O.dpotri(uplo, F)
    assert info == 0 and info2 == 0
    assert rel(O.herm_from(X, uplo), want) < 1e-13
    assert O.inverse_residual(A, X, uplo) < 10 * n * EPS


def test_dlogdet_and_kahan_oracle():
    """The reference's logdet test (test/lapack/cudlogdet.c:54-73) draws
    x_j = (rand()+1)/RAND_MAX and accepts a result within 2 n eps of a
    Kahan-summed value.  Here the Kahan oracle is pinned against the exactly
    rounded sum (math.fsum), and the plain sequential sum the reference CPU
    path uses against the recursive-summation error bound."""
    import math
    for n, incx in [(1, 1), (1000, 1), (4097, 3)]:
        O.srand(0)
        x = O.logdet_input(n, incx)
        logs = np.log(x[::incx][:n])
        kahan = O.kahan_logdet(x, incx, n)
        exact = 2.0 * math.fsum(logs)
        assert abs(kahan - exact) <= 4 * EPS * abs(exact) + 1e-300
        bound = 2.0 * n * EPS * np.sum(np.abs(logs))
        assert abs(O.dlogdet(x, incx, n) - exact) <= bound


@pytest.mark.parametrize("n", [2, 64, 300])
def test_dlogdet_of_factor_closed_form(n):
    O.srand(0)
    A = O.dlatmc(n, 2.0)
    O.srand(0)
    import ctypes
    libc = ctypes.CDLL(None)
    d = [1.0, 2.0] + [libc.rand() / 2147483647.0 + 1.0 for _ in range(n - 2)]
    F, info = O.dpotrf("L", A.copy(order="F"))
    diag = np.ascontiguousarray(F.reshape(-1, order="F"))
    got = O.dlogdet(diag, n + 1, n)
    assert abs(got - np.sum(np.log(d))) < 1e-12 * n


# ---------------------------------------------------------------- complex LAPACK

@pytest.mark.parametrize("n", [1, 2, 16, 17, 33, 64, 65, 100])
@pytest.mark.parametrize("uplo", "LU")
def test_zpotrf(n, uplo):
    A = spd_np(n, n, cplx=True)
    L = np.linalg.cholesky(A)
    F, info = O.zpotrf(uplo, O.with_ld(A, n + 2), n)
    assert info == 0
    got = np.tril(F[:n]) if uplo == "L" else np.triu(F[:n]).conj().T
    assert rel(got, L) < 1e-13
    U, info = O.zpotrf_unblocked(uplo, O.with_ld(A, n), n)
    got = np.tril(U) if uplo == "L" else np.triu(U).conj().T
    assert rel(got, L) < 1e-13


@pytest.mark.parametrize("uplo", "LU")
def test_zpotrf_info(uplo):
    n = 80
    A = spd_np(n, 4, cplx=True)
    L = np.linalg.cholesky(A)
    A[37, 37] = np.vdot(L[37, :37], L[37, :37]).real - 0.5
    for fn in (O.zpotrf, O.zpotrf_unblocked):
        _, info = fn(uplo, O.with_ld(A, n), n)
        assert info == 38


@pytest.mark.parametrize("n", [3, 64, 100])
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_ztrtri(n, uplo, diag):
    T = tri_rand(rng(n), n, True)
    M = tri_op(T, uplo, diag)
    X, info = O.ztrtri(uplo, diag, O.with_ld(T, n), n)
    got = tri_op(X, uplo, diag) if diag == "U" else (np.tril(X) if uplo == "L" else np.triu(X))
    assert info == 0 and rel(got, np.linalg.inv(M)) < 1e-12


@pytest.mark.parametrize("n", [5, 40, 90])
@pytest.mark.parametrize("uplo", "LU")
def test_zlauum_zpotri(n, uplo):
    A = spd_np(n, n + 7, cplx=True)
    F, info = O.zpotrf(uplo, A.copy(order="F"))
    tri = np.tril if uplo == "L" else np.triu
    T = tri(F)
    want = T.conj().T @ T if uplo == "L" else T @ T.conj().T
    X, _ = O.zlauum(uplo, F.copy(order="F"))
    assert rel(tri(X), tri(want)) < 1e-13
    Y, info = O.zpotri(uplo, F.copy(order="F"))
    assert info == 0
    assert rel(O.herm_from(Y, uplo), np.linalg.inv(A)) < 1e-12


def test_zlogdet_forms():
    n = 500
    A = spd_np(n, 9, cplx=True)
    F, _ = O.zpotrf("L", A.copy(order="F"))
    flat = np.ascontiguousarray(F.reshape(-1, order="F"))
    sign, want = np.linalg.slogdet(A)
    assert abs(O.zlogdet(flat, n + 1, n) - want) < 1e-11
    # the reference's single-precision logs are only float-accurate
    assert abs(O.zlogdet_ref(flat, n + 1, n) - want) < 1e-4 * n
