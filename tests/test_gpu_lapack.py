"""GPU LAPACK parity: POTRF / TRTRI / LAUUM / POTRI / LOGDET through the C ABI.

The bar is the north star's: backward error ||A - L L^H||_F / ||A||_F <=
10 n eps and ||I - A inv(A)|| / cond(A) <= 10 n eps, on the same seeded
inputs the oracle (the reference's CPU path, restated) factors; the factor
and inverse also agree with the oracle's to a tolerance set by the
conditioning, `info` matches exactly, and the unreferenced triangle and
padding rows come back bit-identical.  Full-size cases (n = 4096, 2048) are
checked through the size-independent backward errors, computed on the GPU in
FP64.
"""
import numpy as np
import pytest
import torch

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 31, 32, 33, 64, 65, 127, 128, 129, 255, 256, 257, 300, 513, 1024]


def other_mask(n, uplo):
    return np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)


def tri(M, uplo):
    return np.tril(M) if uplo == "L" else np.triu(M)


def rel(a, b):
    d = np.linalg.norm(a - b)
    return 0.0 if d == 0.0 else d / np.linalg.norm(b)


def gpu_call(gpu, fn, lapack, *args):
    st, info = fn(lapack, *args)
    torch.cuda.synchronize()
    return st, info


# ---------------------------------------------------------------- POTRF

@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf(gpu, lapack, uplo, n):
    A = O.spd(n + 1, n, 1.0)            # config-1 form G G^T / n + I
    lda = n + 3
    Ap = np.full((lda, n), -7.0, order="F")
    Ap[:n] = A
    dA = gpu.to_device(Ap)
    st, info = gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), lda)
    assert st == 0 and info == 0
    F = gpu.from_device(dA)
    ref, rinfo = O.dpotrf(uplo, np.asfortranarray(A.copy()))
    assert rinfo == 0
    assert O.potrf_backward_error(A, F[:n], uplo) <= 10 * n * EPS
    assert rel(tri(F[:n], uplo), tri(ref, uplo)) <= 10 * n * EPS * 5   # cond(A) ~ 5
    np.testing.assert_array_equal(F[:n][other_mask(n, uplo)], A[other_mask(n, uplo)])
    assert np.all(F[n:] == -7.0)


@pytest.mark.parametrize("cap", ["100000:8", "100000:64", "1200:448"])
@pytest.mark.parametrize("prec,uplo", [("D", "L"), ("D", "U"), ("Z", "L"), ("Z", "U")])
def test_potrf_trailing_cap(gpu, lapack, monkeypatch, prec, uplo, cap):
    """Trailing updates capped at a few workgroups (each then takes several
    tiles; CUCHOLESKY_TRAILING_CAP_<prec> = "t_max:workgroups", read on every
    call): the same tiles computed by other workgroups, so the factor is
    bit-identical to the uncapped one, and within the backward-error bound."""
    if prec == "D":
        n = 2500
        A, fn = O.spd(n + 1, n, 1.0), gpu.cuDpotrf
    else:
        n = 1500
        A, fn = O.hpd(n, n, 0.1), gpu.cuZpotrf
    monkeypatch.setenv(f"CUCHOLESKY_TRAILING_CAP_{prec}", "0:0")
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, fn, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    F0 = gpu.from_device(dA)
    monkeypatch.setenv(f"CUCHOLESKY_TRAILING_CAP_{prec}", cap)
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, fn, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    F = gpu.from_device(dA)
    np.testing.assert_array_equal(F, F0)
    assert O.potrf_backward_error(A, F, uplo) <= 10 * n * EPS


@pytest.mark.parametrize("n", [2, 100, 1000])
@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_dlatmc(gpu, lapack, uplo, n):
    """The reference's own test matrix (test/lapack/cudpotrf.c:49-92)."""
    O.srand(0)
    A = O.dlatmc(n, 2.0)
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    F = gpu.from_device(dA)
    assert O.potrf_backward_error(A, F, uplo) <= 10 * n * EPS
    ref, _ = O.dpotrf(uplo, A.copy(order="F"))
    assert np.max(np.abs(tri(F, uplo) - tri(ref, uplo))) <= 100 * n * EPS


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("bad", [0, 5, 127, 128, 129, 200, 255, 256, 300, 399])
def test_dpotrf_info_non_spd(gpu, lapack, uplo, bad):
    """Leading minor of order bad+1 not positive: info = bad + 1, the
    reference's pass criterion (test/lapack/cudpotrf.c:100)."""
    n = 400
    A = O.spd(3, n, 1.0)
    L = np.linalg.cholesky(A)
    A[bad, bad] = np.dot(L[bad, :bad], L[bad, :bad]) - 1.0
    _, rinfo = O.dpotrf(uplo, A.copy(order="F"))
    assert rinfo == bad + 1
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == rinfo
    # columns before the failing block are fully factored
    F = gpu.from_device(dA)
    done = (bad // 128) * 128
    if done:
        got = tri(F, uplo)[:done, :done] if uplo == "L" else tri(F, uplo)[:done, :done].T
        np.testing.assert_allclose(np.tril(got), L[:done, :done], rtol=0, atol=1e-12)


@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_nan_and_reuse(gpu, lapack, uplo):
    n = 300
    A = O.spd(5, n, 1.0)
    A[150, 150] = np.nan
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 151
    # the handle is reusable after a failure
    B = O.spd(6, n, 1.0)
    dB = gpu.to_device(B)
    st, info = gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dB), n)
    assert st == 0 and info == 0
    assert O.potrf_backward_error(B, gpu.from_device(dB), uplo) <= 10 * n * EPS


# ---------------------------------------------------------------- TRTRI / LAUUM / POTRI

def tri_rand(seed, n, cplx=False):
    r = np.random.default_rng(seed)
    T = r.standard_normal((n, n))
    if cplx:
        T = T + 1j * r.standard_normal((n, n))
    return np.asfortranarray(T / np.sqrt(n) + 3 * np.eye(n))


@pytest.mark.parametrize("n", [1, 5, 64, 128, 129, 300, 700])
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_dtrtri(gpu, lapack, uplo, diag, n):
    T = tri_rand(n, n)
    dT = gpu.to_device(T)
    st, info = gpu_call(gpu, gpu.cuDtrtri, lapack, ord(uplo), ord(diag), n, gpu.ptr(dT), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dT)
    ref, rinfo = O.dtrtri(uplo, diag, T.copy(order="F"))
    assert rinfo == 0
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    if diag == "U":
        np.fill_diagonal(mask, False)
        assert np.all(np.diag(X) == np.diag(T))     # a unit diagonal is never written
    assert rel(X[mask], ref[mask]) <= 20 * n * EPS
    np.testing.assert_array_equal(X[~tri(np.ones((n, n)), uplo).astype(bool)],
                                  T[~tri(np.ones((n, n)), uplo).astype(bool)])


@pytest.mark.parametrize("uplo", "LU")
def test_dtrtri_info(gpu, lapack, uplo):
    n = 300
    T = tri_rand(1, n)
    T[40, 40] = 0.0
    T[250, 250] = 0.0
    _, rinfo = O.dtrtri(uplo, "N", T.copy(order="F"))
    dT = gpu.to_device(T)
    st, info = gpu_call(gpu, gpu.cuDtrtri, lapack, ord(uplo), ord("N"), n, gpu.ptr(dT), n)
    assert st == 0 and info == rinfo == (251 if uplo == "L" else 41)


@pytest.mark.parametrize("n", [1, 7, 64, 128, 129, 300, 700])
@pytest.mark.parametrize("uplo", "LU")
def test_dlauum(gpu, lapack, uplo, n):
    T = tri_rand(n + 2, n)
    dT = gpu.to_device(T)
    st, info = gpu_call(gpu, gpu.cuDlauum, lapack, ord(uplo), n, gpu.ptr(dT), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dT)
    ref, _ = O.dlauum(uplo, T.copy(order="F"))
    assert rel(tri(X, uplo), tri(ref, uplo)) <= 10 * n * EPS
    np.testing.assert_array_equal(X[other_mask(n, uplo)], T[other_mask(n, uplo)])


@pytest.mark.parametrize("n", [1, 2, 33, 128, 129, 300, 1024])
@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_dpotri(gpu, lapack, uplo, n):
    """Config-1 pipeline: dpotrf -> dlogdet(diag, incx = lda + 1) -> dpotri."""
    A = O.spd(n, n, 1.0)
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    st, ld = gpu.cuDlogdet(lapack, gpu.ptr(dA), n + 1, n)
    assert st == 0
    sign, want = np.linalg.slogdet(A)
    assert abs(ld - want) <= 10 * n * EPS * max(1.0, abs(want))
    st, info = gpu_call(gpu, gpu.cuDpotri, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dA)
    F, _ = O.dpotrf(uplo, A.copy(order="F"))
    ref, _ = O.dpotri(uplo, F)
    condA = np.linalg.cond(A)
    assert O.inverse_residual(A, X, uplo) <= 10 * n * EPS
    assert rel(O.herm_from(X, uplo), O.herm_from(ref, uplo)) <= 10 * n * EPS * condA


@pytest.mark.parametrize("uplo", "LU")
def test_dpotri_closed_form(gpu, lapack, uplo):
    n = 500
    O.srand(0)
    A = O.dlatmc(n, 2.0)
    want = np.linalg.inv(A)
    dA = gpu.to_device(A)
    assert gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n) == (0, 0)
    assert gpu_call(gpu, gpu.cuDpotri, lapack, ord(uplo), n, gpu.ptr(dA), n) == (0, 0)
    assert rel(O.herm_from(gpu.from_device(dA), uplo), want) <= 10 * n * EPS * 2


# ---------------------------------------------------------------- LOGDET

@pytest.mark.parametrize("n,incx", [(1, 1), (1000, 1), (4097, 3), (1 << 20, 1)])
def test_dlogdet_reference_known_answer(gpu, lapack, n, incx):
    """test/lapack/cudlogdet.c: |Kahan sum - result| < 2 n eps."""
    O.srand(0)
    x = O.logdet_input(n, incx)
    dx = torch.from_numpy(x).cuda()
    st, got = gpu.cuDlogdet(lapack, gpu.ptr(dx), incx, n)
    assert st == 0
    assert abs(got - O.kahan_logdet(x, incx, n)) < 2 * n * EPS


def test_zlogdet_double_precision(gpu, lapack):
    n = 3000
    r = np.random.default_rng(0)
    x = (r.random(n) + 0.5) + 1j * r.random(n)
    dx = torch.from_numpy(x).cuda()
    st, got = gpu.cuZlogdet(lapack, gpu.ptr(dx), 1, n)
    assert st == 0
    assert abs(got - O.zlogdet(np.ascontiguousarray(x), 1, n)) < 2 * n * EPS


# ---------------------------------------------------------------- complex

@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 128, 129, 300])
@pytest.mark.parametrize("uplo", "LU")
def test_zpotrf_zpotri(gpu, lapack, uplo, n):
    A = O.hpd(n, n, 0.1)                 # config-2 form H H^H / n + 0.1 I
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, gpu.cuZpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    F = gpu.from_device(dA)
    condA = np.linalg.cond(A)
    assert O.potrf_backward_error(A, F, uplo) <= 10 * n * EPS
    ref, _ = O.zpotrf(uplo, A.copy(order="F"))
    assert rel(tri(F, uplo), tri(ref, uplo)) <= 10 * n * EPS * condA
    np.testing.assert_array_equal(F[other_mask(n, uplo)], A[other_mask(n, uplo)])
    st, ld = gpu.cuZlogdet(lapack, gpu.ptr(dA), n + 1, n)
    assert st == 0 and abs(ld - np.linalg.slogdet(A)[1]) <= 10 * n * EPS * max(1, abs(ld))
    st, info = gpu_call(gpu, gpu.cuZpotri, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dA)
    assert O.inverse_residual(A, X, uplo) <= 10 * n * EPS
    refi, _ = O.zpotri(uplo, ref)
    assert rel(O.herm_from(X, uplo), O.herm_from(refi, uplo)) <= 10 * n * EPS * condA


@pytest.mark.parametrize("uplo", "LU")
def test_zpotrf_info(gpu, lapack, uplo):
    n = 200
    A = O.hpd(2, n, 0.5)
    L = np.linalg.cholesky(A)
    A[77, 77] = np.vdot(L[77, :77], L[77, :77]).real - 0.25
    _, rinfo = O.zpotrf(uplo, A.copy(order="F"))
    dA = gpu.to_device(A)
    st, info = gpu_call(gpu, gpu.cuZpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n)
    assert st == 0 and info == rinfo == 78


@pytest.mark.parametrize("n", [3, 64, 65, 200])
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_ztrtri(gpu, lapack, uplo, diag, n):
    T = tri_rand(n, n, True)
    dT = gpu.to_device(T)
    st, info = gpu_call(gpu, gpu.cuZtrtri, lapack, ord(uplo), ord(diag), n, gpu.ptr(dT), n)
    assert st == 0 and info == 0
    X = gpu.from_device(dT)
    ref, _ = O.ztrtri(uplo, diag, T.copy(order="F"))
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    if diag == "U":
        np.fill_diagonal(mask, False)
    assert rel(X[mask], ref[mask]) <= 20 * n * EPS


@pytest.mark.parametrize("n", [5, 64, 65, 200])
@pytest.mark.parametrize("uplo", "LU")
def test_zlauum(gpu, lapack, uplo, n):
    T = tri_rand(n + 9, n, True)
    dT = gpu.to_device(T)
    st, info = gpu_call(gpu, gpu.cuZlauum, lapack, ord(uplo), n, gpu.ptr(dT), n)
    assert st == 0 and info == 0
    ref, _ = O.zlauum(uplo, T.copy(order="F"))
    X = gpu.from_device(dT)
    assert rel(tri(X, uplo), tri(ref, uplo)) <= 10 * n * EPS
    assert np.all(np.diag(X).imag == 0.0)


# ---------------------------------------------------------------- trti22

@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("diag", "NU")
def test_dtrti22(gpu, lapack, uplo, diag):
    n = 32
    T = tri_rand(77, n)
    dT = gpu.to_device(T)
    dB = torch.zeros_like(dT)
    dinfo = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    assert gpu.cuDtrti22(lapack, ord(uplo), ord(diag), n, gpu.ptr(dT), n, gpu.ptr(dB), n,
                         gpu.ptr(dinfo)) == 0
    torch.cuda.synchronize()
    assert int(dinfo.item()) == 0
    ref, _ = O.dtrtri(uplo, diag, T.copy(order="F"))
    X = gpu.from_device(dB)
    mask = tri(np.ones((n, n)), uplo).astype(bool)
    if diag == "U":
        np.fill_diagonal(mask, False)
    assert rel(X[mask], ref[mask]) <= 20 * n * EPS
    # zero on the diagonal -> info, B untouched
    T[9, 9] = 0.0
    dT = gpu.to_device(T)
    dB = torch.zeros_like(dT)
    assert gpu.cuDtrti22(lapack, ord(uplo), ord("N"), n, gpu.ptr(dT), n, gpu.ptr(dB), n,
                         gpu.ptr(dinfo)) == 0
    torch.cuda.synchronize()
    assert int(dinfo.item()) == 10
    assert torch.count_nonzero(dB).item() == 0


# ---------------------------------------------------------------- full size

def gpu_spd(n, alpha, seed, cplx=False):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if cplx:
        H = torch.complex(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)) / 2 ** 0.5
        A = H @ H.conj().T / n
    else:
        G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
        A = G @ G.T / n
    A += alpha * torch.eye(n, device="cuda", dtype=A.dtype)
    return A


@pytest.mark.parametrize("uplo", "LU")
def test_dpotrf_dpotri_4096(gpu, lapack, uplo):
    n = 4096
    A = gpu_spd(n, 0.1, 11)                  # symmetric: row- and column-major agree
    dA = A.clone()
    assert gpu_call(gpu, gpu.cuDpotrf, lapack, ord(uplo), n, gpu.ptr(dA), n) == (0, 0)
    Ft = dA.T                                 # column-major view as a matrix
    T = torch.tril(Ft) if uplo == "L" else torch.triu(Ft)
    R = T @ T.T if uplo == "L" else T.T @ T
    assert (torch.linalg.norm(A - R) / torch.linalg.norm(A)).item() <= 10 * n * EPS
    assert gpu_call(gpu, gpu.cuDpotri, lapack, ord(uplo), n, gpu.ptr(dA), n) == (0, 0)
    Xt = dA.T
    Tx = torch.tril(Xt) if uplo == "L" else torch.triu(Xt)
    X = Tx + Tx.T - torch.diag(torch.diag(Tx))
    res = torch.linalg.norm(torch.eye(n, device="cuda", dtype=torch.float64) - A @ X)
    assert (res / (torch.linalg.norm(A) * torch.linalg.norm(X))).item() <= 10 * n * EPS


def test_zpotrf_upper_2048(gpu, lapack):
    n = 2048
    A = gpu_spd(n, 0.1, 12, cplx=True)
    dA = A.T.contiguous()                    # column-major storage of A
    assert gpu_call(gpu, gpu.cuZpotrf, lapack, ord("U"), n, gpu.ptr(dA), n) == (0, 0)
    U = torch.triu(dA.T)
    assert (torch.linalg.norm(A - U.conj().T @ U) / torch.linalg.norm(A)).item() <= 10 * n * EPS
