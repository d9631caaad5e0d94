"""GPU parity of the single-precision surface and the solvers, through the C ABI.

* cuSgemm2 / cuSsyrk / cuStrsm: FP32 data and accumulation.  Checked against
  the oracle's FP64 BLAS applied to the same FP32 inputs, with the elementwise
  rounding bound of a length-k FP32 inner product, 2 (k + 2) eps32
  (|alpha| |op(A)| |op(B)| + |beta| |C|).
* cuSpotrf: FP32 Cholesky (diagonal blocks in FP64): backward error
  ||A - L L^T||_F / ||A||_F <= 10 n eps32.
* cuDpotrs: FP64 solve with the cuDpotrf factor against the oracle's two
  dtrsm calls on the oracle's factor; normwise backward error
  ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= 10 n eps.
* cuDsposv / cuDsposvx: the mixed-precision solve must meet LAPACK DSPOSV's
  own stopping test in every column, ||r||_inf <= ||x||_inf ||A||_inf eps
  sqrt(n) (eps = 2^-53), recomputed here in FP64 on the host; the ITER and
  fall-back codes follow DSPOSV.
"""
import numpy as np
import pytest
import torch

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
UNIT = 2.0 ** -53


def rng(*seed):
    return np.random.default_rng(list(seed))


def opm(M, t):
    return M if t == "N" else M.T


def f32(M):
    return np.asfortranarray(M, dtype=np.float32)


def to_dev32(M):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(M, np.float32).T)).cuda()


def from_dev(t):
    return t.detach().cpu().numpy().T


# ------------------------------------------------------------------ FP32 BLAS

# (640, 384, 256): whole 128 x 128 tiles and whole K steps, the unpredicated loads
SGEMM_SHAPES = [(1, 1, 1), (16, 16, 4), (37, 29, 41), (64, 128, 16), (130, 70, 200),
                (300, 257, 129), (640, 384, 256), (1000, 900, 700)]


@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
@pytest.mark.parametrize("shape", SGEMM_SHAPES)
def test_sgemm(gpu, blas, ta, tb, shape):
    m, n, k = shape
    r = rng(1, m, n, k, ord(ta), ord(tb))
    A = f32(r.random((m, k) if ta == "N" else (k, m)))
    B = f32(r.random((k, n) if tb == "N" else (n, k)))
    C = f32(r.random((m, n)))
    alpha, beta = 0.75, -0.5
    pad = 3
    Cp = np.full((m + pad, n), 7.0, np.float32, order="F")
    Cp[:m] = C
    dA, dB, dC = to_dev32(A), to_dev32(B), to_dev32(Cp)
    dD = to_dev32(np.full((m + pad, n), -3.0, np.float32))
    st = gpu.cuSgemm2(blas, ord(ta), ord(tb), m, n, k, alpha, gpu.ptr(dA), A.shape[0], gpu.ptr(dB),
                      B.shape[0], beta, gpu.ptr(dC), m + pad, gpu.ptr(dD), m + pad)
    torch.cuda.synchronize()
    assert st == 0
    D = from_dev(dD)
    ref = O.dgemm(ta, tb, m, n, k, alpha, np.asfortranarray(A, np.float64),
                  np.asfortranarray(B, np.float64), beta, np.asfortranarray(C, np.float64))
    bound = 2 * (k + 2) * EPS32 * (abs(alpha) * np.abs(opm(A, ta)).astype(np.float64)
                                   @ np.abs(opm(B, tb)).astype(np.float64) + abs(beta) * np.abs(C))
    assert np.all(np.abs(D[:m].astype(np.float64) - ref) <= bound + 1e-30)
    assert np.all(D[m:] == -3.0)


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("n,k", [(1, 1), (33, 17), (128, 64), (257, 300), (2000, 384)])
def test_ssyrk(gpu, blas, uplo, trans, n, k):
    r = rng(2, n, k, ord(uplo), ord(trans))
    A = f32(r.random((n, k) if trans == "N" else (k, n)))
    C = f32(r.random((n, n)))
    dA, dC = to_dev32(A), to_dev32(C)
    alpha, beta = -1.0, 1.0
    st = gpu.cuSsyrk(blas, ord(uplo), ord(trans), n, k, alpha, gpu.ptr(dA), A.shape[0], beta,
                     gpu.ptr(dC), n)
    torch.cuda.synchronize()
    assert st == 0
    D = from_dev(dC).astype(np.float64)
    ref = O.dsyrk(uplo, trans, n, k, alpha, np.asfortranarray(A, np.float64), beta,
                  np.asfortranarray(C, np.float64))
    Aop = opm(A, trans).astype(np.float64)
    bound = 2 * (k + 2) * EPS32 * (np.abs(Aop) @ np.abs(Aop).T + np.abs(C))
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.all(np.abs(D - ref)[mask] <= bound[mask] + 1e-30)
    assert np.array_equal(D[~mask], C.astype(np.float64)[~mask])   # other triangle untouched


@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("diag", "NU")
def test_strsm(gpu, blas, side, uplo, trans, diag):
    m, n = 200, 70
    k = m if side == "L" else n
    r = rng(3, ord(side), ord(uplo), ord(trans), ord(diag))
    T = f32(r.standard_normal((k, k)) / np.sqrt(k) + 3 * np.eye(k))
    B = f32(r.random((m, n)))
    dT, dB = to_dev32(T), to_dev32(B)
    st = gpu.cuStrsm(blas, ord(side), ord(uplo), ord(trans), ord(diag), m, n, 1.5, gpu.ptr(dT), k,
                     gpu.ptr(dB), m)
    torch.cuda.synchronize()
    assert st == 0
    X = from_dev(dB).astype(np.float64)
    Tm = np.tril(T) if uplo == "L" else np.triu(T)
    Tm = Tm.astype(np.float64)
    if diag == "U":
        np.fill_diagonal(Tm, 1.0)
    Top = opm(Tm, trans)
    lhs = Top @ X if side == "L" else X @ Top
    rhs = 1.5 * B.astype(np.float64)
    # backward error of a triangular solve in FP32
    res = np.linalg.norm(lhs - rhs) / (np.linalg.norm(Top) * np.linalg.norm(X) + np.linalg.norm(rhs))
    assert res <= 10 * k * EPS32


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n", [1, 31, 128, 129, 500, 1500])
def test_spotrf(gpu, lapack, uplo, n):
    A = O.spd(n + 11, n, 1.0)
    dA = to_dev32(A)
    st, info = gpu.cuSpotrf(lapack, ord(uplo), n, gpu.ptr(dA), n)
    torch.cuda.synchronize()
    assert st == 0 and info == 0
    F = from_dev(dA).astype(np.float64)
    A32 = A.astype(np.float32).astype(np.float64)
    assert O.potrf_backward_error(A32, F, uplo) <= 10 * n * EPS32


def test_spotrf_info(gpu, lapack):
    n = 300
    A = O.spd(5, n, 1.0)
    A[200, 200] = -1.0
    dA = to_dev32(A)
    st, info = gpu.cuSpotrf(lapack, ord("L"), n, gpu.ptr(dA), n)
    torch.cuda.synchronize()
    assert st == 0 and info == 201


# ------------------------------------------------------------------ solvers

def normwise_backward(A, X, B):
    R = B - A @ X
    return (np.abs(R).max(axis=0) / (np.abs(A).sum(axis=1).max() * np.abs(X).max(axis=0)
                                     + np.abs(B).max(axis=0)))


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n,nrhs,pad", [(1, 1, 0), (65, 3, 0), (300, 16, 0), (1000, 5, 0), (700, 40, 0),
                                        (3000, 16, 0), (2049, 17, 1), (1024, 2, 1)])
def test_dpotrs(gpu, lapack, uplo, n, nrhs, pad):
    """pad: lda = n + pad (an odd lda takes the element-wise load path of the chained solve)."""
    A = O.spd(n + 3, n, 0.5)
    B = np.asfortranarray(rng(4, n, nrhs).standard_normal((n, nrhs)))
    lda = n + pad
    Ap = np.full((lda, n), np.nan, order="F")
    Ap[:n] = A
    dA = gpu.to_device(Ap)
    st, info = gpu.cuDpotrf(lapack, ord(uplo), n, gpu.ptr(dA), lda)
    assert st == 0 and info == 0
    ldb = n + 2
    Bp = np.full((ldb, nrhs), 5.0, order="F")
    Bp[:n] = B
    dB = gpu.to_device(Bp)
    st, info = gpu.cuDpotrs(lapack, ord(uplo), n, nrhs, gpu.ptr(dA), lda, gpu.ptr(dB), ldb)
    torch.cuda.synchronize()
    assert st == 0 and info == 0
    Xg = gpu.from_device(dB)
    assert np.all(Xg[n:] == 5.0)
    X = Xg[:n]
    # oracle: its own factor, then the two triangular solves (LAPACK dpotrs)
    F, rinfo = O.dpotrf(uplo, np.asfortranarray(A.copy()))
    assert rinfo == 0
    Xo = np.asfortranarray(B.copy())
    if uplo == "L":
        O.dtrsm("L", "L", "N", "N", n, nrhs, 1.0, F, Xo)
        O.dtrsm("L", "L", "T", "N", n, nrhs, 1.0, F, Xo)
    else:
        O.dtrsm("L", "U", "T", "N", n, nrhs, 1.0, F, Xo)
        O.dtrsm("L", "U", "N", "N", n, nrhs, 1.0, F, Xo)
    assert np.all(normwise_backward(A, X, B) <= 10 * n * EPS)
    cond = np.linalg.cond(A)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 10 * n * EPS * cond


def run_sposv(gpu, lapack, uplo, n, nrhs, alpha, factor, seed=0, A=None, refine="I"):
    A = O.spd(100 + seed, n, alpha) if A is None else A
    B = np.asfortranarray(rng(5, n, nrhs, seed).standard_normal((n, nrhs)))
    Astore = np.asfortranarray(np.tril(A) if uplo == "L" else np.triu(A))  # one triangle only
    Astore[~(np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool)))] = np.nan
    dA = gpu.to_device(Astore)
    dB = gpu.to_device(B)
    dX = torch.zeros_like(dB)
    st, it, used, info = gpu.cuDsposvx(lapack, ord(uplo), n, nrhs, gpu.ptr(dA), n, gpu.ptr(dB), n,
                                       gpu.ptr(dX), n, factor, refine)
    torch.cuda.synchronize()
    assert st == 0 and info == 0
    X = gpu.from_device(dX)
    return A, B, X, it, used, gpu.from_device(dA), Astore


def dsposv_converged(A, X, B):
    n = A.shape[0]
    cte = np.abs(A).sum(axis=1).max() * UNIT * np.sqrt(n)
    R = B - A @ X
    return np.all(np.abs(R).max(axis=0) <= np.abs(X).max(axis=0) * cte * 4)  # host FP64 rounding


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n,nrhs,alpha", [(1, 1, 1.0), (100, 1, 1.0), (700, 16, 1.0),
                                          (1000, 4, 1e-3), (1500, 16, 1e-5)])
def test_dsposv_fp32(gpu, lapack, uplo, n, nrhs, alpha):
    A, B, X, it, used, Aafter, Astore = run_sposv(gpu, lapack, uplo, n, nrhs, alpha, "S")
    assert it >= 0 and used == "S"
    assert dsposv_converged(A, X, B)
    np.testing.assert_array_equal(np.nan_to_num(Aafter, nan=1e300), np.nan_to_num(Astore, nan=1e300))


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n,nrhs", [(1024, 16), (1000, 7), (1001, 16), (2100, 33)])
def test_dsposv_fp32_refinement_steps(gpu, lapack, uplo, n, nrhs):
    """A well-conditioned system (cond ~ 5) refined from the FP32 factor gains
    about seven digits per step, so it meets the stopping test within three
    steps only if both FP32 triangular solves are exact to FP32 rounding: a
    block read from the wrong rows or columns costs convergence.  n = 1024
    (whole 128-row blocks) and 1000 (a partial last block) take the 16-byte
    loads of the FP32 copy (ld = n, a multiple of 4); n = 1001 and 2100 the
    element-wise loads; 33 right-hand sides span three 16-column slices."""
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, uplo, n, nrhs, 1.0, "S", seed=7)
    assert used == "S" and 0 <= it <= 3, (it, used)
    assert dsposv_converged(A, X, B)


@pytest.mark.parametrize("n,nrhs", [(300, 3), (1200, 16)])
def test_dsposv_bf16(gpu, lapack, n, nrhs):
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, "L", n, nrhs, 1.0, "B", seed=1)
    assert it >= 0 and used == "B"
    assert dsposv_converged(A, X, B)


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n,nrhs,alpha", [(300, 3, 1.0), (1200, 16, 1e-3), (2100, 4, 1e-4)])
def test_dsposv_split_bf16(gpu, lapack, uplo, n, nrhs, alpha):
    # three-product split BF16 (~16 bits): refines conditions that plain BF16 cannot (cond ~ 4 / alpha)
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, uplo, n, nrhs, alpha, "X", seed=3)
    assert it >= 0 and used == "X", (it, used)
    assert dsposv_converged(A, X, B)


def test_dsposv_split_bf16_large(gpu, lapack):
    # several panels with look-ahead (n > 2 NB); alpha = 1e-3 as in the n = 32768 workload
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, "L", 4500, 16, 1e-3, "X", seed=4)
    assert it >= 0 and used == "X", (it, used)
    assert dsposv_converged(A, X, B)


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("n,nrhs,alpha", [(300, 3, 1.0), (1200, 16, 1e-2), (2100, 4, 1e-3)])
def test_dsposv_fp16(gpu, lapack, uplo, n, nrhs, alpha):
    """FP16 products (11 significant bits) refine cond ~ 4 / alpha = 4e3, where
    BF16 products (8 bits) do not converge."""
    A, B, X, it, used, Aafter, Astore = run_sposv(gpu, lapack, uplo, n, nrhs, alpha, "H", seed=5)
    assert it >= 0 and used == "H", (it, used)
    assert dsposv_converged(A, X, B)
    np.testing.assert_array_equal(np.nan_to_num(Aafter, nan=1e300), np.nan_to_num(Astore, nan=1e300))


@pytest.mark.parametrize("scale", [1e10, 1e-10])
def test_dsposv_fp16_scaled_matrix(gpu, lapack, scale):
    """Factor entries reach sqrt(1e10) ~ 1e5, beyond FP16's 65504, or sit near
    1e-5, in its subnormal range: the panel copies are scaled by a power of
    two (undone exactly in the products), so the FP16 factor still refines
    in a few steps."""
    n = 1500
    A = np.asfortranarray(O.spd(21, n, 1.0) * scale)
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, "L", n, 8, 1.0, "H", seed=6, A=A)
    assert used == "H" and 0 <= it <= 4, (it, used)
    assert dsposv_converged(A, X, B)


def test_dsposv_fp16_large(gpu, lapack):
    # several panels with look-ahead (n > 2 NB), triangle-grid 128 x 128 FP16 trailing updates
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, "L", 4500, 16, 1e-3, "H", seed=4)
    assert it >= 0 and used == "H", (it, used)
    assert dsposv_converged(A, X, B)


def test_dsposv_fp16_not_spd_falls_back_to_fp64(gpu, lapack):
    n = 200
    A = O.spd(9, n, 1.0)
    A[150, 150] = -5.0
    B = np.ones((n, 1), order="F")
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    dX = torch.zeros_like(dB)
    st, it, used, info = gpu.cuDsposvx(lapack, ord("L"), n, 1, gpu.ptr(dA), n, gpu.ptr(dB), n,
                                       gpu.ptr(dX), n, "H", "I")
    torch.cuda.synchronize()
    assert st == 0 and it == -3 and used == "D" and info == 151


def test_dsposv_bf16_ill_conditioned_falls_back_to_fp32(gpu, lapack):
    # cond ~ 4e5: beyond what a bf16-product factor refines, within FP32's reach
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, "L", 800, 2, 1e-5, "B", seed=2)
    assert used in ("S", "B") and it >= 0
    assert dsposv_converged(A, X, B)


def test_dsposv_falls_back_to_fp64(gpu, lapack):
    # A = G G^T / n + 1e-10 I with G of rank n/2: half the spectrum sits at
    # 1e-10, cond ~ 1e10, beyond what an FP32 factor refines (or it fails
    # outright); the FP64 path overwrites A with its factor, as LAPACK dsposv
    n, nrhs = 600, 2
    G = rng(6).standard_normal((n, n // 2))
    A0 = np.asfortranarray(G @ G.T / n + 1e-10 * np.eye(n))
    A, B, X, it, used, Aafter, _ = run_sposv(gpu, lapack, "L", n, nrhs, 0.0, "S", seed=3, A=A0)
    assert it < 0 and it in (-3, -31) and used == "D"
    assert np.all(normwise_backward(A, X, B) <= 10 * n * EPS * 10)
    assert O.potrf_backward_error(A, Aafter, "L") <= 10 * n * EPS


def test_dsposv_not_spd(gpu, lapack):
    n = 200
    A = O.spd(9, n, 1.0)
    A[150, 150] = -5.0
    B = np.ones((n, 1), order="F")
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    dX = torch.zeros_like(dB)
    st, it, info = gpu.cuDsposv(lapack, ord("L"), n, 1, gpu.ptr(dA), n, gpu.ptr(dB), n, gpu.ptr(dX), n)
    torch.cuda.synchronize()
    assert st == 0 and it == -3 and info == 151


@pytest.mark.parametrize("factor", "SB")
def test_dsposv_large_split_residual(gpu, lapack, factor):
    """n = 16384 takes the split-K residual path (128 x 16 tiles, 16 splits,
    partial sums reduced in a fixed order) and, with factor B, the bfloat16
    panel copies of every trailing update; the solution must still meet
    dsposv's test without falling back."""
    n, nrhs = 16384, 16
    g = torch.Generator(device="cuda").manual_seed(11)
    G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
    A = G @ G.T / n
    A.diagonal().add_(1.0)
    del G
    B = torch.randn(nrhs, n, generator=g, device="cuda", dtype=torch.float64)
    X = torch.zeros_like(B)
    st, it, used, info = gpu.cuDsposvx(lapack, ord("L"), n, nrhs, gpu.ptr(A), n, gpu.ptr(B), n,
                                       gpu.ptr(X), n, factor)
    torch.cuda.synchronize()
    assert st == 0 and info == 0 and it >= 0 and used == factor
    R = B - X @ A
    anorm = A.abs().sum(dim=1).max().item()
    cte = anorm * UNIT * np.sqrt(n)
    assert torch.all(R.abs().max(dim=1).values <= 4 * cte * X.abs().max(dim=1).values)


@pytest.mark.parametrize("uplo", "LU")
def test_dsposv_gmres_bf16(gpu, lapack, uplo):
    """GMRES-IR keeps the bfloat16-product factor where classical refinement
    with it stalls (cond ~ 4e3): the solution still meets dsposv's test."""
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, uplo, 1000, 4, 1e-3, "B", seed=4, refine="G")
    assert it >= 0 and used == "B"
    assert dsposv_converged(A, X, B)


@pytest.mark.parametrize("uplo", "LU")
def test_dsposv_gmres_fp16(gpu, lapack, uplo):
    """GMRES-IR with the FP16-product factor at cond ~ 4e4 (beyond what its
    classical refinement contracts), over several panels with look-ahead."""
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, uplo, 2600, 4, 1e-4, "H", seed=8, refine="G")
    assert it >= 0 and used == "H", (it, used)
    assert dsposv_converged(A, X, B)


def test_dsposv_gmres_fp32_ill_conditioned(gpu, lapack):
    A, B, X, it, used, _, _ = run_sposv(gpu, lapack, "L", 900, 3, 1e-6, "S", seed=5, refine="G")
    assert it >= 0 and used == "S"
    assert dsposv_converged(A, X, B)
