"""GPU parity at the full sizes of the five benchmark workloads.

1. dpotrf (lower, n = 1024, A = G G^T/n + I) -> dlogdet -> dpotri, against
   the oracle on the same matrix (factor, log-determinant and inverse).
2. dpotrf + dpotri (lower, n = 16384, A = G G^T/n + 0.1 I) and zpotrf +
   zpotri (upper, n = 8192, A = H H^H/n + 0.1 I): backward error
   ||A - L L^H||_F / ||A||_F <= 10 n eps and inverse residual
   ||I - A X||_F / (||A||_F ||X||_F) <= 10 n eps, computed in FP64 on the
   GPU (the oracle would take minutes at these sizes); the log-determinant
   against an independent factorisation (torch.linalg.cholesky).
3. Mixed-precision solve, n = 32768, 16 right-hand sides,
   A = G G^T/n + alpha I for alpha in {1, 1e-3, 1e-6}, every factor /
   refinement variant: DSPOSV's stopping test per column
   (||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n)), the normwise backward
   error max_j ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= n eps,
   and the precision actually used.
4. Batched factorisation + logdet (+ inverse) of 20000 matrices,
   n_i ~ U[8, 256], A_i = G_i G_i^T/n_i + 0.5 I, FP64 and FP32 storage:
   every matrix checked on the GPU (backward error, inverse residual, logdet
   against its own factor's diagonal), and a sample of 500 against the
   oracle (factor, logdet, inverse).
5. Batched mixed solve + inverse + logdet of 512 matrices, n_i log-uniform in
   [128, 2048], alpha_i log-uniform in [1e-4, 1], 4 right-hand sides: every
   matrix checked on the GPU, a sample against the oracle.

The Gaussian matrices come from torch's seeded generator on the GPU (the
benchmark's inputs); workload 1 uses the oracle's generator so that the oracle
factors the identical matrix.
"""
import math

import numpy as np
import pytest
import torch

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def gpu_spd(n, alpha, seed, cplx=False):
    """A = G G^H / n + alpha I on the GPU, exactly Hermitian."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    if cplx:
        H = torch.complex(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))
        H /= 2 ** 0.5
        A = H @ H.conj().T.contiguous()
        del H
    else:
        G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
        A = G @ G.T.contiguous()
        del G
    A /= n
    A = (A + A.conj().T) / 2
    A.diagonal().add_(alpha)
    return A.contiguous()


def mat(T):
    """The matrix whose column-major storage is the row-major tensor T."""
    return T.T


def factor_backward_error(A, M, uplo):
    F = torch.tril(M) if uplo == "L" else torch.triu(M)
    R = F @ F.conj().T if uplo == "L" else F.conj().T @ F
    return (torch.linalg.norm(A - R) / torch.linalg.norm(A)).item()


def hermitian_from(M, uplo):
    if uplo == "L":
        T = torch.tril(M)
        return T + torch.tril(M, -1).conj().T
    T = torch.triu(M)
    return T + torch.triu(M, 1).conj().T


def inverse_residual(A, X):
    n = A.shape[0]
    eye = torch.eye(n, dtype=A.dtype, device=A.device)
    return (torch.linalg.norm(eye - A @ X) / (torch.linalg.norm(A) * torch.linalg.norm(X))).item()


# --------------------------------------------------------------------------
# 1. n = 1024 pipeline against the oracle
# --------------------------------------------------------------------------

def test_workload1_pipeline_vs_oracle(gpu, lapack):
    n = 1024
    A = O.spd(1, n, 1.0)
    dA = gpu.to_device(A, "cuda")
    st, info = gpu.cuDpotrf(lapack, gpu.Lower, n, gpu.ptr(dA), n)
    assert (st, info) == (0, 0)
    st, logdet = gpu.cuDlogdet(lapack, gpu.ptr(dA), n + 1, n)
    assert st == 0
    F = gpu.from_device(dA).copy()
    st, info = gpu.cuDpotri(lapack, gpu.Lower, n, gpu.ptr(dA), n)
    assert (st, info) == (0, 0)
    X = gpu.from_device(dA).copy()
    ref, rinfo = O.dpotrf("L", A.copy(order="F"))
    assert rinfo == 0
    assert O.potrf_backward_error(A, F, "L") <= 10 * n * EPS
    diff = np.linalg.norm(np.tril(F) - np.tril(ref)) / np.linalg.norm(np.tril(ref))
    assert diff <= 100 * n * EPS, diff
    want = O.dlogdet(np.ascontiguousarray(ref.reshape(-1, order="F")), n + 1, n)
    assert abs(logdet - want) <= 10 * n * EPS * max(1.0, abs(want)), (logdet, want)
    Xref, rinfo = O.dpotri("L", ref.copy(order="F"))
    assert rinfo == 0
    assert O.inverse_residual(A, X, "L") <= 10 * n * EPS
    cond = np.linalg.cond(A)
    dx = np.linalg.norm(np.tril(X) - np.tril(Xref)) / np.linalg.norm(np.tril(Xref))
    assert dx <= 100 * n * EPS * cond, (dx, cond)


# --------------------------------------------------------------------------
# 2. n = 16384 (real, lower) and n = 8192 (complex, upper)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["d_lower_16384", "z_upper_8192"])
def test_workload2_factor_inverse(gpu, lapack, kind):
    cplx = kind.startswith("z")
    n = 8192 if cplx else 16384
    uplo = "U" if cplx else "L"
    A = gpu_spd(n, 0.1, seed=11 if cplx else 10, cplx=cplx)
    W = A.T.contiguous()                    # column-major storage of A
    potrf = gpu.cuZpotrf if cplx else gpu.cuDpotrf
    potri = gpu.cuZpotri if cplx else gpu.cuDpotri
    logdet_fn = gpu.cuZlogdet if cplx else gpu.cuDlogdet
    ul = gpu.Upper if cplx else gpu.Lower
    st, info = potrf(lapack, ul, n, gpu.ptr(W), n)
    assert (st, info) == (0, 0)
    torch.cuda.synchronize()
    bwd = factor_backward_error(A, mat(W), uplo)
    print(f"{kind}: potrf backward error {bwd:.3e}", flush=True)
    assert bwd <= 10 * n * EPS, bwd
    st, ld = logdet_fn(lapack, gpu.ptr(W), n + 1, n)
    assert st == 0
    Lt = torch.linalg.cholesky(A)           # independent factorisation
    want = 2.0 * torch.log(Lt.diagonal().real).sum().item()
    del Lt
    assert abs(ld - want) <= 10 * n * EPS * max(1.0, abs(want)), (ld, want)
    st, info = potri(lapack, ul, n, gpu.ptr(W), n)
    assert (st, info) == (0, 0)
    torch.cuda.synchronize()
    X = hermitian_from(mat(W), uplo)
    del W
    res = inverse_residual(A, X)
    print(f"{kind}: potri inverse residual {res:.3e}", flush=True)
    assert res <= 10 * n * EPS, res
    del A, X
    torch.cuda.empty_cache()


# --------------------------------------------------------------------------
# 3. mixed-precision solve at n = 32768
# --------------------------------------------------------------------------

VARIANTS = (("S", "I"), ("B", "I"), ("X", "I"), ("H", "I"), ("B", "G"), ("H", "G"))


@pytest.fixture(scope="module")
def base32768(gpu):
    n = 32768
    g = torch.Generator(device="cuda").manual_seed(3)
    G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
    S = G @ G.T.contiguous()
    del G
    S /= n
    S = (S + S.T) / 2
    yield S
    del S
    torch.cuda.empty_cache()


@pytest.mark.parametrize("alpha", [1.0, 1e-3, 1e-6])
def test_workload3_mixed_solve(gpu, lapack, base32768, alpha):
    n, nrhs = 32768, 16
    A = base32768.clone()
    A.diagonal().add_(alpha)
    g = torch.Generator(device="cuda").manual_seed(4)
    B = torch.randn(nrhs, n, generator=g, device="cuda", dtype=torch.float64)  # column-major n x nrhs
    anorm = A.abs().sum(dim=1).max().item()

    def check(X, label):
        R = B - X @ A                       # rows: (b - A x)^T (A symmetric)
        rn = R.abs().max(dim=1).values
        xn = X.abs().max(dim=1).values
        bn = B.abs().max(dim=1).values
        berr = (rn / (anorm * xn + bn)).max().item()
        dsposv = (rn <= xn * anorm * EPS * math.sqrt(n)).all().item()
        print(f"alpha={alpha:g} {label}: backward error {berr:.3e}", flush=True)
        assert berr <= n * EPS, (label, berr)
        return dsposv

    # FP64 reference path: potrf + potrs
    F = A.clone()
    X = B.clone()
    st, info = gpu.cuDpotrf(lapack, gpu.Lower, n, gpu.ptr(F), n)
    assert (st, info) == (0, 0)
    st, info = gpu.cuDpotrs(lapack, gpu.Lower, n, nrhs, gpu.ptr(F), n, gpu.ptr(X), n)
    assert (st, info) == (0, 0)
    check(X, "fp64")
    del F
    for factor, refine in VARIANTS:
        X = torch.zeros_like(B)
        st, it, used, info = gpu.cuDsposvx(lapack, gpu.Lower, n, nrhs, gpu.ptr(A), n, gpu.ptr(B), n,
                                           gpu.ptr(X), n, factor, refine)
        assert (st, info) == (0, 0), (factor, refine, st, info)
        label = f"{factor}{refine} iter={it} used={used}"
        ok = check(X, label)
        # the precision used is the one asked for, or the documented retries
        # (16-bit -> FP32 -> FP64)
        assert used in (factor, "S", "D"), label
        if alpha == 1.0:                    # well conditioned: no retry
            assert used == factor and it >= 0, label
        if it >= 0:
            assert used != "D"
            if refine == "I":               # classical IR stops on DSPOSV's test
                assert ok, label
        else:
            assert used == "D", label
            A = base32768.clone()           # A now holds its FP64 factor
            A.diagonal().add_(alpha)
    del A, B, X
    torch.cuda.empty_cache()


# --------------------------------------------------------------------------
# 4. batched: 20000 matrices, n in [8, 256]
# --------------------------------------------------------------------------

def packed_batch(sizes, alpha, seed):
    """Packed column-major A_i = G_i G_i^T / n_i + alpha I (exactly symmetric)."""
    count = len(sizes)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)[:-1]
    total = int((sizes.astype(np.int64) ** 2).sum())
    buf = torch.empty(total, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        G = torch.randn(len(idx), n, n, generator=g, device="cuda", dtype=torch.float64)
        M = torch.bmm(G, G.transpose(1, 2).contiguous()) / n
        M = (M + M.transpose(1, 2)) / 2
        M += alpha * torch.eye(n, device="cuda", dtype=torch.float64)
        gather = torch.from_numpy(offs[idx]).cuda()[:, None] + torch.arange(n * n, device="cuda")[None, :]
        buf[gather.reshape(-1)] = M.reshape(-1)
        del G, M
    return buf, offs


def groups(sizes, offs, buf):
    """(n, indices, stacked n x n matrices) per distinct size, as matrices."""
    for n in np.unique(sizes):
        if n == 0:
            continue
        idx = np.nonzero(sizes == n)[0]
        gather = torch.from_numpy(offs[idx]).cuda()[:, None] + torch.arange(n * n, device="cuda")[None, :]
        yield int(n), idx, buf[gather].reshape(len(idx), n, n).transpose(1, 2)


@pytest.mark.parametrize("prec", "DS")
@pytest.mark.parametrize("inverse", [False, True], ids=["potrf", "potri"])
def test_workload4_batched_20000(gpu, lapack, prec, inverse):
    count = 20000
    sizes = np.random.default_rng(5).integers(8, 257, size=count)
    A0, offs = packed_batch(sizes, 0.5, seed=6)
    dtype = torch.float64 if prec == "D" else torch.float32
    eps = EPS if prec == "D" else EPS32
    work = A0.to(dtype, copy=True)
    # the input as stored (FP32 rounded for 'S'); a copy, since for 'D' the
    # conversion alone would return the tensor that is factored in place
    Ain = work.to(torch.float64, copy=True)
    logd = torch.zeros(count, dtype=torch.float64, device="cuda")
    info = torch.full((count,), -7, dtype=torch.int32, device="cuda")
    with gpu.BatchedPlan(lapack, gpu.Lower, prec, sizes.tolist(), offs.tolist()) as plan:
        assert plan.status == 0
        fn = plan.run_inverse if inverse else plan.run
        assert fn(gpu.ptr(work), gpu.ptr(logd), gpu.ptr(info)) == 0
        torch.cuda.synchronize()
    assert int((info != 0).sum().item()) == 0
    out = work.double()
    worst_b, worst_r, worst_l = 0.0, 0.0, 0.0
    for (n, idx, Ag), (_, _, Og) in zip(groups(sizes, offs, Ain), groups(sizes, offs, out)):
        Ag = torch.tril(Ag) + torch.tril(Ag, -1).transpose(1, 2)   # from the stored triangle
        if not inverse:
            L = torch.tril(Og)
            R = torch.bmm(L, L.transpose(1, 2).contiguous())
            b = (torch.linalg.norm(Ag - R, dim=(1, 2)) / torch.linalg.norm(Ag, dim=(1, 2))).max().item()
            worst_b = max(worst_b, b / n)
            assert b <= 10 * n * eps, (n, b)
            want = 2.0 * torch.log(torch.diagonal(L, dim1=1, dim2=2)).sum(dim=1)
            got = logd[torch.from_numpy(idx).cuda()]
            dl = ((got - want).abs() / want.abs().clamp(min=1.0)).max().item()
            worst_l = max(worst_l, dl / n)
            assert dl <= 10 * n * eps, (n, dl)
        else:
            X = torch.tril(Og) + torch.tril(Og, -1).transpose(1, 2)
            eye = torch.eye(n, dtype=torch.float64, device="cuda")
            r = (torch.linalg.norm(eye - torch.bmm(Ag, X), dim=(1, 2))
                 / (torch.linalg.norm(Ag, dim=(1, 2)) * torch.linalg.norm(X, dim=(1, 2)))).max().item()
            worst_r = max(worst_r, r / n)
            assert r <= 10 * n * eps, (n, r)
    print(f"batched {prec} inverse={inverse}: worst backward/n {worst_b:.2e}, inverse residual/n "
          f"{worst_r:.2e}, logdet/n {worst_l:.2e}", flush=True)
    # a sample of 500 against the oracle (factor or inverse, and logdet)
    pick = np.random.default_rng(8).choice(count, size=500, replace=False)
    ain = Ain.cpu().numpy()
    res = out.cpu().numpy()
    lg = logd.cpu().numpy()
    for i in pick:
        n, o = int(sizes[i]), int(offs[i])
        A = ain[o:o + n * n].reshape((n, n), order="F")
        A = np.tril(A) + np.tril(A, -1).T
        ref, rinfo = O.dpotrf("L", np.asfortranarray(A.copy()))
        assert rinfo == 0
        want = 2.0 * math.fsum(np.log(np.diag(ref)))
        assert abs(lg[i] - want) <= 10 * n * eps * max(1.0, abs(want)), (i, n, lg[i], want)
        got = res[o:o + n * n].reshape((n, n), order="F")
        if not inverse:
            d = np.linalg.norm(np.tril(got) - np.tril(ref)) / np.linalg.norm(np.tril(ref))
            assert d <= 100 * n * eps, (i, n, d)
        else:
            Xref, rinfo = O.dpotri("L", np.asfortranarray(ref.copy()))
            assert rinfo == 0
            d = np.linalg.norm(np.tril(got) - np.tril(Xref)) / np.linalg.norm(np.tril(Xref))
            assert d <= 100 * n * eps * np.linalg.cond(A), (i, n, d)


# --------------------------------------------------------------------------
# 5. batched mixed solve: 512 matrices, n in [128, 2048]
# --------------------------------------------------------------------------

def test_workload5_batched_mixed_512(gpu, lapack):
    count, nrhs = 512, 4
    rng = np.random.default_rng(9)
    sizes = np.exp(rng.uniform(math.log(128), math.log(2048), size=count)).astype(np.int64)
    alphas = np.exp(rng.uniform(math.log(1e-4), math.log(1.0), size=count))
    offA = np.zeros(count, dtype=np.int64)
    offA[1:] = np.cumsum(sizes ** 2)[:-1]
    offB = np.zeros(count, dtype=np.int64)
    offB[1:] = np.cumsum(sizes * nrhs)[:-1]
    A = torch.empty(int((sizes ** 2).sum()), dtype=torch.float64, device="cuda")
    B = torch.empty(int((sizes * nrhs).sum()), dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(12)
    for i in range(count):
        n = int(sizes[i])
        G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
        M = G @ G.T.contiguous() / n
        M = (M + M.T) / 2
        M.diagonal().add_(float(alphas[i]))
        A[offA[i]:offA[i] + n * n] = M.reshape(-1)
        B[offB[i]:offB[i] + n * nrhs] = torch.randn(n * nrhs, generator=g, device="cuda",
                                                    dtype=torch.float64)
    A0 = A.clone()
    X = torch.zeros_like(B)
    Ainv = torch.zeros_like(A)
    st, it, logdet, info = gpu.cuDsposvxBatched(lapack, gpu.Lower, sizes.tolist(), offA.tolist(),
                                                gpu.ptr(A), nrhs, offB.tolist(), gpu.ptr(B), gpu.ptr(X),
                                                gpu.ptr(Ainv))
    assert st == 0
    torch.cuda.synchronize()
    assert all(v == 0 for v in info), [i for i, v in enumerate(info) if v]
    worst = {"backward": 0.0, "inverse": 0.0, "logdet": 0.0}
    for i in range(count):
        n = int(sizes[i])
        Ai = A0[offA[i]:offA[i] + n * n].reshape(n, n)          # symmetric
        Bi = B[offB[i]:offB[i] + n * nrhs].reshape(nrhs, n)     # rows: right-hand sides
        Xi = X[offB[i]:offB[i] + n * nrhs].reshape(nrhs, n)
        R = Bi - Xi @ Ai
        anorm = Ai.abs().sum(dim=1).max()
        berr = (R.abs().max(dim=1).values
                / (anorm * Xi.abs().max(dim=1).values + Bi.abs().max(dim=1).values)).max().item()
        assert berr <= n * EPS, (i, n, berr)
        Inv = hermitian_from(mat(Ainv[offA[i]:offA[i] + n * n].reshape(n, n)), "L")
        res = inverse_residual(Ai, Inv)
        assert res <= 10 * n * EPS, (i, n, res)
        want = 2.0 * torch.log(torch.linalg.cholesky(Ai).diagonal()).sum().item()
        dl = abs(logdet[i] - want) / max(1.0, abs(want))
        assert dl <= 10 * n * EPS, (i, n, logdet[i], want)
        worst["backward"] = max(worst["backward"], berr)
        worst["inverse"] = max(worst["inverse"], res)
        worst["logdet"] = max(worst["logdet"], dl)
    # A is left unchanged where the mixed solve converged
    ok = [i for i in range(count) if it[i] >= 0]
    for i in ok[:64]:
        n = int(sizes[i])
        assert torch.equal(A[offA[i]:offA[i] + n * n], A0[offA[i]:offA[i] + n * n]), i
    print(f"batched mixed: worst {worst}, iter histogram "
          f"{dict(zip(*np.unique(np.array(it), return_counts=True)))}", flush=True)
    # oracle sample (the smaller matrices keep the CPU work short)
    a0 = A0.cpu().numpy()
    ai = Ainv.cpu().numpy()
    for i in np.argsort(sizes)[::64]:
        n, o = int(sizes[i]), int(offA[i])
        M = a0[o:o + n * n].reshape((n, n), order="F")
        ref, rinfo = O.dpotrf("L", np.asfortranarray(M.copy()))
        assert rinfo == 0
        want = O.dlogdet(np.ascontiguousarray(ref.reshape(-1, order="F")), n + 1, n)
        assert abs(logdet[i] - want) <= 10 * n * EPS * max(1.0, abs(want)), (i, n)
        Xref, rinfo = O.dpotri("L", np.asfortranarray(ref.copy()))
        assert rinfo == 0
        got = ai[o:o + n * n].reshape((n, n), order="F")
        d = np.linalg.norm(np.tril(got) - np.tril(Xref)) / np.linalg.norm(np.tril(Xref))
        assert d <= 100 * n * EPS * np.linalg.cond(M), (i, n, d)
