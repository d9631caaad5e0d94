"""GPU parity of the batched factorisation + log-determinant (cuPotrfBatched*).

Each matrix of the batch is checked against the oracle on the same input:
the factor by backward error ||A - L L^T||_F / ||A||_F <= 10 n eps (eps of
the storage precision), the log-determinant against 2 sum log L_jj of the
oracle's own factor (relative 10 n eps), and info exactly (non-SPD matrices
are planted in the batch).  The unreferenced triangle of every matrix and the
gaps between matrices must come back bit-identical; replays of the captured
graph must reproduce the first run bit for bit.
"""
import math

import numpy as np
import pytest
import torch

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def make_batch(seed, sizes, gap=3):
    """Packed column-major matrices with `gap` sentinel elements between them."""
    offs, pos = [], 0
    for n in sizes:
        offs.append(pos)
        pos += n * n + gap
    buf = np.full(pos, np.nan)
    mats = []
    for i, n in enumerate(sizes):
        A = O.spd(1000 + seed * 97 + i, n, 0.5) if n > 0 else np.zeros((0, 0))
        mats.append(A)
        buf[offs[i]:offs[i] + n * n] = A.ravel(order="F")
    return buf, offs, mats


def triangle_only(buf, offs, sizes, uplo):
    """NaN in the unreferenced triangle of every matrix (it must survive)."""
    out = buf.copy()
    for o, n in zip(offs, sizes):
        M = out[o:o + n * n].reshape((n, n), order="F")
        mask = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
        M[mask] = np.nan
        out[o:o + n * n] = M.ravel(order="F")
    return out


def run(gpu, lapack, uplo, prec, sizes, buf, offs):
    dtype = torch.float64 if prec == "D" else torch.float32
    dA = torch.from_numpy(buf).to(dtype).cuda()
    dlog = torch.zeros(len(sizes), dtype=torch.float64, device="cuda")
    dinfo = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    with gpu.BatchedPlan(lapack, ord(uplo), prec, sizes, offs) as plan:
        assert plan.status == 0
        assert plan.run(gpu.ptr(dA), gpu.ptr(dlog), gpu.ptr(dinfo)) == 0
        torch.cuda.synchronize()
    return dA.double().cpu().numpy(), dlog.cpu().numpy(), dinfo.cpu().numpy()


def check(buf_in, out, logdet, info, offs, sizes, mats, uplo, eps):
    for i, (o, n) in enumerate(zip(offs, sizes)):
        assert info[i] == 0, (i, n, info[i])
        if n == 0:
            assert logdet[i] == 0.0
            continue
        F = out[o:o + n * n].reshape((n, n), order="F")
        A = mats[i] if eps == EPS else mats[i].astype(np.float32).astype(np.float64)
        assert O.potrf_backward_error(A, F, uplo) <= 10 * n * eps, (i, n)
        ref, rinfo = O.dpotrf(uplo, np.asfortranarray(A.copy()))
        assert rinfo == 0
        want = 2.0 * math.fsum(np.log(np.diag(ref)))
        assert abs(logdet[i] - want) <= 10 * n * eps * max(1.0, abs(want)), (i, n, logdet[i], want)
    # unreferenced triangles and gaps untouched (NaN stays NaN)
    np.testing.assert_array_equal(np.isnan(out), np.isnan(buf_in))


SIZES_EDGE = [1, 2, 15, 16, 17, 31, 32, 33, 100, 127, 128, 129, 200, 255, 256, 0, 8]


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("prec", "DS")
def test_batched_edge_sizes(gpu, lapack, uplo, prec):
    buf, offs, mats = make_batch(1, SIZES_EDGE)
    buf = triangle_only(buf, offs, SIZES_EDGE, uplo)
    out, logdet, info = run(gpu, lapack, uplo, prec, SIZES_EDGE, buf, offs)
    check(buf if prec == "D" else buf.astype(np.float32).astype(np.float64), out, logdet, info, offs,
          SIZES_EDGE, mats, uplo, EPS if prec == "D" else EPS32)


@pytest.mark.parametrize("prec", "DS")
def test_batched_random_sizes(gpu, lapack, prec):
    sizes = list(np.random.default_rng(7).integers(8, 257, size=300))
    buf, offs, mats = make_batch(2, sizes)
    buf = triangle_only(buf, offs, sizes, "L")
    out, logdet, info = run(gpu, lapack, "L", prec, sizes, buf, offs)
    check(buf if prec == "D" else buf.astype(np.float32).astype(np.float64), out, logdet, info, offs,
          sizes, mats, "L", EPS if prec == "D" else EPS32)


def test_batched_info_and_neighbours(gpu, lapack):
    """Non-SPD matrices report their failing column; their neighbours are unaffected."""
    sizes = [50, 130, 64, 256, 40]
    bad = {1: 77, 3: 200}           # matrix -> 0-based failing column
    buf, offs, mats = make_batch(3, sizes)
    for i, k in bad.items():
        n = sizes[i]
        A = mats[i]
        L = np.linalg.cholesky(A)
        A[k, k] = np.dot(L[k, :k], L[k, :k]) - 1.0
        buf[offs[i]:offs[i] + n * n] = A.ravel(order="F")
    out, logdet, info = run(gpu, lapack, "L", "D", sizes, buf, offs)
    for i, n in enumerate(sizes):
        if i in bad:
            assert info[i] == bad[i] + 1 and np.isnan(logdet[i])
        else:
            assert info[i] == 0
            F = out[offs[i]:offs[i] + n * n].reshape((n, n), order="F")
            assert O.potrf_backward_error(mats[i], F, "L") <= 10 * n * EPS


def test_batched_graph_replay(gpu, lapack):
    """The captured graph replays bit-identically; new buffers re-capture."""
    sizes = [64, 200, 17, 256, 128]
    buf, offs, _ = make_batch(4, sizes)
    results = []
    with gpu.BatchedPlan(lapack, ord("L"), "D", sizes, offs) as plan:
        dA = torch.from_numpy(buf).cuda()
        dlog = torch.zeros(len(sizes), dtype=torch.float64, device="cuda")
        dinfo = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
        for rep in range(3):
            dA.copy_(torch.from_numpy(buf))          # eager work between replays
            torch.cuda.synchronize()
            assert plan.run(gpu.ptr(dA), gpu.ptr(dlog), gpu.ptr(dinfo)) == 0
            torch.cuda.synchronize()
            results.append((dA.cpu().numpy().copy(), dlog.cpu().numpy().copy()))
        dB = torch.from_numpy(buf).cuda()             # other buffers: a new capture
        dlog2 = torch.zeros_like(dlog)
        assert plan.run(gpu.ptr(dB), gpu.ptr(dlog2), gpu.ptr(dinfo)) == 0
        torch.cuda.synchronize()
        results.append((dB.cpu().numpy().copy(), dlog2.cpu().numpy().copy()))
    for A, ld in results[1:]:
        np.testing.assert_array_equal(np.nan_to_num(A), np.nan_to_num(results[0][0]))
        np.testing.assert_array_equal(ld, results[0][1])


# ---------------------------------------------------------------- inverse (cuPotriBatchedRun)

def run_inverse(gpu, lapack, uplo, prec, sizes, buf, offs):
    dtype = torch.float64 if prec == "D" else torch.float32
    dA = torch.from_numpy(buf).to(dtype).cuda()
    dlog = torch.zeros(len(sizes), dtype=torch.float64, device="cuda")
    dinfo = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    with gpu.BatchedPlan(lapack, ord(uplo), prec, sizes, offs) as plan:
        assert plan.status == 0
        assert plan.run_inverse(gpu.ptr(dA), gpu.ptr(dlog), gpu.ptr(dinfo)) == 0
        torch.cuda.synchronize()
    return dA.double().cpu().numpy(), dlog.cpu().numpy(), dinfo.cpu().numpy()


def check_inverse(buf_in, out, logdet, info, offs, sizes, mats, uplo, eps):
    """Per matrix: ||I - A X|| / cond(A) <= 10 n eps, X against the oracle's
    dpotri of its own factor (relative 10 n eps cond), logdet as for the
    factorisation; unreferenced triangles and gaps untouched."""
    for i, (o, n) in enumerate(zip(offs, sizes)):
        assert info[i] == 0, (i, n, info[i])
        if n == 0:
            continue
        X = out[o:o + n * n].reshape((n, n), order="F")
        A = mats[i] if eps == EPS else mats[i].astype(np.float32).astype(np.float64)
        condA = np.linalg.cond(A)
        assert O.inverse_residual(A, X, uplo) <= 10 * n * eps, (i, n)
        F, rinfo = O.dpotrf(uplo, np.asfortranarray(A.copy()))
        assert rinfo == 0
        want = 2.0 * math.fsum(np.log(np.diag(F)))
        assert abs(logdet[i] - want) <= 10 * n * eps * max(1.0, abs(want)), (i, n)
        ref, _ = O.dpotri(uplo, F.copy(order="F"))   # the oracle inverts in place
        Xf, Rf = O.herm_from(X, uplo), O.herm_from(ref, uplo)
        assert np.linalg.norm(Xf - Rf) <= 10 * n * eps * condA * np.linalg.norm(Rf), (i, n)
    np.testing.assert_array_equal(np.isnan(out), np.isnan(buf_in))


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("prec", "DS")
def test_batched_inverse_edge_sizes(gpu, lapack, uplo, prec):
    buf, offs, mats = make_batch(5, SIZES_EDGE)
    buf = triangle_only(buf, offs, SIZES_EDGE, uplo)
    out, logdet, info = run_inverse(gpu, lapack, uplo, prec, SIZES_EDGE, buf, offs)
    check_inverse(buf if prec == "D" else buf.astype(np.float32).astype(np.float64), out, logdet, info,
                  offs, SIZES_EDGE, mats, uplo, EPS if prec == "D" else EPS32)


def test_batched_inverse_closed_form(gpu, lapack):
    """dlatmc matrices (H D H, cond 2): the inverse is H D^-1 H exactly."""
    sizes = [7, 48, 130, 256]
    offs, pos, mats, want = [], 0, [], []
    for n in sizes:
        offs.append(pos)
        pos += n * n
        O.srand(n)
        A = O.dlatmc(n, 2.0)
        mats.append(A)
        want.append(np.linalg.inv(A))
    buf = np.concatenate([A.ravel(order="F") for A in mats])
    out, logdet, info = run_inverse(gpu, lapack, "L", "D", sizes, buf, offs)
    for i, (o, n) in enumerate(zip(offs, sizes)):
        assert info[i] == 0
        X = O.herm_from(out[o:o + n * n].reshape((n, n), order="F"), "L")
        assert np.linalg.norm(X - want[i]) <= 10 * n * EPS * 2 * np.linalg.norm(want[i]), (i, n)


def test_batched_inverse_skips_failed(gpu, lapack):
    """A non-SPD matrix reports info and is left as the partial factor; the others are inverted."""
    sizes = [40, 90, 33]
    buf, offs, mats = make_batch(6, sizes)
    n, k = sizes[1], 60
    A = mats[1]
    L = np.linalg.cholesky(A)
    A[k, k] = np.dot(L[k, :k], L[k, :k]) - 1.0
    buf[offs[1]:offs[1] + n * n] = A.ravel(order="F")
    out, logdet, info = run_inverse(gpu, lapack, "L", "D", sizes, buf, offs)
    assert info[1] == k + 1 and np.isnan(logdet[1])
    for i in (0, 2):
        m = sizes[i]
        X = out[offs[i]:offs[i] + m * m].reshape((m, m), order="F")
        assert info[i] == 0 and O.inverse_residual(mats[i], X, "L") <= 10 * m * EPS


def test_batched_inverse_then_factor_replays(gpu, lapack):
    """One plan alternates factor-only and factor+inverse runs (each re-captures)."""
    sizes = [64, 200, 17]
    buf, offs, mats = make_batch(7, sizes)
    with gpu.BatchedPlan(lapack, ord("L"), "D", sizes, offs) as plan:
        dA = torch.from_numpy(buf).cuda()
        dlog = torch.zeros(len(sizes), dtype=torch.float64, device="cuda")
        dinfo = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
        for rep in range(2):
            for inverse in (True, False):
                dA.copy_(torch.from_numpy(buf))
                torch.cuda.synchronize()
                fn = plan.run_inverse if inverse else plan.run
                assert fn(gpu.ptr(dA), gpu.ptr(dlog), gpu.ptr(dinfo)) == 0
                torch.cuda.synchronize()
                out = dA.cpu().numpy()
                for i, (o, n) in enumerate(zip(offs, sizes)):
                    M = out[o:o + n * n].reshape((n, n), order="F")
                    if inverse:
                        assert O.inverse_residual(mats[i], M, "L") <= 10 * n * EPS
                    else:
                        assert O.potrf_backward_error(mats[i], M, "L") <= 10 * n * EPS


# ---------------------------------------------------------------- persistent loop

@pytest.mark.parametrize("prec", "DS")
@pytest.mark.parametrize("inverse", [False, True], ids=["potrf", "potri"])
def test_batched_persistent_loop_reuse(gpu, lapack, monkeypatch, prec, inverse):
    """One workgroup per CU (CUCHOLESKY_BATCHED_PER_CU=1, read at plan
    creation) and more than two matrices per workgroup: every workgroup runs
    the persistent loop several times, reusing its LDS tiles and failure flag
    after a matrix that is not positive definite.  Factor or inverse, logdet
    and info of every matrix against the oracle."""
    monkeypatch.setenv("CUCHOLESKY_BATCHED_PER_CU", "1")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    count = 2 * cus + 97
    sizes = [int(v) for v in np.random.default_rng(21).integers(0, 257, size=count)]
    sizes[:6] = [0, 1, 256, 255, 16, 17]
    buf, offs, mats = make_batch(8, sizes)
    bad = {}
    for i in range(10, count, count // 9):   # planted non-SPD matrices, every size class
        n = sizes[i]
        if n < 2:
            continue
        k = int(np.random.default_rng(i).integers(1, n))
        A = mats[i]
        L = np.linalg.cholesky(A)
        A[k, k] = np.dot(L[k, :k], L[k, :k]) - 1.0
        buf[offs[i]:offs[i] + n * n] = A.ravel(order="F")
        bad[i] = k
    eps = EPS if prec == "D" else EPS32
    fn = run_inverse if inverse else run
    out, logdet, info = fn(gpu, lapack, "L", prec, sizes, buf, offs)
    for i, (o, n) in enumerate(zip(offs, sizes)):
        if i in bad:
            assert info[i] == bad[i] + 1 and np.isnan(logdet[i]), (i, n, info[i])
            continue
        assert info[i] == 0, (i, n, info[i])
        if n == 0:
            assert logdet[i] == 0.0
            continue
        A = mats[i] if prec == "D" else mats[i].astype(np.float32).astype(np.float64)
        M = out[o:o + n * n].reshape((n, n), order="F")
        F, rinfo = O.dpotrf("L", np.asfortranarray(A.copy()))
        assert rinfo == 0
        want = 2.0 * math.fsum(np.log(np.diag(F)))
        assert abs(logdet[i] - want) <= 10 * n * eps * max(1.0, abs(want)), (i, n)
        if inverse:
            assert O.inverse_residual(A, M, "L") <= 10 * n * eps, (i, n)
            ref, _ = O.dpotri("L", F.copy(order="F"))
            Rf = O.herm_from(ref, "L")
            assert np.linalg.norm(O.herm_from(M, "L") - Rf) <= 10 * n * eps * np.linalg.cond(A) * np.linalg.norm(Rf)
        else:
            assert O.potrf_backward_error(A, M, "L") <= 10 * n * eps, (i, n)
            d = np.linalg.norm(np.tril(M) - np.tril(F)) / np.linalg.norm(np.tril(F))
            assert d <= 100 * n * eps, (i, n, d)


# ---------------------------------------------------------------- mixed-precision batched solve

def test_batched_mixed_solve_inverse_logdet(gpu, lapack):
    """cuDsposvxBatched on uneven sizes and conditions (A = G G^T / n + alpha I):
    per matrix, X converges to the DSPOSV criterion with an FP64-level
    normwise backward error, logdet matches the oracle's factor, the inverse
    passes ||I - A X|| <= 10 n eps cond and matches the oracle's dpotri."""
    sizes = [1, 7, 129, 300, 520, 64, 1000]
    alphas = [1.0, 1e-2, 1e-4, 1.0, 1e-3, 1e-1, 1e-4]
    nrhs = 4
    offA = list(np.cumsum([0] + [n * n for n in sizes[:-1]]))
    offB = list(np.cumsum([0] + [n * nrhs for n in sizes[:-1]]))
    mats = [O.spd(300 + i, n, a) for i, (n, a) in enumerate(zip(sizes, alphas))]
    rhs = [np.asfortranarray(np.random.default_rng(i).standard_normal((n, nrhs))) for i, n in enumerate(sizes)]
    bufA = np.concatenate([np.tril(A).ravel(order="F") for A in mats])
    bufB = np.concatenate([B.ravel(order="F") for B in rhs])
    dA = torch.from_numpy(bufA).cuda()
    dB = torch.from_numpy(bufB).cuda()
    dX = torch.zeros_like(dB)
    dI = torch.full_like(dA, np.nan)
    st, it, ld, info = gpu.cuDsposvxBatched(lapack, ord("L"), sizes, offA, gpu.ptr(dA), nrhs, offB,
                                            gpu.ptr(dB), gpu.ptr(dX), gpu.ptr(dI))
    torch.cuda.synchronize()
    assert st == 0
    X, Inv = dX.cpu().numpy(), dI.cpu().numpy()
    for i, n in enumerate(sizes):
        A, B = mats[i], rhs[i]
        assert info[i] == 0 and it[i] >= 0, (i, n, info[i], it[i])
        Xi = X[offB[i]:offB[i] + n * nrhs].reshape((n, nrhs), order="F")
        R = B - A @ Xi
        berr = np.abs(R).max(axis=0) / (np.abs(A).sum(axis=1).max() * np.abs(Xi).max(axis=0))
        assert np.all(berr <= 4 * np.sqrt(n) * EPS), (i, n, berr)
        F, rinfo = O.dpotrf("L", np.asfortranarray(A.copy()))
        want = 2.0 * math.fsum(np.log(np.diag(F)))
        assert abs(ld[i] - want) <= 10 * n * EPS * max(1.0, abs(want)), (i, n, ld[i], want)
        Xinv = Inv[offA[i]:offA[i] + n * n].reshape((n, n), order="F")
        assert O.inverse_residual(A, Xinv, "L") <= 10 * n * EPS, (i, n)
        ref, _ = O.dpotri("L", F.copy(order="F"))
        Rf = O.herm_from(ref, "L")
        assert np.linalg.norm(O.herm_from(Xinv, "L") - Rf) <= 10 * n * EPS * np.linalg.cond(A) * np.linalg.norm(Rf)
    # A is untouched where refinement converged (here: all)
    np.testing.assert_array_equal(dA.cpu().numpy(), bufA)


def test_batched_mixed_solve_not_spd(gpu, lapack):
    """A matrix that is not positive definite reports info and NaN logdet; its neighbours solve."""
    sizes, nrhs = [40, 90, 33], 2
    mats = [O.spd(400 + i, n, 0.5) for i, n in enumerate(sizes)]
    k = 50
    L = np.linalg.cholesky(mats[1])
    mats[1][k, k] = np.dot(L[k, :k], L[k, :k]) - 1.0
    offA = list(np.cumsum([0] + [n * n for n in sizes[:-1]]))
    offB = list(np.cumsum([0] + [n * nrhs for n in sizes[:-1]]))
    dA = torch.from_numpy(np.concatenate([A.ravel(order="F") for A in mats])).cuda()
    dB = torch.ones(sum(n * nrhs for n in sizes), dtype=torch.float64, device="cuda")
    dX = torch.zeros_like(dB)
    st, it, ld, info = gpu.cuDsposvxBatched(lapack, ord("L"), sizes, offA, gpu.ptr(dA), nrhs, offB,
                                            gpu.ptr(dB), gpu.ptr(dX), 0)
    assert st == 0
    assert info[1] == k + 1 and np.isnan(ld[1])
    assert info[0] == 0 and info[2] == 0 and it[0] >= 0 and it[2] >= 0


# ---------------------------------------------------------------- task-graph plan (cuDsposvxBatchedPlan*)

def solve_plan_case(gpu, lapack, uplo, sizes, alphas, nrhs, seed, bad=None, reps=1):
    """Run one plan `reps` times (graph replays, the inputs restored between
    runs) and check every matrix against the oracle on its own input."""
    count = len(sizes)
    offA = [int(v) for v in np.cumsum([0] + [n * n + 5 for n in sizes[:-1]])]   # gaps between matrices
    offB = [int(v) for v in np.cumsum([0] + [n * nrhs + 3 for n in sizes[:-1]])]
    mats = [O.spd(700 + seed * 31 + i, n, a) if n > 0 else np.zeros((0, 0))
            for i, (n, a) in enumerate(zip(sizes, alphas))]
    for i, k in (bad or {}).items():
        L = np.linalg.cholesky(mats[i])
        mats[i][k, k] = np.dot(L[k, :k], L[k, :k]) - 1.0
    lenA = offA[-1] + sizes[-1] ** 2 + 5
    lenB = max(offB[-1] + sizes[-1] * nrhs + 3, 1)
    bufA = np.full(lenA, np.nan)
    for i, n in enumerate(sizes):
        M = mats[i].copy()
        mask = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
        M[mask] = np.nan                                  # the unreferenced triangle is never read
        bufA[offA[i]:offA[i] + n * n] = M.ravel(order="F")
    rng = np.random.default_rng(seed)
    bufB = rng.standard_normal(lenB)
    dA = torch.from_numpy(bufA).cuda()
    dB = torch.from_numpy(bufB).cuda()
    dX = torch.full_like(dB, -3.0)
    dI = torch.full_like(dA, -5.0)
    it = torch.full((count,), -99, dtype=torch.int32, device="cuda")
    ld = torch.zeros(count, dtype=torch.float64, device="cuda")
    info = torch.full((count,), -99, dtype=torch.int32, device="cuda")
    with gpu.BatchedSolvePlan(lapack, ord(uplo), sizes, offA, nrhs, offB) as plan:
        assert plan.status == 0
        outs = []
        for rep in range(reps):
            dX.fill_(-3.0)
            dI.fill_(-5.0)
            torch.cuda.synchronize()
            assert plan.run(gpu.ptr(dA), gpu.ptr(dB), gpu.ptr(dX), gpu.ptr(dI), gpu.ptr(it), gpu.ptr(ld),
                            gpu.ptr(info)) == 0
            torch.cuda.synchronize()
            outs.append((dX.cpu().numpy().copy(), dI.cpu().numpy().copy(), it.cpu().numpy().copy(),
                         ld.cpu().numpy().copy(), info.cpu().numpy().copy()))
    np.testing.assert_array_equal(np.isnan(dA.cpu().numpy()), np.isnan(bufA))   # A is never written
    for rep in range(1, reps):                            # replays are bit-identical
        for a, b in zip(outs[0], outs[rep]):
            np.testing.assert_array_equal(a, b)
    X, Inv, its, lds, infos = outs[0]
    for i, n in enumerate(sizes):
        if bad and i in bad:
            assert infos[i] == bad[i] + 1 and np.isnan(lds[i]), (i, infos[i])
            continue
        assert infos[i] == 0 and its[i] >= 0, (i, n, infos[i], its[i])
        if n == 0:
            assert lds[i] == 0.0
            continue
        A = mats[i]
        F, rinfo = O.dpotrf("L", np.asfortranarray(A.copy()))
        assert rinfo == 0
        want = 2.0 * math.fsum(np.log(np.diag(F)))
        assert abs(lds[i] - want) <= 10 * n * EPS * max(1.0, abs(want)), (i, n, lds[i], want)
        Xi = np.asarray(Inv[offA[i]:offA[i] + n * n]).reshape((n, n), order="F")
        assert O.inverse_residual(A, Xi, uplo) <= 10 * n * EPS, (i, n)
        ref, _ = O.dpotri("L", F.copy(order="F"))
        Rf = O.herm_from(ref, "L")
        assert np.linalg.norm(O.herm_from(Xi, uplo) - Rf) <= 10 * n * EPS * np.linalg.cond(A) * np.linalg.norm(Rf)
        # the other triangle of the inverse output is left alone
        other = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
        assert np.all(Xi[other] == -5.0), i
        if nrhs:
            Bi = bufB[offB[i]:offB[i] + n * nrhs].reshape((n, nrhs), order="F")
            Xs = X[offB[i]:offB[i] + n * nrhs].reshape((n, nrhs), order="F")
            R = Bi - A @ Xs
            anorm = np.abs(A).sum(axis=1).max()
            # LAPACK DSPOSV's stopping test, column by column
            assert np.all(np.abs(R).max(axis=0) <= np.abs(Xs).max(axis=0) * anorm * EPS * np.sqrt(n) * 1.01), (i, n)
    return outs[0]


@pytest.mark.parametrize("uplo", "LU")
def test_solve_plan_uneven_sizes(gpu, lapack, uplo):
    sizes = [1, 63, 64, 65, 130, 0, 257, 700, 2, 128]
    alphas = [1.0, 1e-2, 1e-4, 1.0, 1e-3, 1.0, 1e-1, 1e-4, 1.0, 1e-2]
    solve_plan_case(gpu, lapack, uplo, sizes, alphas, nrhs=4, seed=1)


@pytest.mark.parametrize("nrhs", [0, 1, 7])
def test_solve_plan_rhs_counts(gpu, lapack, nrhs):
    """nrhs = 0 (inverse + logdet only), 1, and 7 (two groups of right-hand sides)."""
    sizes = [100, 33, 300, 64]
    solve_plan_case(gpu, lapack, "L", sizes, [1e-3, 1.0, 1e-2, 1e-4], nrhs=nrhs, seed=2 + nrhs)


def test_solve_plan_replay_and_failures(gpu, lapack):
    """Graph replays give identical results; a matrix that is not positive
    definite reports its first failing column (tiles 0 and 2) and NaN logdet,
    its neighbours are solved."""
    sizes = [200, 150, 90, 400, 70]
    bad = {1: 5, 3: 140}
    solve_plan_case(gpu, lapack, "L", sizes, [1.0, 1.0, 1e-2, 1.0, 1e-3], nrhs=4, seed=5, bad=bad, reps=3)


def test_solve_plan_many_small(gpu, lapack):
    """More matrices than persistent workgroups: each workgroup runs many tasks."""
    sizes = [int(v) for v in np.random.default_rng(11).integers(1, 200, size=700)]
    alphas = list(np.exp(np.random.default_rng(12).uniform(np.log(1e-4), 0.0, size=700)))
    solve_plan_case(gpu, lapack, "L", sizes, alphas, nrhs=2, seed=13)
