"""GPU parity of the no-allocation batched factorisation (cuPotrfBatchedNoAlloc)
against the plan path (cuPotrfBatchedPlanCreate + cuPotrfBatchedRun) on the
same inputs: the factors, log-determinants and info words must be identical
(the same kernels factor every matrix; only the entry point differs), on two
calls in a row, the second with a smaller batch so that the shared scratch
still holds the first call's longer order.  Also: a size out of range (info
-1, logdet NaN, matrix untouched), the count limit, and a sample against the
oracle.
"""
import math

import numpy as np
import pytest
import torch

from conftest import EPS
import oracle as O

pytestmark = pytest.mark.gpu

MAX_COUNT = 65536


def packed(sizes, seed, dtype, bad=()):
    """Packed column-major A_i = G_i G_i^T / n_i + 0.5 I on the GPU (matrices
    listed in `bad` get a non-positive pivot planted at column n_i / 2)."""
    sizes = np.asarray(sizes, dtype=np.int64)
    offs = np.zeros(len(sizes), dtype=np.int64)
    offs[1:] = np.cumsum(sizes ** 2)[:-1]
    buf = torch.zeros(int((sizes ** 2).sum()) + 1, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    for n in np.unique(sizes):
        if n == 0 or n > 256:
            continue
        idx = np.nonzero(sizes == n)[0]
        G = torch.randn(len(idx), n, n, generator=g, device="cuda", dtype=torch.float64)
        M = torch.bmm(G, G.transpose(1, 2).contiguous()) / n
        M = (M + M.transpose(1, 2)) / 2
        M += 0.5 * torch.eye(int(n), device="cuda", dtype=torch.float64)
        gather = torch.from_numpy(offs[idx]).cuda()[:, None] + torch.arange(int(n) * int(n), device="cuda")[None, :]
        buf[gather.reshape(-1)] = M.reshape(-1)
    for i in bad:
        n, k = int(sizes[i]), int(sizes[i]) // 2
        buf[int(offs[i]) + k * n + k] = -1.0
    return buf.to(dtype), offs


def run_plan(gpu, lapack, uplo, prec, sizes, offs, src):
    work = src.clone()
    ld = torch.full((len(sizes),), -3.0, dtype=torch.float64, device="cuda")
    info = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    with gpu.BatchedPlan(lapack, uplo, prec, [int(v) for v in sizes], [int(v) for v in offs]) as plan:
        assert plan.status == 0
        assert plan.run(gpu.ptr(work), gpu.ptr(ld), gpu.ptr(info)) == 0
        torch.cuda.synchronize()
    return work, ld, info


def run_noalloc(gpu, lapack, uplo, prec, sizes, offs, src, stream=None):
    work = src.clone()
    ld = torch.full((len(sizes),), -3.0, dtype=torch.float64, device="cuda")
    info = torch.full((len(sizes),), -7, dtype=torch.int32, device="cuda")
    dn = torch.tensor(np.asarray(sizes, dtype=np.int32), device="cuda")
    doff = torch.tensor(np.asarray(offs, dtype=np.int64), device="cuda")
    torch.cuda.synchronize()
    st = gpu.lib().cuPotrfBatchedNoAlloc(gpu._hv(lapack), uplo, prec.encode(), len(sizes), gpu.ptr(dn),
                                         gpu.ptr(doff), gpu.ptr(work), gpu.ptr(ld), gpu.ptr(info), stream)
    assert st == 0
    torch.cuda.synchronize()
    return work, ld, info


def same(a, b):
    return torch.equal(torch.nan_to_num(a, nan=12345.0), torch.nan_to_num(b, nan=12345.0))


@pytest.mark.parametrize("uplo,prec", [("L", "D"), ("U", "S")])
def test_noalloc_matches_plan_on_two_calls(gpu, lapack, uplo, prec):
    dtype = torch.float64 if prec == "D" else torch.float32
    u = gpu.Lower if uplo == "L" else gpu.Upper
    rng = np.random.default_rng(31)
    for call, (count, seed) in enumerate(((3000, 41), (700, 42))):   # the second call is smaller
        sizes = rng.integers(0, 257, size=count)
        sizes[:6] = [0, 1, 256, 64, 65, 128]
        bad = [10, 20, count - 1] if sizes[10] > 1 and sizes[20] > 1 and sizes[count - 1] > 1 else [10]
        src, offs = packed(sizes, seed, dtype, bad=[i for i in bad if sizes[i] > 1])
        pw, pld, pinfo = run_plan(gpu, lapack, u, prec, sizes, offs, src)
        nw, nld, ninfo = run_noalloc(gpu, lapack, u, prec, sizes, offs, src)
        assert torch.equal(pinfo, ninfo), (call, torch.nonzero(pinfo != ninfo)[:5])
        assert same(pld, nld), (call, torch.nonzero(pld != nld)[:5])
        assert torch.equal(pw, nw), (call, torch.nonzero(pw != nw)[:5])
        planted = [i for i in bad if sizes[i] > 1]
        assert all(int(ninfo[i]) == sizes[i] // 2 + 1 for i in planted), [int(ninfo[i]) for i in planted]
        assert int((ninfo != 0).sum()) == len(planted)


def test_noalloc_factor_against_oracle(gpu, lapack):
    sizes = np.array([0, 1, 7, 16, 17, 63, 64, 65, 127, 128, 129, 200, 255, 256])
    src, offs = packed(sizes, 7, torch.float64)
    a0 = src.cpu().numpy()
    w, ld, info = run_noalloc(gpu, lapack, gpu.Lower, "D", sizes, offs, src)
    assert int(info.abs().sum()) == 0
    out, lg = w.cpu().numpy(), ld.cpu().numpy()
    for i, n in enumerate(sizes):
        n, o = int(n), int(offs[i])
        if n == 0:
            assert lg[i] == 0.0
            continue
        A = a0[o:o + n * n].reshape((n, n), order="F")
        F = out[o:o + n * n].reshape((n, n), order="F")
        assert O.potrf_backward_error(A, F, "L") <= 10 * n * EPS, (i, n)
        ref, rinfo = O.dpotrf("L", np.asfortranarray(A.copy()))
        assert rinfo == 0
        want = 2.0 * math.fsum(np.log(np.diag(ref)))
        assert abs(lg[i] - want) <= 10 * n * EPS * max(1.0, abs(want)), (i, n)
        d = np.linalg.norm(np.tril(F) - np.tril(ref)) / np.linalg.norm(np.tril(ref))
        assert d <= 100 * n * EPS, (i, n, d)
        np.testing.assert_array_equal(np.triu(F, 1), np.triu(A, 1))   # the other triangle untouched


def test_noalloc_size_out_of_range(gpu, lapack):
    sizes = np.array([32, 300, 48, -1])
    safe = np.array([32, 0, 48, 0])
    src, offs = packed(safe, 3, torch.float64)
    offs = np.array([0, 32 * 32, 32 * 32, 32 * 32 + 48 * 48])
    w, ld, info = run_noalloc(gpu, lapack, gpu.Lower, "D", sizes, offs, src)
    assert info.tolist() == [0, -1, 0, -1]
    assert math.isnan(float(ld[1])) and math.isnan(float(ld[3]))
    ref, ldr, infor = run_plan(gpu, lapack, gpu.Lower, "D", safe, offs, src)
    assert torch.equal(w, ref)                      # the out-of-range entries touched nothing
    assert float(ld[0]) == float(ldr[0]) and float(ld[2]) == float(ldr[2])


def test_noalloc_rejects_count_above_limit(gpu, lapack):
    st = gpu.lib().cuPotrfBatchedNoAlloc(gpu._hv(lapack), gpu.Lower, b"D", MAX_COUNT + 1, 8, 8, 8, 8, 8, None)
    assert st == 1                                  # CUDA_ERROR_INVALID_VALUE, nothing launched
    st = gpu.lib().cuPotrfBatchedNoAlloc(gpu._hv(lapack), gpu.Lower, b"Q", 4, 8, 8, 8, 8, 8, None)
    assert st == 1
    st = gpu.lib().cuPotrfBatchedNoAlloc(gpu._hv(lapack), gpu.Lower, b"D", 0, 0, 0, 0, 0, 0, None)
    assert st == 0                                  # an empty batch is a no-op
