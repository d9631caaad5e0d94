#!/usr/bin/env python3
"""Benchmark: blocked Cholesky factorisation on one MI355X.

A step is one cuDpotrf (lower, FP64) of an n = 16384 SPD matrix
A = G G^T / n + 0.1 I (G i.i.d. N(0,1), seeded).  The factorisation is in
place, so every step (warm-up and timed) gets its own copy of A, made in HBM
before the timed region.  `value` is POTRF TFLOP/s counted as n^3/3 over the
wall time of the K timed steps.

The JSON line also carries
  * roofline: the dominant kernel -- the trailing-matrix SYRK launched by the
    factorisation on the library's main stream -- timed live during the timed
    steps with HIP event pairs recorded on that stream around each launch
    (cuLAPACKSetTiming); achieved = sum of algorithmic SYRK flops
    n_t (n_t + 1) k over the sum of launch durations, against the FP64
    matrix-core peak; traffic = HBM bytes per launch measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (read from the committed summary under
    profiles/ when present);
  * cpu_baseline: the reference's CPU algorithm (the oracle's restatement,
    OpenMP) timed on this host on a bounded sample, in the same unit;
  * extra: dpotri at the same n, zpotrf / zpotri (upper, n = 8192) and
    dlogdet, each checked for accuracy.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "POTRF TFLOP/s (n³/3): FP64 vs BF16/FP32+IR at FP64 accuracy; batched mats/s"
FP64_PEAK = 78.6   # TFLOP/s, MI355X dense FP64 matrix
EPS = 2.220446049250313e-16


def package():
    name = "cuda_cholesky_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "cuda-cholesky_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_spd(torch, n, alpha, seed, cplx=False):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if cplx:
        H = torch.complex(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))
        H /= 2 ** 0.5
        A = H @ H.conj().T
        A /= n
    else:
        G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
        A = G @ G.T
        A /= n
        del G
    A.diagonal().add_(alpha)
    return A


def backward_error(torch, A, F, uplo):
    """||A - L L^H||_F / ||A||_F with F the column-major factor storage."""
    Ft = F.T if not F.is_complex() else F.T
    T = torch.tril(Ft) if uplo == "L" else torch.triu(Ft)
    R = T @ T.conj().T if uplo == "L" else T.conj().T @ T
    At = A.T if not A.is_complex() else A.T
    return (torch.linalg.norm(At - R) / torch.linalg.norm(At)).item()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "syrk_trailing_update_dpotrf_n16384_pmc.txt")


def measured_traffic():
    """HBM bytes per trailing-SYRK launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def roofline_live(records):
    """Roofline entry from the (ms, flops) event records of the trailing SYRKs."""
    total_t = sum(ms for ms, _ in records) * 1e-3
    total_f = sum(fl for _, fl in records)
    achieved = total_f / total_t / 1e12
    return {
        "kernel": "gemm_kernel<double, N, T, 128, 128, tri-grid> (trailing SYRK, lower)",
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": FP64_PEAK,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK, 4),
        "traffic": measured_traffic(),
        "launches": len(records),
        "avg_launch_us": round(total_t / len(records) * 1e6, 2),
        "flops_per_launch_avg": total_f / len(records),
    }


MIXED_VARIANTS = (("S", "I"), ("B", "I"), ("X", "I"), ("H", "I"), ("B", "G"), ("H", "G"))


def mixed_precision(torch, pkg, h, n, nrhs, alphas, variants=MIXED_VARIANTS):
    """SPD solve A X = B (n x n, nrhs right-hand sides): FP64 cuDpotrf +
    cuDpotrs, and the mixed-precision cuDsposvx with FP32, BF16, split-BF16 or
    FP16 trailing-update products (FP64 diagonal blocks, FP64 residuals)
    refined to LAPACK dsposv's stopping test by classical iterative refinement
    or GMRES-IR (BF16, FP16).  Wall time of each whole
    solve, refinement steps, and the normwise backward error
    max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)."""
    out = {"n": n, "nrhs": nrhs}
    for alpha in alphas:
        A = gpu_spd(torch, n, alpha, seed=3)
        g = torch.Generator(device="cuda").manual_seed(4)
        B = torch.randn(nrhs, n, generator=g, device="cuda", dtype=torch.float64)  # column-major n x nrhs
        anorm = A.abs().sum(dim=1).max().item()

        def berr(X):
            R = B - X @ A          # rows: (b - A x)^T, A symmetric
            num = R.abs().max(dim=1).values
            den = anorm * X.abs().max(dim=1).values + B.abs().max(dim=1).values
            return (num / den).max().item()

        rec = {}
        F = A.clone()
        X = B.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, info = pkg.cuDpotrf(h, pkg.Lower, n, pkg.ptr(F), n)
        assert st == 0 and info == 0, (st, info)
        st, info = pkg.cuDpotrs(h, pkg.Lower, n, nrhs, pkg.ptr(F), n, pkg.ptr(X), n)
        assert st == 0 and info == 0, (st, info)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        rec["fp64_ms"] = round(t64 * 1e3, 2)
        rec["fp64_backward_error"] = berr(X)
        del F
        for factor, refine in variants:
            X = torch.zeros_like(B)
            for rep in range(2):       # the first call also grows the workspace
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st, it, used, info = pkg.cuDsposvx(h, pkg.Lower, n, nrhs, pkg.ptr(A), n, pkg.ptr(B), n,
                                                   pkg.ptr(X), n, factor, refine)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                assert st == 0 and info == 0, (st, info)
                if it < 0:             # A now holds its FP64 factor: regenerate
                    A = gpu_spd(torch, n, alpha, seed=3)
            key = {"S": "fp32", "B": "bf16", "X": "bf16x3", "H": "fp16"}[factor] + ("_gmres" if refine == "G" else "")
            rec[key] = {"ms": round(dt * 1e3, 2), "iter": it, "used": used,
                        "backward_error": berr(X), "speedup_vs_fp64": round(t64 / dt, 2)}
        out[f"alpha={alpha:g}"] = rec
        log(f"[bench] mixed n={n} alpha={alpha:g}: {rec}")
        del A, B, X
        torch.cuda.empty_cache()
    return out


def batched(torch, pkg, h, count, nmin, nmax, reps, loop_sample):
    """Batched POTRF + logdet of `count` SPD matrices, n_i ~ U[nmin, nmax],
    A_i = G_i G_i^T / n_i + 0.5 I, packed contiguously (column-major, ld n_i),
    in FP64 and FP32 storage: one graph launch per batch (cuPotrfBatchedRun;
    and POTRF + logdet + POTRI, cuPotriBatchedRun), timed with HIP events on
    the launch stream over `reps` runs (each on a fresh copy of the batch,
    made outside the timed launch).  Compared with a per-matrix loop of
    cuDpotrf + cuDlogdet (FP64) over the first `loop_sample` matrices."""
    import numpy as np
    rng = np.random.default_rng(5)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)[:-1]
    total = int((sizes.astype(np.int64) ** 2).sum())
    A0 = torch.empty(total, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(6)
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        G = torch.randn(len(idx), n, n, generator=g, device="cuda", dtype=torch.float64)
        M = torch.bmm(G, G.transpose(1, 2).contiguous()) / n
        M += 0.5 * torch.eye(n, device="cuda", dtype=torch.float64)
        # one scatter per size (symmetric: storage order irrelevant)
        gather = torch.from_numpy(offs[idx]).cuda()[:, None] + torch.arange(n * n, device="cuda")[None, :]
        A0[gather.reshape(-1)] = M.reshape(-1)
        del G, M, gather
    flops = float((sizes.astype(np.float64) ** 3).sum() / 3)
    tri_bytes = float((sizes.astype(np.float64) * (sizes + 1) / 2).sum())
    out = {"count": count, "n_range": [nmin, nmax], "flops_n3_3": flops}
    s = torch.cuda.Stream()
    for prec, dtype, esz in (("D", torch.float64, 8), ("S", torch.float32, 4)):
        src = A0.to(dtype)
        work = torch.empty_like(src)
        logd = torch.empty(count, dtype=torch.float64, device="cuda")
        info = torch.empty(count, dtype=torch.int32, device="cuda")
        res = {}
        with pkg.BatchedPlan(h, pkg.Lower, prec, sizes.tolist(), offs.tolist()) as plan:
            assert plan.status == 0
            for inverse in (False, True):
                times = []
                fn = plan.run_inverse if inverse else plan.run
                for r in range(reps + 1):
                    work.copy_(src)
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(s)
                    assert fn(pkg.ptr(work), pkg.ptr(logd), pkg.ptr(info), s.cuda_stream) == 0
                    e1.record(s)
                    s.synchronize()
                    if r > 0:            # the first run captures the graph
                        times.append(e0.elapsed_time(e1) * 1e-3)
                assert int(info.abs().sum().item()) == 0
                res[inverse] = sorted(times)[len(times) // 2]
                if prec == "D" and not inverse:
                    ref_log = logd.clone()
        t, ti = res[False], res[True]
        byts = 2 * tri_bytes * esz            # read and write one triangle per matrix
        out["fp64" if prec == "D" else "fp32"] = {
            "ms": round(t * 1e3, 3), "matrices_per_s": round(count / t), "tflops": round(flops / t / 1e12, 3),
            "hbm_gbs_algorithmic": round(byts / t / 1e9, 1),
            "hbm_frac": round(byts / t / 8e12, 4),
            # POTRF + logdet + POTRI in one launch (cuPotriBatchedRun); flops n^3/3 + n^3
            "with_inverse": {"ms": round(ti * 1e3, 3), "matrices_per_s": round(count / ti),
                             "tflops_potrf_plus_potri": round(4 * flops / ti / 1e12, 3)}}
        del src, work
    # per-matrix FP64 loop: cuDpotrf + cuDlogdet on each matrix
    k = min(loop_sample, count)
    work = A0[: int(offs[k - 1] + sizes[k - 1] ** 2)].clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        n = int(sizes[i])
        ptr = pkg.ptr(work) + int(offs[i]) * 8
        st, inf = pkg.cuDpotrf(h, pkg.Lower, n, ptr, n)
        assert st == 0 and inf == 0
        st, ld = pkg.cuDlogdet(h, ptr, n + 1, n)
        assert abs(ld - ref_log[i].item()) <= 1e-9 * max(1.0, abs(ld)), (i, ld, ref_log[i].item())
    t_loop = time.perf_counter() - t0
    out["per_matrix_loop_fp64"] = {"sample": k, "matrices_per_s": round(k / t_loop)}
    out["speedup_fp64_vs_loop"] = round(out["fp64"]["matrices_per_s"] / (k / t_loop), 1)
    log(f"[bench] batched: {out}")
    del A0, work
    torch.cuda.empty_cache()
    return out


def cpu_baseline(torch, n_cpu):
    """The reference's CPU algorithm (oracle restatement) on this host: one
    timed dpotrf of an n_cpu-order SPD matrix of the same form."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = O.threads()
    A = np.asfortranarray(gpu_spd(torch, n_cpu, 0.1, seed=7).cpu().numpy())
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    _, info = O.dpotrf("L", A)
    dt = time.perf_counter() - t0
    assert info == 0, info
    tflops = n_cpu ** 3 / 3 / dt / 1e12
    return {
        "value": round(tflops, 6),
        "unit": "TFLOP/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": (f"oracle dpotrf (the reference CPU algorithm restated: nb=32 blocked, OpenMP "
                   f"column loops, gcc -O2) lower n={n_cpu} FP64, A = G G^T/n + 0.1 I, one run "
                   f"of {dt:.1f} s on {threads} OpenMP threads"),
        "seconds": round(dt, 2),
    }


def blas3(torch, pkg, n):
    """Standalone level-3 BLAS rates (best of 3 timed calls after one
    warm-up, data in HBM): DGEMM / DSYRK / DTRSM / DTRMM at n x n x n and
    SGEMM / SSYRK (FP32 matrix cores) at n, ZGEMM / ZHERK at n / 2, TFLOP/s by
    the usual counts (GEMM 2n^3, SYRK / TRSM / TRMM n^3; complex 4x; real
    flops for complex)."""
    bh = pkg.BLASHandle()
    out = {"n": n}
    g = torch.Generator(device="cuda").manual_seed(9)
    A = torch.randn(n * n, generator=g, device="cuda", dtype=torch.float64)
    B = torch.randn(n * n, generator=g, device="cuda", dtype=torch.float64)
    C = torch.zeros(n * n, device="cuda", dtype=torch.float64)
    T = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64).tril_() / n
    T.diagonal().add_(1.0)
    T = T.T.contiguous().reshape(-1)          # column-major lower triangle, well conditioned
    P = pkg.ptr

    def best(fn):
        fn()
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            assert st == 0, st
        return min(ts)

    d = best(lambda: pkg.cuDgemm2(bh, pkg.NoTrans, pkg.NoTrans, n, n, n, 1.0, P(A), n, P(B), n, 0.0, P(C), n,
                                  P(C), n))
    out["dgemm_tflops"] = round(2 * n ** 3 / d / 1e12, 2)
    d = best(lambda: pkg.cuDsyrk(bh, pkg.Lower, pkg.NoTrans, n, n, 1.0, P(A), n, 0.0, P(C), n))
    out["dsyrk_tflops"] = round(n ** 3 / d / 1e12, 2)
    d = best(lambda: pkg.cuDtrmm2(bh, pkg.Left, pkg.Lower, pkg.NoTrans, pkg.NonUnit, n, n, 1.0, P(T), n, P(B),
                                  n, P(C), n))
    out["dtrmm_tflops"] = round(n ** 3 / d / 1e12, 2)
    Bw = B.clone()
    d = best(lambda: (Bw.copy_(B), pkg.cuDtrsm(bh, pkg.Left, pkg.Lower, pkg.NoTrans, pkg.NonUnit, n, n, 1.0,
                                               P(T), n, P(Bw), n))[1])
    out["dtrsm_tflops_incl_copy"] = round(n ** 3 / d / 1e12, 2)
    del A, B, C, T, Bw
    As = torch.randn(n * n, generator=g, device="cuda", dtype=torch.float32)
    Bs = torch.randn(n * n, generator=g, device="cuda", dtype=torch.float32)
    Cs = torch.zeros(n * n, device="cuda", dtype=torch.float32)
    d = best(lambda: pkg.cuSgemm2(bh, pkg.NoTrans, pkg.NoTrans, n, n, n, 1.0, P(As), n, P(Bs), n, 0.0, P(Cs), n,
                                  P(Cs), n))
    out["sgemm_tflops"] = round(2 * n ** 3 / d / 1e12, 2)
    d = best(lambda: pkg.cuSsyrk(bh, pkg.Lower, pkg.NoTrans, n, n, 1.0, P(As), n, 0.0, P(Cs), n))
    out["ssyrk_tflops"] = round(n ** 3 / d / 1e12, 2)
    del As, Bs, Cs
    nz = n // 2
    Z = torch.randn(nz * nz * 2, generator=g, device="cuda", dtype=torch.float64)
    W = torch.randn(nz * nz * 2, generator=g, device="cuda", dtype=torch.float64)
    Zc = torch.zeros(nz * nz * 2, device="cuda", dtype=torch.float64)
    d = best(lambda: pkg.cuZgemm2(bh, pkg.NoTrans, pkg.NoTrans, nz, nz, nz, 1.0, P(Z), nz, P(W), nz, 0.0,
                                  P(Zc), nz, P(Zc), nz))
    out["zgemm_tflops_real"] = round(8 * nz ** 3 / d / 1e12, 2)
    d = best(lambda: pkg.cuZherk(bh, pkg.Lower, pkg.NoTrans, nz, nz, 1.0, P(Z), nz, 0.0, P(Zc), nz))
    out["zherk_tflops_real"] = round(4 * nz ** 3 / d / 1e12, 2)
    del Z, W, Zc
    torch.cuda.empty_cache()
    bh.close()
    return out


def host_api(torch, pkg, A0, n, reps=2):
    """cuMultiGPUDpotrf on a host copy of A: the PCIe-inclusive rate of the
    host-pointer API (upload, factor, download), after one call that sizes the
    pinned staging and device buffers."""
    import numpy as np
    host = A0.cpu().numpy().T                  # symmetric: any storage order is A
    out = {}
    with pkg.MultiGPU((torch.cuda.current_device(),)) as ctx, pkg.MultiGPULAPACKHandle(ctx) as hl:
        for rep in range(reps):
            M = np.array(host, order="F")
            t = time.perf_counter()
            st, info = pkg.cuMultiGPUDpotrf(hl, pkg.Lower, n, M, n)
            dt = time.perf_counter() - t
            assert st == 0 and info == 0, (st, info)
        L = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
        out["ms"] = round(dt * 1e3, 2)
        out["tflops_pcie_inclusive"] = round(n ** 3 / 3 / dt / 1e12, 3)
        out["bytes_moved"] = 2 * n * n * 8
        out["backward_error"] = backward_error(torch, A0, L, "L")
    return out


def batched_mixed(torch, pkg, h, count, nmin, nmax, nrhs, reps=3):
    """Batched solve + logdet + inverse of `count` FP64 SPD matrices
    A_i = G_i G_i^T / n_i + alpha_i I with n_i log-uniform in [nmin, nmax] and
    alpha_i log-uniform in [1e-4, 1], `nrhs` N(0,1) right-hand sides each:
    one plan (cuDsposvxBatchedPlanCreate), one graph launch per batch
    (cuDsposvxBatchedRun), timed with HIP events on the launch stream (median
    of `reps` runs after the capturing run).  Per matrix: the normwise
    backward error of X (max over right-hand sides,
    ||b - A x||_inf / (||A||_inf ||x||_inf)), the inverse residual
    ||I - A A^-1||_F / (||A||_F ||A^-1||_F) and the logdet against the FP64
    Cholesky factor of the same matrix (torch.linalg.cholesky).  The same
    work with plain FP64 routines per matrix (cuDpotrf + cuDpotrs + cuDlogdet
    + cuDpotri on a copy) is timed beside it."""
    import numpy as np
    rng = np.random.default_rng(9)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(np.int64)
    alphas = np.exp(rng.uniform(np.log(1e-4), 0.0, size=count))
    offA = np.zeros(count, dtype=np.int64)
    offA[1:] = np.cumsum(sizes ** 2)[:-1]
    offB = np.zeros(count, dtype=np.int64)
    offB[1:] = np.cumsum(sizes * nrhs)[:-1]
    A = torch.empty(int((sizes ** 2).sum()), dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(10)
    for i, n in enumerate(sizes.tolist()):
        G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
        M = G @ G.T / n
        M.diagonal().add_(float(alphas[i]))
        A[offA[i]:offA[i] + n * n] = M.reshape(-1)
    B = torch.randn(int((sizes * nrhs).sum()), generator=g, device="cuda", dtype=torch.float64)
    X = torch.zeros_like(B)
    Ainv = torch.empty_like(A)
    it = torch.zeros(count, dtype=torch.int32, device="cuda")
    ld = torch.zeros(count, dtype=torch.float64, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    t_plan = time.perf_counter()
    plan = pkg.BatchedSolvePlan(h, pkg.Lower, sizes.tolist(), offA.tolist(), nrhs, offB.tolist())
    t_plan = time.perf_counter() - t_plan
    assert plan.status == 0, plan.status
    times = []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        assert plan.run(pkg.ptr(A), pkg.ptr(B), pkg.ptr(X), pkg.ptr(Ainv), pkg.ptr(it), pkg.ptr(ld),
                        pkg.ptr(info), s.cuda_stream) == 0
        e1.record(s)
        s.synchronize()
        if r > 0:                      # the first run captures the graph
            times.append(e0.elapsed_time(e1) * 1e-3)
    plan.close()
    t = sorted(times)[len(times) // 2]
    info_h, it_h, ld_h = info.cpu().numpy(), it.cpu().numpy(), ld.cpu().numpy()
    assert np.all(info_h == 0), info_h[:8]
    berr, ires, lderr = [], [], []
    for i, n in enumerate(sizes.tolist()):
        M = A[offA[i]:offA[i] + n * n].view(n, n)           # symmetric; A is never written
        Xi = X[offB[i]:offB[i] + n * nrhs].view(nrhs, n).T
        Bi = B[offB[i]:offB[i] + n * nrhs].view(nrhs, n).T
        R = Bi - M @ Xi
        berr.append((R.abs().amax(0) / (M.abs().sum(1).max() * Xi.abs().amax(0))).max().item())
        T = Ainv[offA[i]:offA[i] + n * n].view(n, n).T      # column-major lower triangle
        S = torch.tril(T) + torch.tril(T, -1).T
        ires.append((torch.linalg.matrix_norm(torch.eye(n, device="cuda", dtype=torch.float64) - M @ S)
                     / (torch.linalg.matrix_norm(M) * torch.linalg.matrix_norm(S))).item())
        ref = 2.0 * torch.log(torch.diagonal(torch.linalg.cholesky(M))).sum().item()
        lderr.append(abs(float(ld_h[i]) - ref) / max(1.0, abs(ref)))
    # plain FP64 per matrix: factor, solve, logdet, inverse
    A64 = A.clone()
    X64 = B.clone()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i, n in enumerate(sizes.tolist()):
        Ai = pkg.ptr(A64) + int(offA[i]) * 8
        st1, inf1 = pkg.cuDpotrf(h, pkg.Lower, n, Ai, n)
        st2, inf2 = pkg.cuDpotrs(h, pkg.Lower, n, nrhs, Ai, n, pkg.ptr(X64) + int(offB[i]) * 8, n)
        st3, _ = pkg.cuDlogdet(h, Ai, n + 1, n)
        st4, inf4 = pkg.cuDpotri(h, pkg.Lower, n, Ai, n)
        assert st1 == st2 == st3 == st4 == 0 and inf1 == inf2 == inf4 == 0
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t1
    n3 = float((sizes.astype(np.float64) ** 3).sum())
    q = lambda v: {"max": float(np.max(v)), "median": float(np.median(v))}
    # algorithmic work of the task graph: factor n^3/3, L^-1 n^3/3, L^-T L^-1 n^3/3 (FP64 matrix cores)
    return {"count": count, "n_range": [nmin, nmax], "alpha_range": [1e-4, 1.0], "nrhs": nrhs,
            "ms": round(t * 1e3, 3), "matrices_per_s": round(count / t, 1),
            "plan_create_ms": round(t_plan * 1e3, 1),
            "fp64_per_matrix_ms": round(t64 * 1e3, 2), "speedup_vs_fp64_per_matrix": round(t64 / t, 2),
            "flops_potrf_trtri_lauum": n3, "tflops": round(n3 / t / 1e12, 3),
            "mfma_frac": round(n3 / t / 1e12 / FP64_PEAK, 4),
            "iter_histogram": {str(k): int(v) for k, v in zip(*np.unique(it_h, return_counts=True))},
            "backward_error": q(berr), "inverse_residual": q(ires), "logdet_rel_error": q(lderr)}


def pipeline_1024(torch, pkg, h, reps=5):
    """The n = 1024 workload: dpotrf (lower) -> dlogdet -> dpotri of
    A = G G^T / n + I (the oracle's generator, seed 1), on the GPU (best of
    `reps`, wall time per pipeline) and with the oracle (the reference's CPU
    algorithm) on this host; both results checked."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = 1024
    A = O.spd(1, n, 1.0)
    dA0 = pkg.to_device(A, "cuda")
    best = None
    for r in range(reps + 1):
        dA = dA0.clone()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st1, i1 = pkg.cuDpotrf(h, pkg.Lower, n, pkg.ptr(dA), n)
        st2, lg = pkg.cuDlogdet(h, pkg.ptr(dA), n + 1, n)
        st3, i3 = pkg.cuDpotri(h, pkg.Lower, n, pkg.ptr(dA), n)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        assert st1 == st2 == st3 == 0 and i1 == i3 == 0
        if r > 0:
            best = dt if best is None else min(best, dt)
    X = pkg.from_device(dA)
    t = time.perf_counter()
    F, info = O.dpotrf("L", A.copy(order="F"))
    lref = O.dlogdet(np.ascontiguousarray(F.reshape(-1, order="F")), n + 1, n)
    Xref, info2 = O.dpotri("L", F.copy(order="F"))
    tc = time.perf_counter() - t
    assert info == info2 == 0
    return {"n": n, "gpu_ms": round(best * 1e3, 3), "cpu_ms": round(tc * 1e3, 2), "cpu_threads": O.threads(),
            "speedup": round(tc / best, 1), "gpu_tflops_potrf_plus_potri_n3": round((n ** 3 / 3 + n ** 3) / best / 1e12, 3),
            "logdet": lg, "logdet_rel_diff_vs_cpu": abs(lg - lref) / abs(lref),
            "inverse_residual": O.inverse_residual(A, X, "L")}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--n-cpu", type=int, default=12288, help="CPU baseline sample order")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip dpotri / zpotrf / zpotri")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision solve (n = 32768, 16 right-hand sides)")
    ap.add_argument("--n-mixed", type=int, default=32768)
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched factorisation (20000 matrices, n in [8, 256])")
    ap.add_argument("--no-batched-mixed", action="store_true",
                    help="skip the batched mixed-precision solve + logdet + inverse (512 matrices, "
                         "n log-uniform in [128, 2048])")
    ap.add_argument("--no-blas", action="store_true", help="skip the standalone level-3 BLAS rates")
    ap.add_argument("--mixed-alphas", default="1,1e-3,1e-6", help="alphas of the mixed-precision leg")
    ap.add_argument("--mixed-variants", default=",".join(f + r for f, r in MIXED_VARIANTS),
                    help="factor+refine pairs of the mixed-precision leg, e.g. HI,SI")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-pointer cuMultiGPUDpotrf leg (PCIe-inclusive)")
    ap.add_argument("--counter-calibration", action="store_true",
                    help="run one elementwise copy of A (n^2 * 8 bytes read and written) before the "
                         "steps, as the known-traffic reference kernel of a rocprofv3 --pmc run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (--gpus {args.gpus} "
              "requested). The factorisation would shard as a 2-D block-cyclic tile layout with a "
              "per-step panel broadcast; see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = package()
    pkg.lib()
    n, lda = args.n, args.n
    log(f"[bench] generating A (n={n}) and {args.warmup + args.steps} copies in HBM")
    A0 = gpu_spd(torch, n, 0.1, seed=1)        # symmetric: storage order is irrelevant
    copies = [A0.clone() for _ in range(args.warmup + args.steps)]
    torch.cuda.synchronize()
    if args.counter_calibration:
        torch.mul(A0, 1.0, out=copies[0])
        torch.cuda.synchronize()
        copies[0].copy_(A0)
        torch.cuda.synchronize()
    h = pkg.LAPACKHandle()

    def step(A):
        st, info = pkg.cuDpotrf(h, pkg.Lower, n, pkg.ptr(A), lda)
        if st != 0 or info != 0:
            raise RuntimeError(f"cuDpotrf failed: status {st} info {info}")

    for i in range(args.warmup):
        step(copies[i])
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    pkg.cuLAPACKSetTiming(h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(copies[args.warmup + i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st, records = pkg.cuLAPACKGetTiming(h)
    pkg.cuLAPACKSetTiming(h, 0)
    assert st == 0 and records, st
    ms = (t1 - t0) * 1e3 / args.steps
    tflops = n ** 3 / 3 / (ms * 1e-3) / 1e12
    A = copies[-1]
    del copies[:-1]
    bwd = backward_error(torch, A0, A, "L")
    log(f"[bench] dpotrf n={n}: {ms:.2f} ms/step, {tflops:.2f} TFLOP/s, backward error {bwd:.2e}")
    roof = roofline_live(records)
    log(f"[bench] trailing syrk: {roof['launches']} launches, avg {roof['avg_launch_us']} us, "
        f"{roof['achieved']} TFLOP/s ({roof['frac'] * 100:.1f}% of FP64 peak)")

    extra = {"dpotrf_backward_error": bwd, "dpotrf_backward_bound": 10 * n * EPS}
    if not args.no_extra:
        # dpotri on the factor: TFLOP/s with n^3 and with the reference's 2n^3/3 + n^2/2 + 5n/6
        Ainv = A.clone()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st, info = pkg.cuDpotri(h, pkg.Lower, n, pkg.ptr(Ainv), lda)
        dt = time.perf_counter() - t
        assert st == 0 and info == 0
        extra["dpotri_ms"] = round(dt * 1e3, 2)
        extra["dpotri_tflops_n3"] = round(n ** 3 / dt / 1e12, 3)
        extra["dpotri_tflops_ref_formula"] = round((2 * n ** 3 / 3 + n ** 2 / 2 + 5 * n / 6) / dt / 1e12, 3)
        st, ld = pkg.cuDlogdet(h, pkg.ptr(A), lda + 1, n)
        extra["dlogdet_of_factor"] = ld
        del Ainv
        log(f"[bench] dpotri n={n}: {dt * 1e3:.1f} ms")
        nz = 8192
        Z0 = gpu_spd(torch, nz, 0.1, seed=2, cplx=True)
        Zc = Z0.T.contiguous()                      # column-major storage of A
        for rep in range(2):
            Z = Zc.clone()
            torch.cuda.synchronize()
            t = time.perf_counter()
            st, info = pkg.cuZpotrf(h, pkg.Upper, nz, pkg.ptr(Z), nz)
            dtz = time.perf_counter() - t
            assert st == 0 and info == 0
        extra["zpotrf_upper_8192_ms"] = round(dtz * 1e3, 2)
        extra["zpotrf_upper_8192_tflops_real"] = round(4 * nz ** 3 / 3 / dtz / 1e12, 3)
        extra["zpotrf_backward_error"] = backward_error(torch, Zc, Z, "U")
        t = time.perf_counter()
        st, info = pkg.cuZpotri(h, pkg.Upper, nz, pkg.ptr(Z), nz)
        dtz = time.perf_counter() - t
        assert st == 0 and info == 0
        extra["zpotri_upper_8192_ms"] = round(dtz * 1e3, 2)
        extra["zpotri_upper_8192_tflops_real_n3"] = round(4 * nz ** 3 / dtz / 1e12, 3)
        del Z, Z0, Zc
        log("[bench] complex done")

    if not args.no_extra and not args.no_blas:
        extra["blas3"] = blas3(torch, pkg, 8192)
        log(f"[bench] blas3: {extra['blas3']}")
    if not args.no_host:
        extra["host_pointer_dpotrf"] = host_api(torch, pkg, A0, n)
        log(f"[bench] cuMultiGPUDpotrf (host arrays) n={n}: {extra['host_pointer_dpotrf']['ms']} ms")

    if not args.no_mixed:
        del A
        torch.cuda.empty_cache()
        alphas = tuple(float(a) for a in args.mixed_alphas.split(","))
        variants = tuple((v[0], v[1]) for v in args.mixed_variants.split(","))
        extra["mixed_precision_solve"] = mixed_precision(torch, pkg, h, args.n_mixed, 16, alphas, variants)

    if not args.no_batched:
        extra["batched_potrf_logdet"] = batched(torch, pkg, h, 20000, 8, 256, 5, 2000)

    if not args.no_batched_mixed:
        extra["batched_mixed_solve_inverse"] = batched_mixed(torch, pkg, h, 512, 128, 2048, 4)
        log(f"[bench] batched solve+inverse: {extra['batched_mixed_solve_inverse']}")

    if not args.no_extra:
        extra["pipeline_n1024"] = pipeline_1024(torch, pkg, h)
        log(f"[bench] n=1024 pipeline: {extra['pipeline_n1024']}")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline n={args.n_cpu}")
        cpu = cpu_baseline(torch, args.n_cpu)
        log(f"[bench] CPU {cpu['value'] * 1e3:.2f} GFLOP/s on {cpu['cores']} threads")

    h.close()
    out = {
        "metric": METRIC,
        "value": round(tflops, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"dpotrf lower FP64 n={n} lda={lda}, A = G G^T/n + 0.1 I (seeded)",
                   "n": n, "uplo": "L", "step": "cuDpotrf on a fresh HBM-resident copy of A"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
