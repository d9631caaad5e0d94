#!/usr/bin/env python3
"""Benchmark: blocked Cholesky factorisation on one MI355X.

A step is one cuDpotrf (lower, FP64) of an n = 16384 SPD matrix
A = G G^T / n + 0.1 I (G i.i.d. N(0,1), seeded), preceded by the
device-to-device restore of A from a pristine copy already in HBM (the
factorisation is in place).  `value` is POTRF TFLOP/s counted as n^3/3 over
the wall time of the K timed steps.

The JSON line also carries
  * roofline: the dominant kernel -- the trailing-matrix SYRK of the
    factorisation -- replayed through cuDsyrk at the exact shapes the
    factorisation launches it with (n_t = n - j - 256, k = 256), each launch
    bracketed by HIP events on its own stream; achieved = sum of algorithmic
    SYRK flops n_t (n_t + 1) k / sum of launch durations, against the FP64
    matrix-core peak;
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
NB_OUTER = 256     # outer panel width of the library's dpotrf (trailing SYRK inner dimension)


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


def roofline_syrk(torch, pkg, A, n, lda):
    """Replay the factorisation's trailing SYRKs and time each launch."""
    blas = pkg.BLASHandle()
    s = torch.cuda.Stream()
    base = pkg.ptr(A)
    starts, ends, flops = [], [], []
    with torch.cuda.stream(s):
        for j in range(0, n, NB_OUTER):
            t = n - j - NB_OUTER
            if t <= 0:
                break
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            P = base + ((j + NB_OUTER) + j * lda) * 8
            C = base + ((j + NB_OUTER) + (j + NB_OUTER) * lda) * 8
            e0.record(s)
            st = pkg.cuDsyrk(blas, pkg.Lower, pkg.NoTrans, t, NB_OUTER, -1e-30, P, lda, 1.0, C, lda,
                             s.cuda_stream)
            e1.record(s)
            assert st == 0, st
            starts.append(e0)
            ends.append(e1)
            flops.append(t * (t + 1) * NB_OUTER)
    s.synchronize()
    blas.close()
    times = [a.elapsed_time(b) * 1e-3 for a, b in zip(starts, ends)]
    total_t, total_f = sum(times), sum(flops)
    achieved = total_f / total_t / 1e12
    return {
        "kernel": "gemm_kernel<double, N, T, 128, 128> (trailing SYRK, lower, triangle grid)",
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": FP64_PEAK,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK, 4),
        "traffic": None,
        "launches": len(times),
        "avg_launch_us": round(total_t / len(times) * 1e6, 2),
        "flops_per_launch_avg": total_f / len(times),
    }


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
        "kind": "port",
        "sample": (f"oracle dpotrf (the reference CPU algorithm restated: nb=32 blocked, OpenMP "
                   f"column loops, gcc -O2) lower n={n_cpu} FP64, A = G G^T/n + 0.1 I, one run "
                   f"of {dt:.1f} s on {threads} OpenMP threads"),
        "seconds": round(dt, 2),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--n-cpu", type=int, default=12288, help="CPU baseline sample order")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip dpotri / zpotrf / zpotri")
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
    log(f"[bench] generating A (n={n}) on the GPU")
    A0 = gpu_spd(torch, n, 0.1, seed=1)        # symmetric: storage order is irrelevant
    A = torch.empty_like(A0)
    torch.cuda.synchronize()
    h = pkg.LAPACKHandle()

    def step():
        A.copy_(A0)
        torch.cuda.current_stream().synchronize()
        st, info = pkg.cuDpotrf(h, pkg.Lower, n, pkg.ptr(A), lda)
        if st != 0 or info != 0:
            raise RuntimeError(f"cuDpotrf failed: status {st} info {info}")

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    tflops = n ** 3 / 3 / (ms * 1e-3) / 1e12
    bwd = backward_error(torch, A0, A, "L")
    log(f"[bench] dpotrf n={n}: {ms:.2f} ms/step, {tflops:.2f} TFLOP/s, backward error {bwd:.2e}")

    log("[bench] roofline replay of the trailing SYRK")
    roof = roofline_syrk(torch, pkg, A0.clone(), n, lda)
    log(f"[bench] syrk achieved {roof['achieved']} TFLOP/s ({roof['frac'] * 100:.1f}% of FP64 peak)")

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
                   "n": n, "uplo": "L", "step": "restore A (D2D copy) + cuDpotrf"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
