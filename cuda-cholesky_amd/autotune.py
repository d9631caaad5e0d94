#!/usr/bin/env python3
"""Measure the tunable schedule parameters of the library on this GPU.

* Outer block size of the blocked factorisations: for every precision (D:
  cuDpotrf, Z: cuZpotrf, S: cuSpotrf, B: the BF16 factor of cuDsposvx) and
  every order in the sweep, each candidate block size is set with
  cuLAPACKSetBlockSize and the factorisation timed (best of a few runs, fresh
  input each time).  The winner for each order covers all n up to it.
* Persistent schedule of the batched factorisation + inverse
  (cuPotriBatchedRun): workgroups per CU, from 1 up to the kernel's
  occupancy, for FP64 / FP32 storage, timed on a batch of uneven sizes
  (median of a few graph launches, HIP events on the launch stream).  The
  factor-only runs (cuPotrfBatchedRun) take the register-resident kernels,
  whose grid is their occupancy.
* Tile choice of each GEMM kernel family: a launch takes the family's larger
  tile when it has at least `min_tiles` of them.  Each candidate threshold is
  set through $CUCHOLESKY_TILE_MIN_<family> (read by the library on every
  launch) and the routines that use the family are timed at several orders;
  the threshold with the least total time wins.  Families: D (double tile
  kernel, 128 vs 64: cuDpotrf, cuDpotri), Z (complex, 64 vs 32: cuZpotrf),
  S (FP32 kernel, 128 vs 64: cuSpotrf), H and W (16-bit kernel, 128 vs 64 and
  256 vs 128: the FP16-product factor of cuDsposvx).  A threshold above any
  launch's tile count means "always the smaller tile".
* Workgroup cap of the factorisation's late trailing updates: a trailing
  update of order t <= t_max runs on at most `workgroups` workgroups, each
  taking several tiles, so that the look-ahead chain (diagonal blocks and
  panel products on the other stream) finds free room on the CUs.  Each
  candidate (t_max, workgroups) is set through
  $CUCHOLESKY_TRAILING_CAP_<precision> and the factorisation timed
  (median of a few runs, candidates interleaved, since the differences are
  about 1-2 %).

The table goes to tuning.txt beside the library (or --out), where the library
reads it at first use:

    potrf_nb <precision> <n_max> <nb>
    batched_per_cu <precision> <inverse 0|1> <workgroups per CU>
    gemm_tile <family> <min_tiles>
    trailing_cap <precision> <t_max> <workgroups>

    python cuda-cholesky_amd/autotune.py [--quick] [--only nb|batched|tiles|cap] [--out PATH]
"""
import argparse
import importlib.util
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))

SWEEP = {
    # precision: (orders, candidate block sizes)
    "D": ([2048, 4096, 8192, 16384], [256, 384, 512, 768]),
    "Z": ([2048, 4096, 8192], [128, 192, 256, 384]),
    "S": ([4096, 8192, 16384, 32768], [384, 512, 768, 1024]),
    "B": ([4096, 8192, 16384, 32768], [384, 512, 768, 1024]),
}
QUICK = {
    "D": ([4096, 16384], [256, 384, 512]),
    "Z": ([4096], [128, 256]),
    "S": ([8192], [384, 768]),
    "B": ([8192], [384, 768]),
}


def package():
    name = "cuda_cholesky_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def spd(torch, n, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if dtype == torch.complex128:
        H = torch.complex(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))
        A = H @ H.conj().T / n
    else:
        G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
        A = G @ G.T / n
    A.diagonal().add_(1.0)
    return A.to(dtype) if dtype != torch.complex128 else A


def time_one(torch, pkg, h, prec, n, A0, B, reps):
    best = float("inf")
    for _ in range(reps):
        A = A0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if prec == "D":
            st, info = pkg.cuDpotrf(h, pkg.Lower, n, pkg.ptr(A), n)
        elif prec == "Z":
            st, info = pkg.cuZpotrf(h, pkg.Upper, n, pkg.ptr(A), n)
        elif prec == "S":
            st, info = pkg.cuSpotrf(h, pkg.Lower, n, pkg.ptr(A), n)
        else:
            X = torch.zeros_like(B)
            st, _, used, info = pkg.cuDsposvx(h, pkg.Lower, n, B.shape[0], pkg.ptr(A), n, pkg.ptr(B), n,
                                              pkg.ptr(X), n, "B", "I")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if st != 0 or info != 0:
            raise RuntimeError(f"{prec} n={n}: status {st} info {info}")
        best = min(best, dt)
    return best


def tune_batched(torch, pkg, h, count, reps, per_cu_max=6):
    """batched_per_cu lines: a batch of `count` SPD matrices, n_i ~ U[8, 256]."""
    import numpy as np
    rng = np.random.default_rng(5)
    sizes = rng.integers(8, 257, size=count)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)[:-1]
    A0 = torch.empty(int((sizes.astype(np.int64) ** 2).sum()), dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(6)
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        G = torch.randn(len(idx), n, n, generator=g, device="cuda", dtype=torch.float64)
        M = torch.bmm(G, G.transpose(1, 2).contiguous()) / n + 0.5 * torch.eye(n, device="cuda",
                                                                                 dtype=torch.float64)
        for j, i in enumerate(idx):
            A0[offs[i]:offs[i] + n * n] = M[j].reshape(-1)
        del G, M
    s = torch.cuda.Stream()
    logd = torch.empty(count, dtype=torch.float64, device="cuda")
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    lines = []
    for prec, dtype in (("D", torch.float64), ("S", torch.float32)):
        src = A0.to(dtype)
        work = torch.empty_like(src)
        for inverse in (True,):
            times = {}
            for per_cu in range(1, per_cu_max + 1):
                os.environ["CUCHOLESKY_BATCHED_PER_CU"] = str(per_cu)
                with pkg.BatchedPlan(h, pkg.Lower, prec, sizes.tolist(), offs.tolist()) as plan:
                    assert plan.status == 0
                    fn = plan.run_inverse if inverse else plan.run
                    ts = []
                    for r in range(reps + 1):
                        work.copy_(src)
                        torch.cuda.synchronize()
                        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        e0.record(s)
                        assert fn(pkg.ptr(work), pkg.ptr(logd), pkg.ptr(info), s.cuda_stream) == 0
                        e1.record(s)
                        s.synchronize()
                        if r > 0:  # the first launch captures the graph
                            ts.append(e0.elapsed_time(e1))
                    assert int(info.abs().sum().item()) == 0
                times[per_cu] = sorted(ts)[len(ts) // 2]
            os.environ.pop("CUCHOLESKY_BATCHED_PER_CU", None)
            best = min(times, key=times.get)
            detail = " ".join(f"{k}:{v:.3f}ms" for k, v in times.items())
            lines.append(f"batched_per_cu {prec} {int(inverse)} {best}   # {count} matrices n in [8, 256] {detail}")
            print(lines[-1], flush=True)
        del src, work
    return lines


# family: (routine, orders, candidate thresholds); 1 << 30: never the larger tile
TILE_SWEEP = {
    "D": ("D", [4096, 8192, 16384], [120, 240, 480, 960, 1920, 1 << 30]),
    "Z": ("Z", [2048, 4096, 8192], [120, 240, 480, 960, 1440, 1920, 1 << 30]),
    "S": ("S", [4096, 8192, 16384], [60, 120, 240, 480, 1 << 30]),
    "H": ("H", [8192, 16384], [120, 240, 480, 1 << 30]),
    "W": ("H", [16384, 32768], [128, 256, 512, 1024, 1 << 30]),
}
TILE_QUICK = {
    "D": ("D", [8192], [120, 240, 1 << 30]),
    "W": ("H", [16384], [256, 512, 1 << 30]),
}


def time_routine(torch, pkg, h, routine, n, A0, B, reps):
    """Best wall time of one factorisation (D, Z, S) or of the FP16-product
    solve (H) of order n on a fresh copy of A0."""
    best = float("inf")
    for _ in range(reps):
        A = A0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if routine == "H":
            X = torch.zeros_like(B)
            st, _, used, info = pkg.cuDsposvx(h, pkg.Lower, n, B.shape[0], pkg.ptr(A), n, pkg.ptr(B), n,
                                              pkg.ptr(X), n, "H", "I")
        elif routine == "D":
            st, info = pkg.cuDpotrf(h, pkg.Lower, n, pkg.ptr(A), n)
        elif routine == "Z":
            st, info = pkg.cuZpotrf(h, pkg.Upper, n, pkg.ptr(A), n)
        else:
            st, info = pkg.cuSpotrf(h, pkg.Lower, n, pkg.ptr(A), n)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
        if st != 0 or info != 0:
            raise RuntimeError(f"{routine} n={n}: status {st} info {info}")
    return best


def tune_tiles(torch, pkg, h, sweep, reps):
    """gemm_tile lines: per family, the threshold with the least total time
    over the family's orders."""
    lines = []
    for fam, (routine, orders, cands) in sweep.items():
        dtype = {"D": torch.float64, "Z": torch.complex128, "S": torch.float32, "H": torch.float64}[routine]
        total = {c: 0.0 for c in cands}
        detail = []
        for n in orders:
            A0 = spd(torch, n, dtype, seed=n)
            B = torch.randn(4, n, device="cuda", dtype=torch.float64) if routine == "H" else None
            row = []
            for c in cands:
                os.environ[f"CUCHOLESKY_TILE_MIN_{fam}"] = str(c)
                time_routine(torch, pkg, h, routine, n, A0, B, 1)   # workspace, warm-up
                t = time_routine(torch, pkg, h, routine, n, A0, B, reps)
                total[c] += t
                row.append(f"{'never' if c >= 1 << 30 else c}:{t * 1e3:.2f}ms")
            os.environ.pop(f"CUCHOLESKY_TILE_MIN_{fam}", None)
            detail.append(f"n={n} " + " ".join(row))
            del A0, B
            torch.cuda.empty_cache()
        best = min(total, key=total.get)
        lines.append(f"gemm_tile {fam} {best}   # {'; '.join(detail)}")
        print(lines[-1], flush=True)
    return lines


# precision: (order, candidate (t_max, workgroups) pairs; (0, 0) = no cap)
CAP_SWEEP = {
    "D": (16384, [(0, 0), (8192, 384), (8192, 448), (12288, 448), (16384, 448), (16384, 480)]),
    "Z": (8192, [(0, 0), (8192, 384), (8192, 448)]),
}
CAP_QUICK = {"D": (16384, [(0, 0), (12288, 448)])}


def tune_cap(torch, pkg, h, sweep, reps):
    """trailing_cap lines: per precision, the candidate with the least median
    time, the candidates timed in turn reps times."""
    lines = []
    for prec, (n, cands) in sweep.items():
        dtype = {"D": torch.float64, "Z": torch.complex128}[prec]
        A0 = spd(torch, n, dtype, seed=n)
        var = f"CUCHOLESKY_TRAILING_CAP_{prec}"
        times = {c: [] for c in cands}
        for c in cands:  # warm-up
            os.environ[var] = f"{c[0]}:{c[1]}"
            time_routine(torch, pkg, h, prec, n, A0, None, 1)
        for _ in range(max(3, reps)):
            for c in cands:
                os.environ[var] = f"{c[0]}:{c[1]}"
                times[c].append(time_routine(torch, pkg, h, prec, n, A0, None, 1))
        os.environ.pop(var, None)
        med = {c: sorted(v)[len(v) // 2] for c, v in times.items()}
        best = min(med, key=med.get)
        detail = " ".join(f"{c[0]}:{c[1]}:{t * 1e3:.2f}ms" for c, t in med.items())
        if best[1] > 0:
            lines.append(f"trailing_cap {prec} {best[0]} {best[1]}   # n={n} median {detail}")
        else:
            lines.append(f"# trailing_cap {prec}: no cap   # n={n} median {detail}")
        print(lines[-1], flush=True)
        del A0
        torch.cuda.empty_cache()
    return lines


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--quick", action="store_true", help="a reduced sweep (a minute or two)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    ap.add_argument("--only", choices=["nb", "batched", "tiles", "cap"], help="run one of the sweeps")
    args = ap.parse_args()
    import torch
    pkg = package()
    pkg.lib()
    h = pkg.LAPACKHandle()
    sweep = QUICK if args.quick else SWEEP
    lines = ["# schedule parameters measured by autotune.py on " + torch.cuda.get_device_name(0),
             "# potrf_nb <precision> <n_max> <nb>   (seconds per candidate in the trailing comment)",
             "# batched_per_cu <precision> <inverse> <workgroups per CU>",
             "# gemm_tile <family> <min_tiles>   (milliseconds per order and candidate in the trailing comment)",
             "# trailing_cap <precision> <t_max> <workgroups>   (median milliseconds per candidate in the trailing comment)"]
    if args.only in ("batched", "tiles", "cap"):
        sweep = {}
    for prec, (orders, cands) in sweep.items():
        dtype = {"D": torch.float64, "Z": torch.complex128, "S": torch.float32, "B": torch.float64}[prec]
        for i, n in enumerate(orders):
            A0 = spd(torch, n, dtype, seed=n)
            B = torch.randn(4, n, device="cuda", dtype=torch.float64) if prec == "B" else None
            times = {}
            for nb in cands:
                assert pkg.cuLAPACKSetBlockSize(h, prec, nb) == 0
                time_one(torch, pkg, h, prec, n, A0, B, 1)          # grow the workspace first
                times[nb] = time_one(torch, pkg, h, prec, n, A0, B, args.reps)
            pkg.cuLAPACKSetBlockSize(h, prec, 0)
            best = min(times, key=times.get)
            n_max = n if i + 1 < len(orders) else 1 << 40
            detail = " ".join(f"{nb}:{t * 1e3:.2f}ms" for nb, t in times.items())
            lines.append(f"potrf_nb {prec} {n_max} {best}   # n={n} {detail}")
            print(lines[-1], flush=True)
            del A0, B
            torch.cuda.empty_cache()
    if args.only in (None, "batched"):
        lines += tune_batched(torch, pkg, h, 4000 if args.quick else 20000, args.reps + 2)
    if args.only in (None, "tiles"):
        lines += tune_tiles(torch, pkg, h, TILE_QUICK if args.quick else TILE_SWEEP, args.reps)
    if args.only in (None, "cap"):
        lines += tune_cap(torch, pkg, h, CAP_QUICK if args.quick else CAP_SWEEP, args.reps + 2)
    h.close()
    if args.only is not None and os.path.exists(args.out):
        # keep the entries of the sweeps that did not run
        ran = {"nb": "potrf_nb", "batched": "batched_per_cu", "tiles": "gemm_tile", "cap": "trailing_cap"}[args.only]
        with open(args.out) as f:
            lines += [ln.rstrip("\n") for ln in f
                      if not ln.startswith("#") and not ln.startswith(ran) and ln.strip()]
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
