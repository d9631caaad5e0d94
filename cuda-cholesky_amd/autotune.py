#!/usr/bin/env python3
"""Measure the outer block size of the blocked factorisations on this GPU.

For every precision (D: cuDpotrf, Z: cuZpotrf, S: cuSpotrf, B: the BF16
factor of cuDsposvx) and every order in the sweep, each candidate block size
is set with cuLAPACKSetBlockSize and the factorisation timed (best of a few
runs, fresh input each time, HIP events on the device clock).  The winner
for each order covers all n up to it; the table goes to tuning.txt beside
the library (or --out), where the library reads it at first use:

    potrf_nb <precision> <n_max> <nb>

    python cuda-cholesky_amd/autotune.py [--quick] [--out PATH]
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
    "S": ([4096, 8192, 16384], [384, 512, 768, 1024]),
    "B": ([4096, 8192, 16384], [384, 512, 768, 1024]),
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


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--quick", action="store_true", help="a reduced sweep (a minute or two)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    args = ap.parse_args()
    import torch
    pkg = package()
    pkg.lib()
    h = pkg.LAPACKHandle()
    sweep = QUICK if args.quick else SWEEP
    lines = ["# outer block sizes measured by autotune.py on " + torch.cuda.get_device_name(0),
             "# potrf_nb <precision> <n_max> <nb>   (seconds per candidate in the trailing comment)"]
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
    h.close()
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
