#!/usr/bin/env python3
"""Benchmark of the plain FP64 GPU drop-in solve (approxchol_lap's solve phase).

Workload (default): 3-D 128^3 grid Laplacian, n = 2,097,152, m = 6,242,304,
edge weights exp(U(-2,2)) (numpy PCG64 seed 2), one RHS ~ N(0,1) from the same
stream projected to sum zero, tol = 1e-8, maxits = 1000; the LDLinv is built
once on the host (liblapfactor, seed 1) and uploaded.  A step is one full
solve f(b) with b already resident in HBM.  `value` is solves/s.

Prints ONE JSON line (rank 0).  --gpus N > 1 exits: one GPU only.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "laplacians.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(name: str):
    from lapgpu import graphs

    if name == "grid3_128":
        rng = np.random.default_rng(2)
        a = graphs.weighted_grid3(128, 128, 128, rng, "expU2")
        b = rng.standard_normal(a.shape[0])
    elif name == "grid3_64":
        rng = np.random.default_rng(2)
        a = graphs.weighted_grid3(64, 64, 64, rng, "expU2")
        b = rng.standard_normal(a.shape[0])
    elif name == "grid2_100":
        a = graphs.grid2(100)
        b = np.random.default_rng(0).standard_normal(a.shape[0])
    else:
        raise SystemExit(f"unknown workload {name}")
    b = b - b.mean()
    return a, b


def algorithmic_bytes(n: int, nnzL: int, nnzF: int, sv: int = 8, si: int = 4) -> dict:
    """Bytes each kernel class must move per PCG iteration (one RHS)."""
    spmv = nnzL * (sv + si) + (n + 1) * si + 2 * n * sv
    # two sweeps over (index, weight) and the row pointers, the diagonal
    # scale s, and one read + one write of y per sweep (plus r in forward)
    ldl = 2 * (nnzF * (sv + si) + (n + 1) * si) + n * sv + 5 * n * sv
    vec = 10 * n * 8
    return {"spmv": spmv, "ldl": ldl, "vec": vec}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="grid3_128")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-its", type=int, default=0, help="oracle iterations for the CPU baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        return 2

    import torch

    import lapgpu
    from lapgpu import _native

    t0 = time.time()
    a, b = make_workload(args.workload)
    la = lapgpu.lap(a)
    n = a.shape[0]
    log(f"workload {args.workload}: n={n} nnz(L)={la.nnz} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    f = lapgpu.build_ldli(a, seed=1)
    t_factor = time.time() - t0
    log(f"factor nnzF={f.nnz} in {t_factor:.1f}s")

    ctx = _native.Context(0)
    ctx.upload_lap(la)
    ctx.upload_ldli(f)
    info = ctx.factor_info()
    log(f"schedule: {info}")

    dev = torch.device("cuda:0")
    db = torch.from_numpy(b).to(dev)
    dx = torch.empty_like(db)
    torch.cuda.synchronize()

    def step():
        return ctx.solve_device(db.data_ptr(), dx.data_ptr(), n, tol=args.tol, maxits=1000, stag_test=5)

    for _ in range(args.warmup):
        r = step()
        log(f"warmup solve: {r}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for _ in range(args.steps):
        r = step()
        its.append(r["its"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    x = dx.cpu().numpy()
    relres = float(np.linalg.norm(la @ x - b) / np.linalg.norm(b))
    log(f"timed {args.steps} solves in {elapsed:.3f}s its={its} true relres={relres:.3e}")

    # per-kernel-class device times (HIP events on the library's stream)
    prof = ctx.solve_device(db.data_ptr(), dx.data_ptr(), n, tol=args.tol, maxits=1000, stag_test=5,
                            flags=_native.PROFILE)
    tm = ctx.last_timings()
    byts = algorithmic_bytes(n, la.nnz, f.nnz)
    spmv_gbs = byts["spmv"] / (tm["spmv_ms"] * 1e-3) / 1e9 if tm["spmv_ms"] > 0 else None
    ldl_gbs = byts["ldl"] / (tm["ldl_ms"] * 1e-3) / 1e9 if tm["ldl_ms"] > 0 else None
    log(f"profile solve {prof}: {tm} spmv {spmv_gbs} GB/s ldl {ldl_gbs} GB/s")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle

        oracle.build()
        cpu_its = args.cpu_its or (20 if n > 500000 else 200)
        t0 = time.perf_counter()
        _, io = oracle.approxchol_solve(la, f, b, tol=args.tol, maxits=cpu_its, stag_test=5)
        t_cpu = time.perf_counter() - t0
        per_it = t_cpu / max(io["its"], 1)
        its_full = int(np.median(its))
        cpu = {
            "value": 1.0 / (per_it * its_full),
            "unit": "solves/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle pcg+LDLsolver (C, 1 thread) for {io['its']} iterations on the same graph, factor and b "
                      f"({t_cpu:.1f}s, {per_it * 1e3:.1f} ms/iteration), scaled to the GPU's {its_full} iterations",
        }
        log(f"cpu baseline: {cpu}")

    ms = elapsed / args.steps * 1e3
    out = {
        "metric": "solves/s (plain FP64 GPU PCG, relres<=1e-8)",
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: exp(U(-2,2)) weights seed 2, 1 RHS N(0,1) sum-zero, tol {args.tol}",
            "n": n,
            "nnz_L": int(la.nnz),
            "nnzF": int(f.nnz),
            "iterations": int(np.median(its)),
            "ms_per_iteration": ms / max(np.median(its), 1),
            "true_relres": relres,
            "levels": info["nlevels"],
            "spmv_ms": tm["spmv_ms"],
            "ldl_ms": tm["ldl_ms"],
            "vec_ms": tm["vec_ms"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv_dot",
            "achieved": spmv_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (spmv_gbs / HBM_PEAK_GBS) if spmv_gbs else None,
            "traffic": None,
            "ldl_apply_GBps": ldl_gbs,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
