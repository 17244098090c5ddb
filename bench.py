#!/usr/bin/env python3
"""Benchmarks of the MI355X solve phase of approxchol_lap.

Default (--mode plain): the plain FP64 GPU drop-in solve on a 3-D 128^3 grid
Laplacian, n = 2,097,152, m = 6,242,304, edge weights exp(U(-2,2)) (numpy
PCG64 seed 2), one RHS ~ N(0,1) from the same stream projected to sum zero,
tol = 1e-8, maxits = 1000; the LDLinv is built once on the host (liblapfactor,
seed 1) and uploaded.  A step is one full solve f(b) with b already resident
in HBM.  `value` is solves/s.  Prints ONE JSON line.

Other modes (one JSON line each, same keys):
  multi      2-D 1024x1024 grid, exp(N(0,1)) weights, a 64-column RHS block
             (FP32-inner and FP64), next to 256 small unit-weight 2-D grids
             (n log-uniform in [1e3, 1e5], aspect ratio in [1, 8]) solved by
             the batched engine at the same time; seed 4.
  precision  union of a path and 4 random matchings, n = 1e6, weights
             log-uniform in [1e-3, 1e3], seed 3: FP32-inner vs FP64.
  serve      Poisson stream of requests over resident graphs (2-D/3-D grids
             with exp(U(-2,2)) weights and path+matchings graphs, n
             log-uniform) with a share that brings new graphs; seed 5.

--gpus N > 1 exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import threading
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


def group_bytes(gi: dict, sv: int = 8, si: int = 4) -> int:
    """Algorithmic bytes of the level-group launches of one iteration (both
    sweeps): per composite-row entry its index and weight, per row its
    pointer, one read of its input value (r or y0) and its write, and the
    gathered values counted once per row they produce."""
    rows = gi["rows"]
    return (gi["fentries"] + gi["bentries"]) * (sv + si) + 2 * rows * (si + 3 * sv)


def tail_bytes(ti: dict, sv: int = 8) -> int:
    """Algorithmic bytes of one k_tail launch: its slot programs (16 B per
    slot, 8 B per step), the tail inputs and scales read, yF written."""
    return 16 * ti["nslots"] + 8 * ti["nsteps"] + 3 * ti["nt"] * sv


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size), from the committed
    summary profiles/<kernel>_<workload>_traffic.md, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def emit(out: dict) -> None:
    print(json.dumps(out), flush=True)


# ----------------------------------------------------------------- plain ---
def mode_plain(args) -> int:
    import lapgpu
    from lapgpu import _native
    from lapgpu._native import DeviceArray, synchronize

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
    tune = None
    if args.autotune:
        t0 = time.time()
        tune = ctx.autotune(k=0, groups=True)
        tune["seconds"] = time.time() - t0
        log(f"autotune: {tune}")
    info = ctx.factor_info()
    ti = ctx.tail_info()
    gi = ctx.group_info()
    log(f"schedule: {info} tail {ti} groups {gi}")

    db = DeviceArray.from_host(b)
    dx = DeviceArray(n)
    synchronize()

    def step():
        return ctx.solve_device(db.address, dx.address, n, tol=args.tol, maxits=1000, stag_test=5)

    for _ in range(args.warmup):
        r = step()
        log(f"warmup solve: {r}")
    synchronize()
    t0 = time.perf_counter()
    its = []
    for _ in range(args.steps):
        r = step()
        its.append(r["its"])
    synchronize()
    elapsed = time.perf_counter() - t0
    x = dx.download()
    relres = float(np.linalg.norm(la @ x - b) / np.linalg.norm(b))
    log(f"timed {args.steps} solves in {elapsed:.3f}s its={its} true relres={relres:.3e}")

    # per-kernel device times: HIP events recorded on the library's stream
    # around each kernel class of every iteration of one more solve
    prof = ctx.solve_device(db.address, dx.address, n, tol=args.tol, maxits=1000, stag_test=5,
                            flags=_native.PROFILE)
    tm = ctx.last_timings()
    st = ctx.last_sweep_timings()
    byts = algorithmic_bytes(n, la.nnz, f.nnz)
    spmv_gbs = byts["spmv"] / (tm["spmv_ms"] * 1e-3) / 1e9 if tm["spmv_ms"] > 0 else None
    ldl_gbs = byts["ldl"] / (tm["ldl_ms"] * 1e-3) / 1e9 if tm["ldl_ms"] > 0 else None
    tb = tail_bytes(ti)
    tail_gbs = tb / (st["tail_ms"] * 1e-3) / 1e9 if st["tail_ms"] > 0 else None
    gb = group_bytes(gi)
    g_ms = st["fgroups_ms"] + st["bgroups_ms"]
    group_gbs = gb / (g_ms * 1e-3) / 1e9 if g_ms > 0 else None
    log(f"profile solve {prof}: {tm} {st} spmv {spmv_gbs} GB/s ldl {ldl_gbs} GB/s tail {tail_gbs} GB/s")

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
    emit({
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
            "autotune": tune,
            "n": n,
            "nnz_L": int(la.nnz),
            "nnzF": int(f.nnz),
            "iterations": int(np.median(its)),
            "ms_per_iteration": ms / max(np.median(its), 1),
            "true_relres": relres,
            "levels": info["nlevels"],
            "level_groups": info["nwide"],
            "tail_vertices": ti["nt"],
            "tail_steps": ti["nsteps"],
            "spmv_ms": tm["spmv_ms"],
            "ldl_ms": tm["ldl_ms"],
            "vec_ms": tm["vec_ms"],
            "ldl_split_ms": st,
            "spmv_GBps": spmv_gbs,
            "ldl_apply_GBps": ldl_gbs,
            "factor_host_s": t_factor,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_group (the level-group launches of both LDLinv sweeps, one iteration; "
                      "latency-bound by dependent launches)",
            "achieved": group_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (group_gbs / HBM_PEAK_GBS) if group_gbs else None,
            "traffic": read_traffic("k_group", args.workload),
            "bytes_per_iteration": gb,
            "launches_per_iteration": 2 * gi["ngroups"],
            "ms_per_iteration": g_ms,
            "tail": {"kernel": "k_tail", "achieved": tail_gbs,
                     "frac": (tail_gbs / HBM_PEAK_GBS) if tail_gbs else None,
                     "traffic": read_traffic("k_tail", args.workload), "bytes_per_launch": tb,
                     "avg_launch_ms": st["tail_ms"]},
            "spmv": {"kernel": "k_spmv_dot", "achieved": spmv_gbs,
                     "frac": (spmv_gbs / HBM_PEAK_GBS) if spmv_gbs else None,
                     "traffic": read_traffic("k_spmv_dot", args.workload)},
        },
        "cpu_baseline": cpu,
    })
    if args.calib:
        src = DeviceArray(1 << 26)
        dst = DeviceArray(1 << 26)
        dst.copy_from(src)
        synchronize()
        log("calibration copy: 536870912 bytes read, 536870912 bytes written")
    ctx.close()
    return 0


# ----------------------------------------------------------------- multi ---
def _small_grids(rng, count: int):
    from lapgpu import graphs

    out = []
    for _ in range(count):
        nv = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        ar = rng.uniform(1.0, 8.0)
        ny = max(2, int(round(np.sqrt(nv / ar))))
        nx = max(2, int(round(nv / ny)))
        out.append(graphs.grid2(nx, ny))
    return out


def mode_multi(args) -> int:
    import lapgpu
    from lapgpu import _native, graphs
    from lapgpu._native import DeviceArray, synchronize

    rng = np.random.default_rng(4)
    a = graphs.weighted_grid2(1024, 1024, rng, "expN")
    la = lapgpu.lap(a)
    n = a.shape[0]
    k = args.rhs
    B = rng.standard_normal((k, n))
    B -= B.mean(axis=1, keepdims=True)
    smalls = _small_grids(rng, args.small)
    sb = []
    for g in smalls:
        v = rng.standard_normal(g.shape[0])
        sb.append(v - v.mean())
    t0 = time.time()
    f = lapgpu.build_ldli(a, seed=1)
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(8) as ex:
        sf = list(ex.map(lambda g: lapgpu.build_ldli(g, seed=1), smalls))
    log(f"grid2 1024^2 n={n} nnzF={f.nnz}; {len(smalls)} small grids (n {min(g.shape[0] for g in smalls)}.."
        f"{max(g.shape[0] for g in smalls)}); factors in {time.time() - t0:.1f}s")

    ctx = _native.Context(0)
    ctx.upload_lap(la)
    ctx.upload_ldli(f)
    tune = None
    if not args.no_autotune:
        tune = ctx.autotune(k=k, groups=True)
        log(f"autotune (k={k}): {tune}")
    batch = _native.Batch(0)
    gids = [batch.add_graph(lapgpu.lap(g), ff) for g, ff in zip(smalls, sf)]
    dB = DeviceArray.from_host(B)
    dX = DeviceArray(n * k)
    dsb = [DeviceArray.from_host(v) for v in sb]
    dsx = [DeviceArray(len(v)) for v in sb]
    results = {}

    def run_multi(prec):
        ctx.set_precision(prec, 20)
        results["multi"] = ctx.solve_multi_device(dB.address, dX.address, n, k, tol=1e-8, maxits=1000)

    def run_batch():
        results["batch"] = batch.solve_device(gids, [d.address for d in dsb], [d.address for d in dsx],
                                              tol=1e-8, maxits=1000)

    out_modes = {}
    for prec, pname in ((_native.FP32_INNER, "fp32_inner"), (_native.FP64, "fp64")):
        # warm (graph capture, workspace), then each part alone
        run_multi(prec)
        run_batch()
        synchronize()
        t0 = time.perf_counter()
        run_multi(prec)
        synchronize()
        t_multi = time.perf_counter() - t0
        t0 = time.perf_counter()
        run_batch()
        synchronize()
        t_batch = time.perf_counter() - t0
        log(f"{pname}: 64-RHS block alone {t_multi * 1e3:.1f} ms ({max(results['multi']['its'])} its), "
            f"256 small grids alone {t_batch * 1e3:.1f} ms")
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            th = threading.Thread(target=run_multi, args=(prec,))
            th.start()
            run_batch()
            th.join()
            synchronize()
            times.append(time.perf_counter() - t0)
        X = dX.download().reshape(k, n)
        rel = np.linalg.norm(la @ X.T - B.T, axis=0) / np.linalg.norm(B.T, axis=0)
        srel = max(float(np.linalg.norm(lapgpu.lap(g) @ d.download() - v) / np.linalg.norm(v))
                   for g, d, v in zip(smalls, dsx, sb))
        out_modes[pname] = {
            "wall_s": float(np.median(times)),
            "solves_per_s": (k + len(smalls)) / float(np.median(times)),
            "multi_its_max": int(max(results["multi"]["its"])),
            "multi_true_relres_max": float(rel.max()),
            "small_its_max": int(max(i["its"] for i in results["batch"])),
            "small_true_relres_max": srel,
            "multi_alone_ms": t_multi * 1e3,
            "small_alone_ms": t_batch * 1e3,
            "X": X if pname == "fp64" else X.copy(),
        }
        log(f"{pname}: {out_modes[pname]['solves_per_s']:.1f} solves/s wall {out_modes[pname]['wall_s']*1e3:.1f} ms "
            f"relres {rel.max():.2e} small relres {srel:.2e}")
    err32 = float(np.max(np.linalg.norm(out_modes["fp32_inner"]["X"] - out_modes["fp64"]["X"], axis=1)
                         / np.linalg.norm(out_modes["fp64"]["X"], axis=1)))
    for v in out_modes.values():
        v.pop("X")

    # plain one-at-a-time FP64 GPU PCG on the same GPU: one column of the big
    # grid and a sample of the small grids, each solved alone
    ctx.set_precision(_native.FP64)
    d1 = DeviceArray.from_host(B[0])
    dx1 = DeviceArray(n)
    ctx.solve_device(d1.address, dx1.address, n, tol=1e-8, maxits=1000)
    synchronize()
    t0 = time.perf_counter()
    ctx.solve_device(d1.address, dx1.address, n, tol=1e-8, maxits=1000)
    synchronize()
    t_big = time.perf_counter() - t0
    sample = list(range(0, len(smalls), max(1, len(smalls) // 16)))
    t_small = []
    for i in sample:
        c1 = _native.Context(0)
        c1.upload_lap(lapgpu.lap(smalls[i]))
        c1.upload_ldli(sf[i])
        c1.solve_device(dsb[i].address, dsx[i].address, smalls[i].shape[0], tol=1e-8, maxits=1000)
        synchronize()
        t0 = time.perf_counter()
        c1.solve_device(dsb[i].address, dsx[i].address, smalls[i].shape[0], tol=1e-8, maxits=1000)
        synchronize()
        t_small.append(time.perf_counter() - t0)
        c1.close()
    plain_wall = k * t_big + len(smalls) * float(np.mean(t_small))
    plain_sps = (k + len(smalls)) / plain_wall
    # the workload names FP64 outer / FP32 inner; the FP64 leg is reported beside it
    best_name = "fp32_inner"
    best = out_modes[best_name]
    faster = max(out_modes, key=lambda m: out_modes[m]["solves_per_s"])
    emit({
        "metric": "solves/s (64-RHS block + 256 small grids, relres<=1e-8) vs plain FP64 GPU PCG",
        "value": best["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": best["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 inner",
        "data": "synthetic",
        "config": {
            "workload": "grid2 1024x1024 exp(N(0,1)) weights, 64 RHS N(0,1) sum-zero + 256 unit 2-D grids "
                        "(n log-uniform [1e3,1e5], aspect [1,8]), seed 4, tol 1e-8",
            "reported_mode": best_name,
            "faster_mode": faster,
            "autotune": tune,
            "fp32_inner": out_modes["fp32_inner"],
            "fp64": out_modes["fp64"],
            "fp32_vs_fp64_max_rel_err": err32,
            "plain_one_at_a_time_solves_per_s": plain_sps,
            "plain_big_solve_ms": t_big * 1e3,
            "plain_small_solve_ms_mean": float(np.mean(t_small)) * 1e3,
            "speedup_vs_plain": best["solves_per_s"] / plain_sps,
        },
        "roofline": None,
        "cpu_baseline": None,
    })
    ctx.close()
    batch.close()
    return 0


# ------------------------------------------------------------- precision ---
def mode_precision(args) -> int:
    import lapgpu
    from lapgpu import _native, graphs
    from lapgpu._native import DeviceArray, synchronize

    rng = np.random.default_rng(3)
    a = graphs.path_matchings(1_000_000, 4, rng, "logU3")
    la = lapgpu.lap(a)
    n = a.shape[0]
    b = rng.standard_normal(n)
    b -= b.mean()
    t0 = time.time()
    f = lapgpu.build_ldli(a, seed=1)
    log(f"path+4 matchings n={n} nnzL={la.nnz} nnzF={f.nnz} factor {time.time() - t0:.1f}s")
    ctx = _native.Context(0)
    ctx.upload_lap(la)
    ctx.upload_ldli(f)
    db = DeviceArray.from_host(b)
    dx = DeviceArray(n)
    res = {}
    for prec, name in ((_native.FP64, "fp64"), (_native.FP32_INNER, "fp32_inner")):
        ctx.set_precision(prec, 20)
        ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000)
        synchronize()
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            r = ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000)
            synchronize()
            ts.append(time.perf_counter() - t0)
        x = dx.download()
        ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000, flags=_native.PROFILE)
        log(f"{name} schedule {ctx.factor_info()} tail {ctx.tail_info()} per-iteration ms "
            f"{ctx.last_timings()} {ctx.last_sweep_timings()}")
        res[name] = {"its": r["its"], "stop": r["stop"], "solve_ms": float(np.median(ts)) * 1e3,
                     "true_relres": float(np.linalg.norm(la @ x - b) / np.linalg.norm(b)), "x": x}
        log(f"{name}: {res[name]['its']} its {res[name]['solve_ms']:.1f} ms relres {res[name]['true_relres']:.2e}")
    err = float(np.linalg.norm(res["fp32_inner"]["x"] - res["fp64"]["x"]) / np.linalg.norm(res["fp64"]["x"]))
    for v in res.values():
        v.pop("x")
    emit({
        "metric": "FP32-inner vs FP64 solve (relres<=1e-8): error vs FP64 and iteration ratio",
        "value": 1e3 / res["fp32_inner"]["solve_ms"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": res["fp32_inner"]["solve_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 inner",
        "data": "synthetic",
        "config": {
            "workload": "path + 4 random perfect matchings, n=1e6, weights log-uniform [1e-3,1e3], seed 3, "
                        "1 RHS N(0,1) sum-zero, tol 1e-8, FP64 residual replacement every 20 iterations",
            "fp64": res["fp64"],
            "fp32_inner": res["fp32_inner"],
            "rel_err_vs_fp64": err,
            "iteration_ratio": res["fp32_inner"]["its"] / max(res["fp64"]["its"], 1),
        },
        "roofline": None,
        "cpu_baseline": None,
    })
    ctx.close()
    return 0


# ----------------------------------------------------------------- serve ---
def _serve_graph(rng, nv: float):
    from lapgpu import graphs

    kind = rng.integers(3)
    if kind == 0:
        ar = rng.uniform(1.0, 4.0)
        ny = max(2, int(round(np.sqrt(nv / ar))))
        return graphs.weighted_grid2(max(2, int(round(nv / ny))), ny, rng, "expU2")
    if kind == 1:
        s = max(2, int(round(nv ** (1 / 3))))
        return graphs.weighted_grid3(s, s, max(2, int(round(nv / (s * s)))), rng, "expU2")
    return graphs.path_matchings(max(4, int(nv) // 2 * 2), 4, rng, "logU3")


def mode_serve(args) -> int:
    from concurrent.futures import ThreadPoolExecutor

    # every resident graph owns a stream; HIP maps streams round-robin onto
    # its hardware queues (4 unless GPU_MAX_HW_QUEUES says otherwise), and two
    # busy streams on one queue run one after the other.  Read at HIP start-up,
    # so it is set before the library loads.
    if args.hw_queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)

    import lapgpu
    from lapgpu import _native
    from lapgpu.serve import Request, Server, summary

    rng = np.random.default_rng(5)
    lo, hi = np.log(args.serve_min_n), np.log(args.serve_max_n)
    sizes = np.exp(rng.uniform(lo, hi, args.serve_graphs))
    G = [_serve_graph(rng, nv) for nv in sizes]
    t0 = time.time()
    with ThreadPoolExecutor(8) as ex:
        F = list(ex.map(lambda g: lapgpu.build_ldli(g, seed=1), G))
    log(f"{len(G)} resident graphs, n {min(g.shape[0] for g in G)}..{max(g.shape[0] for g in G)} "
        f"(sum {sum(g.shape[0] for g in G)}), factored in {time.time() - t0:.1f}s")
    srv = Server(0, hbm_budget=200 * 2**30, plain_workers=args.serve_workers,
                 factor_workers=args.factor_workers, batch_window=args.batch_window,
                 max_wait=args.max_wait, quick_workers=args.quick_workers,
                 blocking_waits=not args.spin_waits, wg_max_n=args.wg_max_n)
    for i, (g, f) in enumerate(zip(G, F)):
        assert srv.register(i, g, ldli=f)
    # request stream
    nreq = args.requests
    arr = np.cumsum(rng.exponential(1.0 / args.qps, nreq))
    reqs = []
    new_sizes = np.exp(rng.uniform(lo, np.log(args.serve_new_max_n), nreq))
    for i in range(nreq):
        if rng.uniform() < args.new_frac:
            g = _serve_graph(rng, new_sizes[i])
            b = rng.standard_normal(g.shape[0])
            reqs.append(Request(i, float(arr[i]), b - b.mean(), new_graph=g, tol=1e-8))
        else:
            gi = int(rng.integers(len(G)))
            b = rng.standard_normal(G[gi].shape[0])
            reqs.append(Request(i, float(arr[i]), b - b.mean(), graph=gi, tol=1e-8))
    log(f"{nreq} requests at {args.qps} QPS ({sum(r.new_graph is not None for r in reqs)} with new graphs)")
    # warm every resident graph once (graph capture, workspaces)
    wrng = np.random.default_rng(55)
    srv.run([Request(-1 - i, 0.0, wrng.standard_normal(G[i].shape[0]), graph=i) for i in range(len(G))],
            realtime=False)
    import resource

    ru0 = resource.getrusage(resource.RUSAGE_SELF)
    t0 = time.perf_counter()
    res = srv.run(reqs, progress=lambda d, t, e: log(f"  served {d} of {t} requests at {e:.0f} s"))
    wall = time.perf_counter() - t0
    ru1 = resource.getrusage(resource.RUSAGE_SELF)
    cpu_s = (ru1.ru_utime - ru0.ru_utime) + (ru1.ru_stime - ru0.ru_stime)
    log(f"host CPU during the run: {cpu_s:.1f} s over {wall:.1f} s wall ({cpu_s / wall:.1f} cores busy on average)")
    s = summary(res, wall)
    paths = {p: sum(o.path == p for o in res) for p in {o.path for o in res}}
    log(f"served: {s} paths {paths}")
    for pth in sorted(paths):
        sel = [(r, o) for r, o in zip(reqs, res) if o.path == pth]
        lat = np.array([o.latency for _, o in sel])
        done = max(r.arrival + o.latency for r, o in sel)
        newg = sum(r.new_graph is not None for r, _ in sel)
        log(f"  path {pth}: {len(sel)} ({newg} new graphs), latency p50 {np.percentile(lat, 50):.2f} s "
            f"p99 {np.percentile(lat, 99):.2f} s, last done at {done:.1f} s")
    log(f"  dispatch loop: {srv.stats}")
    up = np.array(srv.upload_log).reshape(-1, 4)
    for name, lg in (("factor", np.array(srv.factor_log)), ("context upload", up[:, 0]),
                     ("  create", up[:, 1]), ("  upload la", up[:, 2]), ("  upload ldli", up[:, 3])):
        if len(lg):
            log(f"  new graphs, {name}: {len(lg)} x, median {np.median(lg) * 1e3:.1f} ms, "
                f"max {lg.max() * 1e3:.1f} ms, sum {lg.sum():.1f} s")
    fc = np.array(srv.factor_cpu_log)
    if len(fc):
        log(f"  new graphs, CPU of the calling thread: factor {fc.sum():.1f} s, "
            f"context upload {np.sum(srv.upload_cpu_log):.1f} s (planner helper threads not counted)")
    sl = np.array(srv.solve_log[len(G):], dtype=float) if len(srv.solve_log) > len(G) else np.zeros((0, 5))
    if len(sl):
        log(f"  solve workers: {sl[:, 2].sum():.1f} s waiting in solves, {sl[:, 4].sum():.1f} s of it on the CPU")
    for lo_n, hi_n in ((0, 1e5), (1e5, 1e6), (1e6, 1e8)):
        m = (sl[:, 0] >= lo_n) & (sl[:, 0] < hi_n) if len(sl) else np.zeros(0, bool)
        if m.any():
            f1 = m & (sl[:, 3] > 0)
            log(f"  solves n in [{lo_n:g},{hi_n:g}): {m.sum()} solves, {sl[m, 1].sum():.0f} columns, "
                f"mean columns {sl[m, 1].mean():.1f}, {sl[m, 2].sum():.1f} s; first on a context: "
                f"{f1.sum()}, {sl[f1, 2].sum():.1f} s")
    # plain one-at-a-time FP64 GPU PCG on the same GPU: the service time of
    # one plain solve per resident graph, then the same arrival stream
    # through a single FIFO server (new graphs: their own plain solve time,
    # factoring excluded)
    svc = {}
    for i, (g, f) in enumerate(zip(G, F)):
        c = _native.Context(0)
        c.upload_lap(lapgpu.lap(g))
        c.upload_ldli(f)
        bb = np.random.default_rng(i).standard_normal(g.shape[0])
        bb -= bb.mean()
        c.solve(bb, tol=1e-8, maxits=1000)
        t1 = time.perf_counter()
        c.solve(bb, tol=1e-8, maxits=1000)
        svc[i] = time.perf_counter() - t1
        c.close()
    sizes_arr = np.array([g.shape[0] for g in G], dtype=float)
    svc_arr = np.array([svc[i] for i in range(len(G))])
    coef = np.polyfit(np.log(sizes_arr), np.log(svc_arr), 1)
    t_free = 0.0
    lat_plain = []
    busy = 0.0
    for r in reqs:
        st = svc[r.graph] if r.graph is not None else float(np.exp(np.polyval(coef, np.log(len(r.b)))))
        start = max(t_free, r.arrival)
        t_free = start + st
        busy += st
        lat_plain.append(t_free - r.arrival)
    plain_sps = len(reqs) / busy
    emit({
        "metric": "served solves/s + p50/p99 latency (relres<=1e-8) vs plain one-at-a-time FP64 GPU PCG",
        "value": s["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": nreq,
        "warmup": len(G),
        "ms_per_step": wall / nreq * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{nreq} requests, Poisson {args.qps} QPS, {len(G)} resident graphs n log-uniform "
                        f"[{args.serve_min_n:g},{args.serve_max_n:g}] (2-D/3-D grids exp(U(-2,2)), path+4 matchings "
                        f"log-uniform weights), {args.new_frac:.0%} with a new graph n <= {args.serve_new_max_n:g}, "
                        "seed 5, tol 1e-8",
            "served": s,
            "paths": paths,
            "plain_solves_per_s": plain_sps,
            "plain_p50_ms": float(np.percentile(lat_plain, 50) * 1e3),
            "plain_p99_ms": float(np.percentile(lat_plain, 99) * 1e3),
            "speedup_solves_per_s": s["solves_per_s"] / plain_sps,
            "p99_ratio_plain_over_served": float(np.percentile(lat_plain, 99) * 1e3) / s["p99_ms"],
        },
        "roofline": None,
        "cpu_baseline": None,
    })
    srv.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="plain", choices=["plain", "multi", "precision", "serve"])
    ap.add_argument("--workload", default="grid3_128")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-its", type=int, default=0, help="oracle iterations for the CPU baseline sample")
    ap.add_argument("--rhs", type=int, default=64)
    ap.add_argument("--small", type=int, default=256)
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--qps", type=float, default=200.0)
    ap.add_argument("--serve-graphs", type=int, default=32)
    ap.add_argument("--serve-min-n", type=float, default=1e4)
    ap.add_argument("--serve-max-n", type=float, default=4e6)
    ap.add_argument("--serve-new-max-n", type=float, default=2e5)
    ap.add_argument("--new-frac", type=float, default=0.2)
    ap.add_argument("--serve-workers", type=int, default=10, help="serving: graphs solved concurrently")
    ap.add_argument("--factor-workers", type=int, default=12, help="serving: host threads factoring new graphs")
    ap.add_argument("--wg-max-n", type=int, default=50000,
                    help="serving: graphs up to this size are solved by the batched engine")
    ap.add_argument("--spin-waits", action="store_true",
                    help="serving: solve workers spin while waiting for the device (default: they sleep)")
    ap.add_argument("--quick-workers", type=int, default=6,
                    help="serving: workers for solves of graphs with n <= 3e5 (0: share the others)")
    ap.add_argument("--batch-window", type=float, default=2.0, help="serving: wait for more requests, x last solve time")
    ap.add_argument("--max-wait", type=float, default=3.0, help="serving: longest wait for more requests (s)")
    ap.add_argument("--hw-queues", type=int, default=16, help="serving: HIP hardware queues (0: HIP's default)")
    ap.add_argument("--autotune", action="store_true",
                    help="plain: tune the schedule on the graph first (lapgpu_autotune)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="multi: skip the schedule autotuning (done by default)")
    ap.add_argument("--calib", action="store_true",
                    help="plain mode: end with a device-to-device copy of 512 MiB (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        return 2
    return {"plain": mode_plain, "multi": mode_multi, "precision": mode_precision, "serve": mode_serve}[args.mode](args)


if __name__ == "__main__":
    sys.exit(main())
