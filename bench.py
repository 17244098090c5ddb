#!/usr/bin/env python3
"""Benchmarks of the MI355X solve phase of approxchol_lap.  Each mode prints
ONE JSON line.

Default (--mode serve): the serving workload, a Poisson stream of solve
requests at 300 QPS (seed 5) over 32 device-resident graphs with n
log-uniform in [1e4, 4e6] (2-D / 3-D grids with exp(U(-2,2)) weights, path +
4 matchings graphs with log-uniform weights), 20 % of the requests bringing
a new graph of the same families (n up to 2e5) that is factored on the host
and uploaded while the service runs; tol 1e-8, each right-hand side N(0,1)
projected to sum zero.  A step is 500 requests of the stream: --warmup W
steps of another stream (seed 55) over the same resident graphs run first,
untimed, then --steps K steps (K x 500 requests; K = 20 is the 10,000-request
stream) are timed.  `value` is served solves/s over the timed stream, with
p50/p99 latency beside it, against plain one-at-a-time FP64 GPU PCG on the
same GPU (a FIFO replay of the same arrivals with the measured plain solve
time of every resident graph and a fit for the new graphs: modelled).  The
same line carries: the admission time of a new graph with and without the
factor/plan cache; the precision and batching forms on their workloads (a
few solves each); and the plain FP64 solve of a 3-D 128^3 grid (n =
2,097,152, exp(U(-2,2)) weights, seed 2) with its per-kernel HBM roofline
from HIP events on the library's stream and its CPU baseline (one full solve
of the C restatement on one host core).

Other modes (same keys):
  plain      the 3-D 128^3 grid solve alone (--steps solves with b resident
             in HBM); roofline and CPU baseline.
  multi      2-D 1024x1024 grid, exp(N(0,1)) weights, a 64-column RHS block
             (FP32-inner and FP64), next to 256 small unit-weight 2-D grids
             solved by the batched engine at the same time; seed 4.
  precision  union of a path and 4 random matchings, n = 1e6, weights
             log-uniform in [1e-3, 1e3], seed 3: FP32-inner vs FP64, with each
             solve's FP64 true residual as the device reports it.
  batched    the 256 small grids of the multi workload through the batched
             engine alone (one launch per step).

--gpus N > 1 exits: the benchmark measures one GPU.
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
STEP_REQUESTS = 500


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def emit(out: dict) -> None:
    print(json.dumps(out), flush=True)


def host_info() -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def algorithmic_bytes(n: int, nnzL: int, nnzF: int, sv: int = 8, si: int = 4) -> dict:
    """Bytes each part of one PCG iteration must move (one RHS): the SpMV
    (values and indices of la, row pointers, p read and q written), the
    LDLinv apply (both sweeps over the factor's index and value, their row
    pointers, the diagonal, one read and one write of y per sweep plus r)
    and the vector updates."""
    spmv = nnzL * (sv + si) + (n + 1) * si + 2 * n * sv
    ldl = 2 * (nnzF * (sv + si) + (2 * n - 1) * si) + n * sv + 4 * n * sv
    vec = 10 * n * 8
    return {"spmv": spmv, "ldl": ldl, "vec": vec}


def rows_bytes(entries: int, rows: int, sv: int = 8, si: int = 4) -> int:
    """The same count for a part of the apply: its factor entries (index and
    value) and, per row of both sweeps, its row pointers, its diagonal and
    its y traffic (the apply's 56 bytes per vertex)."""
    return entries * (sv + si) + rows * (4 * si + sv + 4 * sv)


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size), from the committed
    summary profiles/<kernel>_<workload>_traffic.md, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def _gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9 if ms and ms > 0 else None


def _frac(gbs):
    return gbs / HBM_PEAK_GBS if gbs else None


# ----------------------------------------------------------------- plain ---
def plain_leg(args, nsolves: int, warmup: int, cpu: bool) -> dict:
    """The plain FP64 solve of grid3(128): solves/s, the per-kernel roofline
    of one iteration (HIP events on the library's stream) and the CPU
    baseline.  Returns {"line": ..., "roofline": ..., "cpu_baseline": ...}."""
    import lapgpu
    from lapgpu import _native, workloads
    from lapgpu._native import DeviceArray, synchronize

    t0 = time.time()
    a, b = workloads.config2()
    la = lapgpu.lap(a)
    n = a.shape[0]
    log(f"plain: grid3(128) n={n} nnz(L)={la.nnz} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    f = lapgpu.build_ldli(a, seed=1)
    t_factor = time.time() - t0
    ctx = _native.Context(0)
    budget = getattr(args, "budget", None)
    if budget:
        ctx.set_group_budget(*budget)
    t0 = time.time()
    ctx.upload_lap(la)
    ctx.upload_ldli(f)
    t_upload = time.time() - t0
    tune = None
    if args.autotune:
        t0 = time.time()
        tune = ctx.autotune(k=0, groups=True)
        tune["seconds"] = time.time() - t0
    info, ti, gi = ctx.factor_info(), ctx.tail_info(), ctx.group_info()
    se, di = ctx.sweep_entries(), ctx.dense_info()
    log(f"plain: factor nnzF={f.nnz} in {t_factor:.1f}s, upload+plan {t_upload:.1f}s; schedule {info} tail {ti} "
        f"groups {gi} chained {ctx.chain_info()} dense {di}")
    db = DeviceArray.from_host(b)
    dx = DeviceArray(n)
    synchronize()

    def step():
        return ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000, stag_test=5)

    for _ in range(warmup):
        step()
    synchronize()
    t0 = time.perf_counter()
    its = [step()["its"] for _ in range(nsolves)]
    synchronize()
    elapsed = time.perf_counter() - t0
    x = dx.download()
    relres = float(np.linalg.norm(la @ x - b) / np.linalg.norm(b))
    log(f"plain: {nsolves} solves in {elapsed:.3f}s its={its} true relres={relres:.3e}")

    # per-kernel device times of one more solve: HIP events recorded on the
    # library's stream around each kernel class of every iteration
    ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000, stag_test=5, flags=_native.PROFILE)
    tm, st, dt = ctx.last_timings(), ctx.last_sweep_timings(), ctx.last_dense_timings()
    log(f"plain: per iteration {tm} {st} {dt}")
    byts = algorithmic_bytes(n, la.nnz, f.nnz)
    T0, NT, K = info["tail_start"], ti["nt"], di["k"]
    ldd = (K + 255) // 256 * 256
    groups_ms = st["fgroups_ms"] + st["bgroups_ms"]
    g_alg = rows_bytes(se["groups_f"] + se["groups_b"], T0)
    g_comp = rows_bytes(gi["fentries"] + gi["bentries"], T0)
    tail_ms = st["tail_ext_ms"] + st["tail_ms"]
    t_alg = rows_bytes(se["tail_f"] + se["tail_b"], NT)
    dense_bytes = 8 * (K * ldd + 2 * ldd)
    classes = {
        "k_group": {"ms_per_iteration": groups_ms, "launches_per_iteration": 2 * gi["ngroups"],
                    "algorithmic_bytes": g_alg, "achieved": _gbs(g_alg, groups_ms),
                    "composite_bytes": g_comp, "composite_overhead": g_comp / g_alg,
                    "traffic": read_traffic("k_group", "grid3_128")},
        "tail (k_tail_ext + k_tail_fwd + k_dense + k_tail_bwd)": {
            "ms_per_iteration": tail_ms, "algorithmic_bytes": t_alg, "achieved": _gbs(t_alg, tail_ms),
            "tail_vertices": NT, "sparse_steps": ti["nsteps"]},
        "k_dense": {"ms_per_iteration": dt["dense_ms"], "bytes_per_launch": dense_bytes, "K": K,
                    "achieved": _gbs(dense_bytes, dt["dense_ms"]),
                    "traffic": read_traffic("k_dense", "grid3_128")},
        "k_spmv_dot": {"ms_per_iteration": tm["spmv_ms"], "algorithmic_bytes": byts["spmv"],
                       "achieved": _gbs(byts["spmv"], tm["spmv_ms"]),
                       "traffic": read_traffic("k_spmv_dot", "grid3_128")},
    }
    for v in classes.values():
        v["frac"] = _frac(v["achieved"])
    dom = max(("k_group", "k_dense", "k_spmv_dot"), key=lambda k: classes[k]["ms_per_iteration"] or 0.0)
    apply_gbs = _gbs(byts["ldl"], tm["ldl_ms"])
    both_gbs = _gbs(byts["ldl"] + byts["spmv"], tm["ldl_ms"] + tm["spmv_ms"])
    d = classes[dom]
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": d["achieved"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": d["frac"],
        "traffic": d.get("traffic"),
        "workload": "plain FP64 solve of grid3(128), one PCG iteration",
        "classes": classes,
        "ldl_apply": {"ms_per_iteration": tm["ldl_ms"], "algorithmic_bytes": byts["ldl"], "achieved": apply_gbs,
                      "frac": _frac(apply_gbs)},
        "spmv_plus_ldl_apply": {"achieved": both_gbs, "frac": _frac(both_gbs)},
    }
    cpu_b = None
    if cpu:
        from oracle import oracle

        oracle.build()
        t0 = time.perf_counter()
        xo, io = oracle.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
        t_cpu = time.perf_counter() - t0
        cpu_b = {
            "value": 1.0 / t_cpu,
            "unit": "solves/s",
            "cores": 1,
            "kind": "port",
            "sample": f"one full solve of the same grid3(128) workload (graph, factor, b, tol 1e-8) by the oracle, "
                      f"the C restatement of pcg + LDLsolver, on one host core: {io['its']} iterations in "
                      f"{t_cpu:.2f} s ({t_cpu / max(io['its'], 1) * 1e3:.1f} ms/iteration)",
            "host": host_info(),
        }
        log(f"plain: cpu baseline {cpu_b}")
    ms = elapsed / nsolves * 1e3
    line = {
        "solves_per_s": nsolves / elapsed,
        "ms_per_solve": ms,
        "iterations": int(np.median(its)),
        "ms_per_iteration": ms / max(np.median(its), 1),
        "true_relres": relres,
        "n": n, "nnz_L": int(la.nnz), "nnzF": int(f.nnz),
        "levels": info["nlevels"], "level_groups": info["nwide"], "tail_vertices": NT,
        "level_group_budget": list(budget) if budget else "default", "chained": ctx.chain_info(),
        "dense_block_vertices": K, "tail_sparse_steps": ti["nsteps"],
        "per_iteration_ms": {"spmv": tm["spmv_ms"], "ldl_apply": tm["ldl_ms"], "vectors": tm["vec_ms"], **st, **dt},
        "factor_host_s": t_factor, "upload_plan_s": t_upload, "autotune": tune,
    }
    ctx.close()
    return {"line": line, "roofline": roofline, "cpu_baseline": cpu_b}


def mode_plain(args) -> int:
    out = plain_leg(args, args.steps, args.warmup, cpu=not args.no_cpu)
    ln = out["line"]
    emit({
        "metric": "solves/s (plain FP64 GPU PCG, relres<=1e-8)",
        "value": ln["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ln["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "grid3(128): exp(U(-2,2)) weights seed 2, 1 RHS N(0,1) sum-zero, tol 1e-8", **ln},
        "roofline": out["roofline"],
        "cpu_baseline": out["cpu_baseline"],
    })
    if args.calib:
        from lapgpu._native import DeviceArray, synchronize

        src = DeviceArray(1 << 26)
        dst = DeviceArray(1 << 26)
        dst.copy_from(src)
        synchronize()
        log("calibration copy: 536870912 bytes read, 536870912 bytes written")
    return 0


# ----------------------------------------------------------------- serve ---
def serve_leg(args) -> dict:
    from concurrent.futures import ThreadPoolExecutor

    import lapgpu
    from lapgpu import _native, workloads
    from lapgpu.serve import Request, Server, summary

    nreq = args.steps * args.step_requests
    t0 = time.time()
    G, reqs = workloads.config5(nreq=nreq, qps=args.qps, ngraphs=args.serve_graphs, min_n=args.serve_min_n,
                                max_n=args.serve_max_n, new_frac=args.new_frac, new_max_n=args.serve_new_max_n)
    warm = workloads.serve_requests(G, args.warmup * args.step_requests, qps=args.qps, min_n=args.serve_min_n,
                                    new_max_n=args.serve_new_max_n, new_frac=args.new_frac, rid0=10 ** 7)
    log(f"serve: {len(G)} resident graphs n {min(g.shape[0] for g in G)}..{max(g.shape[0] for g in G)} "
        f"(sum {sum(g.shape[0] for g in G)}); {nreq} timed requests ({sum(r.new_graph is not None for r in reqs)} "
        f"with new graphs) and {len(warm)} warm-up requests at {args.qps} QPS; generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    with ThreadPoolExecutor(8) as ex:
        F = list(ex.map(lambda g: lapgpu.build_ldli(g, seed=1), G))
    log(f"serve: resident graphs factored in {time.time() - t0:.1f}s")
    # solutions kept for the true-residual check: an even sample of the stream
    sample = set(r.rid for r in reqs[:: max(1, nreq // args.check)])
    srv = Server(0, hbm_budget=args.hbm_gib * 2**30, plain_workers=args.serve_workers,
                 factor_workers=args.factor_workers, batch_window=args.batch_window,
                 max_wait=args.max_wait, quick_workers=args.quick_workers,
                 blocking_waits=not args.spin_waits, wg_max_n=args.wg_max_n, keep_ids=sample,
                 new_wg_max_n=args.new_wg_max_n, context_precision=args.serve_precision,
                 exact_batches=args.exact_batches, resident_budget=args.resident_budget,
                 idle_sleep=args.idle_sleep)
    t0 = time.time()
    for i, (g, f) in enumerate(zip(G, F)):
        assert srv.register(i, g, ldli=f)
    log(f"serve: resident graphs uploaded in {time.time() - t0:.1f}s, HBM reserved {srv.hbm_used / 2**30:.1f} GiB")
    if args.replica_nnzf > 0:
        t0 = time.time()
        got = srv.add_replicas(args.replica_nnzf)
        log(f"serve: second contexts for graphs {got} (LDLinv >= {args.replica_nnzf} entries) in "
            f"{time.time() - t0:.1f}s, HBM reserved {srv.hbm_used / 2**30:.1f} GiB")
    # warm-up: one solve per resident graph (captures, workspaces), then the
    # warm-up stream
    wrng = np.random.default_rng(55)
    srv.run([Request(-1 - i, 0.0, wrng.standard_normal(G[i].shape[0]), graph=i) for i in range(len(G))],
            realtime=False)
    if warm:
        t0 = time.perf_counter()
        wres = srv.run(warm, progress=lambda d, t, e: log(f"  warm-up: {d} of {t} requests at {e:.0f} s"))
        log(f"serve: warm-up stream {len(warm)} requests in {time.perf_counter() - t0:.1f}s: "
            f"{summary(wres, time.perf_counter() - t0)}")
    import resource

    # the pipeline figures below cover the timed stream only
    for lst in (srv.factor_log, srv.factor_cpu_log, srv.upload_log, srv.upload_cpu_log, srv.wg_upload_log,
                srv.solve_log):
        lst.clear()
    ru0 = resource.getrusage(resource.RUSAGE_SELF)
    loop0 = srv.stats["loop_cpu_s"]
    t0 = time.perf_counter()
    res = srv.run(reqs, progress=lambda d, t, e: log(f"  served {d} of {t} requests at {e:.0f} s"))
    wall = time.perf_counter() - t0
    ru1 = resource.getrusage(resource.RUSAGE_SELF)
    cpu_s = (ru1.ru_utime - ru0.ru_utime) + (ru1.ru_stime - ru0.ru_stime)
    s = summary(res, wall)
    paths = {p: sum(o.path == p for o in res) for p in sorted({o.path for o in res})}
    log(f"serve: {s} paths {paths}; host CPU {cpu_s:.0f} s over {wall:.1f} s (dispatch loop "
        f"{srv.stats['loop_cpu_s'] - loop0:.0f} s, solve workers {sum(e[4] for e in srv.solve_log):.0f} s, "
        f"factoring {sum(srv.factor_cpu_log):.0f} s, uploads {sum(srv.upload_cpu_log):.0f} s)")
    for pth in paths:
        sel = [(r, o) for r, o in zip(reqs, res) if o.path == pth]
        lat = np.array([o.latency for _, o in sel])
        log(f"  path {pth}: {len(sel)} ({sum(r.new_graph is not None for r, _ in sel)} new graphs), latency p50 "
            f"{np.percentile(lat, 50):.2f} s p99 {np.percentile(lat, 99):.2f} s")
    # the registered graphs whose requests waited longest: graph, n,
    # requests, latency p50 / max, and when its last answer came
    per = {}
    for r, o in zip(reqs, res):
        if r.graph is not None:
            per.setdefault(r.graph, []).append((o.latency, r.arrival + o.latency))
    worst = sorted(per.items(), key=lambda kv: -max(lt for lt, _ in kv[1]))[:4]
    for gi, v in worst:
        lat = np.array([lt for lt, _ in v])
        log(f"  graph {gi} n={G[gi].shape[0]}: {len(v)} requests, latency p50 {np.median(lat):.2f} s max "
            f"{lat.max():.2f} s, last answer at {max(e for _, e in v):.1f} s")
    nf = np.array(srv.factor_log)
    up = np.array(srv.upload_log).reshape(-1, 4)
    wgu = np.array(srv.wg_upload_log).reshape(-1, 2)
    new_n = [r.new_graph.shape[0] for r in reqs if r.new_graph is not None]
    pipeline = {"new_graphs": len(new_n), "new_graph_vertices": int(sum(new_n)),
                "factor_s_median": float(np.median(nf)) if len(nf) else None,
                "factor_s_sum": float(nf.sum()) if len(nf) else None,
                "context_upload_s_median": float(np.median(up[:, 0])) if len(up) else None,
                "context_upload_s_sum": float(up[:, 0].sum()) if len(up) else None,
                "contexts_uploaded": int(len(up)),
                "batched_upload_s_median": float(np.median(wgu[:, 0])) if len(wgu) else None,
                "batched_upload_s_sum": float(wgu[:, 0].sum()) if len(wgu) else None,
                "batched_uploads": int(len(wgu)),
                "host_cpu_s": cpu_s, "factor_workers": args.factor_workers}
    log(f"serve: new-graph pipeline {pipeline}")
    # context solves (plain and multi-RHS) of the timed stream: count,
    # columns, and the worker-seconds they held
    sl = np.array([(n, k, dt) for n, k, dt, _, _ in srv.solve_log]).reshape(-1, 3)
    if len(sl):
        padded = sum(max(8, 1 << (int(k) - 1).bit_length()) for k in sl[:, 1] if k > 1)
        log(f"serve: context solves {len(sl)}, columns mean {sl[:, 1].mean():.1f} max {int(sl[:, 1].max())}, "
            f"worker-seconds {sl[:, 2].sum():.0f} over {wall:.1f} s; multi-RHS columns "
            f"{int(sl[sl[:, 1] > 1, 1].sum())} padded to {padded}")
        for lo, hi in ((5e4, 3e5), (3e5, 1e6), (1e6, 5e6)):
            m = (sl[:, 0] > lo) & (sl[:, 0] <= hi)
            if m.any():
                log(f"  n in ({lo:g},{hi:g}]: {int(m.sum())} solves, {int(sl[m, 1].sum())} columns, "
                    f"{sl[m, 2].sum():.0f} worker-s, median {np.median(sl[m, 2]):.2f} s")
    # true residual ||L x - b|| / ||b|| on the kept sample
    byr = {r.rid: r for r in reqs}
    rr = []
    for o in res:
        if o.x is None:
            continue
        r = byr[o.rid]
        la = lapgpu.lap(G[r.graph] if r.graph is not None else r.new_graph)
        rr.append(float(np.linalg.norm(la @ o.x - r.b) / np.linalg.norm(r.b)))
    check = {"checked": len(rr), "max_true_relres": max(rr) if rr else None,
             "above_tol": int(sum(v > 1e-8 for v in rr)), "max_recursive_relres": s["max_nr"]}
    log(f"serve: true residuals of the sample: {check}")
    # plain one-at-a-time FP64 GPU PCG on the same GPU: the measured plain
    # solve time of every resident graph (the faster of the library's
    # default level-group budget and a smaller one, which expanders prefer),
    # the same arrival stream through one FIFO server (new graphs: a log-log
    # fit of solve time against n, factoring excluded)
    svc = {}
    for i, (g, f) in enumerate(zip(G, F)):
        bb = np.random.default_rng(i).standard_normal(g.shape[0])
        bb -= bb.mean()
        best = float("inf")
        for budget in (None, (16, 4.0, 65536)):
            c = _native.Context(0)
            if budget is not None:
                c.set_group_budget(*budget)
            c.upload_lap(lapgpu.lap(g))
            c.upload_ldli(f)
            c.solve(bb, tol=1e-8, maxits=1000)
            t1 = time.perf_counter()
            c.solve(bb, tol=1e-8, maxits=1000)
            best = min(best, time.perf_counter() - t1)
            c.close()
        svc[i] = best
    sizes_arr = np.array([g.shape[0] for g in G], dtype=float)
    svc_arr = np.array([svc[i] for i in range(len(G))])
    coef = np.polyfit(np.log(sizes_arr), np.log(svc_arr), 1)
    t_free, busy, lat_plain = 0.0, 0.0, []
    for r in reqs:
        stt = svc[r.graph] if r.graph is not None else float(np.exp(np.polyval(coef, np.log(len(r.b)))))
        start = max(t_free, r.arrival)
        t_free = start + stt
        busy += stt
        lat_plain.append(t_free - r.arrival)
    plain_sps = len(reqs) / busy
    srv.close()
    return {
        "wall": wall, "nreq": nreq, "summary": s, "paths": paths, "check": check, "pipeline": pipeline,
        "plain": {"solves_per_s": plain_sps, "p50_ms": float(np.percentile(lat_plain, 50) * 1e3),
                  "p99_ms": float(np.percentile(lat_plain, 99) * 1e3),
                  "how": "modelled: FIFO replay of the same arrivals, service time = the measured plain FP64 "
                         "solve (lapgpu_solve, one at a time; the faster of two level-group budgets) of each "
                         "resident graph; new graphs by a log-log fit of those times against n, factoring "
                         "excluded"},
        "resident_n": [int(g.shape[0]) for g in G],
    }


def cache_leg(args) -> list:
    """Admission of a graph that comes with its request, with and without
    the factor/plan cache (lapgpu.cache): host seconds from the adjacency to
    a graph ready to solve, for graphs of the serving families at three
    sizes — without a cache (factor + plan + upload), into an empty cache
    (the same plus the writes) and from a warm one (read factor and plan,
    upload).  Graphs up to the batched-engine bound take the batched engine
    (the factor is cached, its whole-graph plan is rebuilt: tens of ms)."""
    import shutil
    import tempfile

    import lapgpu
    from lapgpu import _native, workloads
    from lapgpu.cache import GraphCache

    root = tempfile.mkdtemp(prefix="lapgpu_cache_")
    out = []
    batch = _native.Batch(0)
    try:
        cache = GraphCache(root)
        for i, nv in enumerate((1e5, 5e5, 2e6)):
            rng = np.random.default_rng([9, i])
            a = workloads.serve_graph(rng, nv)
            n = a.shape[0]
            wg = n <= args.new_wg_max_n
            settings = "new"

            def admit(la, f, use_cache, key):
                if wg:
                    g = batch.add_graph(la, f)
                    batch.remove_graph(g)
                    return
                ctx = _native.Context(0)
                ctx.set_group_budget(8, 2.0, 16384)
                ctx.set_dense_tail(0)
                if use_cache:
                    cache.upload(ctx, key, settings, la, f)
                else:
                    ctx.upload_lap(la)
                    ctx.upload_ldli(f)
                ctx.close()

            t0 = time.perf_counter()
            la = lapgpu.lap(a)
            f = lapgpu.build_ldli(a, seed=1)
            t1 = time.perf_counter()
            admit(la, f, False, None)
            t2 = time.perf_counter()
            key, la, f, hit = cache.factor(a, seed=1)
            t3 = time.perf_counter()
            admit(la, f, True, key)
            t4 = time.perf_counter()
            key, la, f, hit2 = cache.factor(a, seed=1)
            t5 = time.perf_counter()
            admit(la, f, True, key)
            t6 = time.perf_counter()
            row = {"n": int(n), "engine": "batched" if wg else "context",
                   "no_cache_s": {"factor": t1 - t0, "plan_upload": t2 - t1, "total": t2 - t0},
                   "cache_cold_s": {"factor": t3 - t2, "plan_upload": t4 - t3, "total": t4 - t2},
                   "cache_warm_s": {"factor": t5 - t4, "plan_upload": t6 - t5, "total": t6 - t4},
                   "hits": [bool(hit), bool(hit2)]}
            out.append(row)
            log(f"serve: admission n={n} ({row['engine']}): no cache {t2 - t0:.3f}s (factor {t1 - t0:.3f}), "
                f"cold cache {t4 - t2:.3f}s, warm cache {t6 - t4:.3f}s (factor {t5 - t4:.3f}, "
                f"plan+upload {t6 - t5:.3f})")
        out.append({"cache_bytes": cache.bytes()})
    finally:
        batch.close()
        shutil.rmtree(root, ignore_errors=True)
    return out


def mode_serve(args) -> int:
    # every resident graph owns a stream; HIP maps streams round-robin onto
    # its hardware queues (4 unless GPU_MAX_HW_QUEUES says otherwise), and two
    # busy streams on one queue run one after the other.  Read at HIP start-up,
    # so it is set before the library loads.
    if args.hw_queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    sv = serve_leg(args)
    cl = None if args.no_cache_leg else cache_leg(args)
    # the precision and batching forms on their own workloads, a few solves
    # each (their own modes run them at length)
    forms = None
    if not args.no_forms:
        import copy

        fa = copy.copy(args)
        fa.steps = args.form_steps
        forms = {}
        for name, leg in (("precision", precision_leg), ("batching", multi_leg)):
            t0 = time.time()
            ln = leg(fa)
            forms[name] = {"metric": ln["metric"], "value": ln["value"], "unit": ln["unit"],
                           "dtype": ln["dtype"], "steps": fa.steps, "config": ln["config"]}
            log(f"forms: {name} {ln['value']:.2f} {ln['unit']} in {time.time() - t0:.0f}s")
    pl = None if args.no_plain else plain_leg(args, 5, 1, cpu=not args.no_cpu)
    s = sv["summary"]
    emit({
        "metric": "served solves/s at relres<=1e-8 (p50/p99 latency) vs plain one-at-a-time FP64 GPU PCG",
        "value": s["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sv["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.serve_precision == "fp64" else "f64 (multi-RHS context solves: f64 outer / f32 inner)",
        "data": "synthetic",
        "config": {
            "workload": f"serving: {sv['nreq']} requests ({args.steps} steps of {args.step_requests}), Poisson "
                        f"{args.qps:g} QPS, {args.serve_graphs} resident graphs n log-uniform "
                        f"[{args.serve_min_n:g},{args.serve_max_n:g}] (2-D/3-D grids exp(U(-2,2)), path + 4 "
                        f"matchings log-uniform weights), {args.new_frac:.0%} with a new graph of the same "
                        f"families n <= {args.serve_new_max_n:g}, seed 5, tol 1e-8",
            "context_precision": args.serve_precision,
            "served": s,
            "paths": sv["paths"],
            "true_relres_sample": sv["check"],
            "new_graph_pipeline": sv["pipeline"],
            "plain": sv["plain"],
            "speedup_solves_per_s": s["solves_per_s"] / sv["plain"]["solves_per_s"],
            "p99_ratio_plain_over_served": sv["plain"]["p99_ms"] / s["p99_ms"],
            "plain_grid3_128": pl["line"] if pl else None,
            "admission_cache": cl,
            "forms": forms,
        },
        "roofline": pl["roofline"] if pl else None,
        "cpu_baseline": pl["cpu_baseline"] if pl else None,
    })
    return 0


# ----------------------------------------------------------------- multi ---
def multi_leg(args) -> dict:
    """The batching form on its workload (64-RHS block of grid2(1024) and
    256 small grids): the bench line of --mode multi."""
    import lapgpu
    from lapgpu import _native, workloads
    from lapgpu._native import DeviceArray, synchronize

    a, B, smalls, sb = workloads.config4(rhs=args.rhs, small=args.small)
    la = lapgpu.lap(a)
    n = a.shape[0]
    k = args.rhs
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
        ctx.solve_multi_device(dB.address, dX.address, n, k, tol=1e-8, maxits=1000, flags=_native.PROFILE)
        log(f"{pname}: 64-RHS block alone {t_multi * 1e3:.1f} ms ({max(results['multi']['its'])} its), "
            f"256 small grids alone {t_batch * 1e3:.1f} ms; per iteration {ctx.last_timings()} "
            f"{ctx.last_sweep_timings()} {ctx.last_dense_timings()}")
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
            "X": X,
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
    best = out_modes["fp32_inner"]  # the workload names FP64 outer / FP32 inner
    line = {
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
                        "(n log-uniform [1e3,1e5], aspect [1,8]), seed 4, tol 1e-8; the block and the small "
                        "grids run as two concurrent launches streams (not one launch)",
            "reported_mode": "fp32_inner",
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
    }
    ctx.close()
    batch.close()
    return line


def mode_multi(args) -> int:
    emit(multi_leg(args))
    return 0


# ------------------------------------------------------------- precision ---
def precision_leg(args) -> dict:
    """The precision form on its workload (path + 4 matchings, n = 1e6):
    FP32-inner against FP64 on the same factor, with each solve's FP64 true
    residual as the device reports it (lapgpu_last_true_relres): the bench
    line of --mode precision."""
    import lapgpu
    from lapgpu import _native, workloads
    from lapgpu._native import DeviceArray, synchronize

    a, b = workloads.config3()
    la = lapgpu.lap(a)
    n = a.shape[0]
    t0 = time.time()
    f = lapgpu.build_ldli(a, seed=1)
    log(f"path+4 matchings n={n} nnzL={la.nnz} nnzF={f.nnz} factor {time.time() - t0:.1f}s")
    ctx = _native.Context(0)
    # a smaller level-group budget than the default: its many narrow groups
    # near the tail are run as chained launches, the fastest single-RHS
    # schedule of this expander (FP64 63 ms against 75 ms per solve)
    ctx.set_group_budget(16, 4.0, 65536)
    ctx.upload_lap(la)
    ctx.upload_ldli(f)
    db = DeviceArray.from_host(b)
    dx = DeviceArray(n)
    res = {}
    for prec, name in ((_native.FP64, "fp64"), (_native.FP32_INNER, "fp32_inner")):
        ctx.set_precision(prec, 20)
        ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000)
        synchronize()
        ts, trs = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            r = ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000)
            synchronize()
            ts.append(time.perf_counter() - t0)
            trs.append(ctx.last_true_relres(1))
        x = dx.download()
        ctx.solve_device(db.address, dx.address, n, tol=1e-8, maxits=1000, flags=_native.PROFILE)
        log(f"{name} schedule {ctx.factor_info()} tail {ctx.tail_info()} dense {ctx.dense_info()} per-iteration ms "
            f"{ctx.last_timings()} {ctx.last_sweep_timings()} {ctx.last_dense_timings()}")
        res[name] = {"its": r["its"], "stop": r["stop"], "solve_ms": float(np.median(ts)) * 1e3,
                     "true_relres": float(np.linalg.norm(la @ x - b) / np.linalg.norm(b)),
                     "device_true_relres_per_solve": trs, "x": x}
        log(f"{name}: {res[name]['its']} its {res[name]['solve_ms']:.1f} ms relres {res[name]['true_relres']:.2e}")
    err = float(np.linalg.norm(res["fp32_inner"]["x"] - res["fp64"]["x"]) / np.linalg.norm(res["fp64"]["x"]))
    for v in res.values():
        v.pop("x")
    line = {
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
            "level_group_budget": [16, 4.0, 65536],
            "chain": ctx.chain_info(),
            "fp64": res["fp64"],
            "fp32_inner": res["fp32_inner"],
            "rel_err_vs_fp64": err,
            "iteration_ratio": res["fp32_inner"]["its"] / max(res["fp64"]["its"], 1),
            "speedup_vs_fp64": res["fp64"]["solve_ms"] / res["fp32_inner"]["solve_ms"],
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    ctx.close()
    return line


def mode_precision(args) -> int:
    emit(precision_leg(args))
    return 0


# --------------------------------------------------------------- batched ---
def mode_batched(args) -> int:
    """The batched engine alone: the 256 small unit 2-D grids of the
    batching workload (n log-uniform in [1e3, 1e5], seed 4), one request
    each, in one launch per step (k_wg_pcg, one workgroup per request)."""
    import lapgpu
    from lapgpu import _native, workloads
    from lapgpu._native import DeviceArray, synchronize
    from concurrent.futures import ThreadPoolExecutor

    _, _, smalls, sb = workloads.config4(rhs=1, small=args.small)
    with ThreadPoolExecutor(8) as ex:
        sf = list(ex.map(lambda g: lapgpu.build_ldli(g, seed=1), smalls))
    batch = _native.Batch(0)
    gids = [batch.add_graph(lapgpu.lap(g), f) for g, f in zip(smalls, sf)]
    dsb = [DeviceArray.from_host(v) for v in sb]
    dsx = [DeviceArray(len(v)) for v in sb]

    def run():
        return batch.solve_device(gids, [d.address for d in dsb], [d.address for d in dsx], tol=1e-8,
                                  maxits=1000)

    for _ in range(args.warmup):
        run()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos = run()
    synchronize()
    wall = time.perf_counter() - t0
    rel = max(float(np.linalg.norm(lapgpu.lap(g) @ d.download() - v) / np.linalg.norm(v))
              for g, d, v in zip(smalls, dsx, sb))
    n_tot = sum(g.shape[0] for g in smalls)
    emit({
        "metric": "solves/s (batched engine, 256 small grids per launch, relres<=1e-8)",
        "value": len(smalls) * args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(smalls)} unit 2-D grids, n log-uniform [1e3,1e5] (total {n_tot}), aspect "
                               "[1,8], 1 RHS each N(0,1) sum-zero, seed 4, tol 1e-8, one launch per step",
                   "its_max": int(max(i["its"] for i in infos)), "true_relres_max": rel},
        "roofline": None,
        "cpu_baseline": None,
    })
    batch.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="serve: timed steps of --step-requests requests; plain/multi/precision: timed solves")
    ap.add_argument("--warmup", type=int, default=2, help="serve: untimed steps of another stream first")
    ap.add_argument("--mode", default="serve", choices=["serve", "plain", "multi", "precision", "batched"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="serve: skip the plain grid3(128) leg")
    ap.add_argument("--no-cache-leg", action="store_true", help="serve: skip the admission-cache measurement")
    ap.add_argument("--no-forms", action="store_true", help="serve: skip the precision and batching lines")
    ap.add_argument("--form-steps", type=int, default=5, help="serve: timed solves of each form")
    ap.add_argument("--step-requests", type=int, default=STEP_REQUESTS)
    ap.add_argument("--check", type=int, default=200, help="serve: requests whose true residual is checked")
    ap.add_argument("--rhs", type=int, default=64)
    ap.add_argument("--small", type=int, default=256)
    ap.add_argument("--qps", type=float, default=300.0)
    ap.add_argument("--serve-graphs", type=int, default=32)
    ap.add_argument("--serve-min-n", type=float, default=1e4)
    ap.add_argument("--serve-max-n", type=float, default=4e6)
    ap.add_argument("--serve-new-max-n", type=float, default=2e5)
    ap.add_argument("--new-frac", type=float, default=0.2)
    ap.add_argument("--serve-workers", type=int, default=10, help="serving: graphs solved concurrently")
    ap.add_argument("--factor-workers", type=int, default=12, help="serving: host threads factoring new graphs")
    ap.add_argument("--resident-budget", type=lambda v: tuple(t(x) for t, x in zip((int, float, int), v.split(","))),
                    default=(16, 4.0, 65536),
                    help="serving: level-group budget depth,growth,extra of the registered graphs' contexts")
    ap.add_argument("--hbm-gib", type=float, default=225, help="serving: HBM budget of graphs and workspaces (GiB)")
    ap.add_argument("--replica-nnzf", type=int, default=60_000_000,
                    help="serving: registered graphs whose LDLinv has at least this many entries get a second "
                         "context (0: none)")
    ap.add_argument("--idle-sleep", type=float, default=5e-4,
                    help="serving: seconds the dispatch loop pauses after a pass that changed nothing")
    ap.add_argument("--exact-batches", action="store_true",
                    help="serving: multi-RHS batches of exactly 8/16/32/64 requests (no padding columns)")
    ap.add_argument("--serve-precision", default="fp32", choices=["fp64", "fp32"],
                    help="serving: precision of the context (multi-RHS) solves (fp32: the FP32-inner mode); "
                         "both converge on FP64 true residuals")
    ap.add_argument("--new-wg-max-n", type=int, default=200000,
                    help="serving: largest graph that comes with its request served by the batched engine")
    ap.add_argument("--wg-max-n", type=int, default=200000,
                    help="serving: graphs up to this size are solved by the batched engine")
    ap.add_argument("--spin-waits", action="store_true",
                    help="serving: solve workers spin while waiting for the device (default: they sleep)")
    ap.add_argument("--quick-workers", type=int, default=6,
                    help="serving: workers for solves of graphs with n <= 3e5 (0: share the others)")
    ap.add_argument("--batch-window", type=float, default=2.0, help="serving: wait for more requests, x last solve time")
    ap.add_argument("--max-wait", type=float, default=3.0, help="serving: longest wait for more requests (s)")
    ap.add_argument("--hw-queues", type=int, default=16, help="serving: HIP hardware queues (0: HIP's default)")
    ap.add_argument("--budget", type=lambda v: (int(v.split(",")[0]), float(v.split(",")[1]), int(v.split(",")[2])),
                    default=None, help="plain: level-group budget depth,growth,extra (default: the library's)")
    ap.add_argument("--autotune", action="store_true",
                    help="plain: tune the schedule on the graph first (lapgpu_autotune)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="multi: skip the schedule autotuning (done by default)")
    ap.add_argument("--calib", action="store_true",
                    help="plain mode: end with a device-to-device copy of 512 MiB (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark measures one GPU; --gpus must be 1"}))
        return 2
    return {"plain": mode_plain, "multi": mode_multi, "precision": mode_precision, "serve": mode_serve,
            "batched": mode_batched}[args.mode](args)


if __name__ == "__main__":
    sys.exit(main())
