"""Parity at the measured workloads' full sizes (lapgpu.workloads), through
the C-ABI, against the oracle on the same LDLinv:

* config 2, grid3(128) (n = 2,097,152), plain FP64: SpMV and one LDLinv
  apply to 1e-12 relative, the solve with the same stop reason, iterations
  within +-2, x within 1e-6 relative of the oracle's, true relres at the
  bar of tests/_bars.py;
* config 3, path + 4 matchings (n = 1e6), FP32-inner: converged on its
  FP64 true residual, the bar above, error against the oracle's FP64 x at
  most 1e-6, at most 2 iterations more than the oracle; FP64 as config 2;
* config 4, grid2(1024) with a 64-column block in the FP32-inner mode
  after autotuning (the benchmark's path): every column at the bar, five
  columns against the oracle as config 3; the 256 small grids through the
  batched engine, each against the oracle as config 2;
* config 5, a 300-request stream over its families and size range (new
  graphs up to 2e5 vertices): every result at the bar, and a sample served
  alone equal to the mix (batched: same iterations, x to 1e-10; plain / multi-RHS: same stop,
  iterations +-2, x within 1e-6).
"""

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, workloads
from _bars import assert_relres, relres

pytestmark = pytest.mark.gpu

TOL = 1e-8
X_RTOL = 1e-6


def _close(x, xo):
    return np.linalg.norm(x - xo) <= X_RTOL * np.linalg.norm(xo)


def test_config2_grid3_128_plain(gpu_ctx_factory, oracle_mod):
    a, b = workloads.config2()
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    ctx = gpu_ctx_factory(la, f)
    v = np.random.default_rng(7).standard_normal(la.shape[0])
    ref = oracle_mod.spmv(la, v)
    assert np.linalg.norm(ctx.spmv(v) - ref) <= 1e-12 * np.linalg.norm(ref)
    zo = oracle_mod.ldlsolver(f, b)
    assert np.linalg.norm(ctx.apply_ldli(b) - zo) <= 1e-12 * np.linalg.norm(zo)
    x, info = ctx.solve(b, tol=TOL, maxits=1000, stag_test=5)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=TOL, maxits=1000, stag_test=5)
    print(f"\nconfig 2: its {info['its']} (oracle {io['its']}), relres {relres(la, x, b):.3e}")
    assert info["stop"] == io["stop"] == "converged"
    assert abs(info["its"] - io["its"]) <= 2
    assert _close(x, xo)
    assert_relres(la, x, b, TOL, f)


def test_config3_pathmatch_fp32_inner(gpu_ctx_factory, oracle_mod):
    a, b = workloads.config3()
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=TOL, maxits=1000, stag_test=5)
    ctx = gpu_ctx_factory(la, f)
    x64, i64 = ctx.solve(b, tol=TOL, maxits=1000, stag_test=5)
    assert i64["stop"] == io["stop"] == "converged" and abs(i64["its"] - io["its"]) <= 2
    assert _close(x64, xo)
    assert_relres(la, x64, b, TOL, f)
    ctx.set_precision(_native.FP32_INNER, 20)
    x32, i32 = ctx.solve(b, tol=TOL, maxits=1000, stag_test=5)
    err = np.linalg.norm(x32 - xo) / np.linalg.norm(xo)
    print(f"\nconfig 3: FP32-inner its {i32['its']}, oracle FP64 {io['its']}, error vs FP64 {err:.2e}")
    assert i32["stop"] == "converged"
    assert i32["its"] <= io["its"] + 2
    assert err <= X_RTOL
    assert_relres(la, x32, b, TOL, f)


def test_config4_block_and_small_grids(gpu_ctx_factory, oracle_mod):
    a, B, smalls, sb = workloads.config4()
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    ctx = gpu_ctx_factory(la, f)
    ctx.autotune(k=B.shape[0], groups=True)
    ctx.set_precision(_native.FP32_INNER, 20)
    X, info = ctx.solve_multi(B.T, tol=TOL, maxits=1000, stag_test=5)
    assert all(s == "converged" for s in info["stop"]), info["stop"]
    for j in range(B.shape[0]):
        assert_relres(la, X[:, j], B[j], TOL, f, what=f"column {j}")
    cols = (0, 1, 31, 62, 63)
    with ThreadPoolExecutor(len(cols)) as ex:
        outs = list(ex.map(lambda j: oracle_mod.approxchol_solve(la, f, B[j], tol=TOL, maxits=1000, stag_test=5), cols))
    for j, (xo, io) in zip(cols, outs):
        assert info["its"][j] <= io["its"] + 2, (j, info["its"][j], io)
        assert _close(X[:, j], xo), j
    # the small grids, one workgroup each
    sl = [lapgpu.lap(g) for g in smalls]
    with ThreadPoolExecutor(8) as ex:
        sf = list(ex.map(lambda g: lapgpu.build_ldli(g, seed=1), smalls))
    batch = lapgpu.Batch(0)
    try:
        gids = [batch.add_graph(l, ff) for l, ff in zip(sl, sf)]
        xs, infos = batch.solve(gids, sb, tol=TOL, maxits=1000, stag_test=5)
    finally:
        batch.close()
    with ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(lambda i: oracle_mod.approxchol_solve(sl[i], sf[i], sb[i], tol=TOL, maxits=1000,
                                                                stag_test=5), range(len(smalls))))
    for i, ((xo, io), x, inf) in enumerate(zip(ref, xs, infos)):
        assert inf["stop"] == io["stop"] and abs(inf["its"] - io["its"]) <= 2, (i, inf, io)
        assert _close(x, xo), i
        assert_relres(sl[i], x, sb[i], TOL, sf[i], what=f"small grid {i}")


def test_config5_short_stream():
    from lapgpu.serve import Request, Server

    G, reqs = workloads.config5(nreq=300, ngraphs=8, new_max_n=2e5)
    with ThreadPoolExecutor(8) as ex:
        F = list(ex.map(lambda g: lapgpu.build_ldli(g, seed=1), G))
    srv = Server(0, keep_x=True, plain_workers=4, quick_workers=4, factor_workers=8)
    try:
        for i, (g, ff) in enumerate(zip(G, F)):
            assert srv.register(i, g, ldli=ff)
        res = srv.run(reqs)
        assert len(res) == len(reqs)
        for r, out in zip(reqs, res):
            g = G[r.graph] if r.graph is not None else r.new_graph
            assert out.stop == "converged", out
            assert_relres(lapgpu.lap(g), out.x, r.b, TOL,
                          F[r.graph] if r.graph is not None else (lambda g=g: lapgpu.build_ldli(g, seed=1)),
                          what=f"request {r.rid} ({out.path})")
        paths = {o.path for o in res}
        assert {"batched", "multi"} <= paths, paths
        # a sample served alone: the first request of every (graph, path)
        seen = set()
        for r, out in zip(reqs, res):
            if r.graph is None or (r.graph, out.path) in seen:
                continue
            seen.add((r.graph, out.path))
            alone = srv.run([Request(10_000 + r.rid, 0.0, r.b, graph=r.graph, tol=TOL)])[0]
            assert alone.stop == out.stop
            if out.path == "batched":
                assert alone.its == out.its, r.rid
                assert np.linalg.norm(alone.x - out.x) <= 1e-10 * np.linalg.norm(out.x), r.rid
            else:
                assert abs(alone.its - out.its) <= 2, (r.rid, alone.its, out.its)
                assert _close(out.x, alone.x), r.rid
    finally:
        srv.close()
