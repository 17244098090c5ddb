"""Cancellation: a request withdrawn while it runs stops at an iteration
boundary with stop "cancelled" and its best iterate, and every other
request of the same launch or multi-RHS solve is unchanged — bit for bit in
x and iterations against the same requests run without the cancel (each
batched request has its own workgroup; each multi-RHS column its own
scalars, and the columns' shared launches skip stopped columns).  There is
no reference counterpart (the reference's only interruption is maxtime,
src/pcg.jl:425); the expected values are the engine's own uncancelled
results on the same inputs."""

import threading
import time

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs
from lapgpu.serve import Request, Server

pytestmark = pytest.mark.gpu


def _rhs(n, rng):
    b = rng.standard_normal(n)
    return b - b.mean()


def _relres(la, x, b):
    return float(np.linalg.norm(la @ x - b) / np.linalg.norm(b))


def test_batch_cancel_one_request_others_bit_exact():
    rng = np.random.default_rng(31)
    big = graphs.weighted_grid2(300, 300, rng)  # 90,000 vertices: y in global memory, tens of ms
    small = graphs.weighted_grid2(60, 50, rng)
    B = _native.Batch(0)
    try:
        gb = B.add_graph(lapgpu.lap(big), lapgpu.build_ldli(big, seed=1))
        gs = B.add_graph(lapgpu.lap(small), lapgpu.build_ldli(small, seed=1))
        ids = [gs, gb, gb, gs, gb, gs]
        bs = [_rhs(small.shape[0] if g == gs else big.shape[0], rng) for g in ids]
        xs0, inf0 = B.solve(ids, bs, tol=1e-12)
        assert all(i["stop"] == "converged" for i in inf0)
        h = B.submit_host(ids, bs, tol=1e-12)
        B.cancel(h, 2)
        with pytest.raises(_native.LapGPUError):
            B.cancel(h, 6)  # out of range
        xs, inf = B.collect(h)
        assert inf[2]["stop"] == "cancelled", inf[2]
        assert inf[2]["its"] < inf0[2]["its"]
        assert np.all(np.isfinite(xs[2]))
        # the best iterate: no worse than x = 0
        assert _relres(lapgpu.lap(big), xs[2], bs[2]) <= 1.0 + 1e-12
        for i in (0, 1, 3, 4, 5):
            assert inf[i]["its"] == inf0[i]["its"] and inf[i]["stop"] == inf0[i]["stop"]
            assert np.array_equal(xs[i], xs0[i]), i
        # a cancel after the request finished leaves its result
        h = B.submit_host(ids[:2], bs[:2], tol=1e-12)
        while not B.query(h):
            time.sleep(1e-3)
        B.cancel(h, 0)
        xs2, inf2 = B.collect(h)
        assert inf2[0]["stop"] == "converged" and np.array_equal(xs2[0], xs0[0])
    finally:
        B.close()


def _cancel_when_running(ctx, column, out):
    # spin until the solve is running, then cancel
    t_end = time.perf_counter() + 30.0
    while time.perf_counter() < t_end:
        if ctx.cancel(column):
            out.append(True)
            return
    out.append(False)


def test_context_cancel_single_solve():
    rng = np.random.default_rng(32)
    a = graphs.weighted_grid3(64, 64, 64, rng)
    la = lapgpu.lap(a)
    b = _rhs(a.shape[0], rng)
    ctx = lapgpu.Context(0)
    try:
        ctx.upload_lap(la)
        ctx.upload_ldli(lapgpu.build_ldli(a, seed=1))
        assert ctx.cancel() is False  # nothing running
        x0, i0 = ctx.solve(b, tol=1e-12, maxits=1000)
        assert i0["stop"] == "converged"
        out = []
        th = threading.Thread(target=_cancel_when_running, args=(ctx, 0, out))
        th.start()
        x, info = ctx.solve(b, tol=1e-12, maxits=1000)
        th.join()
        assert out == [True]
        assert info["stop"] == "cancelled", info
        assert info["its"] < i0["its"]
        assert _relres(la, x, b) <= 1.0 + 1e-12
        with pytest.raises(ValueError):
            ctx.cancel(-2)
        # the context solves normally afterwards
        x1, i1 = ctx.solve(b, tol=1e-12, maxits=1000)
        assert i1["its"] == i0["its"] and np.array_equal(x1, x0)
    finally:
        ctx.close()


def test_context_cancel_one_column_others_bit_exact():
    rng = np.random.default_rng(33)
    a = graphs.weighted_grid3(48, 48, 48, rng)
    la = lapgpu.lap(a)
    n = a.shape[0]
    Bm = np.asfortranarray(np.stack([_rhs(n, rng) for _ in range(8)], axis=1))
    ctx = lapgpu.Context(0)
    try:
        ctx.upload_lap(la)
        ctx.upload_ldli(lapgpu.build_ldli(a, seed=1))
        X0, i0 = ctx.solve_multi(Bm, tol=1e-12, maxits=1000)
        assert set(i0["stop"]) == {"converged"}
        out = []
        th = threading.Thread(target=_cancel_when_running, args=(ctx, 3, out))
        th.start()
        X, info = ctx.solve_multi(Bm, tol=1e-12, maxits=1000)
        th.join()
        assert out == [True]
        assert info["stop"][3] == "cancelled" and info["its"][3] < i0["its"][3]
        assert _relres(la, X[:, 3], Bm[:, 3]) <= 1.0 + 1e-12
        for j in range(8):
            if j == 3:
                continue
            assert info["stop"][j] == "converged" and info["its"][j] == i0["its"][j]
            assert np.array_equal(X[:, j], X0[:, j]), j
    finally:
        ctx.close()


def test_server_cancel_queued_and_running():
    """Server.cancel: a request cancelled before it arrives is answered at
    once; one cancelled while its batched launch runs stops early; the
    other requests are the same as served alone."""
    rng = np.random.default_rng(34)
    a = graphs.weighted_grid2(300, 300, rng)
    srv = Server(0, wg_max_n=100_000, keep_x=True)
    try:
        assert srv.register("g", a)
        reqs = [Request(i, 0.0, _rhs(a.shape[0], rng), graph="g", tol=1e-12) for i in range(4)]
        reqs.append(Request(4, 0.2, _rhs(a.shape[0], rng), graph="g", tol=1e-12))
        srv.cancel(4)  # before it arrives

        out = []

        def cancel_rid1():
            t_end = time.perf_counter() + 30.0
            while time.perf_counter() < t_end:
                if srv.on_device(1):
                    srv.cancel(1)
                    out.append(True)
                    return
                time.sleep(1e-4)
            out.append(False)

        th = threading.Thread(target=cancel_rid1)
        th.start()
        res = srv.run(reqs)
        th.join()
        assert out == [True]
        assert res[4].stop == "cancelled" and res[4].path == "queued" and res[4].x is None
        assert res[1].stop == "cancelled" and res[1].path == "batched"
        for i in (0, 2, 3):
            alone = srv.run([Request(100 + i, 0.0, reqs[i].b, graph="g", tol=1e-12)])[0]
            assert res[i].stop == "converged" and alone.its == res[i].its
            assert np.array_equal(alone.x, res[i].x)
        assert res[1].its < res[0].its
    finally:
        srv.close()
