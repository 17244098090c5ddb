"""Serving loop (lapgpu.serve.Server): requests of uneven sizes arrive over
time, are admitted while others run, and overlap on the device through
batched launches, multi-RHS solves and the plain path.  Each result must
meet tol on its true residual and be the same as the request served alone:
bit-for-bit in iterations and to 1e-10 relative for the batched path (one
workgroup per request, nothing shared), and for multi-RHS solves within the
stated tolerance — same stop reason, iterations within +-2 and
||x_mix - x_alone|| <= 1e-6 ||x|| (the columns share launches, and the
reductions group their terms differently than the one-RHS kernels)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from lapgpu.serve import Request, Server, summary
from _bars import assert_relres

pytestmark = pytest.mark.gpu


def _rhs(n, rng):
    b = rng.standard_normal(n)
    return b - b.mean()


def _graphs():
    rng = np.random.default_rng(0)
    return {
        "g2": graphs.weighted_grid2(40, 30, rng),
        "g3": graphs.weighted_grid3(14, 14, 14, rng),
        "pm": graphs.path_matchings(5000, 4, rng),
        "big": graphs.weighted_grid3(32, 32, 32, rng),  # above wg_max_n below: plain path
    }


def test_serve_mixed_stream_matches_alone():
    G = _graphs()
    srv = Server(0, wg_max_n=20000, keep_x=True)
    try:
        for k, a in G.items():
            assert srv.register(k, a)
        rng = np.random.default_rng(5)
        keys = list(G)
        reqs = []
        t = 0.0
        for rid in range(40):
            t += rng.exponential(0.002)
            k = keys[rng.integers(len(keys))]
            reqs.append(Request(rid, t, _rhs(G[k].shape[0], rng), graph=k, tol=1e-8))
        # one request brings a new graph that needs host factoring
        newg = graphs.weighted_grid2(25, 25, np.random.default_rng(9))
        reqs.append(Request(40, t + 0.001, _rhs(625, rng), new_graph=newg, tol=1e-8))
        res = srv.run(reqs)
        s = summary(res, max(r.arrival for r in reqs))
        assert s["solves"] == 41
        for r, out in zip(reqs, res):
            a = G[r.graph] if r.graph is not None else newg
            la = lapgpu.lap(a)
            assert out.stop == "converged", out
            assert_relres(la, out.x, r.b, 1e-8, lambda: lapgpu.build_ldli(a, seed=1))
        paths = {o.path for o in res}
        assert "batched" in paths and paths & {"plain", "multi"}, paths
        # the same requests alone
        checked = set()
        for r, out in zip(reqs, res):
            if r.graph is None or (out.path in checked and out.path != "multi"):
                continue
            checked.add(out.path)
            alone = srv.run([Request(1000 + r.rid, 0.0, r.b, graph=r.graph, tol=1e-8)])[0]
            assert alone.stop == out.stop
            if out.path == "multi":
                assert abs(alone.its - out.its) <= 2
                assert np.linalg.norm(alone.x - out.x) <= 1e-6 * np.linalg.norm(out.x)
            else:
                assert alone.its == out.its
                assert np.linalg.norm(alone.x - out.x) <= 1e-10 * np.linalg.norm(out.x)
    finally:
        srv.close()


def test_serve_burst_uses_multi_rhs():
    """Requests queued on one large graph are solved together (multi-RHS)
    and each matches its solve alone within the stated tolerance."""
    rng = np.random.default_rng(8)
    a = graphs.weighted_grid3(30, 30, 30, rng)
    srv = Server(0, wg_max_n=20000, keep_x=True)
    try:
        assert srv.register("g", a)
        reqs = [Request(i, 0.0, _rhs(a.shape[0], rng), graph="g", tol=1e-8) for i in range(20)]
        res = srv.run(reqs)
        assert sum(o.path == "multi" for o in res) >= 10
        la = lapgpu.lap(a)
        for r, out in zip(reqs, res):
            assert out.stop == "converged"
            assert_relres(la, out.x, r.b, 1e-8, lambda: lapgpu.build_ldli(a, seed=1))
        for r, out in list(zip(reqs, res))[:4]:
            alone = srv.run([Request(100 + r.rid, 0.0, r.b, graph="g", tol=1e-8)])[0]
            assert alone.path == "plain"
            assert abs(alone.its - out.its) <= 2
            assert np.linalg.norm(alone.x - out.x) <= 1e-6 * np.linalg.norm(out.x)
    finally:
        srv.close()


def test_admission_by_hbm_budget():
    G = _graphs()
    srv = Server(0, wg_max_n=20000, hbm_budget=4 * 2**20)
    try:
        assert srv.register("g2", G["g2"], resident=True)
        # far too large for the budget
        assert not srv.register("big", G["big"], resident=False)
        assert srv.hbm_used <= srv.hbm_budget
    finally:
        srv.close()


def test_serve_mixed_tolerances_each_meet_their_own():
    """Requests with different tol and maxits queued together for batched
    launches: each is solved to its own tol (its true residual, the bar of tests/_bars.py) and with its
    own maxits, as if served alone."""
    rng = np.random.default_rng(17)
    a = graphs.weighted_grid2(40, 30, rng)
    la = lapgpu.lap(a)
    srv = Server(0, wg_max_n=20000, keep_x=True)
    try:
        assert srv.register("g", a)
        tols = [1e-4, 1e-10, 1e-6, 1e-10, 1e-4, 1e-8]
        reqs = [Request(i, 0.0, _rhs(1200, rng), graph="g", tol=t) for i, t in enumerate(tols)]
        reqs.append(Request(len(tols), 0.0, _rhs(1200, rng), graph="g", tol=1e-12, maxits=3))
        res = srv.run(reqs, realtime=False)
        assert {o.path for o in res} == {"batched"}
        for r, out in zip(reqs, res):
            if r.maxits == 3:
                assert out.stop == "maxits" and out.its == 3, out
                continue
            assert out.stop == "converged", out
            assert_relres(la, out.x, r.b, r.tol, lambda: lapgpu.build_ldli(a, seed=1))
            alone = srv.run([Request(100 + r.rid, 0.0, r.b, graph="g", tol=r.tol)])[0]
            assert alone.its == out.its and np.array_equal(alone.x, out.x), (r.tol, alone.its, out.its)
    finally:
        srv.close()


def test_serve_new_graphs_take_both_engines():
    """Graphs that come with their request: up to new_wg_max_n vertices on
    the batched engine (above the resident threshold wg_max_n), larger ones
    on a context of their own; both are factored and uploaded on worker
    threads, and each result meets its tol on the true residual."""
    rng = np.random.default_rng(21)
    mid = graphs.weighted_grid2(60, 50, rng)  # 3,000: batched as a new graph, a context if resident
    big = graphs.weighted_grid3(20, 20, 20, rng)  # 8,000: a context either way
    srv = Server(0, wg_max_n=2000, new_wg_max_n=5000, keep_x=True)
    try:
        reqs = [Request(0, 0.0, _rhs(3000, rng), new_graph=mid, tol=1e-8),
                Request(1, 0.0, _rhs(8000, rng), new_graph=big, tol=1e-8),
                Request(2, 0.001, _rhs(3000, rng), new_graph=mid, tol=1e-8)]
        res = srv.run(reqs)
        assert [o.path for o in res] == ["batched", "plain", "batched"], [o.path for o in res]
        for r, out in zip(reqs, res):
            assert out.stop == "converged", out
            a = r.new_graph
            assert_relres(lapgpu.lap(a), out.x, r.b, 1e-8, lambda: lapgpu.build_ldli(a, seed=1))
        # the one-off graphs are dropped after their solves
        assert not srv.graphs and srv.hbm_used == 0
        # the batched result is the engine's own answer for that factor
        g = srv.batch.add_graph(lapgpu.lap(mid), lapgpu.build_ldli(mid, seed=1))
        xs, infos = srv.batch.solve([g], [reqs[0].b], tol=1e-8)
        assert infos[0]["its"] == res[0].its and np.array_equal(xs[0], res[0].x)
        srv.batch.remove_graph(g)
    finally:
        srv.close()


def test_serve_replica_contexts():
    """A registered graph with a large factor gets a second context
    (add_replicas): a burst of requests for it is served by two multi-RHS
    solves at once, each answer at the bar and, like any multi-RHS answer,
    the same as the request alone to the stated tolerance."""
    rng = np.random.default_rng(21)
    a = graphs.path_matchings(30000, 4, rng)
    srv = Server(0, wg_max_n=1000, keep_x=True, replica_batch=4, max_rhs=8, batch_window=0.0, max_wait=0.0)
    try:
        assert srv.register("pm", a)
        assert srv.add_replicas(1) == ["pm"]
        g = srv.graphs["pm"]
        assert len(g.replicas) == 2 and g.kind == "plain"
        reqs = [Request(rid, 0.0, _rhs(a.shape[0], rng), graph="pm", tol=1e-8) for rid in range(40)]
        res = srv.run(reqs, realtime=False)
        la = lapgpu.lap(a)
        for r, out in zip(reqs, res):
            assert out.stop == "converged", out
            assert_relres(la, out.x, r.b, 1e-8, lambda: lapgpu.build_ldli(a, seed=1))
        # two solves of the graph ran at once, one on each context
        assert srv.stats["max_busy_contexts"] == 2
        alone = srv.run([Request(100, 0.0, reqs[7].b, graph="pm", tol=1e-8)])[0]
        assert alone.stop == res[7].stop and abs(alone.its - res[7].its) <= 2
        assert np.linalg.norm(alone.x - res[7].x) <= 1e-6 * np.linalg.norm(res[7].x)
    finally:
        srv.close()
