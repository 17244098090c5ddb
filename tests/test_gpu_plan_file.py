"""Plan files on the device (lapgpu_save_plan / lapgpu_load_plan) and the
serving cache: a context loaded from a plan file solves like the one that
wrote it — bit for bit (the same plan, fixed-order sums) and against the
oracle at the parity bar — and a graph that comes back to the server is
admitted from its cached factor and plan with the same result."""

import os

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from lapgpu.serve import Request, Server
from _bars import assert_relres

pytestmark = pytest.mark.gpu


def test_solve_from_loaded_plan_matches_oracle(tmp_path, oracle_mod):
    rng = np.random.default_rng(51)
    a = graphs.weighted_grid3(30, 30, 30, rng)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    b = rng.standard_normal(a.shape[0])
    b -= b.mean()
    p = str(tmp_path / "g.plan")
    c1 = lapgpu.Context(0)
    c2 = lapgpu.Context(0)
    try:
        c1.upload_lap(la)
        c1.upload_ldli(f)
        x1, i1 = c1.solve(b, tol=1e-8, maxits=1000)
        c1.save_plan(p)
        c2.load_plan(p)
        assert c2.n == a.shape[0]
        assert c2.dense_info() == c1.dense_info() and c2.group_info() == c1.group_info()
        x2, i2 = c2.solve(b, tol=1e-8, maxits=1000)
        assert i2 == i1 and np.array_equal(x2, x1)
        B = np.asfortranarray(np.stack([b, -b], axis=1))
        X1, j1 = c1.solve_multi(B, tol=1e-8, maxits=1000)
        X2, j2 = c2.solve_multi(B, tol=1e-8, maxits=1000)
        assert j2 == j1 and np.array_equal(X2, X1)
        xo, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
        assert i2["stop"] == io["stop"] == "converged" and abs(i2["its"] - io["its"]) <= 2
        assert np.linalg.norm(x2 - xo) <= 1e-4 * np.linalg.norm(xo)
        assert_relres(la, x2, b, 1e-8, f)
        # a damaged file is refused and leaves the context as it was
        raw = bytearray(open(p, "rb").read())
        raw[len(raw) // 3] ^= 1
        bad = str(tmp_path / "bad.plan")
        open(bad, "wb").write(bytes(raw))
        with pytest.raises(lapgpu.LapGPUError, match="checksum"):
            c2.load_plan(bad)
        x3, i3 = c2.solve(b, tol=1e-8, maxits=1000)
        assert np.array_equal(x3, x1)
    finally:
        c1.close()
        c2.close()


def test_server_readmits_from_cache(tmp_path):
    """A graph that comes with two requests, far enough apart that it is
    dropped in between, is factored and planned once: the second admission
    loads both from the cache, and both answers are the same."""
    rng = np.random.default_rng(52)
    big = graphs.weighted_grid3(24, 24, 24, rng)  # 13,824: a context (new_wg_max_n below)
    small = graphs.weighted_grid2(40, 30, rng)  # batched engine
    n1, n2 = big.shape[0], small.shape[0]
    b1 = rng.standard_normal(n1)
    b2 = rng.standard_normal(n2)
    srv = Server(0, wg_max_n=2000, new_wg_max_n=5000, keep_x=True, cache_dir=str(tmp_path / "cache"))
    try:
        res1 = srv.run([Request(0, 0.0, b1 - b1.mean(), new_graph=big),
                        Request(1, 0.0, b2 - b2.mean(), new_graph=small)])
        assert not srv.graphs  # dropped after their solves
        st = dict(srv.cache.stats)
        assert st["factor_misses"] == 2 and st["plan_misses"] == 1 and st["factor_hits"] == 0
        res2 = srv.run([Request(2, 0.0, b1 - b1.mean(), new_graph=big),
                        Request(3, 0.0, b2 - b2.mean(), new_graph=small)])
        st = srv.cache.stats
        assert st["factor_hits"] == 2 and st["plan_hits"] == 1
        assert [o.path for o in res2] == [o.path for o in res1] == ["plain", "batched"]
        for o1, o2 in zip(res1, res2):
            assert o1.stop == o2.stop == "converged" and o1.its == o2.its
            assert np.array_equal(o1.x, o2.x)
        files = sorted(os.listdir(srv.cache.root))
        assert sum(f.endswith(".ldli.npz") for f in files) == 2 and sum(f.endswith(".plan") for f in files) == 1
    finally:
        srv.close()
