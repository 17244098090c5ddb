"""Batched solves (lapgpu_batch_*): many graphs of uneven sizes and many
right-hand sides in one launch, one workgroup per request, each checked
against the CPU oracle on the same factor (iterations within +-2, same stop
reason, true relres <= tol as tests/_bars.py states), and each result the same alone as in the mix
(to 1e-12 relative: the only order-dependent sums are LDS atomics for rows
split over several 64-lane chunks)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


@pytest.fixture(scope="module")
def store():
    gens = [
        graphs.grid2(20),
        graphs.random_tree(300, np.random.default_rng(1)),
        graphs.path_matchings(2000, 4, np.random.default_rng(2)),
        graphs.weighted_grid3(12, 12, 12, np.random.default_rng(3)),
        graphs.weighted_grid2(150, 150, np.random.default_rng(4), "expN"),  # beyond the LDS cap
        graphs.grid2(7, 3),
    ]
    B = lapgpu.Batch(0)
    items = []
    for a in gens:
        la = lapgpu.lap(a)
        f = lapgpu.build_ldli(a, seed=1)
        gid = B.add_graph(la, f)
        items.append((gid, la, f))
    yield B, items
    B.close()


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_batch_matches_oracle(store, oracle_mod, tol):
    B, items = store
    ids, bs = [], []
    for k, (gid, la, f) in enumerate(items):
        for rep in range(2):  # two right-hand sides per graph
            ids.append(gid)
            bs.append(_rhs(la.shape[0], 10 * k + rep))
    xs, infos = B.solve(ids, bs, tol=tol, maxits=1000, stag_test=5)
    for gid, b, x, info in zip(ids, bs, xs, infos):
        la, f = items[gid][1], items[gid][2]
        xo, io = oracle_mod.approxchol_solve(la, f, b, tol=tol, maxits=1000, stag_test=5)
        assert info["stop"] == io["stop"], (gid, info, io)
        assert abs(info["its"] - io["its"]) <= 2, (gid, info, io)
        assert_relres(la, x, b, tol, f)
        assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo)


def test_alone_equals_in_mix(store):
    B, items = store
    gid, la, f = items[3]
    b = _rhs(la.shape[0], 99)
    (x_alone,), (i_alone,) = B.solve([gid], [b], tol=1e-8)
    ids = [it[0] for it in items] * 3 + [gid]
    bs = [_rhs(items[g][1].shape[0], 1000 + k) for k, g in enumerate(ids[:-1])] + [b]
    xs, infos = B.solve(ids, bs, tol=1e-8)
    assert infos[-1]["its"] == i_alone["its"]
    assert np.linalg.norm(xs[-1] - x_alone) <= 1e-12 * np.linalg.norm(x_alone)


def test_batch_equals_plain_path(store, gpu_ctx_factory):
    B, items = store
    gid, la, f = items[2]
    b = _rhs(la.shape[0], 5)
    (xb,), (ib,) = B.solve([gid], [b], tol=1e-8)
    ctx = gpu_ctx_factory(la, f)
    xp, ip = ctx.solve(b, tol=1e-8)
    assert abs(ib["its"] - ip["its"]) <= 2
    assert np.linalg.norm(xb - xp) <= 1e-6 * np.linalg.norm(xp)


def test_batch_edge_cases(store, oracle_mod):
    B, items = store
    gid, la, f = items[0]
    n = la.shape[0]
    xs, infos = B.solve([gid, gid, gid], [np.zeros(n), np.full(n, 4.0), _rhs(n, 1)], tol=1e-8, maxits=0)
    assert infos[0]["stop"] == "zero_rhs" and not xs[0].any()
    assert infos[1]["stop"] == "zero_rhs" and not xs[1].any()
    assert infos[2]["stop"] == "maxits" and infos[2]["its"] == 0 and not xs[2].any()
    xs, infos = B.solve([gid], [_rhs(n, 2)], tol=1e-12, maxits=100, stag_test=1)
    _, io = oracle_mod.approxchol_solve(la, f, _rhs(n, 2), tol=1e-12, maxits=100, stag_test=1)
    assert infos[0]["stop"] == io["stop"] == "stag_test" and infos[0]["its"] == io["its"] == 2


def test_graph_info_and_remove(store):
    B, items = store
    info = B.graph_info(items[0][0])
    assert info["n"] == 400 and info["device_bytes"] > 0
    a = graphs.grid2(5)
    gid = B.add_graph(lapgpu.lap(a), lapgpu.build_ldli(a))
    B.remove_graph(gid)
    with pytest.raises(lapgpu.LapGPUError):
        B.solve([gid], [_rhs(25, 0)])


def test_remove_during_launch_and_id_reuse(store):
    # a graph removed while a launch that reads it is in flight: the frees
    # are ordered after that launch, whose result is unchanged; the id is
    # then reused by the next graph, which solves correctly
    B, items = store
    a = graphs.weighted_grid2(40, 30, np.random.default_rng(7), "expU2")
    la, f = lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)
    gid = B.add_graph(la, f)
    other = items[3][0]
    bs = [_rhs(la.shape[0], 11), _rhs(B.sizes[other], 12)]
    ref, _ = B.solve([gid, other], bs, tol=1e-8)
    h = B.submit_host([gid, other], bs, tol=1e-8)
    assert h is not None
    B.remove_graph(gid)
    a2 = graphs.grid2(25)
    la2, f2 = lapgpu.lap(a2), lapgpu.build_ldli(a2, seed=1)
    gid2 = B.add_graph(la2, f2)
    assert gid2 == gid
    xs, infos = B.collect(h)
    for x, xr in zip(xs, ref):
        assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)
    b2 = _rhs(la2.shape[0], 13)
    x2, i2 = B.solve([gid2], [b2], tol=1e-8)
    assert i2[0]["stop"] == "converged"
    assert_relres(la2, x2[0], b2, 1e-8, f2)
    B.remove_graph(gid2)
