"""GPU parity: liblapgpu (through its C-ABI) against the CPU oracle on the
same seeded inputs and the same LDLinv.

Bars: SpMV and one LDLinv apply to 1e-12 relative (floating point; the
device sums in a different order than the reference's loops); full solves:
true relres ||Lx-b||/||b|| <= tol, iteration count within +-2 of the oracle's
pcg, and the same stop reason.
"""

import math

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu

SPMV_RTOL = 1e-12
APPLY_RTOL = 1e-12


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


CASES = {
    "grid2_100": lambda: graphs.grid2(100),
    "grid3_16_w": lambda: graphs.weighted_grid3(16, 16, 16, np.random.default_rng(2)),
    "pathmatch_5000": lambda: graphs.path_matchings(5000, 4, np.random.default_rng(3)),
    "grid2_37x23_expN": lambda: graphs.weighted_grid2(37, 23, np.random.default_rng(4), "expN"),
    "tree_500": lambda: graphs.random_tree(500, np.random.default_rng(5)),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    a = CASES[request.param]()
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    return request.param, a, la, f


def test_spmv(case, gpu_ctx_factory, oracle_mod):
    name, a, la, f = case
    ctx = gpu_ctx_factory(la, f)
    x = np.random.default_rng(7).standard_normal(la.shape[0])
    y = ctx.spmv(x)
    ref = oracle_mod.spmv(la, x)
    assert np.linalg.norm(y - ref) <= SPMV_RTOL * np.linalg.norm(ref) + 1e-300, name


def test_apply_ldli(case, gpu_ctx_factory, oracle_mod):
    name, a, la, f = case
    ctx = gpu_ctx_factory(la, f)
    r = _rhs(la.shape[0], 8)
    z = ctx.apply_ldli(r)
    ref = oracle_mod.ldlsolver(f, r)
    assert np.linalg.norm(z - ref) <= APPLY_RTOL * np.linalg.norm(ref), name


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_solve(case, gpu_ctx_factory, oracle_mod, tol):
    name, a, la, f = case
    ctx = gpu_ctx_factory(la, f)
    b = _rhs(la.shape[0], 0)
    x, info = ctx.solve(b, tol=tol, maxits=1000, stag_test=5)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=tol, maxits=1000, stag_test=5)
    relres = np.linalg.norm(la @ x - b) / np.linalg.norm(b)
    assert info["stop"] == io["stop"], (name, info, io)
    assert abs(info["its"] - io["its"]) <= 2, (name, info, io)
    assert_relres(la, x, b, tol, f, what=name)
    assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo), name


def test_zero_rhs(gpu_ctx_factory):
    a = graphs.grid2(10)
    la = lapgpu.lap(a)
    ctx = gpu_ctx_factory(la, lapgpu.build_ldli(a))
    x, info = ctx.solve(np.zeros(100), tol=1e-8)
    assert info["stop"] == "zero_rhs" and info["its"] == 0
    assert np.all(x == 0)
    # a constant b is zero after the closure's mean removal
    x, info = ctx.solve(np.full(100, 3.0), tol=1e-8)
    assert info["stop"] == "zero_rhs" and np.all(x == 0)


@pytest.mark.parametrize("maxits", [0, 1, 2, 7])
def test_maxits(gpu_ctx_factory, oracle_mod, maxits):
    a = graphs.weighted_grid2(30, 30, np.random.default_rng(1))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    ctx = gpu_ctx_factory(la, f)
    b = _rhs(900, 3)
    x, info = ctx.solve(b, tol=1e-14, maxits=maxits)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-14, maxits=maxits)
    assert info["its"] == io["its"] == maxits
    assert info["stop"] == "maxits"
    assert np.linalg.norm(x - xo) <= 1e-10 * max(np.linalg.norm(xo), 1e-300)


def test_maxtime(gpu_ctx_factory):
    a = graphs.weighted_grid2(40, 40, np.random.default_rng(1))
    la = lapgpu.lap(a)
    ctx = gpu_ctx_factory(la, lapgpu.build_ldli(a))
    x, info = ctx.solve(_rhs(1600, 4), tol=1e-14, maxits=1000, maxtime=1e-9)
    # like test/testPCG.jl's maxtime=1e-10 runs: stops after one iteration
    assert info["stop"] == "maxtime" and info["its"] == 1


def test_closure_dropin(gpu_ctx_factory):
    """approxchol_lap_gpu returns f(b) with the reference's kwargs and pcgIts."""
    a = graphs.weighted_grid2(50, 40, np.random.default_rng(9), "expN")
    la = lapgpu.lap(a)
    f = lapgpu.approxchol_lap_gpu(a, tol=1e-8, maxits=500)
    b = _rhs(2000, 1)
    its = [0]
    x = f(b, pcgIts=its)
    assert its[0] > 0
    assert_relres(la, x, b, 1e-8, f.ldli)
    x2 = f(b, tol=1e-3, pcgIts=its)
    assert np.linalg.norm(la @ x2 - b) / np.linalg.norm(b) <= 1e-2


def test_components(gpu_ctx_factory):
    """Disconnected graph: lapWrapComponents + blockSolver semantics."""
    import scipy.sparse as sp

    g1 = graphs.grid2(20)
    g2 = graphs.weighted_grid3(5, 6, 7, np.random.default_rng(2))
    iso = sp.csc_matrix((3, 3))
    a = sp.block_diag([g1, iso, g2], format="csc")
    la = lapgpu.lap(a)
    f = lapgpu.approxchol_lap_gpu(a, tol=1e-9)
    n = a.shape[0]
    b = np.random.default_rng(0).standard_normal(n)
    # project b onto the range of la: zero mean per component
    comps = [np.arange(0, 400), np.arange(400, 403), np.arange(403, n)]
    b[comps[1]] = 0
    for c in (comps[0], comps[2]):
        b[c] -= b[c].mean()
    its = [0]
    x = f(b, pcgIts=its)
    assert its[0] > 0
    assert np.linalg.norm(la @ x - b) / np.linalg.norm(b) <= 1e-8
    assert np.all(x[comps[1]] == 0)


def test_large_grid3_true_residual(gpu_ctx_factory):
    """grid3(48) with exp(U(-2,2)) weights, tol 1e-8: true relres holds."""
    a = graphs.weighted_grid3(48, 48, 48, np.random.default_rng(2))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    ctx = gpu_ctx_factory(la, f)
    b = _rhs(a.shape[0], 0)
    x, info = ctx.solve(b, tol=1e-8, maxits=1000)
    assert info["stop"] == "converged"
    assert_relres(la, x, b, 1e-8, f)
    assert math.isfinite(info["nr"]) and info["nr"] < 1e-8


def test_stag_test_one_stops_at_two(gpu_ctx_factory, oracle_mod):
    a = graphs.grid2(12)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    ctx = gpu_ctx_factory(la, f)
    b = _rhs(144, 0)
    _, info = ctx.solve(b, tol=1e-12, maxits=100, stag_test=1)
    _, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-12, maxits=100, stag_test=1)
    assert info["stop"] == io["stop"] == "stag_test" and info["its"] == io["its"] == 2


def test_tol_zero_ends_by_stagnation(gpu_ctx_factory, oracle_mod):
    a = graphs.random_tree(200, np.random.default_rng(8))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    ctx = gpu_ctx_factory(la, f)
    b = _rhs(200, 0)
    x, info = ctx.solve(b, tol=0.0, maxits=3000, stag_test=5)
    _, io = oracle_mod.approxchol_solve(la, f, b, tol=0.0, maxits=3000, stag_test=5)
    assert info["stop"].startswith("stag")
    if info["stop"] != "stag_test":  # the |p|, |q| and beta checks run every 10th iteration
        assert info["its"] % 10 == 0
    assert abs(info["its"] - io["its"]) <= 10
    assert np.linalg.norm(la @ x - b) <= 1e-12 * np.linalg.norm(b)


def test_zero_rhs_leaves_pcgits():
    """pcg returns before it writes pcgIts when b is zero after the
    closure's mean removal (src/pcg.jl:269-271): the caller's value stays."""
    a = graphs.grid2(12)
    f = lapgpu.approxchol_lap_gpu(a, tol=1e-8)
    its = [7]
    x = f(np.full(144, 2.0), pcgIts=its)
    assert its == [7] and np.all(x == 0)
    x = f(_rhs(144, 1), pcgIts=its)
    assert its[0] > 0 and its[0] != 7


def test_apply_with_nonzero_root_d(gpu_ctx_factory, oracle_mod):
    """A caller's LDLinv with d[root] != 0: LDLsolver divides y[root] by it
    (src/approxChol.jl:1351-1355), and so does the device apply, on a
    right-hand side that does not sum to zero (y[root] gathers its sum)."""
    import dataclasses

    a = graphs.weighted_grid2(30, 20, np.random.default_rng(13), "expU2")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    root = int(np.setdiff1d(np.arange(a.shape[0]), f.col)[0])
    d = f.d.copy()
    d[root] = 2.5
    g = dataclasses.replace(f, d=d)
    ctx = gpu_ctx_factory(la, g)
    r = np.random.default_rng(8).standard_normal(la.shape[0]) + 0.5
    z = ctx.apply_ldli(r)
    ref = oracle_mod.ldlsolver(g, r)
    # the root's scale adds the same constant to every entry (backward rows
    # are convex combinations), which the closing mean removal cancels, so
    # the output equals the builders' factor's up to rounding; the plan's
    # root scale itself is pinned by tests/test_plan_host.py
    assert np.linalg.norm(z - ref) <= APPLY_RTOL * np.linalg.norm(ref)
    assert np.linalg.norm(ref - oracle_mod.ldlsolver(f, r)) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("which", ["inf_rho", "inf_beta"])
def test_overflow_exits_on_every_device_path(which, gpu_ctx_factory, oracle_mod):
    """pcg's isinf(rho) and isinf(beta) exits (src/pcg.jl:385-402) on the
    hand-derived inputs of tests/_exits.py (raw pcg: no mean removal of b),
    on the single-RHS, multi-RHS and batched paths: the oracle's stop at
    iteration 1 and its final rho."""
    import _exits as E

    dd, b = (E.D_INF_RHO, E.B_INF_RHO) if which == "inf_rho" else (E.D_INF_BETA, E.B_INF_BETA)
    a, f = E.factor_with_d(lapgpu, *dd)
    la = lapgpu.lap(a)
    with np.errstate(all="ignore"):
        xo, io = oracle_mod.pcg_ldl(la, f, b, tol=1e-8, maxits=30, stag_test=5)
    assert io["stop"] == which and io["its"] == 1
    from lapgpu import _native

    ctx = gpu_ctx_factory(la, f)
    x, info = ctx.solve(b, tol=1e-8, maxits=30, stag_test=5, flags=_native.NO_CENTER)
    assert info["stop"] == which and info["its"] == 1, info
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    X, mi = ctx.solve_multi(np.stack([b, b], axis=1), tol=1e-8, maxits=30, stag_test=5, flags=_native.NO_CENTER)
    assert mi["stop"] == [which, which] and mi["its"] == [1, 1], mi
    assert np.linalg.norm(X[:, 0] - xo) <= 1e-12 * np.linalg.norm(xo)
    batch = lapgpu.Batch(0)
    try:
        gid = batch.add_graph(la, f)
        xs, infos = batch.solve([gid], [b], tol=1e-8, maxits=30, stag_test=5, flags=_native.NO_CENTER)
    finally:
        batch.close()
    assert infos[0]["stop"] == which and infos[0]["its"] == 1, infos
    assert np.linalg.norm(xs[0] - xo) <= 1e-12 * np.linalg.norm(xo)
