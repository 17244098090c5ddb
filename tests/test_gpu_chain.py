"""Runs of narrow level groups in one launch (lapgpu_set_sweep_mode): in the
single right-hand-side sweeps the workgroups of a run take its groups'
rows in sweep order and wait for each gathered value to be written.  Each
row is summed as the one-launch-per-group kernels sum it, so both modes
must give the same bits — one LDLinv apply, whole solves (FP64 and
FP32-inner; multi-RHS solves, which always launch per group, alongside),
after the autotuner changed lanes per row and after a regroup, and over
many iterations (a stale value read by a wait would change them).  Graphs: a path + 4 matchings expander (long
runs of narrow groups near the tail), a random graph of mean degree 30
(long composite rows, run one wave or one workgroup each), a 3-D grid, and
the same graphs with every group counted narrow (LAPGPU_CHAIN_NARROW), so
that whole sweeps run chained."""

import os

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs
from _bars import assert_relres

pytestmark = pytest.mark.gpu


def _rhs(n, rng, k=None):
    b = rng.standard_normal((n, k) if k else n)
    return b - b.mean(axis=0)


@pytest.fixture(scope="module")
def cases():
    rng = np.random.default_rng(71)
    out = []
    # the last item: right-hand sides of the multi-RHS solve
    for name, a, budget, k in (
        ("pathmatch_250k", graphs.path_matchings(250000, 4, rng, "logU3"), (8, 2.0, 16384), 40),
        ("dense_40k", graphs.random_dense(40000, 30, rng), None, 20),
        ("grid3_32", graphs.weighted_grid3(32, 32, 32, rng, "expU2"), (8, 2.0, 16384), 5),
    ):
        out.append((name, a, lapgpu.lap(a), lapgpu.build_ldli(a, seed=1), budget, k))
    return out


def _ctx(la, f, budget, chain, prec="fp64", narrow=None, env=None):
    # runs of any length (the library chains only long ones by default)
    env = dict(env or {}, LAPGPU_CHAIN_MIN_RUN="2")
    if narrow is not None:
        env["LAPGPU_CHAIN_NARROW"] = str(narrow)
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = lapgpu.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    if budget:
        ctx.set_group_budget(*budget)
    ctx.upload_lap(la)
    ctx.upload_ldli(f)
    ctx.set_sweep_mode(chain)
    if prec == "fp32":
        ctx.set_precision(_native.FP32_INNER, 20)
    return ctx


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("narrow", [None, 10**9])
def test_chain_matches_group_launches_bit_for_bit(cases, prec, narrow):
    rng = np.random.default_rng(72)
    for name, a, la, f, budget, k in cases:
        n = a.shape[0]
        b = _rhs(n, rng)
        B = np.asfortranarray(_rhs(n, rng, k))
        res = {}
        for chain in (True, False):
            ctx = _ctx(la, f, budget, chain, prec, narrow)
            try:
                info = ctx.chain_info()
                groups = ctx.group_info()["ngroups"]
                z = ctx.apply_ldli(b)
                x, i = ctx.solve(b, tol=1e-10, maxits=1000)
                x2, _ = ctx.solve(b, tol=1e-10, maxits=1000)  # the captured iteration again
                X, j = ctx.solve_multi(B, tol=1e-10, maxits=1000)
                res[chain] = (z, x, i["its"], x2, X, j["its"])
            finally:
                ctx.close()
        if narrow is not None or name == "pathmatch_250k":
            assert info["runs"] >= 1 and info["chained_groups"] >= 2, (name, info, groups)
        if narrow is not None:
            assert info["chained_groups"] == groups, (name, info, groups)
        z1, x1, its1, x12, X1, jt1 = res[True]
        z0, x0, its0, x02, X0, jt0 = res[False]
        assert np.array_equal(z1, z0), f"{name}: apply differs"
        assert its1 == its0 and np.array_equal(x1, x0), f"{name}: solve differs"
        assert np.array_equal(x12, x1), f"{name}: repeated solve differs"
        assert jt1 == jt0 and np.array_equal(X1, X0), f"{name}: multi-RHS solve differs"
        assert_relres(la, x1, b, 1e-10, f, what=f"{name} {prec}")


def _both_modes(ctx, b, B=None):
    """Solve in chained mode, then one launch per group, on one context."""
    out = []
    for chain in (True, False):
        ctx.set_sweep_mode(chain)
        x, i = ctx.solve(b, tol=1e-10, maxits=1000)
        X, j = ctx.solve_multi(B, tol=1e-10, maxits=1000) if B is not None else (None, {"its": None})
        out.append((x, i["its"], X, j["its"]))
    ctx.set_sweep_mode(True)
    (x1, i1, X1, j1), (x0, i0, X0, j0) = out
    assert i1 == i0 and np.array_equal(x1, x0)
    if B is not None:
        assert j1 == j0 and np.array_equal(X1, X0)
    return x1


def test_chain_after_autotune_and_regroup(cases):
    rng = np.random.default_rng(73)
    name, a, la, f, budget, _ = cases[0]
    b = _rhs(a.shape[0], rng)
    B = np.asfortranarray(_rhs(a.shape[0], rng, 16))
    ctx = _ctx(la, f, budget, True)
    try:
        ctx.autotune(16)  # lanes and slices per row: the tickets follow
        assert ctx.chain_info()["runs"] >= 1
        assert_relres(la, _both_modes(ctx, b, B), b, 1e-10, f, what="tuned")
        ctx.set_group_budget(4, 1.5, 4096)  # regroups the uploaded factor
        assert ctx.chain_info()["runs"] >= 1
        assert_relres(la, _both_modes(ctx, b), b, 1e-10, f, what="regrouped")
        ctx.autotune(0, groups=True)  # whichever budget it keeps
        assert_relres(la, _both_modes(ctx, b), b, 1e-10, f, what="budget tuned")
    finally:
        ctx.close()


def test_chain_over_many_iterations(cases):
    """A solve at tol 0 runs until one of pcg's stagnation rules stops it
    (tens of iterations, at most 400): each iteration runs the chained runs
    of both sweeps on the same device arrays, so a wait that returned a
    value of an earlier iteration would show."""
    rng = np.random.default_rng(74)
    name, a, la, f, budget, _ = cases[0]
    b = _rhs(a.shape[0], rng)
    out = {}
    for chain in (True, False):
        ctx = _ctx(la, f, budget, chain)
        try:
            out[chain] = ctx.solve(b, tol=0.0, maxits=400, stag_test=0)
        finally:
            ctx.close()
    (x1, i1), (x0, i0) = out[True], out[False]
    assert i1["its"] == i0["its"] and i1["stop"] == i0["stop"]
    assert np.array_equal(x1, x0)


def test_chain_wait_timeout_falls_back(cases):
    """With one poll per wait (LAPGPU_CHAIN_POLL_MAX=1) a wait that finds a
    value still pending gives up at once: the launch flags it, every wait
    after it gives up too (every kernel exits normally), the call is run
    again with one launch per group and the context stays so.  The answers
    equal those of a context that launches per group from the start, bit
    for bit, for solves (host and device vectors) and applies."""
    rng = np.random.default_rng(76)
    name, a, la, f, budget, _ = cases[0]
    n = a.shape[0]
    b = _rhs(n, rng)
    ref = _ctx(la, f, budget, False)
    ctx = _ctx(la, f, budget, True, env={"LAPGPU_CHAIN_POLL_MAX": "1"})
    try:
        assert ctx.chain_info()["active"] and ctx.chain_info()["runs"] >= 1
        xr, ir = ref.solve(b, tol=1e-10, maxits=1000)
        zr = ref.apply_ldli(b)
        for _ in range(3):
            x, i = ctx.solve(b, tol=1e-10, maxits=1000)
            assert i["its"] == ir["its"] and np.array_equal(x, xr)
        assert not ctx.chain_info()["active"]
        assert np.array_equal(ctx.apply_ldli(b), zr)
        db = _native.DeviceArray.from_host(b)
        dx = _native.DeviceArray(n)
        ctx.solve_device(db.address, dx.address, n, tol=1e-10, maxits=1000)
        assert np.array_equal(dx.download(), xr)
    finally:
        ctx.close()
        ref.close()
    # a fresh context applies with the waits bounded the same way: the apply
    # itself falls back
    ctx = _ctx(la, f, budget, True, env={"LAPGPU_CHAIN_POLL_MAX": "1"})
    try:
        for _ in range(3):
            assert np.array_equal(ctx.apply_ldli(b), zr)
    finally:
        ctx.close()
