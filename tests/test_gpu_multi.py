"""Many right-hand sides of one graph in one solve (lapgpu_solve_multi).

Each column must be solved as the closure would solve it alone: against the
CPU oracle's pcg with the same LDLinv, the same stop reason, iteration count
within +-2, true relres <= tol (tests/_bars.py) and the solution within 1e-4 relative (the
bars of the single-RHS parity tests).  Columns that stop early (zero column,
maxits) must not disturb the others.
"""

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu


def _rhs_block(n, k, seed):
    B = np.random.default_rng(seed).standard_normal((n, k))
    return B - B.mean(axis=0)


CASES = {
    "grid2_60x50": lambda: graphs.weighted_grid2(60, 50, np.random.default_rng(11), "expN"),
    "grid3_16_w": lambda: graphs.weighted_grid3(16, 16, 16, np.random.default_rng(2)),
    "pathmatch_5000": lambda: graphs.path_matchings(5000, 4, np.random.default_rng(3)),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    a = CASES[request.param]()
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    return request.param, a, la, f


@pytest.mark.parametrize("k", [1, 5, 16, 40, 70])
def test_multi_matches_oracle(case, gpu_ctx_factory, oracle_mod, k):
    name, a, la, f = case
    ctx = gpu_ctx_factory(la, f)
    n = la.shape[0]
    B = _rhs_block(n, k, 100 + k)
    X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
    assert X.shape == (n, k)
    # every column against the oracle for small k, a spread of columns else
    cols = range(k) if k <= 5 else sorted({0, 1, k // 2, k - 2, k - 1})
    for j in cols:
        xo, io = oracle_mod.approxchol_solve(la, f, B[:, j], tol=1e-8, maxits=1000, stag_test=5)
        assert info["stop"][j] == io["stop"], (name, j, info["stop"][j], io)
        assert abs(info["its"][j] - io["its"]) <= 2, (name, j, info["its"][j], io)
        assert np.linalg.norm(X[:, j] - xo) <= 1e-4 * np.linalg.norm(xo), (name, j)
    assert_relres_cols(la, X, B, 1e-8, f, what=name)


def test_multi_equals_single(case, gpu_ctx_factory):
    """The multi path and the single-RHS path agree column by column."""
    name, a, la, f = case
    ctx = gpu_ctx_factory(la, f)
    n = la.shape[0]
    B = _rhs_block(n, 8, 7)
    X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    for j in range(8):
        x1, i1 = ctx.solve(B[:, j], tol=1e-8, maxits=1000)
        assert abs(i1["its"] - info["its"][j]) <= 2, (name, j)
        assert np.linalg.norm(X[:, j] - x1) <= 1e-6 * np.linalg.norm(x1), (name, j)


def test_multi_mixed_stops(gpu_ctx_factory, oracle_mod):
    """A zero column, a tiny-maxits run and normal columns side by side."""
    a = graphs.weighted_grid2(40, 40, np.random.default_rng(3), "expN")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    ctx = gpu_ctx_factory(la, f)
    n = la.shape[0]
    B = _rhs_block(n, 6, 9)
    B[:, 2] = 0.0
    X, info = ctx.solve_multi(B, tol=1e-10, maxits=1000)
    assert info["stop"][2] == "zero_rhs" and info["its"][2] == 0
    assert np.all(X[:, 2] == 0.0)
    for j in (0, 1, 3, 4, 5):
        xo, io = oracle_mod.approxchol_solve(la, f, B[:, j], tol=1e-10, maxits=1000, stag_test=5)
        assert info["stop"][j] == io["stop"]
        assert abs(info["its"][j] - io["its"]) <= 2
    # maxits = 3 for all: everyone stops at 3 (zero column at 0)
    X3, i3 = ctx.solve_multi(B, tol=1e-10, maxits=3)
    for j in (0, 1, 3, 4, 5):
        xo, io = oracle_mod.approxchol_solve(la, f, B[:, j], tol=1e-10, maxits=3, stag_test=5)
        assert i3["its"][j] == 3 and i3["stop"][j] == "maxits" == io["stop"]
        assert np.linalg.norm(X3[:, j] - xo) <= 1e-8 * np.linalg.norm(xo)


def test_multi_device_pointers(gpu_ctx_factory):
    """B and X resident in device memory (column-major n-by-k)."""
    from lapgpu._native import DeviceArray

    a = graphs.grid2(50)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    ctx = gpu_ctx_factory(la, f)
    n = la.shape[0]
    B = _rhs_block(n, 12, 4)
    dB = DeviceArray.from_host(B.T)  # row j of B.T = column j of B
    dX = DeviceArray(n * 12)
    info = ctx.solve_multi_device(dB.address, dX.address, n, 12, tol=1e-8, maxits=500)
    X = dX.download().reshape(12, n).T
    Xh, ih = ctx.solve_multi(B, tol=1e-8, maxits=500)
    assert info["its"] == ih["its"]
    assert np.array_equal(X, Xh)
    dB.free()
    dX.free()


def test_reupload_after_multi_sizes_workspace_again(gpu_ctx_factory, oracle_mod):
    """A multi-RHS solve sizes its tail-input buffers from the tail of the
    plan it runs on.  Uploading another graph of the same n whose tail is
    larger (grid 150x200 -> path + matchings, both n = 30,000), then one
    with a smaller n, must size them again: each later multi-RHS solve is
    checked against the oracle."""
    ga = graphs.weighted_grid2(150, 200, np.random.default_rng(21), "expU2")
    gb = graphs.path_matchings(30000, 4, np.random.default_rng(22), "logU3")
    gc = graphs.weighted_grid2(100, 100, np.random.default_rng(24), "expU2")
    la_a = lapgpu.lap(ga)
    ctx = gpu_ctx_factory(la_a, lapgpu.build_ldli(ga, seed=1))
    nt_a = ctx.tail_info()["nt"]
    B = _rhs_block(30000, 8, 31)
    ctx.solve_multi(B, tol=1e-8)
    prev_nt = nt_a
    for a in (gb, gc):
        la = lapgpu.lap(a)
        f = lapgpu.build_ldli(a, seed=1)
        ctx.upload_lap(la)
        ctx.upload_ldli(f)
        nt = ctx.tail_info()["nt"]
        if a is gb:
            assert nt > prev_nt, (nt, prev_nt)  # the case the workspace must follow
        n = la.shape[0]
        B = _rhs_block(n, 8, 32)
        X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
        for j in (0, 7):
            xo, io = oracle_mod.approxchol_solve(la, f, B[:, j], tol=1e-8, maxits=1000, stag_test=5)
            assert info["stop"][j] == io["stop"] and abs(info["its"][j] - io["its"]) <= 2, (j, info, io)
            assert np.linalg.norm(X[:, j] - xo) <= 1e-4 * np.linalg.norm(xo)
        prev_nt = nt


@pytest.mark.parametrize("budget", [None, (16, 4.0, 65536)])
@pytest.mark.parametrize("k", [12, 64])
def test_multi_long_composite_rows(budget, k, gpu_ctx_factory, oracle_mod):
    """Level groups whose composite rows are longer than 64 entries (on this
    path + matchings graph of 1e5 vertices: thousands of such rows in every
    group, under the library's budget and under 16 / 4 / 65,536): those rows
    take one workgroup each in the same launch as the group's other rows.
    Columns against the oracle, every column on its true residual, FP64 and
    FP32-inner, and a second solve gives the same bits."""
    a = graphs.path_matchings(100000, 4, np.random.default_rng(3))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    n = la.shape[0]
    B = _rhs_block(n, k, 40 + k)
    ctx = lapgpu.Context(0)
    try:
        if budget is not None:
            ctx.set_group_budget(*budget)
        ctx.upload_lap(la)
        ctx.upload_ldli(f)
        assert ctx.group_info()["ngroups"] >= 3
        X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
        for j in (0, k // 2, k - 1):
            xo, io = oracle_mod.approxchol_solve(la, f, B[:, j], tol=1e-8, maxits=1000, stag_test=5)
            assert info["stop"][j] == io["stop"], (j, info["stop"][j], io)
            assert abs(info["its"][j] - io["its"]) <= 2, (j, info["its"][j], io)
            assert np.linalg.norm(X[:, j] - xo) <= 1e-4 * np.linalg.norm(xo), j
        assert_relres_cols(la, X, B, 1e-8, f, what="long rows fp64")
        X2, _ = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
        assert np.array_equal(X, X2)
        ctx.set_precision(lapgpu._native.FP32_INNER, 20)
        X32, i32 = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
        assert all(s == "converged" for s in i32["stop"]), i32
        rel = np.linalg.norm(la @ X32 - B, axis=0) / np.linalg.norm(B, axis=0)
        assert np.all(rel <= 1e-8), rel
        for j in range(k):
            assert np.linalg.norm(X32[:, j] - X[:, j]) <= 1e-5 * np.linalg.norm(X[:, j]), j
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_multi_alternating_k(prec, gpu_ctx_factory):
    """One context serving blocks of 5, 12, 40, 5, 20 and 12 columns (K = 8,
    16, 64, 8, 32, 16: the captured iteration is made again whenever K
    changes) gives every block the bits a fresh context gives it."""
    a = graphs.path_matchings(30000, 4, np.random.default_rng(8))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    n = la.shape[0]
    B = _rhs_block(n, 40, 12)

    def make():
        c = gpu_ctx_factory(la, f)
        if prec == "fp32":
            c.set_precision(lapgpu._native.FP32_INNER, 20)
        return c

    ctx = make()
    fresh = {}
    for k in (5, 12, 40, 5, 20, 12):
        X, info = ctx.solve_multi(B[:, :k], tol=1e-8, maxits=1000)
        if k not in fresh:
            fresh[k] = make().solve_multi(B[:, :k], tol=1e-8, maxits=1000)
        Xf, infof = fresh[k]
        assert np.array_equal(X, Xf), k
        assert list(info["its"]) == list(infof["its"]), k
