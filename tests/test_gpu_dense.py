"""The tail's dense block (lapgpu_set_dense_tail): the deepest tail levels
applied as the matrix D = (I - W^T)^-1 S (I - W)^-1, built on the device from
the block's own sweeps.  Each layout is checked against the oracle on the
same factor, at the bars of the other parity tests: the tail with no block,
a block after a sparse part (forward, D u, backward), and a block that is
the whole tail; single-RHS and multi-RHS (8 and 64 columns: the matrix-core
product), FP64 and FP32-inner.  One LDLinv apply in FP64 to 1e-12 relative
(D sums the block's terms in another order than the sweeps; its constant
part s_root 1 1^T, which the mean removal cancels, is left out of D, or its
rounding would dominate: on the dense random graph it is 190 times the rest
of D u's terms)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs
from _bars import assert_relres

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


CASES = {
    "grid2_60x50": lambda: graphs.weighted_grid2(60, 50, np.random.default_rng(11), "expN"),
    "pathmatch_6000": lambda: graphs.path_matchings(6000, 4, np.random.default_rng(3), "logU3"),
    "grid3_40": lambda: graphs.weighted_grid3(40, 40, 40, np.random.default_rng(2), "expU2"),
    "random_dense_1500": lambda: graphs.random_dense(1500, 120, np.random.default_rng(12)),
}
# dense_max: 0 = no block; 700 = a block after a sparse part; None = the planner's choice
LAYOUTS = {"none": 0, "split": 700, "model": None}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    a = CASES[request.param]()
    return request.param, lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)


@pytest.mark.parametrize("layout", sorted(LAYOUTS))
def test_dense_layouts_match_oracle(case, layout, gpu_ctx_factory, oracle_mod):
    name, la, f = case
    n = la.shape[0]
    ctx = gpu_ctx_factory(la, f, dense_max=LAYOUTS[layout])
    info = ctx.dense_info()
    if layout == "none":
        assert info["k"] == 0
    elif layout == "split":
        assert 0 < info["k"] <= 700 and info["sparse_vertices"] > 0, info
    else:
        assert info["k"] > 0, info  # every case here pays for a block
    r = _rhs(n, 8)
    zo = oracle_mod.ldlsolver(f, r)
    assert np.linalg.norm(ctx.apply_ldli(r) - zo) <= 1e-12 * np.linalg.norm(zo), (name, layout)
    b = _rhs(n, 0)
    x, si = ctx.solve(b, tol=1e-8, maxits=1000, stag_test=5)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert si["stop"] == io["stop"] and abs(si["its"] - io["its"]) <= 2, (name, layout, si, io)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert_relres(la, x, b, 1e-8, f)
    for k in (8, 64):
        B = np.stack([_rhs(n, 40 + j) for j in range(k)], axis=1)
        X, mi = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
        for j in (0, k - 1):
            xo, io = oracle_mod.approxchol_solve(la, f, B[:, j], tol=1e-8, maxits=1000, stag_test=5)
            assert mi["stop"][j] == io["stop"] and abs(mi["its"][j] - io["its"]) <= 2, (name, layout, k, j)
            assert np.linalg.norm(X[:, j] - xo) <= 1e-6 * np.linalg.norm(xo), (name, layout, k, j)
        for j in range(k):
            assert_relres(la, X[:, j], B[:, j], 1e-8, f)
    # FP32-inner: the FP32 copy of D, single and 16 columns
    ctx.set_precision(_native.FP32_INNER, 20)
    zf = ctx.apply_ldli(r)
    assert np.linalg.norm(zf - zo) <= 1e-4 * np.linalg.norm(zo)
    x32, i32 = ctx.solve(b, tol=1e-8, maxits=1000, stag_test=5)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert i32["stop"] == "converged" and i32["its"] <= io["its"] + 2, (name, layout, i32, io)
    assert np.linalg.norm(x32 - xo) <= 1e-6 * np.linalg.norm(xo)
    assert_relres(la, x32, b, 1e-8, f)
    B = np.stack([_rhs(n, 90 + j) for j in range(16)], axis=1)
    X, mi = ctx.solve_multi(B, tol=1e-8, maxits=1000, stag_test=5)
    assert all(s == "converged" for s in mi["stop"])
    for j in range(16):
        assert_relres(la, X[:, j], B[:, j], 1e-8, f)


def test_dense_block_bytes_and_limits(gpu_ctx_factory):
    a = graphs.weighted_grid2(60, 50, np.random.default_rng(11), "expN")
    la, f = lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)
    ctx = gpu_ctx_factory(la, f)
    info = ctx.dense_info()
    K = info["k"]
    ldd = (K + 255) // 256 * 256
    assert info["bytes"] == 8 * K * ldd
    ctx.set_precision(_native.FP32_INNER, 20)
    assert ctx.dense_info()["bytes"] == 12 * K * ldd
    with pytest.raises(lapgpu.LapGPUError):
        ctx.set_dense_tail(-1)
    with pytest.raises(lapgpu.LapGPUError):
        ctx.set_dense_tail(10 ** 6)


def test_wide_scale_spread_keeps_sweeps(gpu_ctx_factory):
    """D u cancels terms of size max(s) |u| down to the result, the sweeps
    difference first: a block whose diagonal scales span more than 1e12 is
    left to the sweeps (the factor of tests/_exits.py spans 1e-150 .. 1e200)."""
    import _exits as E

    a, f = E.factor_with_d(lapgpu, *E.D_INF_BETA)
    ctx = gpu_ctx_factory(lapgpu.lap(a), f)
    assert ctx.dense_info()["k"] == 0
    a, f = E.factor_with_d(lapgpu, 1.0, 1.0)
    ctx = gpu_ctx_factory(lapgpu.lap(a), f)
    assert ctx.dense_info()["k"] > 0
