"""The schedule autotuner (lapgpu_autotune): after tuning the lanes per row
of every level group (and the multi-RHS slices, and the level-group
budget), solves still match the CPU oracle on the same factor: one LDLinv
apply to 1e-12 relative (the lane split only reorders a row's sums), full
solves with the same stop reason, iterations within 2 and true relres <=
tol (tests/_bars.py), in FP64 and in the FP32-inner mode, single and multi-RHS."""

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


@pytest.fixture(scope="module")
def case():
    a = graphs.weighted_grid3(40, 36, 32, np.random.default_rng(21), "expU2")
    return lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)


def test_autotune_single_with_groups(case, gpu_ctx_factory, oracle_mod):
    la, f = case
    ctx = gpu_ctx_factory(la, f)
    t = ctx.autotune(k=0, groups=True)
    assert t["apply_ms_before"] > 0 and t["apply_ms_after"] > 0
    assert t["multi_apply_ms_before"] == 0 and t["multi_apply_ms_after"] == 0
    r = _rhs(la.shape[0], 3)
    z = ctx.apply_ldli(r)
    zo = oracle_mod.ldlsolver(f, r)
    assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)
    b = _rhs(la.shape[0], 4)
    x, info = ctx.solve(b, tol=1e-8, maxits=1000, stag_test=5)
    _, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert info["stop"] == io["stop"] and abs(info["its"] - io["its"]) <= 2
    assert_relres(la, x, b, 1e-8, f)


@pytest.mark.parametrize("k", [8, 16, 64])
def test_autotune_multi(case, gpu_ctx_factory, oracle_mod, k):
    la, f = case
    n = la.shape[0]
    ctx = gpu_ctx_factory(la, f)
    t = ctx.autotune(k=k)
    assert t["multi_apply_ms_before"] > 0 and t["multi_apply_ms_after"] > 0
    B = np.stack([_rhs(n, 10 + j) for j in range(k)], axis=1)
    X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    assert_relres_cols(la, X, B, 1e-8, f)
    _, io = oracle_mod.approxchol_solve(la, f, B[:, 0], tol=1e-8, maxits=1000, stag_test=5)
    assert info["stop"][0] == io["stop"] and abs(info["its"][0] - io["its"]) <= 2


def test_autotune_fp32_inner(case, gpu_ctx_factory):
    la, f = case
    n = la.shape[0]
    ctx = gpu_ctx_factory(la, f)
    ctx.set_precision(_native.FP32_INNER, 20)
    ctx.autotune(k=8, groups=True)
    b = _rhs(n, 5)
    x, info = ctx.solve(b, tol=1e-8, maxits=1000)
    assert info["stop"] == "converged"
    assert_relres(la, x, b, 1e-8, f)
    B = np.stack([_rhs(n, 30 + j) for j in range(8)], axis=1)
    X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    assert_relres_cols(la, X, B, 1e-8, f)


def test_group_budget(case, gpu_ctx_factory, oracle_mod):
    """lapgpu_set_group_budget before an upload and on an uploaded factor:
    one group per level at 1 / 1 / 0 (no composite entries beyond the
    rows' own), and every budget applies the same LDLinv as the oracle."""
    la, f = case
    n = la.shape[0]
    r = _rhs(n, 5)
    zo = oracle_mod.ldlsolver(f, r)
    ref = gpu_ctx_factory(la, f)
    g_def = ref.group_info()
    ctx = lapgpu.Context(0)
    try:
        ctx.set_group_budget(1, 1.0, 0)
        ctx.upload_lap(la)
        ctx.upload_ldli(f)
        g_one = ctx.group_info()
        assert g_one["rows"] == g_def["rows"]
        assert g_one["ngroups"] >= g_def["ngroups"]
        assert g_one["fentries"] <= g_def["fentries"] and g_one["bentries"] <= g_def["bentries"]
        assert g_one["ngroups"] == ctx.factor_info()["nwide"]
        for c in (ref, ctx):
            z = c.apply_ldli(r)
            assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)
        ctx.set_group_budget(8, 2.0, 16384)
        g_mid = ctx.group_info()
        assert g_def["ngroups"] <= g_mid["ngroups"] <= g_one["ngroups"]
        z = ctx.apply_ldli(r)
        assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)
        b = _rhs(n, 6)
        x, info = ctx.solve(b, tol=1e-8, maxits=1000, stag_test=5)
        _, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
        assert info["stop"] == io["stop"] and abs(info["its"] - io["its"]) <= 2
        assert_relres(la, x, b, 1e-8, f)
        for bad in ((0, 2.0, 0), (4, 0.5, 0), (4, 2.0, -1)):
            with pytest.raises(_native.LapGPUError):
                ctx.set_group_budget(*bad)
    finally:
        ctx.close()
