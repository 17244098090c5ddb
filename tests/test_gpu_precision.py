"""FP32-inner mode (lapgpu_set_precision(LAPGPU_FP32_INNER)): the SpMV and
the LDLinv apply run in FP32; dots, alpha/beta and a true-residual
replacement every k iterations run in FP64, and convergence is accepted only
on an FP64 true residual.  Checked: true relres <= tol, the error against the
FP64 path's solution, and the iteration ratio."""

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


@pytest.mark.parametrize("name", ["pathmatch", "grid3"])
def test_fp32_inner_reaches_tol(name, gpu_ctx_factory):
    rng = np.random.default_rng(3)
    a = graphs.path_matchings(20000, 4, rng) if name == "pathmatch" else graphs.weighted_grid3(24, 24, 24, rng)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    b = _rhs(a.shape[0], 0)
    c64 = gpu_ctx_factory(la, f)
    x64, i64 = c64.solve(b, tol=1e-8, maxits=1000)
    c32 = gpu_ctx_factory(la, f)
    c32.set_precision(_native.FP32_INNER, 20)
    x32, i32 = c32.solve(b, tol=1e-8, maxits=1000)
    rel = np.linalg.norm(la @ x32 - b) / np.linalg.norm(b)
    assert i32["stop"] == "converged", i32
    assert rel <= 1e-8, rel
    assert np.linalg.norm(x32 - x64) <= 1e-5 * np.linalg.norm(x64)
    # an FP32 preconditioner is a perturbed one: on badly scaled graphs it
    # costs iterations (the ratio is reported by the precision benchmark)
    assert i32["its"] <= 4 * i64["its"] + 10


def test_fp32_apply_close_to_oracle(gpu_ctx_factory, oracle_mod):
    a = graphs.weighted_grid2(60, 50, np.random.default_rng(1), "expN")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    ctx = gpu_ctx_factory(la, f)
    ctx.set_precision(_native.FP32_INNER, 20)
    r = _rhs(3000, 2)
    z = ctx.apply_ldli(r)
    ref = oracle_mod.ldlsolver(f, r)
    assert np.linalg.norm(z - ref) <= 1e-4 * np.linalg.norm(ref)
    y = ctx.spmv(r)
    assert np.linalg.norm(y - la @ r) <= 1e-5 * np.linalg.norm(la @ r)
    ctx.set_precision(_native.FP64)
    assert np.linalg.norm(ctx.apply_ldli(r) - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("k", [3, 16])
def test_fp32_inner_multi(k, gpu_ctx_factory):
    """FP32-inner mode with k right-hand sides at once: every column meets
    tol on its FP64 true residual and stays close to its FP64 solution."""
    rng = np.random.default_rng(5)
    a = graphs.path_matchings(20000, 4, rng, "logU3")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    n = a.shape[0]
    B = rng.standard_normal((n, k))
    B -= B.mean(axis=0)
    c64 = gpu_ctx_factory(la, f)
    X64, i64 = c64.solve_multi(B, tol=1e-8, maxits=1000)
    c32 = gpu_ctx_factory(la, f)
    c32.set_precision(_native.FP32_INNER, 20)
    X32, i32 = c32.solve_multi(B, tol=1e-8, maxits=1000)
    rel = np.linalg.norm(la @ X32 - B, axis=0) / np.linalg.norm(B, axis=0)
    assert all(s == "converged" for s in i32["stop"]), i32
    assert np.all(rel <= 1e-8), rel
    for j in range(k):
        assert np.linalg.norm(X32[:, j] - X64[:, j]) <= 1e-5 * np.linalg.norm(X64[:, j])
        assert i32["its"][j] <= 4 * i64["its"][j] + 10
    # one column alone in FP32 mode agrees with its multi-RHS result
    x1, i1 = c32.solve(B[:, 0], tol=1e-8, maxits=1000)
    assert np.linalg.norm(x1 - X32[:, 0]) <= 1e-5 * np.linalg.norm(x1)
