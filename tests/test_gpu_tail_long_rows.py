"""Tail programs with long rows: a dense random graph (1500 vertices, ~120
random neighbours each; mean tail row ~150 entries, packed 4 entries per
lane, as tests/test_plan_host.py checks on the host, so most rows are split
into segments added with LDS atomics).  Bars as in the parity tests:
one LDLinv apply to 1e-12 relative of the oracle's, solves with the same
stop reason, iterations within 2 and true relres <= tol (tests/_bars.py), single and
multi-RHS, FP64 and FP32-inner.  The dense tail block is switched off
here, so the slot programs run every level (tests/test_gpu_dense.py covers
the block on this graph)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    a = graphs.random_dense(1500, 120, np.random.default_rng(12))
    return lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


def test_apply_and_solve(case, gpu_ctx_factory, oracle_mod):
    la, f = case
    n = la.shape[0]
    ctx = gpu_ctx_factory(la, f, dense_max=0)
    assert ctx.tail_info()["nt"] == n  # the whole sweep is the tail program
    r = _rhs(n, 1)
    z = ctx.apply_ldli(r)
    zo = oracle_mod.ldlsolver(f, r)
    assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)
    b = _rhs(n, 2)
    x, info = ctx.solve(b, tol=1e-8, maxits=1000, stag_test=5)
    _, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert info["stop"] == io["stop"] and abs(info["its"] - io["its"]) <= 2
    assert_relres(la, x, b, 1e-8, f)


def test_multi_and_fp32(case, gpu_ctx_factory, oracle_mod):
    la, f = case
    n = la.shape[0]
    ctx = gpu_ctx_factory(la, f, dense_max=0)
    B = np.stack([_rhs(n, 10 + j) for j in range(8)], axis=1)
    X, info = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    _, io = oracle_mod.approxchol_solve(la, f, B[:, 3], tol=1e-8, maxits=1000, stag_test=5)
    assert info["stop"][3] == io["stop"] and abs(info["its"][3] - io["its"]) <= 2
    for j in range(B.shape[1]):
        assert_relres(la, X[:, j], B[:, j], 1e-8, f)
    ctx.set_precision(_native.FP32_INNER, 20)
    x32, i32 = ctx.solve(B[:, 0], tol=1e-8, maxits=1000)
    assert_relres(la, x32, B[:, 0], 1e-8, f)
    X32, j32 = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    for j in range(B.shape[1]):
        assert_relres(la, X32[:, j], B[:, j], 1e-8, f)
