"""LAPGPU_EAGER: a solve whose iterations are launched directly instead of
replaying a captured HIP graph runs the same kernels in the same order
(single and multi-RHS, before and after a capture exists, and for a
multi-RHS width other than the captured one).  The tail adds the chunks of
a split row into LDS in whatever order they finish, so two solves of one b
agree to rounding, not bit for bit: the same stop reason, iterations within
1, true relres <= tol (tests/_bars.py), and solutions within 1e-6 of each other."""

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    a = graphs.weighted_grid3(24, 20, 18, np.random.default_rng(31), "expU2")
    return lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)


def _rhs(n, k, seed):
    B = np.random.default_rng(seed).standard_normal((n, k))
    return B - B.mean(axis=0)


def _close(la, f, B, Xe, Xc, ie, ic):
    for j in range(B.shape[1]):
        assert ie["stop"][j] == ic["stop"][j] and abs(ie["its"][j] - ic["its"][j]) <= 1
        for X in (Xe, Xc):
            assert_relres(la, X[:, j], B[:, j], 1e-8, f)
        assert np.linalg.norm(Xe[:, j] - Xc[:, j]) <= 1e-6 * np.linalg.norm(Xc[:, j])


def test_eager_single_matches_captured(case, gpu_ctx_factory, oracle_mod):
    la, f = case
    n = la.shape[0]
    b = _rhs(n, 1, 1)[:, 0]
    ctx = gpu_ctx_factory(la, f)
    xe, ie = ctx.solve(b, tol=1e-8, maxits=1000, flags=_native.EAGER)  # no capture yet
    xc, ic = ctx.solve(b, tol=1e-8, maxits=1000)  # captures
    xe2, ie2 = ctx.solve(b, tol=1e-8, maxits=1000, flags=_native.EAGER)  # capture exists
    assert ie["stop"] == ic["stop"] == ie2["stop"]
    assert abs(ie["its"] - ic["its"]) <= 1 and abs(ie2["its"] - ic["its"]) <= 1
    for x in (xe, xc, xe2):
        assert_relres(la, x, b, 1e-8, f)
    assert np.linalg.norm(xe - xc) <= 1e-6 * np.linalg.norm(xc)
    assert np.linalg.norm(xe2 - xc) <= 1e-6 * np.linalg.norm(xc)
    _, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert ic["stop"] == io["stop"] and abs(ic["its"] - io["its"]) <= 2


def test_eager_multi_matches_captured(case, gpu_ctx_factory):
    la, f = case
    n = la.shape[0]
    ctx = gpu_ctx_factory(la, f)
    B8 = _rhs(n, 8, 2)
    Xc, ic = ctx.solve_multi(B8, tol=1e-8, maxits=1000)  # captures K = 8
    Xe, ie = ctx.solve_multi(B8, tol=1e-8, maxits=1000, flags=_native.EAGER)
    _close(la, f, B8, Xe, Xc, ie, ic)
    B16 = _rhs(n, 12, 3)  # K = 16 while the K = 8 capture is kept
    Ye, je = ctx.solve_multi(B16, tol=1e-8, maxits=1000, flags=_native.EAGER)
    Yc, jc = ctx.solve_multi(B16, tol=1e-8, maxits=1000)
    _close(la, f, B16, Ye, Yc, je, jc)


def test_blocking_wait_mode(case, gpu_ctx_factory, oracle_mod):
    """lapgpu_set_wait_mode: solves that sleep on blocking-sync events give
    the spin-wait solves' results (to rounding), single and multi-RHS, and
    switching back works; an unknown mode is rejected."""
    la, f = case
    n = la.shape[0]
    b = _rhs(n, 1, 4)[:, 0]
    B = _rhs(n, 8, 5)
    ctx = gpu_ctx_factory(la, f)
    xs, is_ = ctx.solve(b, tol=1e-8, maxits=1000)
    Xs, js = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    ctx.set_wait_mode(True)
    xb, ib = ctx.solve(b, tol=1e-8, maxits=1000)
    xb2, ib2 = ctx.solve(b, tol=1e-8, maxits=1000, flags=_native.EAGER)
    Xb, jb = ctx.solve_multi(B, tol=1e-8, maxits=1000)
    for x, i in ((xb, ib), (xb2, ib2)):
        assert i["stop"] == is_["stop"] and abs(i["its"] - is_["its"]) <= 1
        assert_relres(la, x, b, 1e-8, f)
        assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)
    _close(la, f, B, Xb, Xs, jb, js)
    _, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert ib["stop"] == io["stop"] and abs(ib["its"] - io["its"]) <= 2
    ctx.set_wait_mode(False)
    xs2, _ = ctx.solve(b, tol=1e-8, maxits=1000)
    assert np.linalg.norm(xs2 - xs) <= 1e-6 * np.linalg.norm(xs)
    with pytest.raises(_native.LapGPUError):
        ctx._check(ctx._lib.lapgpu_set_wait_mode(ctx._h, 7), "lapgpu_set_wait_mode")
