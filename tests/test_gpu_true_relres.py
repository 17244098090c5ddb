"""The FP64 true residual each solve reports (lapgpu_last_true_relres and
the closure's trueRelres output): the device's ||(b - mean b) - la x|| /
||b - mean b|| for the returned x equals the host's, for single and
multi-RHS solves, FP64 and FP32-inner, and through lapWrapComponents'
block solver over several components.  In FP32-inner mode it is the
per-solve error against FP64 the precision form states, and it meets tol
when the solve converged."""

import numpy as np
import pytest
import scipy.sparse as sp

import lapgpu
from lapgpu import _native, graphs

pytestmark = pytest.mark.gpu


def _host(la, x, b):
    bc = b - b.mean()
    return float(np.linalg.norm(bc - la @ x) / np.linalg.norm(bc))


def _close(dev, host):
    assert abs(dev - host) <= 1e-6 * host + 1e-14, (dev, host)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_true_relres_single_and_multi(prec):
    rng = np.random.default_rng(61)
    a = graphs.path_matchings(20000, 4, rng, "logU3")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1)
    ctx = lapgpu.Context(0)
    try:
        ctx.upload_lap(la)
        ctx.upload_ldli(f)
        if prec == "fp32":
            ctx.set_precision(_native.FP32_INNER, 20)
        b = rng.standard_normal(a.shape[0]) + 3.0  # a mean for the closure to remove
        x, info = ctx.solve(b, tol=1e-8, maxits=1000)
        t = ctx.last_true_relres(1)
        _close(t, _host(la, x, b))
        assert info["stop"] == "converged" and t <= 1e-8
        with pytest.raises(lapgpu.LapGPUError):
            ctx.last_true_relres(3)
        B = np.asfortranarray(rng.standard_normal((a.shape[0], 70)))
        B[:, 5] = 0.0  # a zero column reports 0
        X, inf = ctx.solve_multi(B, tol=1e-8, maxits=1000)
        ts = ctx.last_true_relres(70)
        for j in range(70):
            if j == 5:
                assert inf["stop"][j] == "zero_rhs" and ts[j] == 0.0
                continue
            _close(ts[j], _host(la, X[:, j], B[:, j]))
            assert inf["stop"][j] == "converged" and ts[j] <= 1e-8
        # a loose tolerance: the recursive and the true residual part ways
        x, info = ctx.solve(b, tol=1e-3, maxits=1000)
        _close(ctx.last_true_relres(1), _host(la, x, b))
    finally:
        ctx.close()


def test_closure_reports_true_relres_over_components():
    rng = np.random.default_rng(62)
    g1 = graphs.weighted_grid2(40, 30, rng)
    g2 = graphs.weighted_grid3(10, 10, 10, rng)
    a = sp.block_diag([g1, g2, sp.csc_matrix((1, 1))], format="csc")  # two components and an isolated vertex
    la = lapgpu.lap(a)
    for prec in ("fp64", "fp32"):
        solve = lapgpu.approxchol_lap_gpu(a, tol=1e-8, precision=prec)
        b = rng.standard_normal(a.shape[0])
        tr, its = [0.0], [0]
        x = solve(b, pcgIts=its, trueRelres=tr)
        # the block solver solves each component on its own centred b
        num = den = 0.0
        for lo, hi in ((0, 1200), (1200, 2200)):
            bc = b[lo:hi] - b[lo:hi].mean()
            r = bc - (la[lo:hi, lo:hi] @ x[lo:hi])
            num += float(r @ r)
            den += float(bc @ bc)
        _close(tr[0], float(np.sqrt(num / den)))
        assert tr[0] <= 1e-8 and its[0] > 0
    # one component: the closure itself
    solve = lapgpu.approxchol_lap_gpu(g1, tol=1e-8)
    b = rng.standard_normal(1200)
    tr = [0.0]
    x = solve(b, trueRelres=tr)
    _close(tr[0], _host(lapgpu.lap(g1), x, b))
    assert solve.last["true_relres"] == tr[0]
