"""The reference's other factor builders on the GPU solve: approxchol_lap
with ApproxCholParams orders :given, :wdeg, :random (approxchol_lap1,
src/approxChol.jl:1791-1851) and approxchol_lap2 (src/approxChol.jl:
1559-1579), each against the CPU oracle's pcg on the same factor: same stop
reason, iterations within 2, true relres <= tol (tests/_bars.py)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


@pytest.mark.parametrize("order", ["given", "wdeg", "random"])
@pytest.mark.parametrize("gen", ["grid2", "pm"])
def test_orders_match_oracle(order, gen, oracle_mod):
    rng = np.random.default_rng(12)
    a = graphs.weighted_grid2(90, 70, rng, "expU2") if gen == "grid2" else graphs.path_matchings(8000, 4, rng)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1, order=order)
    its = [0]
    solver = lapgpu.approxchol_lap_gpu(a, tol=1e-8, params=lapgpu.ApproxCholParams(order=order), ldli=f)
    b = _rhs(a.shape[0], 1)
    x = solver(b, pcgIts=its)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert solver_stop(solver) == io["stop"]
    assert abs(its[0] - io["its"]) <= 2
    assert_relres(la, x, b, 1e-8, f)


def solver_stop(solver):
    # the closure of a connected graph is the GPU closure itself
    return solver.last["stop"]


def test_lap2_matches_oracle(oracle_mod):
    rng = np.random.default_rng(13)
    a = graphs.weighted_grid3(18, 16, 14, rng, "expU2")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1, order="deg", split=5, merge=2)
    its = [0]
    solver = lapgpu.approxchol_lap2_gpu(a, tol=1e-8, ldli=f)
    b = _rhs(a.shape[0], 2)
    x = solver(b, pcgIts=its)
    xo, io = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
    assert solver.last["stop"] == io["stop"]
    assert abs(its[0] - io["its"]) <= 2
    assert_relres(la, x, b, 1e-8, f)
    # built without a given factor, the split-edge elimination runs here
    s2 = lapgpu.approxchol_lap2_gpu(a, tol=1e-8)
    x2 = s2(b)
    assert_relres(la, x2, b, 1e-8, f)
