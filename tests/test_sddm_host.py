"""SDDM wrapper host logic (sddmWrapLap, src/solverInterface.jl:324-352;
adjValAndExcess / extendMatrix, src/graphOps.jl:43-78) with a dense
pseudo-inverse Laplacian solver standing in for the GPU solver, so that
only the wrapper's algebra is under test: M x = b to round-off."""

import numpy as np
import scipy.sparse as sp

from lapgpu import graphs
from lapgpu.solver import adj_val_and_excess, extend_matrix, sddm_wrap_lap


def _dense_lap_solver(a, tol=1e-6, maxits=1000, maxtime=float("inf"), verbose=False, pcgIts=None, **kw):
    P = np.linalg.pinv(graphs.lap(a).toarray())

    def f(b, **kw):
        b = np.asarray(b, dtype=np.float64)
        return P @ (b - b.mean())

    f.n = a.shape[0]
    return f


def _sddm(excess_at, val=0.1, seed=0):
    a = graphs.weighted_grid2(8, 6, np.random.default_rng(seed), "expU2")
    ex = np.zeros(a.shape[0])
    ex[excess_at] = val
    return sp.csc_matrix(graphs.lap(a) + sp.diags(ex)), a


def test_sddm_solve_with_ground_vertex():
    M, _ = _sddm([0, 47])
    b = np.random.default_rng(1).random(48)
    x = sddm_wrap_lap(_dense_lap_solver, M)(b)
    assert np.linalg.norm(M @ x - b) <= 1e-10 * np.linalg.norm(b)


def test_extension_adds_one_vertex_joined_to_excess():
    M, a = _sddm([3, 10], val=0.25)
    a0, dval, dex = adj_val_and_excess(M)
    assert abs(a0 - a).max() < 1e-14
    a1 = extend_matrix(a0, dval, dex)
    assert a1.shape == (49, 49)
    col = a1[:, 48].toarray().ravel()
    assert set(np.flatnonzero(col)) == {3, 10} and np.allclose(col[[3, 10]], 0.25)
    assert abs(a1[:48, :48] - a).max() < 1e-14


def test_laplacian_input_is_not_extended():
    M, a = _sddm([])
    a0, dval, dex = adj_val_and_excess(M)
    assert extend_matrix(a0, dval, dex).shape == M.shape
    calls = []

    def solver(adj, **kw):
        calls.append(adj.shape)
        return _dense_lap_solver(adj, **kw)

    b = np.random.default_rng(2).standard_normal(48)
    b -= b.mean()
    x = sddm_wrap_lap(solver, M)(b)
    assert calls == [(48, 48)]
    assert np.linalg.norm(M @ x - b) <= 1e-10 * np.linalg.norm(b)


def test_excess_below_relative_threshold_is_ignored():
    # 100 eps * dVal is the reference's threshold; excess below it does not
    # add a ground vertex
    M, a = _sddm([5], val=1e-16)
    a0, dval, dex = adj_val_and_excess(M)
    assert extend_matrix(a0, dval, dex).shape == M.shape
