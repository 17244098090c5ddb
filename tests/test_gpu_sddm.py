"""SDDM callers on the GPU solve: approxchol_sddm (src/approxChol.jl:1862 =
sddmWrapLap(approxchol_lap)) and harmonic_interp (src/harmonic.jl:16-33),
checked as the reference's own tests check them (test/testSolvers.jl:
119-137 for an SDDM with diagonal excess at the first and last vertex;
test/testByExport.jl:397-404 for harmonic_interp: x[S] == vals and L x = 0
off S) and against scipy's sparse direct solve."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import lapgpu
from lapgpu import graphs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["grid2", "path_matchings"])
def test_approxchol_sddm_excess_first_last(kind):
    rng = np.random.default_rng(3)
    a = (graphs.weighted_grid2(60, 50, rng, "expU2") if kind == "grid2"
         else graphs.path_matchings(4000, 4, rng, "logU3"))
    n = a.shape[0]
    ex = np.zeros(n)
    ex[0] = ex[-1] = 0.1
    M = sp.csc_matrix(graphs.lap(a) + sp.diags(ex))
    b = rng.random(n)
    b -= b.mean()
    its = [0]
    f = lapgpu.approxchol_sddm_gpu(M, tol=1e-8)
    x = f(b, tol=1e-8, maxits=1000, pcgIts=its)
    assert 0 < its[0] < 1000
    # M x - b is the extended Laplacian system's residual without its last
    # row, and the extended right-hand side has the norm of b (b sums to 0)
    assert np.linalg.norm(M @ x - b) <= 1e-8 * np.linalg.norm(b)
    xr = spla.spsolve(M, b)
    assert np.linalg.norm(x - xr) <= 1e-4 * np.linalg.norm(xr)


def test_approxchol_sddm_laplacian_input():
    a = graphs.weighted_grid2(40, 30, np.random.default_rng(4), "expU2")
    L = graphs.lap(a)
    b = np.random.default_rng(5).standard_normal(a.shape[0])
    b -= b.mean()
    x = lapgpu.approxchol_sddm_gpu(L, tol=1e-8)(b)
    assert np.linalg.norm(L @ x - b) <= 1e-8 * np.linalg.norm(b)


def test_harmonic_interp():
    a = graphs.weighted_grid2(40, 35, np.random.default_rng(6), "expU2")
    S = np.array([0, 1, 2])
    vals = np.array([1.0, 2.0, 3.0])
    x = lapgpu.harmonic_interp(a, S, vals, tol=1e-10)
    assert np.array_equal(x[S], vals)
    r = graphs.lap(a) @ x
    r[S] = 0.0
    assert np.abs(r).sum() < 1e-6
    assert x.min() >= 1.0 - 1e-8 and x.max() <= 3.0 + 1e-8  # maximum principle
