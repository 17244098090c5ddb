"""Error of the GPU solve against an FP64 sparse direct solve on small
graphs (no CHOLMOD here: scipy's SuperLU on the Laplacian grounded at
vertex 0, then both solutions centred, since the solve's x has mean zero).

The bound that relres <= tol gives is ||x - x*|| / ||x*|| <= kappa * tol;
the test asserts a loose 1e-5 and prints the measured error of each case
(pytest -s), next to the true relres and the iteration count."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import lapgpu
from lapgpu import graphs
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu

CASES = {
    # the unweighted 100 x 100 grid with b from seed 0, factor seed 1
    "grid2_100": lambda: graphs.grid2(100),
    "grid3_16_expU2": lambda: graphs.weighted_grid3(16, 16, 16, np.random.default_rng(2), "expU2"),
    "pathmatch_3000_logU3": lambda: graphs.path_matchings(3000, 4, np.random.default_rng(3), "logU3"),
    "grid2_64x40_expN": lambda: graphs.weighted_grid2(64, 40, np.random.default_rng(4), "expN"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_error_vs_direct(name):
    a = sp.csc_matrix(CASES[name]())
    la = lapgpu.lap(a).tocsc()
    n = la.shape[0]
    b = np.random.default_rng(0).standard_normal(n)
    b -= b.mean()
    solve = lapgpu.approxchol_lap_gpu(a, tol=1e-8, maxits=1000)
    its = [0]
    x = solve(b, pcgIts=its)
    xd = np.zeros(n)
    xd[1:] = spla.spsolve(la[1:, 1:].tocsc(), b[1:])
    xd -= xd.mean()
    xc = x - x.mean()
    relres = np.linalg.norm(la @ x - b) / np.linalg.norm(b)
    err = np.linalg.norm(xc - xd) / np.linalg.norm(xd)
    print(f"\n{name}: n={n} its={its[0]} relres={relres:.2e} error vs direct={err:.2e}")
    assert_relres(la, x, b, 1e-8, solve.ldli, what=name)
    assert err <= 1e-5
