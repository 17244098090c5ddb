"""Pins the CPU oracle (oracle/lap_oracle.c) against facts that do not come
from the oracle itself, so that a plausible restatement mistake (a dropped
term, a wrong sign or index, a transposed operand, a wrong loop order) fails:

* SpMV: scipy's CSR product and a dense product; L*1 = 0; grid2 interior
  rows are 4 x_i - sum of the neighbours.
* forward!/backward!/LDLsolver (src/approxChol.jl:1347-1408):
  - on trees and cycles the randomized elimination is exact (every column
    has at most two entries), so LDLsolver(b) = pinv(L) b;
  - on any factor, backward! is the transpose of forward! (both built as
    dense matrices column by column), LDLsolver is symmetric, positive on
    mean-zero vectors and returns mean-zero vectors;
  - the closed form of one column (y[j_k] += w_k y_i, y_i <- P y_i).
* pcg (src/pcg.jl:256-445): iterates and iteration count of a textbook PCG
  and of scipy's cg on the same operator and preconditioner; an exact
  preconditioner converges in one iteration; b = 0 returns zeros; maxits and
  maxtime stop where the reference stops; stag_test and the bestx rule.
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import lapgpu
from lapgpu import graphs


def _mean0(v):
    return v - v.mean()


# ---------------------------------------------------------------- SpMV ------

@pytest.mark.parametrize("seed", [0, 1])
def test_spmv_against_scipy_and_dense(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    a = graphs.path_matchings(300, 3, rng)
    la = lapgpu.lap(a)
    x = rng.standard_normal(300)
    y = oracle_mod.spmv(la, x)
    assert np.allclose(y, la.tocsr() @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(y, la.toarray() @ x, rtol=1e-13, atol=1e-13)
    assert np.abs(oracle_mod.spmv(la, np.ones(300))).max() < 1e-9


def test_spmv_grid2_closed_form(oracle_mod):
    a = graphs.grid2(7)
    la = lapgpu.lap(a)
    x = np.random.default_rng(3).standard_normal(49)
    y = oracle_mod.spmv(la, x)
    X = x.reshape(7, 7)
    Y = y.reshape(7, 7)
    for i in range(1, 6):
        for j in range(1, 6):
            expect = 4 * X[i, j] - X[i - 1, j] - X[i + 1, j] - X[i, j - 1] - X[i, j + 1]
            assert abs(Y[i, j] - expect) < 1e-12


# ------------------------------------------------------- LDLinv apply -------

EXACT_GRAPHS = {
    "tree_w": lambda: graphs.random_tree(120, np.random.default_rng(1)),
    "tree_unit": lambda: graphs.random_tree(90, np.random.default_rng(2), weighted=False),
    "path": lambda: graphs.path_graph(50),
    "star": lambda: graphs.star_graph(40),
    "cycle_w": lambda: graphs.cycle_graph(60, np.random.default_rng(3)),
    "cycle_unit": lambda: graphs.cycle_graph(33),
}


@pytest.mark.parametrize("name", sorted(EXACT_GRAPHS))
def test_ldlsolver_is_pinv_on_exact_factors(oracle_mod, name):
    a = EXACT_GRAPHS[name]()
    la = lapgpu.lap(a).toarray()
    f = lapgpu.build_ldli(a, seed=7)
    assert f.max_col_len() <= 2  # precondition for exactness
    Lp = np.linalg.pinv(la)
    rng = np.random.default_rng(11)
    for _ in range(3):
        b = _mean0(rng.standard_normal(la.shape[0]))
        z = oracle_mod.ldlsolver(f, b)
        ref = Lp @ b
        assert np.linalg.norm(z - ref) <= 1e-10 * np.linalg.norm(ref)


def _dense_map(fn, n):
    M = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = fn(e)
    return M


@pytest.mark.parametrize("gen", ["grid2", "pm", "grid3w"])
def test_backward_is_transpose_of_forward(oracle_mod, gen):
    rng = np.random.default_rng(5)
    a = {"grid2": lambda: graphs.grid2(9),
         "pm": lambda: graphs.path_matchings(80, 4, rng),
         "grid3w": lambda: graphs.weighted_grid3(4, 5, 4, rng)}[gen]()
    f = lapgpu.build_ldli(a, seed=3)
    n = a.shape[0]
    F = _dense_map(lambda e: oracle_mod.forward(f, e), n)
    B = _dense_map(lambda e: oracle_mod.backward(f, e), n)
    assert np.allclose(B, F.T, rtol=1e-12, atol=1e-14)
    # forward is unit-triangular in elimination order up to the diagonal P
    order = np.concatenate([f.col, np.setdiff1d(np.arange(n), f.col)])
    Fo = F[np.ix_(order, order)]
    assert np.allclose(np.triu(Fo, 1), 0.0, atol=1e-14)


def test_forward_single_column_closed_form(oracle_mod):
    """One column with entries (j1,f1),(j2,f2),(root,1): y[j1] += f1 y_i,
    y[j2] += f2 (1-f1) y_i, y[root] += (1-f1)(1-f2) y_i, y_i <- (1-f1)(1-f2) y_i."""
    f1, f2 = 0.3, 0.55
    fac = lapgpu.LDLinv(col=np.array([0, 1, 2]), colptr=np.array([0, 3, 4, 5]),
                        rowval=np.array([1, 2, 3, 3, 3]), fval=np.array([f1, f2, 1.0, 1.0, 1.0]),
                        d=np.ones(4))
    y = np.array([2.0, 0.0, 0.0, 0.0])
    out = oracle_mod.forward(fac, y)
    P = (1 - f1) * (1 - f2)
    # column 0 then columns 1 and 2 (each a single sentinel) push into 3
    y1 = f1 * 2.0
    y2 = f2 * (1 - f1) * 2.0
    expect = np.array([P * 2.0, y1, y2, P * 2.0 + y1 + y2])
    assert np.allclose(out, expect, rtol=1e-15)


@pytest.mark.parametrize("seed", [1, 2])
def test_ldlsolver_symmetric_positive_meanzero(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    a = graphs.weighted_grid2(12, 9, rng, "expN")
    f = lapgpu.build_ldli(a, seed=seed)
    n = a.shape[0]
    u = _mean0(rng.standard_normal(n))
    v = _mean0(rng.standard_normal(n))
    Mu = oracle_mod.ldlsolver(f, u)
    Mv = oracle_mod.ldlsolver(f, v)
    assert abs(Mu.mean()) < 1e-13 * np.abs(Mu).max()
    assert abs(u @ Mv - Mu @ v) <= 1e-12 * abs(u @ Mv)
    assert u @ Mu > 0


# ---------------------------------------------------------------- pcg -------

def _textbook_pcg(A, b, pre, tol, maxits):
    """Hestenes-Stiefel PCG with the relative residual test ||r||/||b|| < tol."""
    x = np.zeros_like(b)
    r = b.copy()
    z = pre(r)
    p = z.copy()
    rho = r @ z
    nb = np.linalg.norm(b)
    for it in range(1, maxits + 1):
        q = A @ p
        al = rho / (p @ q)
        x = x + al * p
        r = r - al * q
        if np.linalg.norm(r) / nb < tol:
            return x, it
        z = pre(r)
        rho, old = r @ z, rho
        p = z + (rho / old) * p
    return x, maxits


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_pcg_matches_textbook_and_scipy(oracle_mod, tol):
    # the setup of test/testPCG.jl:1-11: mapweight(grid2(20), 1/rand), Jacobi
    rng = np.random.default_rng(1)
    n0, ei, ej = graphs.grid2_edges(20, 20)
    a = graphs.from_edges(n0, ei, ej, 1.0 / rng.uniform(size=len(ei)))
    la = lapgpu.lap(a)
    b = _mean0(rng.standard_normal(400))
    dg = la.diagonal()
    x, info = oracle_mod.pcg_diag(la, dg, b, tol=tol, maxits=2000)
    xt, itt = _textbook_pcg(la, b, lambda r: r / dg, tol, 2000)
    assert info["stop"] == "converged"
    assert abs(info["its"] - itt) <= 1
    assert np.linalg.norm(x - xt) <= 1e-6 * np.linalg.norm(xt)
    cnt = [0]
    M = spla.LinearOperator((400, 400), matvec=lambda r: r / dg)
    xs, flag = spla.cg(la, b, rtol=tol, atol=0.0, maxiter=2000, M=M, callback=lambda _: cnt.__setitem__(0, cnt[0] + 1))
    assert flag == 0
    assert abs(info["its"] - cnt[0]) <= 1
    assert np.linalg.norm(la @ x - b) / np.linalg.norm(b) < tol * 1.0001


def test_pcg_ldl_matches_textbook(oracle_mod):
    a = graphs.weighted_grid3(8, 9, 7, np.random.default_rng(4))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=2)
    b = _mean0(np.random.default_rng(0).standard_normal(a.shape[0]))
    x, info = oracle_mod.pcg_ldl(la, f, b, tol=1e-10, maxits=500)
    xt, itt = _textbook_pcg(la, b, lambda r: oracle_mod.ldlsolver(f, r), 1e-10, 500)
    assert info["stop"] == "converged" and abs(info["its"] - itt) <= 1
    assert np.linalg.norm(x - xt) <= 1e-7 * np.linalg.norm(xt)


def test_exact_preconditioner_one_iteration(oracle_mod):
    a = graphs.random_tree(200, np.random.default_rng(8))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    b = _mean0(np.random.default_rng(1).standard_normal(200))
    x, info = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=100)
    assert info["its"] == 1 and info["stop"] == "converged"
    assert np.linalg.norm(la @ x - b) <= 1e-9 * np.linalg.norm(b)


def test_zero_rhs_returns_zeros(oracle_mod):
    a = graphs.grid2(6)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    x, info = oracle_mod.approxchol_solve(la, f, np.zeros(36), tol=1e-8)
    assert info["stop"] == "zero_rhs" and info["its"] == 0 and not x.any()
    x, info = oracle_mod.approxchol_solve(la, f, np.full(36, 2.5), tol=1e-8)  # b .- mean(b) == 0
    assert info["stop"] == "zero_rhs" and not x.any()


def test_maxits_and_bestx(oracle_mod):
    a = graphs.weighted_grid2(15, 15, np.random.default_rng(2))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    b = _mean0(np.random.default_rng(3).standard_normal(225))
    x0, i0 = oracle_mod.approxchol_solve(la, f, b, tol=1e-14, maxits=0)
    assert i0["its"] == 0 and not x0.any()
    x1, i1 = oracle_mod.approxchol_solve(la, f, b, tol=1e-14, maxits=1)
    assert i1["its"] == 1 and i1["stop"] == "maxits"
    # one iteration: x = al * z0 with z0 = pre(b), al = rho/<p, L p>
    z0 = oracle_mod.ldlsolver(f, b)
    al = (b @ z0) / (z0 @ (la @ z0))
    if i1["bestnr"] < 1.0:
        assert np.allclose(x1, al * z0, rtol=1e-12, atol=1e-15)
    else:  # the reference returns bestx = zeros when nr never drops below 1
        assert not x1.any()


def test_maxtime_stops_after_one_iteration(oracle_mod):
    a = graphs.weighted_grid2(30, 30, np.random.default_rng(2))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    b = _mean0(np.random.default_rng(3).standard_normal(900))
    _, info = oracle_mod.approxchol_solve(la, f, b, tol=1e-14, maxits=100, maxtime=1e-10)
    assert info["its"] == 1 and info["stop"] == "maxtime"


def test_stag_test_counter(oracle_mod):
    """stag_test = 1 gives the factor 1 - 1/1 = 0: no positive rho is below
    best_rho * 0, so every iteration counts and pcg stops at itcnt = 2
    (src/pcg.jl:370-383)."""
    a = graphs.grid2(12)
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    b = _mean0(np.random.default_rng(0).standard_normal(144))
    _, info = oracle_mod.approxchol_solve(la, f, b, tol=1e-12, maxits=100, stag_test=1)
    assert info["stop"] == "stag_test" and info["its"] == 2
    # stag_test = 0 disables the counter (1 - 1/0 = -Inf in Julia)
    _, info0 = oracle_mod.approxchol_solve(la, f, b, tol=1e-12, maxits=100, stag_test=0)
    assert info0["stop"] == "converged"


def test_stagnation_checks_every_tenth_iteration(oracle_mod):
    """With tol = 0 the solve can only end by a stagnation rule; the al*|p|
    check runs only when itcnt % 10 == 0 (src/pcg.jl:308-317)."""
    a = graphs.random_tree(200, np.random.default_rng(8))
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a)
    b = _mean0(np.random.default_rng(0).standard_normal(200))
    x, info = oracle_mod.approxchol_solve(la, f, b, tol=0.0, maxits=3000, stag_test=5)
    assert info["stop"] == "stag_al_p" and info["its"] % 10 == 0
    assert np.linalg.norm(la @ x - b) <= 1e-12 * np.linalg.norm(b)
