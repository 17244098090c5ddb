"""The other factor builders of the reference on this host (liblapfactor):
the fixed-order elimination approxChol(LLMatOrd(a, p)) for ApproxCholParams
orders :given, :wdeg and :random (src/approxChol.jl:203-296, 758-862,
1628-1790), and the split-edge eliminations approxChol(LLmatp(a, split),
split[, merge]) (src/approxChol.jl:138-186, 497-603, 979-1339) behind
approxchol_lap2 (src/approxChol.jl:1559-1579).

No golden vectors exist for them (the reference's RNG stream is Julia's), so
each is pinned by what the mathematics fixes: the LDLinv layout invariants;
exactness where the elimination has no choice (every vertex of a path or a
cycle has at most two live neighbours when it is eliminated, in any order,
and min-degree eliminates a tree leaf first), checked as L x = b through the
oracle's LDLsolver; and, on grids, preconditioned CG convergence."""

import numpy as np
import pytest
import scipy.sparse as sp

import lapgpu
from lapgpu import graphs

BUILDERS = {
    "given": dict(order="given"),
    "wdeg": dict(order="wdeg"),
    "random": dict(order="random"),
    "split5_merge2": dict(order="deg", split=5, merge=2),
    "split2": dict(order="deg", split=2, merge=0),
    "split3_merge1": dict(order="deg", split=3, merge=1),
}


def _check_structure(f, n):
    assert len(f.col) == n - 1 and len(f.colptr) == n and len(f.d) == n
    assert len(np.unique(f.col)) == n - 1
    root = np.setdiff1d(np.arange(n), f.col)
    assert len(root) == 1 and f.d[root[0]] == 0.0
    assert np.all(f.d[f.col] > 0)
    pos = np.empty(n, dtype=np.int64)
    pos[f.col] = np.arange(n - 1)
    pos[root] = n - 1
    assert f.colptr[0] == 0 and f.colptr[-1] == f.nnz and np.all(np.diff(f.colptr) >= 1)
    for ii in range(n - 1):
        k0, k1 = f.colptr[ii], f.colptr[ii + 1]
        rows = f.rowval[k0:k1]
        assert np.all(pos[rows] > ii)
        assert len(np.unique(rows)) == len(rows)
        assert f.fval[k1 - 1] == 1.0
        assert np.all((f.fval[k0:k1 - 1] > 0) & (f.fval[k0:k1 - 1] < 1))


def _cycle(n, rng):
    i = np.arange(n)
    return graphs.from_edges(n, i, (i + 1) % n, 10.0 ** rng.uniform(-2, 2, n))


def _path(n, rng):
    i = np.arange(n - 1)
    return graphs.from_edges(n, i, i + 1, 10.0 ** rng.uniform(-2, 2, n - 1))


@pytest.mark.parametrize("name", list(BUILDERS))
@pytest.mark.parametrize("gen", ["grid2", "pm", "tree"])
def test_structure(name, gen):
    rng = np.random.default_rng(1)
    a = {"grid2": lambda: graphs.weighted_grid2(20, 18, rng),
         "pm": lambda: graphs.path_matchings(1000, 4, rng),
         "tree": lambda: graphs.random_tree(300, rng)}[gen]()
    f = lapgpu.build_ldli(a, seed=3, **BUILDERS[name])
    _check_structure(f, a.shape[0])


@pytest.mark.parametrize("name", list(BUILDERS))
@pytest.mark.parametrize("gen", ["path", "cycle"])
def test_exact_on_paths_and_cycles(name, gen, oracle_mod):
    rng = np.random.default_rng(5)
    a = (_path if gen == "path" else _cycle)(200, rng)
    f = lapgpu.build_ldli(a, seed=2, **BUILDERS[name])
    L = lapgpu.lap(a)
    b = rng.standard_normal(200)
    b -= b.mean()
    x = oracle_mod.ldlsolver(f, b)
    assert np.linalg.norm(L @ x - b) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("name", ["split5_merge2", "split2", "split3_merge1"])
def test_split_exact_on_trees(name, oracle_mod):
    rng = np.random.default_rng(7)
    a = graphs.random_tree(400, rng)
    f = lapgpu.build_ldli(a, seed=4, **BUILDERS[name])
    assert f.max_col_len() <= 2
    L = lapgpu.lap(a)
    b = rng.standard_normal(400)
    b -= b.mean()
    x = oracle_mod.ldlsolver(f, b)
    assert np.linalg.norm(L @ x - b) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("name", list(BUILDERS))
def test_preconditions_pcg_on_grids(name, oracle_mod):
    rng = np.random.default_rng(9)
    a = graphs.weighted_grid2(30, 30, rng, "expU2")
    la = lapgpu.lap(a)
    f = lapgpu.build_ldli(a, seed=1, **BUILDERS[name])
    b = rng.standard_normal(900)
    b -= b.mean()
    x, info = oracle_mod.approxchol_solve(la, f, b, tol=1e-8, maxits=300, stag_test=5)
    assert info["stop"] == "converged" and info["its"] < 150
    assert np.linalg.norm(la @ x - b) <= 1e-7 * np.linalg.norm(b)


def test_elimination_orders():
    rng = np.random.default_rng(2)
    a = graphs.weighted_grid2(12, 10, rng)
    n = a.shape[0]
    assert lapgpu.elimination_order(a, "given") is None
    p = lapgpu.elimination_order(a, "random", seed=4)
    assert np.array_equal(np.sort(p), np.arange(n))
    p = lapgpu.elimination_order(a, "wdeg", seed=4)
    assert np.array_equal(np.sort(p), np.arange(n))
    deg = np.asarray(a.sum(axis=0)).ravel()
    # sorted by degree times a factor in [1, 2): a vertex never comes after
    # one of more than twice its degree
    assert np.all(deg[p][:, None] <= 2 * deg[p][None, :] + np.tril(np.full((n, n), np.inf), -1))


def test_given_order_follows_perm_and_rejects_bad_perm():
    rng = np.random.default_rng(3)
    a = graphs.weighted_grid2(8, 7, rng)
    p = rng.permutation(56)
    f = lapgpu.build_ldli(a, seed=1, order="given", perm=p)
    assert np.array_equal(f.col, p[:-1])
    with pytest.raises(RuntimeError, match="permutation"):
        lapgpu.build_ldli(a, order="given", perm=np.zeros(56, dtype=np.int64))


def test_seed_determinism():
    a = graphs.weighted_grid2(15, 15, np.random.default_rng(4))
    for kw in BUILDERS.values():
        f1 = lapgpu.build_ldli(a, seed=8, **kw)
        f2 = lapgpu.build_ldli(a, seed=8, **kw)
        assert np.array_equal(f1.rowval, f2.rowval) and np.array_equal(f1.fval, f2.fval)


def test_lap2_reads_params_as_the_reference(monkeypatch):
    from lapgpu import solver

    seen = {}

    def fake(s, a, **kw):
        seen.update(kw["params"].__dict__)
        return None

    monkeypatch.setattr(solver, "lap_wrap_components", fake)
    solver.approxchol_lap2_gpu(graphs.grid2(3))
    assert seen == {"order": "deg", "stag_test": 5, "split": 5, "merge": 2}
    with pytest.raises(ValueError, match="negative"):
        solver.approxchol_lap2_gpu(sp.csc_matrix(-graphs.grid2(3)))
