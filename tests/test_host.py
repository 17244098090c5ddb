"""CPU tests of the host side: the LDLinv builder (liblapfactor), the golden
fixtures, the Python mirror of the reference's solver interface, and the
C-ABI library's symbol table (no GPU calls)."""

import ctypes
import glob
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import lapgpu
from lapgpu import graphs
from lapgpu import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))


def _load(path):
    z = np.load(path, allow_pickle=False)
    n = len(z["d"])
    a = sp.csc_matrix((z["a_data"], z["a_indices"], z["a_indptr"]), shape=(n, n))
    f = lapgpu.LDLinv(z["col"], z["colptr"], z["rowval"], z["fval"], z["d"])
    return z, a, f


# ------------------------------------------------------------ golden --------

@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_factor_is_reproduced(path):
    z, a, f = _load(path)
    g = lapgpu.build_ldli(a, seed=1)
    for k in ("col", "colptr", "rowval"):
        assert np.array_equal(getattr(g, k), z[k]), k
    assert np.array_equal(g.fval, z["fval"]) and np.array_equal(g.d, z["d"])


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_oracle_outputs(path, oracle_mod):
    z, a, f = _load(path)
    la = lapgpu.lap(a)
    assert np.allclose(oracle_mod.ldlsolver(f, z["b"]), z["ldl_b"], rtol=0, atol=1e-12 * np.abs(z["ldl_b"]).max())
    assert np.allclose(oracle_mod.spmv(la, z["b"]), z["spmv_b"], rtol=0, atol=1e-12 * np.abs(z["spmv_b"]).max())
    x, info = oracle_mod.approxchol_solve(la, f, z["b"], tol=1e-8, maxits=1000, stag_test=5)
    assert info["its"] == int(z["its"]) and info["stop"] == str(z["stop"])
    assert np.linalg.norm(la @ z["x"] - z["b"]) / np.linalg.norm(z["b"]) < 1e-8 * 1.01


# ------------------------------------------------------- factor build -------

@pytest.mark.parametrize("gen", ["grid2", "grid3w", "pm", "tree"])
def test_factor_structure(gen):
    rng = np.random.default_rng(0)
    a = {"grid2": lambda: graphs.grid2(30),
         "grid3w": lambda: graphs.weighted_grid3(10, 8, 9, rng),
         "pm": lambda: graphs.path_matchings(2000, 4, rng),
         "tree": lambda: graphs.random_tree(500, rng)}[gen]()
    n = a.shape[0]
    f = lapgpu.build_ldli(a, seed=4)
    assert len(f.col) == n - 1 and len(f.colptr) == n and len(f.d) == n
    assert len(np.unique(f.col)) == n - 1
    root = np.setdiff1d(np.arange(n), f.col)
    assert len(root) == 1 and f.d[root[0]] == 0.0
    assert np.all(f.d[f.col] > 0)
    pos = np.empty(n, dtype=np.int64)
    pos[f.col] = np.arange(n - 1)
    pos[root] = n - 1
    lens = np.diff(f.colptr)
    assert np.all(lens >= 1) and f.colptr[0] == 0 and f.colptr[-1] == f.nnz
    for ii in range(n - 1):
        k0, k1 = f.colptr[ii], f.colptr[ii + 1]
        rows = f.rowval[k0:k1]
        assert np.all(pos[rows] > ii)           # every row is eliminated later
        assert len(np.unique(rows)) == len(rows)  # multi-edges merged
        assert f.fval[k1 - 1] == 1.0             # sentinel
        assert np.all((f.fval[k0:k1 - 1] > 0) & (f.fval[k0:k1 - 1] < 1))


def test_factor_seed_determinism_and_variation():
    a = graphs.weighted_grid2(20, 20, np.random.default_rng(1))
    f1 = lapgpu.build_ldli(a, seed=1)
    f2 = lapgpu.build_ldli(a, seed=1)
    f3 = lapgpu.build_ldli(a, seed=2)
    assert np.array_equal(f1.rowval, f2.rowval) and np.array_equal(f1.fval, f2.fval)
    assert not (len(f1.rowval) == len(f3.rowval) and np.array_equal(f1.rowval, f3.rowval))


def test_factor_quality_condition_number(oracle_mod):
    """The property behind test/testSolvers.jl:103-107 (condNumber < 20), on
    the library's own factor: the extreme eigenvalues of LDLsolver * L on the
    mean-zero subspace have a ratio well below 20."""
    a = graphs.weighted_grid2(14, 13, np.random.default_rng(3), "expN")
    la = lapgpu.lap(a).toarray()
    f = lapgpu.build_ldli(a, seed=1)
    n = la.shape[0]
    M = np.column_stack([oracle_mod.ldlsolver(f, la[:, j]) for j in range(n)])
    ev = np.sort(np.linalg.eigvals(M).real)[1:]  # drop the null direction
    assert ev[0] > 0 and ev[-1] / ev[0] < 20


def test_negative_weight_is_an_error():
    a = graphs.grid2(5).tolil()
    a[0, 1] = a[1, 0] = -1.0
    with pytest.raises(ValueError, match="negative edge weights"):
        lapgpu.build_ldli(sp.csc_matrix(a))
    with pytest.raises(ValueError, match="negative edge weights"):
        lapgpu.approxchol_lap_gpu(sp.csc_matrix(a))


def test_isolated_vertex_is_an_error():
    a = sp.block_diag([graphs.grid2(3), sp.csc_matrix((1, 1))], format="csc")
    with pytest.raises(RuntimeError, match="isolated"):
        lapgpu.build_ldli(a)


def test_nonsymmetric_is_an_error():
    a = sp.csc_matrix(np.array([[0, 1.0, 0], [1.0, 0, 1.0], [0, 0, 0]]))
    with pytest.raises(RuntimeError):
        lapgpu.build_ldli(a)


# ---------------------------------------------------------- C-ABI -----------

def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:lapgpu|lapfactor)_[a-z0-9_]+)\s*\(", txt)))


def test_lapgpu_exports_every_declared_symbol():
    lib = _native.load()
    names = _declared("lapgpu.h")
    assert "lapgpu_solve" in names and "lapgpu_upload_ldli" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTS)


def test_lapfactor_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "laplacians.jl_amd", "lapgpu", "liblapfactor.so"))
    for name in _declared("lapfactor.h"):
        assert hasattr(lib, name), name


def test_create_reports_errors_through_last_error():
    lib = _native.load()
    h = ctypes.c_void_p()
    rc = lib.lapgpu_create(ctypes.byref(h), 99)
    assert rc != 0
    msg = lib.lapgpu_last_error(None).decode()
    assert "device" in msg


# ------------------------------------------------- interface mirror ---------

def test_params_defaults_match_reference():
    p = lapgpu.ApproxCholParams()
    assert (p.order, p.stag_test, p.split, p.merge) == ("deg", 5, 0, 0)

