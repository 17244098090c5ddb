"""GPU against the committed golden fixtures (tests/golden/*.npz): the factor
stored in the fixture is uploaded as is, so these also cover a factor that
did not come from this process's builder run."""

import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import lapgpu
from _bars import assert_relres, assert_relres_cols

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_gpu_matches_golden(path, gpu_ctx_factory):
    z = np.load(path, allow_pickle=False)
    n = len(z["d"])
    a = sp.csc_matrix((z["a_data"], z["a_indices"], z["a_indptr"]), shape=(n, n))
    la = lapgpu.lap(a)
    f = lapgpu.LDLinv(z["col"], z["colptr"], z["rowval"], z["fval"], z["d"])
    ctx = gpu_ctx_factory(la, f)
    b = z["b"]
    assert np.linalg.norm(ctx.apply_ldli(b) - z["ldl_b"]) <= 1e-12 * np.linalg.norm(z["ldl_b"])
    assert np.linalg.norm(ctx.spmv(b) - z["spmv_b"]) <= 1e-12 * np.linalg.norm(z["spmv_b"])
    x, info = ctx.solve(b, tol=1e-8, maxits=1000, stag_test=5)
    assert abs(info["its"] - int(z["its"])) <= 2 and info["stop"] == str(z["stop"])
    assert np.linalg.norm(x - z["x"]) <= 1e-6 * np.linalg.norm(z["x"])
    assert_relres(la, x, b, 1e-8, f)
