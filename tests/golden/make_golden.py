"""Writes the small golden fixtures used by tests/test_host.py and
tests/test_gpu_golden.py.

The reference ships no golden vectors for the solve path, so these are the
build's own: graphs from lapgpu.graphs, the LDLinv from liblapfactor (seed 1),
a sum-zero N(0,1) right-hand side, and the CPU oracle's outputs (one
LDLsolver application, one SpMV, the full solve closure at tol 1e-8).  The
oracle itself is pinned independently by tests/test_oracle.py.

Run: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "laplacians.jl_amd")]

import lapgpu  # noqa: E402
from lapgpu import graphs  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = {
    "grid2_10": lambda: graphs.grid2(10),
    "tree_50": lambda: graphs.random_tree(50, np.random.default_rng(5)),
    "pathmatch_300": lambda: graphs.path_matchings(300, 4, np.random.default_rng(3)),
    "grid3_6w": lambda: graphs.weighted_grid3(6, 6, 6, np.random.default_rng(2)),
}


def main():
    oracle.build()
    for name, gen in CASES.items():
        a = gen()
        la = lapgpu.lap(a)
        f = lapgpu.build_ldli(a, seed=1)
        n = a.shape[0]
        b = np.random.default_rng(0).standard_normal(n)
        b -= b.mean()
        z = oracle.ldlsolver(f, b)
        y = oracle.spmv(la, b)
        x, info = oracle.approxchol_solve(la, f, b, tol=1e-8, maxits=1000, stag_test=5)
        np.savez(os.path.join(HERE, f"{name}.npz"),
                 a_indptr=a.indptr.astype(np.int64), a_indices=a.indices.astype(np.int64), a_data=a.data,
                 col=f.col, colptr=f.colptr, rowval=f.rowval, fval=f.fval, d=f.d,
                 b=b, ldl_b=z, spmv_b=y, x=x, its=np.int64(info["its"]), stop=info["stop"])
        print(name, n, f.nnz, info)


if __name__ == "__main__":
    main()
