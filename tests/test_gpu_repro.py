"""Run-to-run reproducibility: every row sum of the sweeps is formed in a
fixed order (rows split in chunks store partial sums that a fix-up step
adds in chunk order, no atomic adds), the reductions are fixed-order two-
stage sums and the dense block's product is fixed-order, so solving the
same right-hand side twice, on one context or on two contexts built from
the same factor, gives the same bits — for the plain, multi-RHS, FP32-inner
and batched paths, on graphs whose factor rows are long enough to be split
(a dense random graph, a path + 4 matchings expander)."""

import numpy as np
import pytest

import lapgpu
from lapgpu import _native, graphs

pytestmark = pytest.mark.gpu


def _rhs(n, rng, k=None):
    b = rng.standard_normal((n, k) if k else n)
    return b - b.mean(axis=0)


@pytest.fixture(scope="module")
def cases():
    rng = np.random.default_rng(41)
    out = []
    for a in (graphs.random_dense(1500, 120, rng), graphs.path_matchings(40000, 4, rng, "logU3")):
        out.append((a, lapgpu.lap(a), lapgpu.build_ldli(a, seed=1)))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_context_solves_repeat_bit_for_bit(cases, prec):
    rng = np.random.default_rng(42)
    for a, la, f in cases:
        n = a.shape[0]
        b = _rhs(n, rng)
        B = np.asfortranarray(_rhs(n, rng, 8))
        xs, Xs = [], []
        for _ in range(2):
            ctx = lapgpu.Context(0)
            try:
                ctx.upload_lap(la)
                ctx.upload_ldli(f)
                if prec == "fp32":
                    ctx.set_precision(_native.FP32_INNER, 20)
                for _ in range(2):
                    x, i = ctx.solve(b, tol=1e-10, maxits=1000)
                    xs.append((x, i["its"]))
                    X, j = ctx.solve_multi(B, tol=1e-10, maxits=1000)
                    Xs.append((X, j["its"]))
            finally:
                ctx.close()
        for x, its in xs[1:]:
            assert its == xs[0][1] and np.array_equal(x, xs[0][0])
        for X, its in Xs[1:]:
            assert its == Xs[0][1] and np.array_equal(X, Xs[0][0])


def test_batched_solves_repeat_bit_for_bit(cases):
    rng = np.random.default_rng(43)
    B = _native.Batch(0)
    try:
        ids, bs = [], []
        for a, la, f in cases:
            g = B.add_graph(la, f)
            for _ in range(3):
                ids.append(g)
                bs.append(_rhs(a.shape[0], rng))
        x0, i0 = B.solve(ids, bs, tol=1e-10)
        # another launch, the requests in another order and mixed with others
        order = list(reversed(range(len(ids))))
        x1, i1 = B.solve([ids[k] for k in order] + ids[:2], [bs[k] for k in order] + bs[:2], tol=1e-10)
        for pos, k in enumerate(order):
            assert i1[pos]["its"] == i0[k]["its"] and np.array_equal(x1[pos], x0[k])
    finally:
        B.close()
