"""Factor cache (lapgpu.cache.GraphCache) on the host: the key follows the
adjacency and the seed; a miss factors and stores, a hit returns the very
same LDLinv (array for array) from a pickle-free .npz without factoring."""

import os

import numpy as np

from lapgpu import graphs
from lapgpu.cache import GraphCache, graph_key


def test_factor_cache_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    a = graphs.weighted_grid3(12, 12, 12, rng)
    b = graphs.weighted_grid3(12, 12, 12, rng)  # same shape, other weights
    assert graph_key(a) == graph_key(a.copy())
    assert graph_key(a) != graph_key(b)
    assert graph_key(a, seed=1) != graph_key(a, seed=2)
    c = GraphCache(str(tmp_path / "c"))
    k1, la1, f1, hit1 = c.factor(a)
    k2, la2, f2, hit2 = c.factor(a)
    assert (hit1, hit2) == (False, True) and k1 == k2
    for name in ("col", "colptr", "rowval", "fval", "d"):
        assert np.array_equal(getattr(f1, name), getattr(f2, name)), name
    assert (la1 != la2).nnz == 0
    assert c.stats["factor_hits"] == 1 and c.stats["factor_misses"] == 1
    # only whole files under their final names
    assert sorted(os.listdir(c.root)) == [f"{k1}.ldli.npz"]
    assert c.bytes() > 0
