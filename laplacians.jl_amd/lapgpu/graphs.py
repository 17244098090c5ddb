"""Graph generators and the Laplacian for the solve-phase workloads.

The reference builds its test graphs with `grid2`, `grid3`, `path_graph`,
`rand_matching` and `mapweight` (src/graphGenerators.jl:21-31, 219-241,
288-302; src/graphOps.jl:118-127, 157-184) and the Laplacian with
`lap(a) = Diagonal(a*1) - a` (src/graphOps.jl:25).  These are this package's
own generators with the same vertex numbering (last grid dimension fastest,
as `product_graph` numbers vertices) and their own documented RNG streams
(numpy PCG64 via `numpy.random.default_rng(seed)`); they are not
bit-identical to Julia's random streams.

Every generator returns a symmetric `scipy.sparse.csc_matrix` adjacency
(float64 values, zero diagonal, rows sorted in each column).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

__all__ = [
    "from_edges",
    "path_graph",
    "grid2",
    "grid3",
    "grid2_edges",
    "grid3_edges",
    "path_matchings",
    "lap",
    "weighted_grid2",
    "weighted_grid3",
    "random_tree",
    "cycle_graph",
    "star_graph",
]


def from_edges(n: int, i: np.ndarray, j: np.ndarray, w: np.ndarray | None = None) -> sp.csc_matrix:
    """Symmetric adjacency from undirected edges (i, j, w); duplicates are summed."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    if w is None:
        w = np.ones(len(i))
    w = np.asarray(w, dtype=np.float64)
    keep = i != j
    i, j, w = i[keep], j[keep], w[keep]
    a = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([i, j]), np.concatenate([j, i]))), shape=(n, n))
    a = a.tocsc()
    a.sum_duplicates()
    a.sort_indices()
    return a


def path_graph(n: int) -> sp.csc_matrix:
    """Path 0-1-...-(n-1) with unit weights (src/graphGenerators.jl:21-31)."""
    v = np.arange(n - 1)
    return from_edges(n, v, v + 1)


def grid2_edges(nx: int, ny: int) -> tuple[int, np.ndarray, np.ndarray]:
    """Undirected edges of the nx-by-ny grid, vertex (x, y) -> x*ny + y."""
    idx = np.arange(nx * ny).reshape(nx, ny)
    ei = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    ej = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    return nx * ny, ei, ej


def grid3_edges(n1: int, n2: int, n3: int) -> tuple[int, np.ndarray, np.ndarray]:
    """Undirected edges of the n1-by-n2-by-n3 grid, vertex -> (x*n2 + y)*n3 + z."""
    idx = np.arange(n1 * n2 * n3).reshape(n1, n2, n3)
    ei = np.concatenate([idx[:, :, :-1].ravel(), idx[:, :-1, :].ravel(), idx[:-1, :, :].ravel()])
    ej = np.concatenate([idx[:, :, 1:].ravel(), idx[:, 1:, :].ravel(), idx[1:, :, :].ravel()])
    return n1 * n2 * n3, ei, ej


def grid2(nx: int, ny: int | None = None) -> sp.csc_matrix:
    """Unit-weight 2-D grid (src/graphGenerators.jl:219-226)."""
    n, ei, ej = grid2_edges(nx, nx if ny is None else ny)
    return from_edges(n, ei, ej)


def grid3(n1: int, n2: int | None = None, n3: int | None = None) -> sp.csc_matrix:
    """Unit-weight 3-D grid (src/graphGenerators.jl:234-241)."""
    n2 = n1 if n2 is None else n2
    n3 = n1 if n3 is None else n3
    n, ei, ej = grid3_edges(n1, n2, n3)
    return from_edges(n, ei, ej)


def weighted_grid2(nx: int, ny: int, rng: np.random.Generator, weights: str = "expU2") -> sp.csc_matrix:
    """2-D grid with one random weight per undirected edge (as `mapweight` does)."""
    n, ei, ej = grid2_edges(nx, ny)
    return from_edges(n, ei, ej, _weights(len(ei), rng, weights))


def weighted_grid3(n1: int, n2: int, n3: int, rng: np.random.Generator, weights: str = "expU2") -> sp.csc_matrix:
    """3-D grid with one random weight per undirected edge (as `mapweight` does)."""
    n, ei, ej = grid3_edges(n1, n2, n3)
    return from_edges(n, ei, ej, _weights(len(ei), rng, weights))


def _weights(m: int, rng: np.random.Generator, kind: str) -> np.ndarray:
    if kind == "unit":
        return np.ones(m)
    if kind == "expU2":  # exp(U(-2, 2))
        return np.exp(rng.uniform(-2.0, 2.0, m))
    if kind == "expN":  # exp(N(0, 1))
        return np.exp(rng.standard_normal(m))
    if kind == "logU3":  # 10^U(-3, 3): log-uniform in [1e-3, 1e3]
        return 10.0 ** rng.uniform(-3.0, 3.0, m)
    raise ValueError(f"unknown weight kind {kind!r}")


def path_matchings(n: int, k: int, rng: np.random.Generator, weights: str = "logU3") -> sp.csc_matrix:
    """Union of the path 0..n-1 and k random perfect matchings
    (src/graphGenerators.jl:26-31, 288-302: each matching pairs p[:n/2] with
    p[n/2:] for a random permutation p).  Duplicate edges are merged and one
    weight is drawn per distinct undirected edge."""
    v = np.arange(n - 1)
    ei = [v]
    ej = [v + 1]
    h = n // 2
    for _ in range(k):
        p = rng.permutation(n)
        ei.append(p[:h])
        ej.append(p[h : 2 * h])
    ei = np.concatenate(ei)
    ej = np.concatenate(ej)
    lo = np.minimum(ei, ej)
    hi = np.maximum(ei, ej)
    key = np.unique(lo.astype(np.int64) * n + hi)
    lo = key // n
    hi = key % n
    return from_edges(n, lo, hi, _weights(len(lo), rng, weights))


def random_dense(n: int, deg: int, rng: np.random.Generator) -> sp.csc_matrix:
    """About n * deg / 2 uniformly random edges (self-pairs dropped, repeats
    summed) plus the path 0-1-...-(n-1), which keeps it connected; weights
    U(0.5, 2).  Its factor's last columns are long (tens to hundreds of
    entries), the case of long tail rows."""
    m = n * deg // 2
    i = rng.integers(0, n, m)
    j = rng.integers(0, n, m)
    keep = i != j
    a = from_edges(n, i[keep], j[keep], rng.uniform(0.5, 2.0, int(keep.sum())))
    return sp.csc_matrix(a + path_graph(n))


def random_tree(n: int, rng: np.random.Generator, weighted: bool = True) -> sp.csc_matrix:
    """Random recursive tree: vertex v > 0 attaches to a uniform earlier vertex."""
    v = np.arange(1, n)
    parent = np.floor(rng.uniform(0, 1, n - 1) * v).astype(np.int64)
    w = np.exp(rng.uniform(-2, 2, n - 1)) if weighted else None
    return from_edges(n, v, parent, w)


def cycle_graph(n: int, rng: np.random.Generator | None = None) -> sp.csc_matrix:
    """Cycle 0-1-...-(n-1)-0, random weights exp(U(-2,2)) if rng is given."""
    v = np.arange(n)
    w = None if rng is None else np.exp(rng.uniform(-2, 2, n))
    return from_edges(n, v, (v + 1) % n, w)


def star_graph(n: int) -> sp.csc_matrix:
    """Star with centre 0."""
    v = np.arange(1, n)
    return from_edges(n, np.zeros(n - 1, dtype=np.int64), v)


def lap(a: sp.spmatrix) -> sp.csc_matrix:
    """L = Diagonal(a*1) - a as CSC with an explicit diagonal (src/graphOps.jl:25)."""
    a = sp.csc_matrix(a)
    d = np.asarray(a.sum(axis=1)).ravel()
    L = (sp.diags(d, format="csc") - a).tocsc()
    L.sum_duplicates()
    L.sort_indices()
    return L
