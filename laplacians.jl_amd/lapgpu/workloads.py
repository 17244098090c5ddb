"""The measured workloads, one generator each, shared by the benchmark and
the parity tests.  Streams are numpy PCG64 (`default_rng(seed)`), drawn in
the order written here; the graphs follow the reference's generators
(grid2/grid3 `src/graphGenerators.jl:219-245`, weights as `mapweight`
`src/graphOps.jl:118-127`, path `src/graphGenerators.jl:26-31` and random
perfect matchings `:290-302`), not Julia's random streams.

1. grid2(100), unit weights, b ~ N(0,1) from seed 0;
2. grid3(128), weights exp(U(-2,2)) and then b from seed 2;
3. path + 4 random perfect matchings, n = 1e6, weights 10^U(-3,3), then b,
   seed 3;
4. grid2(1024) with weights exp(N(0,1)), a 64-column block B, and 256 small
   unit-weight grids (n log-uniform in [1e3, 1e5], aspect ratio U(1, 8))
   with one right-hand side each, seed 4;
5. serving: resident and new graphs drawn from 2-D / 3-D grids with
   exp(U(-2,2)) weights and path + matchings graphs (`serve_graph`), seed 5.

Every right-hand side is projected to sum zero.
"""

from __future__ import annotations

import numpy as np

from . import graphs


def _center(b):
    return b - b.mean(axis=-1, keepdims=True)


def config1():
    a = graphs.grid2(100)
    b = np.random.default_rng(0).standard_normal(a.shape[0])
    return a, _center(b)


def config2(side: int = 128):
    rng = np.random.default_rng(2)
    a = graphs.weighted_grid3(side, side, side, rng, "expU2")
    b = rng.standard_normal(a.shape[0])
    return a, _center(b)


def config3(n: int = 1_000_000):
    rng = np.random.default_rng(3)
    a = graphs.path_matchings(n, 4, rng, "logU3")
    b = rng.standard_normal(a.shape[0])
    return a, _center(b)


def small_grids(rng, count: int):
    """`count` unit-weight 2-D grids, n log-uniform in [1e3, 1e5], aspect
    ratio uniform in [1, 8]."""
    out = []
    for _ in range(count):
        nv = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        ar = rng.uniform(1.0, 8.0)
        ny = max(2, int(round(np.sqrt(nv / ar))))
        nx = max(2, int(round(nv / ny)))
        out.append(graphs.grid2(nx, ny))
    return out


def config4(rhs: int = 64, small: int = 256, side: int = 1024):
    """(a, B with rhs rows of length n, the small grids, their rhs)."""
    rng = np.random.default_rng(4)
    a = graphs.weighted_grid2(side, side, rng, "expN")
    B = _center(rng.standard_normal((rhs, a.shape[0])))
    smalls = small_grids(rng, small)
    sb = [_center(rng.standard_normal(g.shape[0])) for g in smalls]
    return a, B, smalls, sb


def serve_graph(rng, nv: float):
    """One graph of about nv vertices from the serving families: a 2-D grid
    (aspect ratio U(1, 4)) or a 3-D grid with exp(U(-2,2)) weights, or a
    path + 4 matchings graph with 10^U(-3,3) weights, each with
    probability 1/3."""
    kind = rng.integers(3)
    if kind == 0:
        ar = rng.uniform(1.0, 4.0)
        ny = max(2, int(round(np.sqrt(nv / ar))))
        return graphs.weighted_grid2(max(2, int(round(nv / ny))), ny, rng, "expU2")
    if kind == 1:
        s = max(2, int(round(nv ** (1 / 3))))
        return graphs.weighted_grid3(s, s, max(2, int(round(nv / (s * s)))), rng, "expU2")
    return graphs.path_matchings(max(4, int(nv) // 2 * 2), 4, rng, "logU3")


def config5(nreq: int = 10000, qps: float = 200.0, ngraphs: int = 32, min_n: float = 1e4,
            max_n: float = 4e6, new_frac: float = 0.2, new_max_n: float | None = None, tol: float = 1e-8,
            seed: int = 5):
    """The serving stream: `ngraphs` resident graphs with n log-uniform in
    [min_n, max_n], then `nreq` requests with exponential inter-arrival
    times at `qps`; each brings a new graph (n log-uniform in [min_n,
    new_max_n], default max_n) with probability new_frac, else names a
    resident graph uniformly.  Returns (resident graphs, requests)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(min_n), np.log(max_n)
    sizes = np.exp(rng.uniform(lo, hi, ngraphs))
    G = [serve_graph(rng, nv) for nv in sizes]
    return G, _requests(rng, G, nreq, qps, min_n, new_max_n if new_max_n else max_n, new_frac, tol, 0)


def serve_requests(G, nreq: int, qps: float = 200.0, min_n: float = 1e4, new_max_n: float = 4e6,
                   new_frac: float = 0.2, tol: float = 1e-8, seed: int = 55, rid0: int = 0):
    """Another stream of the serving kind over the resident graphs G (its own
    seed; request ids from rid0)."""
    return _requests(np.random.default_rng(seed), G, nreq, qps, min_n, new_max_n, new_frac, tol, rid0)


def _requests(rng, G, nreq, qps, min_n, new_max_n, new_frac, tol, rid0):
    """Arrival times, kinds and sizes from `rng` in request order; each
    request's new graph and right-hand side from its own stream
    default_rng([entropy, request id]) (entropy drawn from `rng`), built on
    host threads (numpy fills arrays without holding the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    from .serve import Request

    arr = np.cumsum(rng.exponential(1.0 / qps, nreq))
    new_sizes = np.exp(rng.uniform(np.log(min_n), np.log(new_max_n), nreq))
    is_new = rng.uniform(size=nreq) < new_frac
    which = rng.integers(len(G), size=nreq)
    entropy = int(rng.integers(2**62))

    def make(i):
        r = np.random.default_rng([entropy, rid0 + i])
        if is_new[i]:
            g = serve_graph(r, new_sizes[i])
            return Request(rid0 + i, float(arr[i]), _center(r.standard_normal(g.shape[0])), new_graph=g, tol=tol)
        gi = int(which[i])
        return Request(rid0 + i, float(arr[i]), _center(r.standard_normal(G[gi].shape[0])), graph=gi, tol=tol)

    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(make, range(nreq)))
