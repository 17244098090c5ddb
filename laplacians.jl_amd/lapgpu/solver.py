"""The drop-in: `approxchol_lap_gpu(a; tol, maxits, maxtime, verbose, pcgIts,
params)` returning the solver closure `f(b; ...)`, with the same names,
argument meaning and error behaviour as the reference's `approxchol_lap`
(src/approxChol.jl:1522-1545) and its wrappers (src/solverInterface.jl).

The build of the factor stays on the host (here liblapfactor, the C++
restatement of `approxChol(LLmatp(a))`, or a caller-supplied LDLinv from the
reference's Julia); every solve `f(b)` runs entirely on the GPU through
liblapgpu (include/lapgpu.h).  There is no CPU solve path.
"""

from __future__ import annotations

import math
import time
import warnings
from dataclasses import dataclass
from typing import Callable

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from . import _native
from .factor import LDLinv, build_ldli
from .graphs import lap

__all__ = [
    "ApproxCholParams",
    "approxchol_lap_gpu",
    "approxchol_lap1_gpu",
    "approxchol_lap2_gpu",
    "lap_wrap_components",
    "block_solver",
    "null_solver",
    "approxchol_sddm_gpu",
    "sddm_wrap_lap",
    "adj_val_and_excess",
    "extend_matrix",
    "harmonic_interp",
]


@dataclass
class ApproxCholParams:
    """ApproxCholParams (src/approxChol.jl:80-90); default (:deg, 5, 0, 0)."""

    order: str = "deg"
    stag_test: int = 5
    split: int = 0
    merge: int = 0


def null_solver(b, tol=1e-6, maxits=1000, maxtime=math.inf, verbose=False, pcgIts=None):
    """nullSolver: zeros for singleton components (src/solverInterface.jl)."""
    return np.zeros(len(b))


def _set_its(pcgIts, its):
    if pcgIts is not None and len(pcgIts) > 0:
        pcgIts[0] = its


class GpuLapSolver:
    """The solve closure of approxchol_lapGreedy (src/approxChol.jl:1624):
    f(b; tol, maxits, maxtime, verbose, pcgIts) = pcg(la, b .- mean(b), F; ...)
    with keyword defaults captured at construction.  One more optional
    output beside pcgIts: trueRelres[0] receives the FP64 true relative
    residual ||(b .- mean(b)) - la x|| / ||b .- mean(b)|| of the returned x
    (lapgpu_last_true_relres; in FP32-inner mode the solve's error against
    FP64), also kept in `last["true_relres"]`."""

    def __init__(self, ctx: _native.Context, ldli: LDLinv, tol, maxits, maxtime, verbose, pcgIts, stag_test):
        self.ctx = ctx
        self.ldli = ldli
        self.defaults = dict(tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts)
        self.stag_test = stag_test
        self.last = None

    def __call__(self, b, tol=None, maxits=None, maxtime=None, verbose=None, pcgIts=None, trueRelres=None):
        d = self.defaults
        tol = d["tol"] if tol is None else tol
        maxits = d["maxits"] if maxits is None else maxits
        maxtime = d["maxtime"] if maxtime is None else maxtime
        verbose = d["verbose"] if verbose is None else verbose
        pcgIts = d["pcgIts"] if pcgIts is None else pcgIts
        b = np.asarray(b, dtype=np.float64)
        if b.shape != (self.ctx.n,):
            raise ValueError(f"b has shape {b.shape}, expected ({self.ctx.n},)")
        mi = -1 if (maxits is None or (isinstance(maxits, float) and math.isinf(maxits))) else int(maxits)
        flags = _native.VERBOSE if verbose else 0
        x, info = self.ctx.solve(b, tol=tol, maxits=mi, maxtime=float(maxtime), stag_test=self.stag_test, flags=flags)
        info["true_relres"] = self.ctx.last_true_relres(1)
        self.last = info
        if trueRelres is not None and len(trueRelres) > 0:
            trueRelres[0] = info["true_relres"]
        # a zero right-hand side returns before pcg writes pcgIts
        # (src/pcg.jl:269-271)
        if info["stop"] != "zero_rhs":
            _set_its(pcgIts, info["its"])
        return x


_ORDER_NAMES = {"deg": "greedy degree", "wdeg": "wted degree", "random": "random"}


def approxchol_lap1_gpu(a, tol=1e-6, maxits=1000, maxtime=math.inf, verbose=False, pcgIts=None,
                        params: ApproxCholParams | None = None, seed: int = 1, device: int = 0,
                        ldli: LDLinv | None = None, precision: str = "fp64", replace_every: int = 20):
    """approxchol_lap1 (src/approxChol.jl:1791-1851) for a connected graph:
    la = lap(a) and the LDLinv of params.order on the host (approxchol_lapGreedy
    :deg with its split/merge forms, approxchol_lapWdeg, approxchol_lapRandom,
    and approxchol_lapGiven for any other order, src/approxChol.jl:1581-1790),
    both uploaded once; returns the GPU solve closure.  precision="fp32"
    selects the FP32-inner form (lapgpu_set_precision: SpMV and LDLinv apply
    in FP32, FP64 scalars, an FP64 residual replacement every replace_every
    iterations); the closure's trueRelres then reports each solve's FP64 true
    residual."""
    params = params or ApproxCholParams()
    t1 = time.time()
    a = sp.csc_matrix(a)
    la = lap(a)
    order = params.order if params.order in ("deg", "wdeg", "random") else "given"
    if ldli is None:
        ldli = build_ldli(a, seed=seed, order=order, split=params.split if order == "deg" else 0,
                          merge=params.merge if order == "deg" else 0)
    if precision not in ("fp64", "fp32"):
        raise ValueError(f"precision must be 'fp64' or 'fp32', not {precision!r}")
    ctx = _native.Context(device)
    ctx.upload_lap(la)
    ctx.upload_ldli(ldli)
    if precision == "fp32":
        ctx.set_precision(_native.FP32_INNER, replace_every)
    if verbose:
        print(f"Using {_ORDER_NAMES.get(order, 'given')} ordering. Factorization time: {time.time() - t1}")
        print(f"Factorization size: {2 * ldli.nnz}")
        print(f"Ratio of operator edges to original edges: {2 * ldli.nnz / a.nnz}")
        dg = la.diagonal()
        print(f"ratio of max to min diagonal of laplacian : {dg.max() / dg.min()}")
    return GpuLapSolver(ctx, ldli, tol, maxits, maxtime, verbose, pcgIts, params.stag_test)


def block_solver(comps, solvers, tol=1e-6, maxits=1000, maxtime=math.inf, verbose=False, pcgIts=None):
    """blockSolver (src/solverInterface.jl:152-188): apply the i-th solver on
    the i-th component; pcgIts[1] receives the max over components, and
    trueRelres[1] (a GPU closure's extra output) the true relative residual
    of the whole solution, sqrt(sum_c (t_c |b_c|)^2 / sum_c |b_c|^2) over the
    components' centred right-hand sides b_c."""
    defaults = dict(tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts)

    def f(b, tol=None, maxits=None, maxtime=None, verbose=None, pcgIts=None, trueRelres=None):
        tol = defaults["tol"] if tol is None else tol
        maxits = defaults["maxits"] if maxits is None else maxits
        maxtime = defaults["maxtime"] if maxtime is None else maxtime
        verbose = defaults["verbose"] if verbose is None else verbose
        pcgIts = defaults["pcgIts"] if pcgIts is None else pcgIts
        if pcgIts is not None and len(pcgIts) > 0:
            pcgIts[0] = 0
            tmp = [0]
        else:
            tmp = []
        b = np.asarray(b, dtype=np.float64)
        x = np.zeros(len(b))
        num = den = 0.0
        for i, ind in enumerate(comps):
            if verbose:
                print(f"BlockSolver component index {i + 1}")
            if trueRelres is not None and isinstance(solvers[i], GpuLapSolver):
                tr = [0.0]
                x[ind] = solvers[i](b[ind], tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=tmp,
                                    trueRelres=tr)
                bc = b[ind] - b[ind].mean()
                nb2 = float(bc @ bc)
                num += tr[0] ** 2 * nb2
                den += nb2
            else:
                x[ind] = solvers[i](b[ind], tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=tmp)
            if pcgIts is not None and len(pcgIts) > 0:
                pcgIts[0] = max(pcgIts[0], tmp[0])
        if trueRelres is not None and len(trueRelres) > 0:
            trueRelres[0] = math.sqrt(num / den) if den > 0 else 0.0
        return x

    return f


def lap_wrap_components(solver: Callable, a, tol=1e-6, maxits=1000, maxtime=math.inf, verbose=False,
                        pcgIts=None, **params):
    """lapWrapComponents (src/solverInterface.jl:212-275).  Deviation: the
    reference solves components of fewer than 50 vertices with CHOLMOD on the
    host; here every component with two or more vertices goes to the GPU
    solver (same tolerance contract), so no solve runs on the CPU."""
    t1 = time.time()
    a = sp.csc_matrix(a)
    if np.any(a.diagonal() != 0):
        warnings.warn("The matrix should not have any nonzero diagonal entries.")
        a = sp.csc_matrix(a - sp.diags(a.diagonal()))
        a.eliminate_zeros()
    ncomp, labels = connected_components(a, directed=False)
    if ncomp == 1:
        s = solver(a, tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts, **params)
        if verbose:
            print(f"Solver build time: {round(time.time() - t1, 3)} seconds.")
        return s
    order = np.argsort(labels, kind="stable")
    bounds = np.searchsorted(labels[order], np.arange(ncomp + 1))
    comps = [order[bounds[i]:bounds[i + 1]] for i in range(ncomp)]
    singles = [c for c in comps if len(c) == 1]
    others = [c for c in comps if len(c) > 1]
    solvers = []
    out_comps = []
    if singles:
        out_comps.append(np.concatenate(singles))
        solvers.append(null_solver)
    for ind in others:
        asub = a[ind][:, ind]
        solvers.append(solver(asub, tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts, **params))
        out_comps.append(ind)
    if verbose:
        print(f"Solver build time: {round(time.time() - t1, 3)} seconds.")
    return block_solver(out_comps, solvers, tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts)


def approxchol_lap_gpu(a, tol=1e-6, maxits=1000, maxtime=math.inf, verbose=False, pcgIts=None,
                       params: ApproxCholParams | None = None, **kw):
    """approxchol_lap (src/approxChol.jl:1522-1545) with the solve on the GPU.
    Raises ValueError for negative weights, as the reference's error()."""
    a = sp.csc_matrix(a)
    if a.nnz and a.data.min() < 0:
        raise ValueError("Adjacency matrix can not have negative edge weights")
    return lap_wrap_components(approxchol_lap1_gpu, a, tol=tol, maxits=maxits, maxtime=maxtime,
                               verbose=verbose, pcgIts=pcgIts, params=params or ApproxCholParams(), **kw)


def approxchol_lap2_gpu(a, tol=1e-6, maxits=1000, maxtime=math.inf, verbose=False, pcgIts=None,
                        params: ApproxCholParams | None = None, **kw):
    """approxchol_lap2 (src/approxChol.jl:1559-1579): the split-edge
    elimination under lapWrapComponents.  The reference passes
    ApproxCholParams(:deg, params.stag_test, 2), which its three-argument
    constructor (src/approxChol.jl:90) reads as stag_test 5, split =
    params.stag_test (5 by default) and merge 2; this does the same."""
    a = sp.csc_matrix(a)
    if a.nnz and a.data.min() < 0:
        raise ValueError("Adjacency matrix can not have negative edge weights")
    params = params or ApproxCholParams()
    p2 = ApproxCholParams("deg", 5, int(params.stag_test), 2)
    return lap_wrap_components(approxchol_lap1_gpu, a, tol=tol, maxits=maxits, maxtime=maxtime,
                               verbose=verbose, pcgIts=pcgIts, params=p2, **kw)


def adj_val_and_excess(sddm):
    """adjValAndExcess (src/graphOps.jl:43-48): a = Diagonal(diag(M)) - M,
    dVal = diag(M), dExcess = M * ones."""
    M = sp.csc_matrix(sddm, dtype=np.float64)
    dval = M.diagonal()
    a = sp.csc_matrix(sp.diags(dval) - M)
    a.eliminate_zeros()
    dexcess = np.asarray(M @ np.ones(M.shape[0])).ravel()
    return a, dval, dexcess


def extend_matrix(a, dval, dexcess):
    """extendMatrix (src/graphOps.jl:56-78): [a d; d' 0] with d the diagonal
    excess of the vertices whose excess is above 100 eps times their
    diagonal, or `a` itself when there is no excess at all."""
    a = sp.csc_matrix(a)
    n = a.shape[0]
    assert n == len(dexcess)
    if np.abs(dexcess).sum() == 0:
        return a
    pos = np.flatnonzero(dexcess > 100 * np.finfo(np.float64).eps * dval)
    if len(pos) == 0:
        return a
    ai, aj = a.nonzero()
    av = np.asarray(a[ai, aj]).ravel()
    dpos = dexcess[pos]
    i2 = np.concatenate([ai, pos, np.full(len(pos), n)])
    j2 = np.concatenate([aj, np.full(len(pos), n), pos])
    v2 = np.concatenate([av, dpos, dpos])
    return sp.csc_matrix((v2, (i2, j2)), shape=(n + 1, n + 1))


def sddm_wrap_lap(lap_solver: Callable, sddm, tol=1e-6, maxits=math.inf, maxtime=math.inf, verbose=False,
                  pcgIts=None, **params):
    """sddmWrapLap (src/solverInterface.jl:324-352): solve an SDDM system M x =
    b with a Laplacian solver on the graph extended by one ground vertex
    joined to every vertex with diagonal excess, on [b; -sum(b)], returning
    (x_aug - x_aug[end])[:n].  Without excess, M is a Laplacian and its
    adjacency is solved directly."""
    M = sp.csc_matrix(sddm, dtype=np.float64)
    n = M.shape[0]
    a, dval, dexcess = adj_val_and_excess(M)
    a1 = extend_matrix(a, dval, dexcess)
    if a1.shape == M.shape:
        return lap_solver(a, tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts, **params)
    F = lap_solver(a1, tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts, **params)

    def f(b, tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose, pcgIts=pcgIts):
        b = np.asarray(b, dtype=np.float64)
        xaug = F(np.concatenate([b, [-b.sum()]]), tol=tol, maxits=maxits, maxtime=maxtime, verbose=verbose,
                 pcgIts=pcgIts)
        xaug = xaug - xaug[-1]
        return xaug[:n]

    return f


def approxchol_sddm_gpu(sddm, **kw):
    """approxchol_sddm = sddmWrapLap(approxchol_lap) (src/approxChol.jl:1862)."""
    return sddm_wrap_lap(approxchol_lap_gpu, sddm, **kw)


def harmonic_interp(adj_mat, S, vals, tol=1e-6):
    """harmonic_interp (src/harmonic.jl:16-33): the x that minimises the
    Laplacian quadratic form of `adj_mat` subject to x[S[i]] = vals[i]
    (indices 0-based here), by an SDDM solve on the vertices outside S with
    approxchol_sddm_gpu.  As in the reference, it may fail for disconnected
    graphs."""
    a = sp.csc_matrix(adj_mat, dtype=np.float64)
    n = a.shape[0]
    S = np.asarray(S, dtype=np.int64)
    b = np.zeros(n)
    b[S] = np.asarray(vals, dtype=np.float64)
    inds = np.ones(n, dtype=bool)
    inds[S] = False
    la = lap(a)
    la_sub = la[inds][:, inds]
    b_sub = (-(la @ b))[inds]
    f = approxchol_sddm_gpu(la_sub, tol=tol)
    x_sub = f(b_sub)
    x = b.copy()
    x[inds] = x_sub
    return x
