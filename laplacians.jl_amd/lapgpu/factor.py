"""LDLinv container and the host factor builder binding.

`LDLinv` mirrors the reference's factor type (src/approxCholTypes.jl:73-79):
columns in elimination order, `col[ii]` the eliminated vertex, entries
`colptr[ii]:colptr[ii+1]` of `(rowval, fval)`, the last entry of each column a
sentinel with fval = 1, and `d` indexed by vertex.  Arrays here are 0-based.

`build_ldli` calls liblapfactor (C++), the restatement of the reference's
factor builders used where no Julia host is present: `approxChol(LLmatp(a))`
(src/approxChol.jl:864-977, ApproxCholParams order :deg), its split-edge
forms (split >= 1, src/approxChol.jl:979-1339) and the fixed-order
`approxChol(LLMatOrd(a, p))` (src/approxChol.jl:758-862) for the orders
:given, :wdeg and :random.  Its RNG streams are the library's own
(xoshiro256++ for the elimination, numpy's PCG64 for the :wdeg and :random
orders), so factors are statistically, not bitwise, equivalent to the
reference's for a seed.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBF = None


def _libfactor():
    global _LIBF
    if _LIBF is None:
        path = os.path.join(_HERE, "liblapfactor.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        i64 = ctypes.c_int64
        p64 = ctypes.POINTER(ctypes.c_int64)
        pd = ctypes.POINTER(ctypes.c_double)
        lib.lapfactor_build.argtypes = [i64, p64, p64, pd, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        lib.lapfactor_build.restype = ctypes.c_int
        lib.lapfactor_build_ordered.argtypes = [i64, p64, p64, pd, ctypes.c_int, p64, ctypes.c_uint64,
                                                ctypes.POINTER(ctypes.c_void_p)]
        lib.lapfactor_build_ordered.restype = ctypes.c_int
        lib.lapfactor_build_split.argtypes = [i64, p64, p64, pd, ctypes.c_int, i64, i64, ctypes.c_uint64,
                                              ctypes.POINTER(ctypes.c_void_p)]
        lib.lapfactor_build_split.restype = ctypes.c_int
        lib.lapfactor_n.argtypes = [ctypes.c_void_p]
        lib.lapfactor_n.restype = i64
        lib.lapfactor_nnz.argtypes = [ctypes.c_void_p]
        lib.lapfactor_nnz.restype = i64
        lib.lapfactor_export.argtypes = [ctypes.c_void_p, ctypes.c_int, p64, p64, p64, pd, pd]
        lib.lapfactor_export.restype = ctypes.c_int
        lib.lapfactor_free.argtypes = [ctypes.c_void_p]
        lib.lapfactor_free.restype = None
        lib.lapfactor_last_error.argtypes = []
        lib.lapfactor_last_error.restype = ctypes.c_char_p
        _LIBF = lib
    return _LIBF


def _p64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class LDLinv:
    """0-based LDLinv (src/approxCholTypes.jl:73-79)."""

    col: np.ndarray  # int64, n-1
    colptr: np.ndarray  # int64, n
    rowval: np.ndarray  # int64, nnzF
    fval: np.ndarray  # float64, nnzF
    d: np.ndarray  # float64, n

    @property
    def n(self) -> int:
        return len(self.d)

    @property
    def nnz(self) -> int:
        return len(self.fval)

    def max_col_len(self) -> int:
        return int(np.max(np.diff(self.colptr))) if len(self.colptr) > 1 else 0


def elimination_order(a: sp.spmatrix, order: str, seed: int = 1) -> np.ndarray | None:
    """The fixed elimination order of ApproxCholParams `order`
    (src/approxChol.jl:1628-1790): None for :given (0..n-1); for :wdeg the
    vertices sorted by weighted degree times 1 + U(0,1) (sortperm of
    sum(a, dims=1) .* (1 .+ rand(n))); for :random a random permutation
    (randperm).  The draws come from numpy's PCG64 seeded with `seed`."""
    n = a.shape[0]
    rng = np.random.default_rng(seed)
    if order == "wdeg":
        v = np.asarray(sp.csc_matrix(a).sum(axis=0)).ravel() * (1.0 + rng.random(n))
        return np.argsort(v, kind="stable").astype(np.int64)
    if order == "random":
        return rng.permutation(n).astype(np.int64)
    return None


def build_ldli(a: sp.spmatrix, seed: int = 1, order: str = "deg", split: int = 0, merge: int = 0,
               perm: np.ndarray | None = None) -> LDLinv:
    """Randomized approximate elimination of a connected graph's adjacency,
    as approxchol_lap1 builds it for ApproxCholParams(order, _, split, merge)
    (src/approxChol.jl:1581-1851): order "deg" eliminates by approximate
    minimum degree (with split >= 1, every edge as `split` copies, at most
    `merge` kept per row, merge < 1 for all); any other order eliminates in
    a fixed order: `perm` if given, else elimination_order(a, order, seed)."""
    a = sp.csc_matrix(a)
    n = a.shape[0]
    colptr = np.ascontiguousarray(a.indptr, dtype=np.int64)
    rowval = np.ascontiguousarray(a.indices, dtype=np.int64)
    nzval = np.ascontiguousarray(a.data, dtype=np.float64)
    lib = _libfactor()
    h = ctypes.c_void_p()
    if order == "deg" and perm is None:
        if split >= 1:
            rc = lib.lapfactor_build_split(n, _p64(colptr), _p64(rowval), _pd(nzval), 0, int(split), int(merge),
                                           seed, ctypes.byref(h))
        else:
            rc = lib.lapfactor_build(n, _p64(colptr), _p64(rowval), _pd(nzval), 0, seed, ctypes.byref(h))
    else:
        p = perm if perm is not None else elimination_order(a, order, seed)
        pp = None
        if p is not None:
            p = np.ascontiguousarray(p, dtype=np.int64)
            if p.shape != (n,):
                raise ValueError(f"perm has shape {p.shape}, expected ({n},)")
            pp = _p64(p)
        rc = lib.lapfactor_build_ordered(n, _p64(colptr), _p64(rowval), _pd(nzval), 0, pp, seed, ctypes.byref(h))
    if rc != 0:
        msg = lib.lapfactor_last_error().decode()
        raise ValueError(msg) if rc == -2 else RuntimeError(f"lapfactor build failed ({rc}): {msg}")
    try:
        nnz = lib.lapfactor_nnz(h)
        col = np.empty(max(n - 1, 0), dtype=np.int64)
        cp = np.empty(n, dtype=np.int64)
        rv = np.empty(nnz, dtype=np.int64)
        fv = np.empty(nnz, dtype=np.float64)
        d = np.empty(n, dtype=np.float64)
        rc = lib.lapfactor_export(h, 0, _p64(col), _p64(cp), _p64(rv), _pd(fv), _pd(d))
        if rc != 0:
            raise RuntimeError(lib.lapfactor_last_error().decode())
    finally:
        lib.lapfactor_free(h)
    return LDLinv(col, cp, rv, fv, d)


def save_ldli(path: str, f: LDLinv) -> None:
    """Write a factor as an uncompressed .npz (no pickled objects)."""
    np.savez(path, col=f.col, colptr=f.colptr, rowval=f.rowval, fval=f.fval, d=f.d)


def load_ldli(path: str) -> LDLinv:
    z = np.load(path, allow_pickle=False)
    return LDLinv(z["col"], z["colptr"], z["rowval"], z["fval"], z["d"])
