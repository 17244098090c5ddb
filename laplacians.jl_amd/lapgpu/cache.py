"""Factor and plan cache: what a graph costs on the host before its first
solve — the randomized factorization (LDLinv, src/approxCholTypes.jl:73-79)
and the device plan (lapgpu_save_plan / lapgpu_load_plan) — kept in a
directory, so that a graph seen before is admitted without either.

Files under the cache directory, named by the graph's key (a 64-bit hash of
the adjacency's CSC arrays, its shape and the factor seed):

* `<key>.ldli.npz`: the LDLinv (col, colptr, rowval, fval, d) as plain
  arrays, read with numpy's loader with pickles refused;
* `<key>.<settings>.plan`: the plan a context built from that factor with
  the given settings (level-group budget, dense block limit).

Entries are written to a temporary name and renamed, so a concurrent reader
sees a whole file or none; a plan file that fails its checks is rebuilt.
The cache keeps everything it writes (no eviction); `bytes()` reports the
total.
"""

from __future__ import annotations

import os
import threading

import numpy as np
import scipy.sparse as sp

from .factor import LDLinv, build_ldli
from .graphs import lap

try:
    import xxhash

    def _hasher():
        return xxhash.xxh3_64()
except ImportError:  # the standard library's hash when xxhash is absent
    import hashlib

    def _hasher():
        return hashlib.blake2b(digest_size=8)


def graph_key(a: sp.spmatrix, seed: int = 1) -> str:
    """Key of an adjacency (its CSC arrays, shape and dtype) and factor seed."""
    a = sp.csc_matrix(a)
    h = _hasher()
    h.update(np.array([a.shape[0], a.shape[1], a.nnz, seed], dtype=np.int64).tobytes())
    for arr, dt in ((a.indptr, np.int64), (a.indices, np.int64), (a.data, np.float64)):
        h.update(np.ascontiguousarray(arr, dtype=dt).data)
    return h.hexdigest()


class GraphCache:
    def __init__(self, root: str):
        self.root = root
        os.makedirs(root, exist_ok=True)
        self._lock = threading.Lock()
        self.stats = {"factor_hits": 0, "factor_misses": 0, "plan_hits": 0, "plan_misses": 0,
                      "plan_rejected": 0}

    def _count(self, k: str) -> None:
        with self._lock:
            self.stats[k] += 1

    def _path(self, key: str, suffix: str) -> str:
        return os.path.join(self.root, f"{key}.{suffix}")

    def factor(self, a: sp.spmatrix, seed: int = 1, key: str | None = None):
        """(key, la, LDLinv, hit): the factor of `a` from the cache, or built
        (build_ldli with `seed`) and stored."""
        a = sp.csc_matrix(a)
        key = key or graph_key(a, seed)
        la = lap(a)
        p = self._path(key, "ldli.npz")
        if os.path.exists(p):
            with np.load(p, allow_pickle=False) as z:
                f = LDLinv(z["col"], z["colptr"], z["rowval"], z["fval"], z["d"])
            self._count("factor_hits")
            return key, la, f, True
        f = build_ldli(a, seed=seed)
        tmp = p + f".{os.getpid()}.{threading.get_ident()}.tmp.npz"
        np.savez(tmp, col=f.col, colptr=f.colptr, rowval=f.rowval, fval=f.fval, d=f.d)
        os.replace(tmp, p)
        self._count("factor_misses")
        return key, la, f, False

    def plan_path(self, key: str, settings: str) -> str:
        return self._path(key, f"{settings}.plan")

    def upload(self, ctx, key: str, settings: str, la, f) -> bool:
        """Give ctx the graph: from the plan file when there is a good one
        (True), else by uploading la and the factor and saving the plan."""
        p = self.plan_path(key, settings)
        if os.path.exists(p):
            try:
                ctx.load_plan(p)
                self._count("plan_hits")
                return True
            except Exception:
                self._count("plan_rejected")
        ctx.upload_lap(la)
        ctx.upload_ldli(f)
        tmp = p + f".{os.getpid()}.{threading.get_ident()}.tmp"
        ctx.save_plan(tmp)
        os.replace(tmp, p)
        self._count("plan_misses")
        return False

    def bytes(self) -> int:
        return sum(e.stat().st_size for e in os.scandir(self.root) if e.is_file())
