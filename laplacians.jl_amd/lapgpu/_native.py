"""ctypes binding of liblapgpu.so (include/lapgpu.h).

The library must be built in-tree (laplacians.jl_amd/build.py); there is no
CPU fallback: every entry point raises if the library or a GPU is missing.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblapgpu.so")

STOP_NAMES = {
    0: "converged",
    1: "maxits",
    2: "maxtime",
    3: "stag_test",
    4: "stag_al_p",
    5: "stag_al_q",
    6: "stag_beta",
    7: "inf_rho",
    8: "inf_beta",
    9: "zero_rhs",
}

VERBOSE = 1
NO_CENTER = 2
PROFILE = 4
FP64 = 0
FP32_INNER = 1

# every symbol declared in include/lapgpu.h
EXPORTS = (
    "lapgpu_create",
    "lapgpu_upload_csc",
    "lapgpu_upload_ldli",
    "lapgpu_solve",
    "lapgpu_solve_device",
    "lapgpu_apply_ldli",
    "lapgpu_spmv",
    "lapgpu_set_precision",
    "lapgpu_factor_info",
    "lapgpu_last_timings",
    "lapgpu_last_error",
    "lapgpu_destroy",
)

_LIB = None


def load() -> ctypes.CDLL:
    """Load liblapgpu.so and declare the C-ABI signatures."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with laplacians.jl_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    dbl = ctypes.c_double
    ci = ctypes.c_int
    pi64 = ctypes.POINTER(ctypes.c_int64)
    pdbl = ctypes.POINTER(ctypes.c_double)
    pint = ctypes.POINTER(ctypes.c_int)
    L.lapgpu_create.argtypes = [ctypes.POINTER(P), ci]
    L.lapgpu_upload_csc.argtypes = [P, i64, P, P, P, ci]
    L.lapgpu_upload_ldli.argtypes = [P, i64, i64, P, P, P, P, P, ci]
    L.lapgpu_solve.argtypes = [P, P, P, i64, dbl, i64, dbl, ci, ci, pi64, pdbl, pint]
    L.lapgpu_solve_device.argtypes = [P, P, P, i64, dbl, i64, dbl, ci, ci, pi64, pdbl, pint]
    L.lapgpu_apply_ldli.argtypes = [P, P, P, i64]
    L.lapgpu_spmv.argtypes = [P, P, P, i64]
    L.lapgpu_set_precision.argtypes = [P, ci, ci]
    L.lapgpu_factor_info.argtypes = [P, pi64, pi64, pi64, pi64]
    L.lapgpu_last_timings.argtypes = [P, pdbl, pdbl, pdbl, pdbl, pi64]
    L.lapgpu_last_error.argtypes = [P]
    L.lapgpu_last_error.restype = ctypes.c_char_p
    L.lapgpu_destroy.argtypes = [P]
    L.lapgpu_destroy.restype = None
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("lapgpu_last_error", "lapgpu_destroy"):
            f.restype = ci
    _LIB = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class LapGPUError(RuntimeError):
    pass


class Context:
    """One device context: a stream, one uploaded graph and its workspace."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.lapgpu_create(ctypes.byref(h), device)
        if rc != 0:
            raise LapGPUError(f"lapgpu_create: {self._lib.lapgpu_last_error(None).decode()}")
        self._h = h
        self.n = 0

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self._lib.lapgpu_last_error(self._h).decode()
            raise LapGPUError(f"{what} failed ({rc}): {msg}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lapgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_lap(self, la) -> None:
        """Upload la = lap(a) as CSC (scipy), 0-based."""
        n = la.shape[0]
        cp = np.ascontiguousarray(la.indptr, dtype=np.int64)
        rv = np.ascontiguousarray(la.indices, dtype=np.int64)
        nz = np.ascontiguousarray(la.data, dtype=np.float64)
        self._check(self._lib.lapgpu_upload_csc(self._h, n, _p(cp), _p(rv), _p(nz), 0), "lapgpu_upload_csc")
        self.n = n

    def upload_ldli(self, f) -> None:
        """Upload a 0-based LDLinv (lapgpu.factor.LDLinv)."""
        col = np.ascontiguousarray(f.col, dtype=np.int64)
        cp = np.ascontiguousarray(f.colptr, dtype=np.int64)
        rv = np.ascontiguousarray(f.rowval, dtype=np.int64)
        fv = np.ascontiguousarray(f.fval, dtype=np.float64)
        d = np.ascontiguousarray(f.d, dtype=np.float64)
        self._check(
            self._lib.lapgpu_upload_ldli(self._h, len(d), len(fv), _p(col), _p(cp), _p(rv), _p(fv), _p(d), 0),
            "lapgpu_upload_ldli",
        )

    def solve(self, b, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5, flags=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        its = ctypes.c_int64(0)
        nr = ctypes.c_double(0)
        stop = ctypes.c_int(0)
        self._check(
            self._lib.lapgpu_solve(self._h, _p(b), _p(x), len(b), float(tol), int(maxits), float(maxtime),
                                   int(stag_test), int(flags), ctypes.byref(its), ctypes.byref(nr), ctypes.byref(stop)),
            "lapgpu_solve",
        )
        return x, {"its": its.value, "nr": nr.value, "stop": STOP_NAMES.get(stop.value, str(stop.value))}

    def solve_device(self, d_b: int, d_x: int, n: int, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5, flags=0):
        """Solve with device pointers (e.g. torch tensors' data_ptr())."""
        its = ctypes.c_int64(0)
        nr = ctypes.c_double(0)
        stop = ctypes.c_int(0)
        self._check(
            self._lib.lapgpu_solve_device(self._h, ctypes.c_void_p(d_b), ctypes.c_void_p(d_x), n, float(tol),
                                          int(maxits), float(maxtime), int(stag_test), int(flags),
                                          ctypes.byref(its), ctypes.byref(nr), ctypes.byref(stop)),
            "lapgpu_solve_device",
        )
        return {"its": its.value, "nr": nr.value, "stop": STOP_NAMES.get(stop.value, str(stop.value))}

    def apply_ldli(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        self._check(self._lib.lapgpu_apply_ldli(self._h, _p(r), _p(z), len(r)), "lapgpu_apply_ldli")
        return z

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._check(self._lib.lapgpu_spmv(self._h, _p(x), _p(y), len(x)), "lapgpu_spmv")
        return y

    def set_precision(self, mode: int, replace_every: int = 20) -> None:
        self._check(self._lib.lapgpu_set_precision(self._h, int(mode), int(replace_every)), "lapgpu_set_precision")

    def factor_info(self) -> dict:
        a, b, c, d = (ctypes.c_int64() for _ in range(4))
        self._check(self._lib.lapgpu_factor_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)),
                    "lapgpu_factor_info")
        return {"nlevels": a.value, "nwide": b.value, "tail_start": c.value, "nnzF": d.value}

    def last_timings(self) -> dict:
        v = [ctypes.c_double() for _ in range(4)]
        its = ctypes.c_int64()
        self._check(self._lib.lapgpu_last_timings(self._h, *(ctypes.byref(t) for t in v), ctypes.byref(its)),
                    "lapgpu_last_timings")
        return {"spmv_ms": v[0].value, "ldl_ms": v[1].value, "vec_ms": v[2].value, "total_ms": v[3].value,
                "its": its.value}
