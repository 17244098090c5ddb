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
    10: "cancelled",
}

VERBOSE = 1
NO_CENTER = 2
PROFILE = 4
EAGER = 8
WAIT_SPIN = 0
WAIT_BLOCKING = 1
SWEEP_GROUPS = 0  # lapgpu_set_sweep_mode: one launch per level group
SWEEP_CHAIN = 1   # runs of narrow level groups chained in one launch (default)
TUNE_GROUPS = 1  # lapgpu_autotune: also tune the level-group budget
FP64 = 0
FP32_INNER = 1

# every symbol declared in include/lapgpu.h
EXPORTS = (
    "lapgpu_create",
    "lapgpu_upload_csc",
    "lapgpu_upload_ldli",
    "lapgpu_solve",
    "lapgpu_solve_device",
    "lapgpu_solve_multi",
    "lapgpu_solve_multi_device",
    "lapgpu_apply_ldli",
    "lapgpu_spmv",
    "lapgpu_set_precision",
    "lapgpu_set_group_budget",
    "lapgpu_set_wait_mode",
    "lapgpu_set_sweep_mode",
    "lapgpu_chain_info",
    "lapgpu_factor_info",
    "lapgpu_last_timings",
    "lapgpu_last_sweep_timings",
    "lapgpu_tail_info",
    "lapgpu_group_info",
    "lapgpu_autotune",
    "lapgpu_last_error",
    "lapgpu_destroy",
    "lapgpu_set_dense_tail",
    "lapgpu_sweep_entries",
    "lapgpu_last_dense_timings",
    "lapgpu_dense_info",
    "lapgpu_cancel",
    "lapgpu_save_plan",
    "lapgpu_last_true_relres",
    "lapgpu_load_plan",
    "lapgpu_batch_create",
    "lapgpu_batch_add_graph",
    "lapgpu_batch_remove_graph",
    "lapgpu_batch_graph_info",
    "lapgpu_batch_solve",
    "lapgpu_batch_solve_device",
    "lapgpu_batch_solve_device_noalloc",
    "lapgpu_batch_submit",
    "lapgpu_batch_submit_host",
    "lapgpu_batch_query",
    "lapgpu_batch_cancel",
    "lapgpu_batch_collect",
    "lapgpu_batch_last_error",
    "lapgpu_batch_destroy",
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
    L.lapgpu_solve_multi.argtypes = [P, P, P, i64, i64, dbl, i64, dbl, ci, ci, P, P, P]
    L.lapgpu_solve_multi_device.argtypes = [P, P, P, i64, i64, dbl, i64, dbl, ci, ci, P, P, P]
    L.lapgpu_apply_ldli.argtypes = [P, P, P, i64]
    L.lapgpu_spmv.argtypes = [P, P, P, i64]
    L.lapgpu_set_precision.argtypes = [P, ci, ci]
    L.lapgpu_set_group_budget.argtypes = [P, ci, dbl, i64]
    L.lapgpu_set_wait_mode.argtypes = [P, ci]
    L.lapgpu_set_sweep_mode.argtypes = [P, ci]
    L.lapgpu_chain_info.argtypes = [P, pi64, pi64, pint]
    L.lapgpu_factor_info.argtypes = [P, pi64, pi64, pi64, pi64]
    L.lapgpu_last_timings.argtypes = [P, pdbl, pdbl, pdbl, pdbl, pi64]
    L.lapgpu_last_sweep_timings.argtypes = [P, pdbl, pdbl, pdbl, pdbl]
    L.lapgpu_tail_info.argtypes = [P, pi64, pi64, pi64, pi64]
    L.lapgpu_group_info.argtypes = [P, pi64, pi64, pi64, pi64]
    L.lapgpu_autotune.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.lapgpu_last_error.argtypes = [P]
    L.lapgpu_last_error.restype = ctypes.c_char_p
    L.lapgpu_destroy.argtypes = [P]
    L.lapgpu_destroy.restype = None
    pi32 = ctypes.POINTER(ctypes.c_int32)
    L.lapgpu_batch_create.argtypes = [ctypes.POINTER(P), ci]
    L.lapgpu_set_dense_tail.argtypes = [P, i64]
    L.lapgpu_sweep_entries.argtypes = [P, pi64, pi64, pi64, pi64, pi64]
    L.lapgpu_last_dense_timings.argtypes = [P, ctypes.POINTER(dbl), ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
    L.lapgpu_dense_info.argtypes = [P, pi64, pi64, pi64]
    L.lapgpu_batch_add_graph.argtypes = [P, i64, P, P, P, i64, P, P, P, P, P, ci, pi32]
    L.lapgpu_batch_remove_graph.argtypes = [P, ctypes.c_int32]
    L.lapgpu_batch_graph_info.argtypes = [P, ctypes.c_int32, pi64, pi64]
    L.lapgpu_batch_solve.argtypes = [P, i64, P, P, P, dbl, i64, dbl, ci, ci, P, P, P]
    L.lapgpu_batch_solve_device.argtypes = [P, i64, P, P, P, dbl, i64, dbl, ci, ci, P, P, P]
    L.lapgpu_batch_solve_device_noalloc.argtypes = [P, i64, P, P, P, dbl, i64, dbl, ci, ci, P, P, P]
    L.lapgpu_batch_submit.argtypes = [P, i64, P, P, P, dbl, i64, dbl, ci, ci, pi32]
    L.lapgpu_batch_query.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ci)]
    L.lapgpu_batch_cancel.argtypes = [P, ctypes.c_int32, i64]
    L.lapgpu_cancel.argtypes = [P, i64]
    L.lapgpu_save_plan.argtypes = [P, ctypes.c_char_p]
    L.lapgpu_last_true_relres.argtypes = [P, P, i64]
    L.lapgpu_load_plan.argtypes = [P, ctypes.c_char_p, pi64]
    L.lapgpu_batch_submit_host.argtypes = [P, i64, P, P, dbl, i64, dbl, ci, ci, pi32]
    L.lapgpu_batch_collect.argtypes = [P, ctypes.c_int32, P, P, P, P]
    L.lapgpu_batch_last_error.argtypes = [P]
    L.lapgpu_batch_last_error.restype = ctypes.c_char_p
    L.lapgpu_batch_destroy.argtypes = [P]
    L.lapgpu_batch_destroy.restype = None
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("lapgpu_last_error", "lapgpu_destroy", "lapgpu_batch_last_error", "lapgpu_batch_destroy"):
            f.restype = ci
    _LIB = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


_HIP = None


def _hip() -> ctypes.CDLL:
    """The HIP runtime liblapgpu.so links (the same soname resolves to the
    already loaded copy), for device buffers owned by that runtime."""
    global _HIP
    if _HIP is None:
        load()
        try:
            H = ctypes.CDLL("libamdhip64.so.7")
        except OSError:
            H = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipSetDevice.argtypes = [ctypes.c_int]
        for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipSetDevice, H.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _HIP = H
    return _HIP


class DeviceArray:
    """A float64 array in device memory (HBM) allocated by liblapgpu's own HIP
    runtime: the resident inputs and outputs of the *_device entry points.

    torch bundles a second HIP runtime; device pointers of either work with
    the library, but a process that initialises liblapgpu's runtime first
    may not get a device from torch's afterwards, so tests and the benchmark
    use this class."""

    def __init__(self, count: int, device: int = 0):
        H = _hip()
        self.count = int(count)
        self.device = device
        self.ptr = ctypes.c_void_p()
        if H.hipSetDevice(device) != 0 or H.hipMalloc(ctypes.byref(self.ptr), max(8 * self.count, 8)) != 0:
            raise LapGPUError(f"hipMalloc of {self.count} doubles failed")

    @classmethod
    def from_host(cls, a: np.ndarray, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size, device)
        d.upload(a)
        return d

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.count:
            raise ValueError("DeviceArray.upload: size mismatch")
        if _hip().hipMemcpy(self.ptr, _p(a), 8 * self.count, 1) != 0:
            raise LapGPUError("hipMemcpy host to device failed")

    def copy_from(self, other: "DeviceArray") -> None:
        """Device-to-device copy (a plain copy kernel of known size)."""
        if other.count != self.count:
            raise ValueError("DeviceArray.copy_from: size mismatch")
        if _hip().hipMemcpy(self.ptr, other.ptr, 8 * self.count, 3) != 0:
            raise LapGPUError("hipMemcpy device to device failed")

    def download(self) -> np.ndarray:
        out = np.empty(self.count, dtype=np.float64)
        if _hip().hipMemcpy(_p(out), self.ptr, 8 * self.count, 2) != 0:
            raise LapGPUError("hipMemcpy device to host failed")
        return out

    def free(self) -> None:
        if self.ptr.value:
            _hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(device: int = 0) -> None:
    H = _hip()
    H.hipSetDevice(device)
    if H.hipDeviceSynchronize() != 0:
        raise LapGPUError("hipDeviceSynchronize failed")


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

    def last_true_relres(self, k: int = 1):
        """FP64 true relative residuals of the last solve's k columns
        (lapgpu_last_true_relres): a float for k = 1, else a list."""
        out = np.zeros(k)
        self._check(self._lib.lapgpu_last_true_relres(self._h, _p(out), int(k)), "lapgpu_last_true_relres")
        return float(out[0]) if k == 1 else out.tolist()

    def save_plan(self, path) -> None:
        """Write the context's plan (lapgpu_save_plan) to `path`."""
        self._check(self._lib.lapgpu_save_plan(self._h, os.fsencode(path)), "lapgpu_save_plan")

    def load_plan(self, path) -> None:
        """Replace upload_lap + upload_ldli by a plan file (lapgpu_load_plan)."""
        n = ctypes.c_int64(0)
        self._check(self._lib.lapgpu_load_plan(self._h, os.fsencode(path), ctypes.byref(n)), "lapgpu_load_plan")
        self.n = n.value

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

    def cancel(self, column: int = -1) -> bool:
        """From another thread: cancel column `column` (-1: all) of the
        solve running on this context; those columns stop at the next
        iteration boundary with stop "cancelled" and their best iterate.
        Returns False when no solve was running."""
        rc = self._lib.lapgpu_cancel(self._h, int(column))
        if rc < 0:
            raise ValueError(f"lapgpu_cancel: column {column} out of range")
        return rc == 0

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

    def solve_multi(self, B, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5, flags=0):
        """Solve for every column of the n-by-k matrix B (each column its own
        pcg, all sharing the launches).  Returns X (n-by-k) and per-column
        its / nr / stop lists."""
        B = np.asfortranarray(B, dtype=np.float64)
        if B.ndim != 2:
            raise ValueError("solve_multi: B must be n-by-k")
        n, k = B.shape
        X = np.empty((n, k), dtype=np.float64, order="F")
        its = np.zeros(k, dtype=np.int64)
        nr = np.zeros(k, dtype=np.float64)
        stop = np.zeros(k, dtype=np.int32)
        self._check(
            self._lib.lapgpu_solve_multi(self._h, _p(B), _p(X), n, k, float(tol), int(maxits), float(maxtime),
                                         int(stag_test), int(flags), _p(its), _p(nr), _p(stop)),
            "lapgpu_solve_multi",
        )
        return X, {"its": its.tolist(), "nr": nr.tolist(), "stop": [STOP_NAMES.get(int(s), str(s)) for s in stop]}

    def solve_multi_device(self, d_B: int, d_X: int, n: int, k: int, tol=1e-6, maxits=1000,
                           maxtime=float("inf"), stag_test=5, flags=0):
        """Multi-RHS solve with device pointers to column-major n-by-k B and X."""
        its = np.zeros(k, dtype=np.int64)
        nr = np.zeros(k, dtype=np.float64)
        stop = np.zeros(k, dtype=np.int32)
        self._check(
            self._lib.lapgpu_solve_multi_device(self._h, ctypes.c_void_p(d_B), ctypes.c_void_p(d_X), n, k,
                                                float(tol), int(maxits), float(maxtime), int(stag_test),
                                                int(flags), _p(its), _p(nr), _p(stop)),
            "lapgpu_solve_multi_device",
        )
        return {"its": its.tolist(), "nr": nr.tolist(), "stop": [STOP_NAMES.get(int(s), str(s)) for s in stop]}

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

    def set_group_budget(self, max_depth: int = 32, max_growth: float = 16.0, max_extra: int = 196608) -> None:
        """Level-group budget of the LDLinv schedule (applies to the next
        upload_ldli; regroups an uploaded factor at once)."""
        self._check(self._lib.lapgpu_set_group_budget(self._h, int(max_depth), float(max_growth), int(max_extra)),
                    "lapgpu_set_group_budget")

    def set_dense_tail(self, max_k: int) -> None:
        """Largest dense tail block the planner may choose (0: none); takes
        effect at the next upload_ldli."""
        self._check(self._lib.lapgpu_set_dense_tail(self._h, int(max_k)), "lapgpu_set_dense_tail")

    def last_dense_timings(self) -> dict:
        v = [ctypes.c_double() for _ in range(3)]
        self._check(self._lib.lapgpu_last_dense_timings(self._h, *(ctypes.byref(t) for t in v)),
                    "lapgpu_last_dense_timings")
        return {"tail_fwd_ms": v[0].value, "dense_ms": v[1].value, "tail_bwd_ms": v[2].value}

    def sweep_entries(self) -> dict:
        v = [ctypes.c_int64() for _ in range(5)]
        self._check(self._lib.lapgpu_sweep_entries(self._h, *(ctypes.byref(t) for t in v)), "lapgpu_sweep_entries")
        return {"groups_f": v[0].value, "groups_b": v[1].value, "tail_f": v[2].value, "tail_b": v[3].value,
                "dense_k": v[4].value}

    def dense_info(self) -> dict:
        v = [ctypes.c_int64() for _ in range(3)]
        self._check(self._lib.lapgpu_dense_info(self._h, *(ctypes.byref(t) for t in v)), "lapgpu_dense_info")
        return {"k": v[0].value, "sparse_vertices": v[1].value, "bytes": v[2].value}

    def set_wait_mode(self, blocking: bool) -> None:
        """Wait for the device by sleeping (blocking=True) or spinning."""
        self._check(self._lib.lapgpu_set_wait_mode(self._h, WAIT_BLOCKING if blocking else WAIT_SPIN),
                    "lapgpu_set_wait_mode")

    def set_sweep_mode(self, chain: bool) -> None:
        """Runs of narrow level groups chained in one launch each (chain=True,
        the default) or one launch per group; the results are the same bits."""
        self._check(self._lib.lapgpu_set_sweep_mode(self._h, SWEEP_CHAIN if chain else SWEEP_GROUPS),
                    "lapgpu_set_sweep_mode")

    def chain_info(self) -> dict:
        """Chained runs of narrow level groups in the uploaded schedule."""
        a, b, on = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        self._check(self._lib.lapgpu_chain_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(on)),
                    "lapgpu_chain_info")
        return {"runs": a.value, "chained_groups": b.value, "active": bool(on.value)}

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

    def tail_info(self) -> dict:
        v = [ctypes.c_int64() for _ in range(4)]
        self._check(self._lib.lapgpu_tail_info(self._h, *(ctypes.byref(t) for t in v)), "lapgpu_tail_info")
        return {"nt": v[0].value, "nsteps": v[1].value, "nslots": v[2].value, "nedges_in": v[3].value}

    def autotune(self, k: int = 0, groups: bool = False) -> dict:
        """Tune the schedule on this graph (lapgpu_autotune): lanes per row of
        every level group, slices per row of k-column solves when k > 1, and
        with groups=True the level-group budget.  Returns the device ms of
        one LDLinv apply before/after (single and k-column)."""
        t = (ctypes.c_double * 4)()
        self._check(self._lib.lapgpu_autotune(self._h, int(k), TUNE_GROUPS if groups else 0, t),
                    "lapgpu_autotune")
        return {"apply_ms_before": t[0], "apply_ms_after": t[1], "multi_apply_ms_before": t[2],
                "multi_apply_ms_after": t[3]}

    def group_info(self) -> dict:
        v = [ctypes.c_int64() for _ in range(4)]
        self._check(self._lib.lapgpu_group_info(self._h, *(ctypes.byref(t) for t in v)), "lapgpu_group_info")
        return {"ngroups": v[0].value, "rows": v[1].value, "fentries": v[2].value, "bentries": v[3].value}

    def last_sweep_timings(self) -> dict:
        """Per-iteration ms of the sweep kernels of the last PROFILE solve."""
        v = [ctypes.c_double() for _ in range(4)]
        self._check(self._lib.lapgpu_last_sweep_timings(self._h, *(ctypes.byref(t) for t in v)),
                    "lapgpu_last_sweep_timings")
        return {"fgroups_ms": v[0].value, "tail_ext_ms": v[1].value, "tail_ms": v[2].value,
                "bgroups_ms": v[3].value}


class Batch:
    """Device-resident graph store; solves many (graph, b) requests in one
    launch, one workgroup per request (lapgpu_batch_* in include/lapgpu.h)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.lapgpu_batch_create(ctypes.byref(h), device)
        if rc != 0:
            raise LapGPUError(f"lapgpu_batch_create: {self._lib.lapgpu_batch_last_error(None).decode()}")
        self._h = h
        self.sizes = {}

    def _check(self, rc, what):
        if rc != 0:
            raise LapGPUError(f"{what} failed ({rc}): {self._lib.lapgpu_batch_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lapgpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_graph(self, la, f) -> int:
        n = la.shape[0]
        cp = np.ascontiguousarray(la.indptr, dtype=np.int64)
        rv = np.ascontiguousarray(la.indices, dtype=np.int64)
        nz = np.ascontiguousarray(la.data, dtype=np.float64)
        col = np.ascontiguousarray(f.col, dtype=np.int64)
        fcp = np.ascontiguousarray(f.colptr, dtype=np.int64)
        frv = np.ascontiguousarray(f.rowval, dtype=np.int64)
        fv = np.ascontiguousarray(f.fval, dtype=np.float64)
        d = np.ascontiguousarray(f.d, dtype=np.float64)
        gid = ctypes.c_int32(-1)
        self._check(self._lib.lapgpu_batch_add_graph(self._h, n, _p(cp), _p(rv), _p(nz), len(fv), _p(col), _p(fcp),
                                                     _p(frv), _p(fv), _p(d), 0, ctypes.byref(gid)),
                    "lapgpu_batch_add_graph")
        self.sizes[gid.value] = n
        return gid.value

    def remove_graph(self, gid: int) -> None:
        self._check(self._lib.lapgpu_batch_remove_graph(self._h, int(gid)), "lapgpu_batch_remove_graph")
        self.sizes.pop(gid, None)

    def graph_info(self, gid: int) -> dict:
        n = ctypes.c_int64()
        by = ctypes.c_int64()
        self._check(self._lib.lapgpu_batch_graph_info(self._h, int(gid), ctypes.byref(n), ctypes.byref(by)),
                    "lapgpu_batch_graph_info")
        return {"n": n.value, "device_bytes": by.value}

    def solve(self, graph_ids, bs, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5, flags=0):
        """Solve request i = (graph_ids[i], bs[i]); returns (xs, infos)."""
        ids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        bcat = np.ascontiguousarray(np.concatenate([np.asarray(b, dtype=np.float64) for b in bs]) if len(bs) else
                                    np.zeros(0))
        xcat = np.empty_like(bcat)
        k = len(ids)
        its = np.zeros(k, dtype=np.int64)
        nr = np.zeros(k)
        stop = np.zeros(k, dtype=np.int32)
        self._check(self._lib.lapgpu_batch_solve(self._h, k, _p(ids), _p(bcat), _p(xcat), float(tol), int(maxits),
                                                 float(maxtime), int(stag_test), int(flags), _p(its), _p(nr), _p(stop)),
                    "lapgpu_batch_solve")
        xs, off = [], 0
        for b in bs:
            xs.append(xcat[off:off + len(b)].copy())
            off += len(b)
        infos = [{"its": int(its[i]), "nr": float(nr[i]), "stop": STOP_NAMES.get(int(stop[i]), str(stop[i]))}
                 for i in range(k)]
        return xs, infos

    def solve_device(self, graph_ids, b_ptrs, x_ptrs, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5,
                     flags=0):
        ids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        k = len(ids)
        pb = np.ascontiguousarray(b_ptrs, dtype=np.uint64)
        px = np.ascontiguousarray(x_ptrs, dtype=np.uint64)
        its = np.zeros(k, dtype=np.int64)
        nr = np.zeros(k)
        stop = np.zeros(k, dtype=np.int32)
        self._check(self._lib.lapgpu_batch_solve_device(self._h, k, _p(ids), _p(pb), _p(px), float(tol), int(maxits),
                                                        float(maxtime), int(stag_test), int(flags), _p(its), _p(nr),
                                                        _p(stop)),
                    "lapgpu_batch_solve_device")
        return [{"its": int(its[i]), "nr": float(nr[i]), "stop": STOP_NAMES.get(int(stop[i]), str(stop[i]))}
                for i in range(k)]

    def solve_device_noalloc(self, graph_ids, b_ptrs, x_ptrs, tol=1e-6, maxits=1000, maxtime=float("inf"),
                             stag_test=5, flags=0):
        """solve_device through lapgpu_batch_solve_device_noalloc: no device
        allocation; at most 1024 requests and 64 MiB of work vectors."""
        ids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        k = len(ids)
        pb = np.ascontiguousarray(b_ptrs, dtype=np.uint64)
        px = np.ascontiguousarray(x_ptrs, dtype=np.uint64)
        its = np.zeros(k, dtype=np.int64)
        nr = np.zeros(k)
        stop = np.zeros(k, dtype=np.int32)
        self._check(self._lib.lapgpu_batch_solve_device_noalloc(self._h, k, _p(ids), _p(pb), _p(px), float(tol),
                                                                int(maxits), float(maxtime), int(stag_test),
                                                                int(flags), _p(its), _p(nr), _p(stop)),
                    "lapgpu_batch_solve_device_noalloc")
        return [{"its": int(its[i]), "nr": float(nr[i]), "stop": STOP_NAMES.get(int(stop[i]), str(stop[i]))}
                for i in range(k)]

    def submit(self, graph_ids, b_ptrs, x_ptrs, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5,
               flags=0):
        """Enqueue a launch (device pointers); returns its slot, or None when
        every slot holds an uncollected launch."""
        ids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        pb = np.ascontiguousarray(b_ptrs, dtype=np.uint64)
        px = np.ascontiguousarray(x_ptrs, dtype=np.uint64)
        slot = ctypes.c_int32(-1)
        rc = self._lib.lapgpu_batch_submit(self._h, len(ids), _p(ids), _p(pb), _p(px), float(tol), int(maxits),
                                           float(maxtime), int(stag_test), int(flags), ctypes.byref(slot))
        if rc == 1:
            return None
        self._check(rc, "lapgpu_batch_submit")
        return (slot.value, len(ids), None)

    def submit_host(self, graph_ids, bs, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5, flags=0):
        """Enqueue a launch with host right-hand sides; returns a handle for
        query/collect, or None when every launch slot is occupied."""
        ids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        bcat = np.ascontiguousarray(np.concatenate([np.asarray(b, dtype=np.float64) for b in bs]))
        slot = ctypes.c_int32(-1)
        rc = self._lib.lapgpu_batch_submit_host(self._h, len(ids), _p(ids), _p(bcat), float(tol), int(maxits),
                                                float(maxtime), int(stag_test), int(flags), ctypes.byref(slot))
        if rc == 1:
            return None
        self._check(rc, "lapgpu_batch_submit_host")
        return (slot.value, len(ids), [len(b) for b in bs])

    def cancel(self, handle, index: int) -> None:
        """Cancel request `index` (its position in the submit call) of the
        launch `handle`: it stops at its next iteration boundary with stop
        "cancelled" (unless it has finished); the others run on unchanged."""
        self._check(self._lib.lapgpu_batch_cancel(self._h, handle[0], int(index)), "lapgpu_batch_cancel")

    def query(self, handle) -> bool:
        done = ctypes.c_int(0)
        self._check(self._lib.lapgpu_batch_query(self._h, handle[0], ctypes.byref(done)), "lapgpu_batch_query")
        return bool(done.value)

    def collect(self, handle):
        """Wait for a launch; returns (xs or None, infos)."""
        k = handle[1]
        lens = handle[2]
        its = np.zeros(k, dtype=np.int64)
        nr = np.zeros(k)
        stop = np.zeros(k, dtype=np.int32)
        xcat = np.empty(sum(lens)) if lens is not None else None
        xp = _p(xcat) if xcat is not None else None
        self._check(self._lib.lapgpu_batch_collect(self._h, handle[0], xp, _p(its), _p(nr), _p(stop)),
                    "lapgpu_batch_collect")
        infos = [{"its": int(its[i]), "nr": float(nr[i]), "stop": STOP_NAMES.get(int(stop[i]), str(stop[i]))}
                 for i in range(k)]
        if lens is None:
            return None, infos
        xs, off = [], 0
        for ln in lens:
            xs.append(xcat[off:off + ln])
            off += ln
        return xs, infos
