"""ctypes front-end of the CPU oracle (lap_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module.  The product never
imports it.

All arrays are 0-based numpy arrays; `la` is a scipy CSC Laplacian and the
factor is an object with .col/.colptr/.rowval/.fval/.d (0-based), i.e. the
reference's LDLinv (src/approxCholTypes.jl:73-79) shifted to 0-based.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblaporacle.so")
SRC_PATH = os.path.join(HERE, "lap_oracle.c")

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


class PcgResult(ctypes.Structure):
    _fields_ = [
        ("its", ctypes.c_int64),
        ("stop", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("nr", ctypes.c_double),
        ("bestnr", ctypes.c_double),
        ("rho", ctypes.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (a checker, not the product)."""
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC_PATH):
        subprocess.check_call(
            ["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fPIC", "-shared", "-o", LIB_PATH, SRC_PATH, "-lm"]
        )
    return LIB_PATH


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        dbl = ctypes.c_double
        L.orc_spmv_csc.argtypes = [i64, P, P, P, P, P]
        L.orc_forward.argtypes = [i64, P, P, P, P, P]
        L.orc_backward.argtypes = [i64, P, P, P, P, P]
        L.orc_ldlsolver.argtypes = [i64, P, P, P, P, P, P, P]
        L.orc_pcg_ldl.argtypes = [i64, P, P, P, P, P, P, P, P, P, dbl, i64, dbl, i64, P, ctypes.POINTER(PcgResult)]
        L.orc_pcg_diag.argtypes = [i64, P, P, P, P, P, dbl, i64, dbl, i64, P, ctypes.POINTER(PcgResult)]
        L.orc_approxchol_solve.argtypes = [i64, P, P, P, P, P, P, P, P, P, dbl, i64, dbl, i64, P, ctypes.POINTER(PcgResult)]
        for f in (L.orc_pcg_ldl, L.orc_pcg_diag, L.orc_approxchol_solve):
            f.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _csc(la):
    return la.shape[0], _i64(la.indptr), _i64(la.indices), _f64(la.data)


def _fac(f):
    return _i64(f.col), _i64(f.colptr), _i64(f.rowval), _f64(f.fval), _f64(f.d)


def spmv(la, x):
    n, cp, rv, nz = _csc(la)
    x = _f64(x)
    y = np.empty(n)
    lib().orc_spmv_csc(n, _p(cp), _p(rv), _p(nz), _p(x), _p(y))
    return y


def forward(f, y):
    col, cp, rv, fv, _ = _fac(f)
    y = _f64(y).copy()
    lib().orc_forward(len(f.d), _p(col), _p(cp), _p(rv), _p(fv), _p(y))
    return y


def backward(f, y):
    col, cp, rv, fv, _ = _fac(f)
    y = _f64(y).copy()
    lib().orc_backward(len(f.d), _p(col), _p(cp), _p(rv), _p(fv), _p(y))
    return y


def ldlsolver(f, b):
    col, cp, rv, fv, d = _fac(f)
    b = _f64(b)
    y = np.empty_like(b)
    lib().orc_ldlsolver(len(d), _p(col), _p(cp), _p(rv), _p(fv), _p(d), _p(b), _p(y))
    return y


def _res(r):
    return {"its": int(r.its), "stop": STOP_NAMES[int(r.stop)], "nr": r.nr, "bestnr": r.bestnr, "rho": r.rho}


def pcg_ldl(la, f, b, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=0):
    """pcg(la, b, LDLsolver) (src/pcg.jl:256-445); returns (bestx, info)."""
    n, cp, rv, nz = _csc(la)
    col, fcp, frv, fv, d = _fac(f)
    b = _f64(b)
    x = np.empty(n)
    r = PcgResult()
    rc = lib().orc_pcg_ldl(n, _p(cp), _p(rv), _p(nz), _p(col), _p(fcp), _p(frv), _p(fv), _p(d), _p(b),
                           tol, int(maxits), maxtime, int(stag_test), _p(x), ctypes.byref(r))
    if rc != 0:
        raise MemoryError("oracle pcg")
    return x, _res(r)


def pcg_diag(la, dg, b, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=0):
    """pcg(la, b, x -> x ./ dg) (the Jacobi setup of test/testPCG.jl:1-11)."""
    n, cp, rv, nz = _csc(la)
    dg = _f64(dg)
    b = _f64(b)
    x = np.empty(n)
    r = PcgResult()
    rc = lib().orc_pcg_diag(n, _p(cp), _p(rv), _p(nz), _p(dg), _p(b), tol, int(maxits), maxtime,
                            int(stag_test), _p(x), ctypes.byref(r))
    if rc != 0:
        raise MemoryError("oracle pcg")
    return x, _res(r)


def approxchol_solve(la, f, b, tol=1e-6, maxits=1000, maxtime=float("inf"), stag_test=5):
    """The solve closure f(b) of src/approxChol.jl:1624; returns (x, info)."""
    n, cp, rv, nz = _csc(la)
    col, fcp, frv, fv, d = _fac(f)
    b = _f64(b)
    x = np.empty(n)
    r = PcgResult()
    rc = lib().orc_approxchol_solve(n, _p(cp), _p(rv), _p(nz), _p(col), _p(fcp), _p(frv), _p(fv), _p(d), _p(b),
                                    tol, int(maxits), maxtime, int(stag_test), _p(x), ctypes.byref(r))
    if rc != 0:
        raise MemoryError("oracle pcg")
    return x, _res(r)
