"""The solve bar of the GPU tests, in one place.

A solve's true relative residual ||L x - b|| / ||b|| must be at most tol.
pcg stops on its recursive residual (src/pcg.jl:348-361), which rounding
moves away from the true one; where the reference's own loop (the oracle)
on the same LDLinv ends with a true residual above tol, the device may end
within 1 % of that value instead:

    relres <= max(tol, 1.01 * relres_oracle)

The oracle solve is only run when the device's residual is above tol.
"""

from __future__ import annotations

import numpy as np

RELRES_ORACLE_SLACK = 1.01


def relres(la, x, b) -> float:
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(la @ x - b) / nb) if nb > 0 else 0.0


def assert_relres(la, x, b, tol, f=None, maxits=1000, stag_test=5, what=""):
    """x solves la x = b to the bar above; f is the LDLinv the device used,
    or a callable that returns it (needed only when the true residual lands
    above tol)."""
    rr = relres(la, x, b)
    if rr <= tol:
        return rr
    assert f is not None, f"{what}: true relres {rr:.3e} > tol {tol:.1e} and no factor to compare with"
    from oracle import oracle

    if callable(f):
        f = f()
    xo, _ = oracle.approxchol_solve(la, f, b, tol=tol, maxits=maxits, stag_test=stag_test)
    ro = relres(la, xo, b)
    assert rr <= max(tol, RELRES_ORACLE_SLACK * ro), f"{what}: true relres {rr:.3e}, tol {tol:.1e}, oracle {ro:.3e}"
    return rr


def assert_relres_cols(la, X, B, tol, f=None, maxits=1000, stag_test=5, what=""):
    for j in range(B.shape[1]):
        assert_relres(la, X[:, j], B[:, j], tol, f, maxits, stag_test, what=f"{what} column {j}")
