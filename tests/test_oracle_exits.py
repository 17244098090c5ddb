"""The oracle's overflow exits of pcg (src/pcg.jl:385-402), pinned by the
hand derivation of tests/_exits.py: on the 3-vertex path with crafted
diagonal scales, one pcg step is computed in exact rational arithmetic, and
the oracle must stop on exactly the branch it predicts, at iteration 1,
with no NaN on the way (every intermediate product finite)."""

import numpy as np
import pytest

import lapgpu
import _exits as E


def _check_finite_step(ex):
    for v in ex["r1"] + ex["z1"] + ex["terms"]:
        assert abs(v) < E.DBL_MAX


def test_inf_rho_closed_form(oracle_mod):
    ex = E.exact_first_step(*E.D_INF_RHO, E.B_INF_RHO)
    assert ex["rho0"] < E.DBL_MAX and ex["rho1"] > E.DBL_MAX
    _check_finite_step(ex)
    a, f = E.factor_with_d(lapgpu, *E.D_INF_RHO)
    la = lapgpu.lap(a)
    with np.errstate(all="ignore"):
        x, info = oracle_mod.pcg_ldl(la, f, E.B_INF_RHO, tol=1e-8, maxits=30, stag_test=5)
    assert info["stop"] == "inf_rho" and info["its"] == 1, info
    assert np.isinf(info["rho"])
    # the first iterate improved on nr = 1 and is returned (src/pcg.jl:348-352)
    assert info["bestnr"] < 1.0 and np.all(np.isfinite(x))


def test_inf_beta_closed_form(oracle_mod):
    ex = E.exact_first_step(*E.D_INF_BETA, E.B_INF_BETA)
    assert ex["rho1"] < E.DBL_MAX and ex["rho1"] / ex["rho0"] > E.DBL_MAX
    _check_finite_step(ex)
    a, f = E.factor_with_d(lapgpu, *E.D_INF_BETA)
    la = lapgpu.lap(a)
    with np.errstate(all="ignore"):
        x, info = oracle_mod.pcg_ldl(la, f, E.B_INF_BETA, tol=1e-8, maxits=30, stag_test=5)
    assert info["stop"] == "inf_beta" and info["its"] == 1, info
    assert info["rho"] == pytest.approx(float(ex["rho1"]), rel=1e-12)


def test_branch_order_on_the_same_inputs(oracle_mod):
    """isinf(rho) is tested before beta is formed: on the inf-rho input,
    beta would be inf too, and the stop must name rho."""
    ex = E.exact_first_step(*E.D_INF_RHO, E.B_INF_RHO)
    assert ex["rho1"] / ex["rho0"] > 1e10
    a, f = E.factor_with_d(lapgpu, *E.D_INF_RHO)
    with np.errstate(all="ignore"):
        _, info = oracle_mod.pcg_ldl(lapgpu.lap(a), f, E.B_INF_RHO, tol=1e-8, maxits=30, stag_test=5)
    assert info["stop"] == "inf_rho"
