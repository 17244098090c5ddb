"""Inputs that end pcg (src/pcg.jl:256-445) by its overflow exits, with
their outcome derived by hand.

The graph is the 3-vertex path 0 - 1 - 2 with unit weights; min-degree
elimination takes vertex 2, then 1 (root 0), each column holding one
(sentinel) entry, so P = 1 and the diagonal scale of vertex v is
s_v = 1/d_v.  Write r by its edge flows u_1 = r_1 + r_2, u_2 = r_2
(r_0 = -u_1, the vector sums to zero).  The forward sweep computes exactly
these flows, the backward sweep integrates s_v u_v into potentials from the
root, so LDLsolver(r) has potential jumps dz_v = s_v u_v, L p has flows
dp_v (unit weights), and by summation by parts

    rho = r . z = sum_v s_v u_v^2,     p . L p = sum_v dp_v^2.

One pcg iteration from p = z therefore acts on each flow alone:

    al = sum s u^2 / sum s^2 u^2,    u_v' = u_v (1 - al s_v),
    rho' = sum s_v u_v'^2.

With s_1 = 1, s_2 = sig and flows (c, c/sig), rho' ~ rho sig / 4: one
step multiplies rho by about sig/4.  `exact_first_step` evaluates these
formulas in exact rational arithmetic on the float inputs; the cases below
put rho' beyond DBL_MAX (isinf(rho)) or rho'/rho beyond DBL_MAX with rho'
finite (isinf(beta)), every vertex value and product of the step staying
finite.
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np

DBL_MAX = Fraction(np.finfo(float).max)


def rhs_from_flows(u1: float, u2: float) -> np.ndarray:
    return np.array([-u1, u1 - u2, u2])


def factor_with_d(lapgpu, d1: float, d2: float):
    import dataclasses

    from lapgpu import graphs

    a = graphs.path_graph(3)
    f = lapgpu.build_ldli(a, seed=1)
    assert list(f.col) == [2, 1] and np.all(f.fval == 1.0) and f.d[0] == 0.0, f
    d = f.d.copy()
    d[1], d[2] = d1, d2
    return a, dataclasses.replace(f, d=d)


def exact_first_step(d1: float, d2: float, b: np.ndarray) -> dict:
    """rho, rho', the vertex values of r' and z' and their products, exactly."""
    s = [Fraction(1) / Fraction(d1), Fraction(1) / Fraction(d2)]
    r = [Fraction(v) for v in b]
    u = [r[1] + r[2], r[2]]
    rho0 = sum(si * ui * ui for si, ui in zip(s, u))
    al = rho0 / sum(si * si * ui * ui for si, ui in zip(s, u))
    u1 = [ui * (1 - al * si) for si, ui in zip(s, u)]
    rho1 = sum(si * ui * ui for si, ui in zip(s, u1))
    r1 = [-u1[0], u1[0] - u1[1], u1[1]]
    zt = [Fraction(0), s[0] * u1[0], s[0] * u1[0] + s[1] * u1[1]]
    mu = sum(zt) / 3
    z1 = [v - mu for v in zt]
    return {"rho0": rho0, "rho1": rho1, "r1": r1, "z1": z1, "terms": [a * b_ for a, b_ in zip(z1, r1)]}


# isinf(rho) at iteration 1: s = (1, 1e12), flows (c, c / 1e12), sig c^2 = 8.5e308
D_INF_RHO = (1.0, 1e-12)
C_INF_RHO = float(np.sqrt(8.5 / 1e12) * 1e154)
B_INF_RHO = rhs_from_flows(C_INF_RHO, C_INF_RHO / 1e12)

# isinf(beta) at iteration 1: s = (1e-150, 1e200), flows (c, c * 1e-350),
# c = 2e53: rho = 4e-44, rho' = 1e306, rho' / rho = 2.5e349
D_INF_BETA = (1e150, 1e-200)
B_INF_BETA = rhs_from_flows(2e53, 2e53 * 1e-175 * 1e-175)
