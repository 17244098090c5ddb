"""lapgpu — the MI355X (gfx950) solve phase of Laplacians.jl's approxchol_lap.

Public surface (mirrors the reference's names):
  approxchol_lap_gpu(a; tol, maxits, maxtime, verbose, pcgIts, params) -> f(b)
  approxchol_sddm_gpu(M; ...) -> f(b), harmonic_interp(a, S, vals; tol)
  ApproxCholParams, lap, build_ldli (host factor), Context (raw C-ABI handle)
"""

from .factor import LDLinv, build_ldli, elimination_order, load_ldli, save_ldli  # noqa: F401
from .graphs import lap  # noqa: F401
from ._native import Batch, Context, LapGPUError  # noqa: F401
from .solver import (  # noqa: F401
    ApproxCholParams,
    approxchol_lap1_gpu,
    approxchol_lap2_gpu,
    approxchol_lap_gpu,
    approxchol_sddm_gpu,
    block_solver,
    harmonic_interp,
    lap_wrap_components,
    sddm_wrap_lap,
)
