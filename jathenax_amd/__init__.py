"""MI355X-native per-RK-stage update for JathenAX's periodic isothermal hydro solver.

Same module surface as the reference package (jathenax/__init__.py:3-6 exports
RK_STAGES, rk_stage, stress_face_fluxes and flux_divergence); the hot path runs as
hand-written HIP kernels for gfx950 in libjathenax_amd.so.
"""

from .fluxes import RECONSTRUCTIONS, RIEMANN_SOLVERS, flux_divergence, stress_face_fluxes
from .hydro import (RK_STAGES, isothermal_step, isothermal_timestep, recover_isothermal,
                    rk_stage)
from .run import conserved_sums, evolve_isothermal

__all__ = ["RK_STAGES", "rk_stage", "stress_face_fluxes", "flux_divergence",
           "RECONSTRUCTIONS", "RIEMANN_SOLVERS", "isothermal_step", "isothermal_timestep",
           "recover_isothermal", "evolve_isothermal", "conserved_sums"]
