"""Face-operator registries and the standalone flux operators (jathenax/fluxes.py).

The reconstructions and Riemann solvers themselves run inside the fused stage kernel
(csrc/jx_ops.hpp, csrc/jx_stage.hpp); the registries map the reference's names to the
kernel's template ids, in the reference's order, so option strings and dict lookups
(KeyError on unknown names) behave as in fluxes.py:120-123 and fluxes.py:204.
"""

import torch

from . import _lib
from .device import as_device, context_for, dtype_id, ptr

RECONSTRUCTIONS = {"dc": 0, "plm": 1, "ppm4": 2, "ppmx": 3, "wenoz": 4}
RIEMANN_SOLVERS = {"llf": 0, "hlle": 1, "roe": 2}


def stress_face_fluxes(tau):
    """Average cell-centred (tau_xx, tau_xy, tau_yy) onto periodic faces (fluxes.py:269-280)."""
    tau = as_device(tau)
    if tau.ndim != 3 or tau.shape[-1] != 3:
        raise ValueError("Expected stress with shape (ny, nx, 3)")
    ny, nx, _ = tau.shape
    fx = torch.empty((ny, nx, 2), dtype=tau.dtype, device=tau.device)
    fy = torch.empty_like(fx)
    ctx = context_for(tau)
    ctx.check(_lib.load().jx_stress_face_fluxes(ctx.handle, dtype_id(tau), ny, nx, ptr(tau),
                                                ptr(fx), ptr(fy)), "stress_face_fluxes")
    return fx, fy


def flux_divergence(flux_x, flux_y, dx, dy):
    """(F[i+1]-F[i])/dx + (G[j+1]-G[j])/dy with periodic wrap (fluxes.py:283-287)."""
    fx = as_device(flux_x)
    fy = as_device(flux_y, dtype=fx.dtype)
    if fx.shape != fy.shape or fx.ndim != 3:
        raise ValueError("flux_x and flux_y must both have shape (ny, nx, nvar)")
    ny, nx, nv = fx.shape
    out = torch.empty_like(fx)
    ctx = context_for(fx)
    ctx.check(_lib.load().jx_flux_divergence(ctx.handle, dtype_id(fx), ny, nx, nv, ptr(fx),
                                             ptr(fy), float(dx), float(dy), ptr(out)),
              "flux_divergence")
    return out
