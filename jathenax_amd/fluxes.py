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


def _axis(axis):
    if axis not in (0, 1):
        raise ValueError("axis must be 0 (y faces) or 1 (x faces)")
    return axis


def reconstruct_faces(W, axis, reconstruct, dfloor):
    """(left, right) states at each cell's low face along `axis` (fluxes.py:124-131)."""
    recon = RECONSTRUCTIONS[reconstruct]
    W = as_device(W)
    if W.ndim != 3:
        raise ValueError("W must have shape (ny, nx, nvar)")
    ny, nx, nv = W.shape
    left, right = torch.empty_like(W), torch.empty_like(W)
    ctx = context_for(W)
    ctx.check(_lib.load().jx_reconstruct_faces(ctx.handle, dtype_id(W), ny, nx, nv, ptr(W),
                                               _axis(axis), recon, float(dfloor), ptr(left),
                                               ptr(right)), "reconstruct_faces")
    return left, right


def isothermal_flux(left, right, cs, axis, rsolver="roe"):
    """Gas and passive-scalar fluxes in (x, y) component order (fluxes.py:207-219)."""
    rs = RIEMANN_SOLVERS[rsolver]
    left = as_device(left)
    right = as_device(right, dtype=left.dtype, device=left.device)
    if left.shape != right.shape or left.ndim != 3 or left.shape[-1] < 3:
        raise ValueError("left and right must both have shape (ny, nx, nvar>=3)")
    ny, nx, nv = left.shape
    flux = torch.empty_like(left)
    ctx = context_for(left)
    ctx.check(_lib.load().jx_isothermal_flux(ctx.handle, dtype_id(left), ny, nx, nv, ptr(left),
                                             ptr(right), float(cs), _axis(axis), rs, ptr(flux)),
              "isothermal_flux")
    return flux


def viscous_momentum_flux(W, dx, dy, viscosity, smagorinsky_cs, axis):
    """-mu * stress at each low face along `axis`, shape (ny, nx, 2) (fluxes.py:248-266)."""
    W = as_device(W)
    if W.ndim != 3 or W.shape[-1] < 3:
        raise ValueError("W must have shape (ny, nx, nvar>=3)")
    ny, nx, nv = W.shape
    out = torch.empty((ny, nx, 2), dtype=W.dtype, device=W.device)
    ctx = context_for(W)
    ctx.check(_lib.load().jx_viscous_momentum_flux(ctx.handle, dtype_id(W), ny, nx, nv, ptr(W),
                                                   float(dx), float(dy), float(viscosity),
                                                   float(smagorinsky_cs), _axis(axis),
                                                   ptr(out)), "viscous_momentum_flux")
    return out


def add_momentum_flux(flux, add):
    """flux[..., 1:3] + add as a new array (hydro.py:106-110, 117-119)."""
    flux = as_device(flux)
    add = as_device(add, dtype=flux.dtype, device=flux.device)
    ny, nx, nv = flux.shape
    if tuple(add.shape) != (ny, nx, 2):
        raise ValueError("momentum flux increment must have shape (ny, nx, 2)")
    out = flux.clone()
    ctx = context_for(flux)
    ctx.check(_lib.load().jx_add_momentum_flux(ctx.handle, dtype_id(flux), ny * nx, nv, ptr(out),
                                               ptr(add)), "add_momentum_flux")
    return out
