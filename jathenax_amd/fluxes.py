"""Face operators, their registries and the standalone flux functions (jathenax/fluxes.py).

Every operator runs as a HIP kernel on the argument's device.  The registries map the
reference's names to callables with the reference's signatures, in the reference's order
(fluxes.py:120-123, fluxes.py:204): RECONSTRUCTIONS[name](W, axis) -> (left, right) and
RIEMANN_SOLVERS[name](rho_l, rho_r, vn_l, vn_r, vt_l, vt_r, cs) -> flux (..., 3) in
(mass, normal, tangential) order.  The fused stage kernel (csrc/jx_stage.hpp) evaluates the
built-in operators itself; a registry entry replaced by the caller's own function is called
as the reference calls it, on the unfused stage path (hydro.isothermal_step,
run.evolve_isothermal), and unknown names raise KeyError as the reference's dict lookups do.
"""

import math

import torch

from . import _lib
from .device import as_device, context_for, dtype_id, ptr


def _axis(axis):
    if axis not in (0, 1):
        raise ValueError("axis must be 0 (y faces) or 1 (x faces)")
    return axis


def _faces(W, axis, recon_id, dfloor):
    """Faces along axis 1 (x) or 0 (y) of a (ny, nx, nvar) state or a (ny, nx) field."""
    W = as_device(W)
    if W.ndim not in (2, 3):
        raise ValueError("W must have shape (ny, nx, nvar) or (ny, nx)")
    field = W.ndim == 2
    if field:
        W = W.unsqueeze(-1).contiguous()
    ny, nx, nv = W.shape
    left, right = torch.empty_like(W), torch.empty_like(W)
    ctx = context_for(W)
    ctx.check(_lib.load().jx_reconstruct_faces(ctx.handle, dtype_id(W), ny, nx, nv, ptr(W),
                                               _axis(axis), recon_id, float(dfloor), ptr(left),
                                               ptr(right)), "reconstruct_faces")
    if field:
        return left[..., 0], right[..., 0]
    return left, right


def dc_face_states(W, axis):
    """The two adjacent cell averages at each face (fluxes.py:6-8)."""
    return _faces(W, axis, 0, math.nan)


def plm_face_states(W, axis):
    """Piecewise linear edges with the harmonic limiter (fluxes.py:11-18)."""
    return _faces(W, axis, 1, math.nan)


def ppm4_face_states(W, axis):
    """Fourth-order PPM edges with AthenaK's PPM4 limiter (fluxes.py:21-42)."""
    return _faces(W, axis, 2, math.nan)


def ppmx_face_states(W, axis):
    """PPM with the smooth-extremum limiter (fluxes.py:45-84)."""
    return _faces(W, axis, 3, math.nan)


def wenoz_face_states(W, axis):
    """Fifth-order WENO-Z edges (fluxes.py:87-117)."""
    return _faces(W, axis, 4, math.nan)


RECONSTRUCTIONS = {
    "dc": dc_face_states, "plm": plm_face_states, "ppm4": ppm4_face_states,
    "ppmx": ppmx_face_states, "wenoz": wenoz_face_states,
}
_RECON_IDS = {dc_face_states: 0, plm_face_states: 1, ppm4_face_states: 2,
              ppmx_face_states: 3, wenoz_face_states: 4}


def _solver(rs_id, name, ref):
    def flux(rho_l, rho_r, vn_l, vn_r, vt_l, vt_r, cs):
        parts = [as_device(x) for x in (rho_l, rho_r, vn_l, vn_r, vt_l, vt_r)]
        dtype = torch.float32 if all(t.dtype == torch.float32 for t in parts) else torch.float64
        parts = torch.broadcast_tensors(*[t.to(dtype) for t in parts])
        shape = parts[0].shape
        left = torch.stack((parts[0], parts[2], parts[4]), dim=-1).reshape(1, -1, 3)
        right = torch.stack((parts[1], parts[3], parts[5]), dim=-1).reshape(1, -1, 3)
        if left.shape[1] == 0:
            return torch.empty((*shape, 3), dtype=dtype, device=left.device)
        out = _flux_kernel(left, right, cs, 1, rs_id)
        return out.reshape(*shape, 3)

    flux.__name__ = name
    flux.__doc__ = (f"Isothermal {name[1:-5].upper()} flux in (mass, normal, tangential) "
                    f"order ({ref}).")
    return flux


def _flux_kernel(left, right, cs, axis, rs_id):
    ny, nx, nv = left.shape
    flux = torch.empty_like(left)
    ctx = context_for(left)
    ctx.check(_lib.load().jx_isothermal_flux(ctx.handle, dtype_id(left), ny, nx, nv, ptr(left),
                                             ptr(right), float(cs), _axis(axis), rs_id,
                                             ptr(flux)), "isothermal_flux")
    return flux


_llf_flux = _solver(0, "_llf_flux", "fluxes.py:175-183")
_hlle_flux = _solver(1, "_hlle_flux", "fluxes.py:186-201")
_roe_flux = _solver(2, "_roe_flux", "fluxes.py:134-172")

RIEMANN_SOLVERS = {"llf": _llf_flux, "hlle": _hlle_flux, "roe": _roe_flux}
_RSOLVER_IDS = {_llf_flux: 0, _hlle_flux: 1, _roe_flux: 2}


def fused_operator_ids(reconstruct, rsolver):
    """Kernel ids of the registry entries `reconstruct` and `rsolver` (KeyError for unknown
    names), or None when either entry is not a built-in operator."""
    rid = _RECON_IDS.get(RECONSTRUCTIONS[reconstruct])
    sid = _RSOLVER_IDS.get(RIEMANN_SOLVERS[rsolver])
    if rid is None or sid is None:
        return None
    return rid, sid


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


def reconstruct_faces(W, axis, reconstruct, dfloor):
    """(left, right) states at each cell's low face along `axis`; PPMX and WENO-Z floor the
    face density (fluxes.py:124-131)."""
    fn = RECONSTRUCTIONS[reconstruct]
    rid = _RECON_IDS.get(fn)
    if rid is not None:
        return _faces(W, axis, rid, dfloor)
    left, right = fn(W, axis)
    left, right = as_device(left), as_device(right)
    if reconstruct in ("ppmx", "wenoz"):
        left, right = left.clone(), right.clone()
        left[..., 0] = torch.fmax(left[..., 0], torch.tensor(dfloor, dtype=left.dtype,
                                                             device=left.device))
        right[..., 0] = torch.fmax(right[..., 0], torch.tensor(dfloor, dtype=right.dtype,
                                                               device=right.device))
    return left, right


def isothermal_flux(left, right, cs, axis, rsolver="roe"):
    """Gas and passive-scalar fluxes in (x, y) component order (fluxes.py:207-219)."""
    fn = RIEMANN_SOLVERS[rsolver]
    left = as_device(left)
    right = as_device(right, dtype=left.dtype, device=left.device)
    if left.shape != right.shape or left.ndim != 3 or left.shape[-1] < 3:
        raise ValueError("left and right must both have shape (ny, nx, nvar>=3)")
    rs_id = _RSOLVER_IDS.get(fn)
    if rs_id is not None:
        return _flux_kernel(left, right, cs, _axis(axis), rs_id)
    # A caller's solver, applied as fluxes.py:209-219 applies the registry entry.
    normal, tangent = (1, 2) if _axis(axis) == 1 else (2, 1)
    flux = as_device(fn(left[..., 0], right[..., 0], left[..., normal], right[..., normal],
                        left[..., tangent], right[..., tangent], cs), dtype=left.dtype,
                     device=left.device)
    if axis == 0:
        flux = torch.stack((flux[..., 0], flux[..., 2], flux[..., 1]), dim=-1)
    if left.shape[-1] > 3:
        mass = flux[..., :1]
        upwind = torch.where(mass >= 0, left[..., 3:], right[..., 3:])
        flux = torch.cat((flux, mass * upwind), dim=-1)
    return flux.contiguous()


def roe_isothermal_flux(left, right, cs, axis):
    """AthenaK's isothermal Roe flux, including any passive scalars (fluxes.py:222-224)."""
    return isothermal_flux(left, right, cs, axis)


def face_velocity_gradients(W, dx, dy, axis):
    """(dv/dx, dv/dy), each (ny, nx, 2), at each low face along `axis` (fluxes.py:227-238)."""
    W = as_device(W)
    if W.ndim != 3 or W.shape[-1] < 3:
        raise ValueError("W must have shape (ny, nx, nvar>=3)")
    ny, nx, nv = W.shape
    dv_dx = torch.empty((ny, nx, 2), dtype=W.dtype, device=W.device)
    dv_dy = torch.empty_like(dv_dx)
    ctx = context_for(W)
    ctx.check(_lib.load().jx_face_velocity_gradients(ctx.handle, dtype_id(W), ny, nx, nv, ptr(W),
                                                     float(dx), float(dy), _axis(axis),
                                                     ptr(dv_dx), ptr(dv_dy)),
              "face_velocity_gradients")
    return dv_dx, dv_dy


def smagorinsky_viscosity(dv_dx, dv_dy, dx, dy, smagorinsky_cs):
    """Kinematic eddy viscosity (Cs sqrt(dx dy))^2 |S| at each face (fluxes.py:241-245)."""
    gx = as_device(dv_dx)
    gy = as_device(dv_dy, dtype=gx.dtype, device=gx.device)
    if gx.shape != gy.shape or gx.ndim < 1 or gx.shape[-1] != 2:
        raise ValueError("dv_dx and dv_dy must both have shape (..., 2)")
    nu = torch.empty(gx.shape[:-1], dtype=gx.dtype, device=gx.device)
    ctx = context_for(gx)
    ctx.check(_lib.load().jx_smagorinsky_viscosity(ctx.handle, dtype_id(gx), nu.numel(), ptr(gx),
                                                   ptr(gy), float(dx), float(dy),
                                                   float(smagorinsky_cs), ptr(nu)),
              "smagorinsky_viscosity")
    return nu


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
