"""Primitive recovery, RK stages, timestep selection and the fused step (jathenax/hydro.py).

`isothermal_step` keeps the reference call signature (hydro.py:92-95) and runs every RK
stage as one fused HIP kernel on the state's device.  Inputs are never modified.
"""

import ctypes
import math

import torch

from . import _lib
from .device import as_device, context_for, dtype_id, ptr
from .fluxes import fused_operator_ids

# (gam0, gam1, beta, delta) rows; the table is part of the interface (hydro.py:20-36).
RK_STAGES = {
    "rk2": ((0.0, 1.0, 1.0, None), (0.5, 0.5, 0.5, None)),
    "rk3": ((0.0, 1.0, 1.0, None), (0.25, 0.75, 0.25, None),
            (2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, None)),
    "rk4": ((0.0, 1.0, 1.193743905974738, None),
            (0.121098479554482, 0.721781678111411, 0.099279895495783, 0.217683334308543),
            (-3.843833699660025, 2.121209265338722, 1.131678018054042, 1.065841341361089),
            (0.546370891121863, 0.198653035682705, 0.310665766509336, 0.0)),
}
MAX_NVAR = 5
MAX_RK_STAGES = 8


def rk_table(integrator):
    """The integrator's rows of RK_STAGES as a flat float64 ctypes array, looked up on every
    call as hydro.py:100 does (KeyError for an unknown name); delta None becomes NaN."""
    rows = RK_STAGES[integrator]
    if not 1 <= len(rows) <= MAX_RK_STAGES:
        raise ValueError(f"RK_STAGES[{integrator!r}] must have 1 to {MAX_RK_STAGES} rows")
    flat = []
    for row in rows:
        gam0, gam1, beta, delta = row
        flat += [float(gam0), float(gam1), float(beta),
                 math.nan if delta is None else float(delta)]
    return len(rows), (ctypes.c_double * len(flat))(*flat)


def run_params(U, dx, dy, cs, dfloor, integrator, *, reconstruct="ppm4", rsolver="roe",
               viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0, cfl=0.0, tlim=0.0, nlim=0,
               dt_update=1.0, tdriv_start=0.0, dedt=0.0):
    """Validate options like the reference (KeyError for unknown names) and pack them.
    The RK rows travel with the parameters, so an edited or added RK_STAGES entry is
    applied as the reference applies it."""
    stages, rows = rk_table(integrator)
    ids = fused_operator_ids(reconstruct, rsolver)
    if ids is None:
        raise ValueError(f"RECONSTRUCTIONS[{reconstruct!r}] or RIEMANN_SOLVERS[{rsolver!r}] is "
                         "not a built-in operator; only isothermal_step and evolve_isothermal "
                         "run replaced operators (on the unfused stage path)")
    recon, rs = ids
    if U.ndim != 3:
        raise ValueError("state must have shape (ny, nx, nvar)")
    ny, nx, nvar = U.shape
    if not 3 <= nvar <= MAX_NVAR:
        raise ValueError(f"nvar must be in [3, {MAX_NVAR}] (hydro plus up to two scalars)")
    p = _lib.RunParams(
        ny=ny, nx=nx, nvar=nvar, dtype=dtype_id(U), reconstruct=recon, rsolver=rs,
        integrator=stages, nlim=int(nlim), dx=float(dx), dy=float(dy), cs=float(cs),
        dfloor=float(dfloor), cfl=float(cfl), tlim=float(tlim), viscosity=float(viscosity),
        smagorinsky_cs=float(smagorinsky_cs), drag_rate=float(drag_rate),
        dt_update=float(dt_update), tdriv_start=float(tdriv_start), dedt=float(dedt),
        rk_rows=ctypes.cast(rows, ctypes.POINTER(ctypes.c_double)))
    p._rk_rows = rows   # keeps the table alive as long as the parameters
    return p


def rk_stage(U, U_saved, div_flux, dt, coefficients):
    """Apply one RK row to U and the saved register (hydro.py:39-51)."""
    gam0, gam1, beta, delta = coefficients
    U = as_device(U)
    Us = as_device(U_saved, dtype=U.dtype)
    div = as_device(div_flux, dtype=U.dtype)
    if not (U.shape == Us.shape == div.shape):
        raise ValueError("U, U_saved and div_flux must have the same shape")
    out = torch.empty_like(U)
    us_out = torch.empty_like(U) if delta is not None else None
    ctx = context_for(U)
    ctx.check(_lib.load().jx_rk_stage(
        ctx.handle, dtype_id(U), U.numel(), ptr(U), ptr(Us), ptr(div), float(dt), float(gam0),
        float(gam1), float(beta), float(delta or 0.0), int(delta is not None), ptr(out),
        ptr(us_out)), "rk_stage")
    return out, (us_out if delta is not None else Us)


def recover_isothermal(U, dfloor):
    """Floor the density into U and return (U, W) with W = (rho, m/rho...) (hydro.py:54-60)."""
    U = as_device(U)
    Uf = torch.empty_like(U)
    W = torch.empty_like(U)
    nv = U.shape[-1]
    ctx = context_for(U)
    ctx.check(_lib.load().jx_recover_isothermal(ctx.handle, dtype_id(U), U.numel() // nv, nv,
                                                ptr(U), float(dfloor), ptr(Uf), ptr(W)),
              "recover_isothermal")
    return Uf, W


def isothermal_timestep(W, dx, dy, cs, cfl, previous_dt, time, tlim, *,
                        viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0):
    """Hydro, diffusion, drag, growth and end-time limits from primitives (hydro.py:63-89)."""
    W = as_device(W)
    if W.ndim != 3 or W.shape[-1] < 3:
        raise ValueError("W must have shape (ny, nx, nvar>=3)")
    ny, nx, nv = W.shape
    p = _lib.RunParams(ny=ny, nx=nx, nvar=nv, dtype=dtype_id(W), dx=float(dx), dy=float(dy),
                       cs=float(cs), cfl=float(cfl), viscosity=float(viscosity),
                       smagorinsky_cs=float(smagorinsky_cs), drag_rate=float(drag_rate))
    out = ctypes.c_double()
    ctx = context_for(W)
    ctx.check(_lib.load().jx_isothermal_timestep(ctx.handle, ctypes.byref(p), ptr(W),
                                                 float(previous_dt), float(time), float(tlim),
                                                 ctypes.byref(out)), "isothermal_timestep")
    return out.value


def isothermal_step(U, dt, dx, dy, cs, dfloor, integrator, *,
                    viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0,
                    raw_force=None, dedt=0.0, trace=False, reconstruct="ppm4", rsolver="roe",
                    predict_stress=None, sgs_weights=None):
    """Advance one periodic isothermal step; every RK stage is one fused HIP launch.

    trace=True runs the stages as separate launches and returns (U, stages) with the
    reference's dictionaries; a registry entry replaced by the caller's own operator also
    takes that path.  predict_stress(sgs_weights, primitives) is called before
    each stage with the stage's (rho, vx, vy) as a (ny, nx, 3) tensor view and must return
    a floating (ny, nx, 3) stress, which the fused stage adds to the momentum fluxes."""
    RK_STAGES[integrator]
    if trace or fused_operator_ids(reconstruct, rsolver) is None:
        from .trace import traced_step
        return traced_step(U, dt, dx, dy, cs, dfloor, integrator, viscosity=viscosity,
                           smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate,
                           raw_force=raw_force, dedt=dedt, trace=trace,
                           reconstruct=reconstruct, rsolver=rsolver,
                           predict_stress=predict_stress, sgs_weights=sgs_weights)
    U = as_device(U)
    p = run_params(U, dx, dy, cs, dfloor, integrator, reconstruct=reconstruct, rsolver=rsolver,
                   viscosity=viscosity, smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate)
    raw = None
    if raw_force is not None:
        raw = as_device(raw_force, dtype=U.dtype, device=U.device)
        if tuple(raw.shape) != (U.shape[0], U.shape[1], 2):
            raise ValueError("raw_force must have shape (ny, nx, 2)")
    out = torch.empty_like(U)
    ctx = context_for(U)
    from . import tuning
    tuning.apply_single(ctx, p)
    if predict_stress is None:
        ctx.check(_lib.load().jx_isothermal_step(ctx.handle, ctypes.byref(p), ptr(U), ptr(out),
                                                 float(dt), ptr(raw), float(dedt)),
                  "isothermal_step")
        return out
    from .trace import stress_callback
    W = torch.empty_like(U)
    tau = torch.empty((U.shape[0], U.shape[1], 3), dtype=U.dtype, device=U.device)
    fn, errors = stress_callback(predict_stress, sgs_weights, W, tau)
    rc = _lib.load().jx_isothermal_step_sgs(ctx.handle, ctypes.byref(p), ptr(U), ptr(out),
                                            float(dt), ptr(raw), float(dedt), ptr(W), ptr(tau),
                                            fn, None)
    if errors:
        raise errors[0]
    ctx.check(rc, "isothermal_step")
    return out


def _finite_nonnegative(**values):
    for name, value in values.items():
        if not math.isfinite(value) or value < 0.0:
            raise ValueError(f"{name} must be finite and nonnegative")
