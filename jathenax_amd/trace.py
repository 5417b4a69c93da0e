"""Traced and learned-stress stage paths (jathenax/hydro.py:92-146, jathenax/run.py:20-69).

The fused stage kernel keeps face states and fluxes on chip. Two callers need them in HBM:
`isothermal_step(trace=True)` returns them, and `predict_stress` calls back into Python
between stages. For these, each stage runs as separate HIP launches, one per reference
array function, keeping every intermediate array: reconstruction, Riemann fluxes, viscous
and learned-stress face fluxes, divergence, RK combine, drag, kick and recovery. The
launches use the fused kernel's scalar operators, so in FP64 a traced step is bitwise
equal to a fused one.

`predict_stress(sgs_weights, primitives)` receives the stage's primitives (rho, vx, vy)
as a (ny, nx, 3) CUDA tensor view and must return a floating (ny, nx, 3) array of
(tau_xx, tau_xy, tau_yy): a tensor on the same device, or anything `torch.as_tensor`
accepts. The stress is averaged onto faces and added to the momentum fluxes with a plus
sign (hydro.py:111-119).
"""

import math

import numpy as np
import torch

from . import _lib
from .device import as_device, context_for, dtype_id, ptr
from .fluxes import (RECONSTRUCTIONS, RIEMANN_SOLVERS, add_momentum_flux, flux_divergence,
                     isothermal_flux, reconstruct_faces, stress_face_fluxes,
                     viscous_momentum_flux)
from .forcing import normalize_force, update_force
from .hydro import RK_STAGES, isothermal_timestep, recover_isothermal, rk_stage


def _drag(U, W, beta, dt, rate):
    out = U.clone()
    ny, nx, nv = U.shape
    ctx = context_for(U)
    ctx.check(_lib.load().jx_apply_drag(ctx.handle, dtype_id(U), ny * nx, nv, ptr(out), ptr(W),
                                        float(beta), float(dt), float(rate)), "drag")
    return out


def _kick(U, acceleration, dt):
    out = U.clone()
    ny, nx, nv = U.shape
    ctx = context_for(U)
    ctx.check(_lib.load().jx_apply_kick(ctx.handle, dtype_id(U), ny * nx, nv, ptr(out),
                                        ptr(acceleration), float(dt)), "kick")
    return out


def checked_stress(predict_stress, sgs_weights, W, dtype):
    """The caller's stress for primitives W[..., :3], validated as in hydro.py:113-115."""
    primitives = W[..., :3]
    tau = predict_stress(sgs_weights, primitives)
    if not isinstance(tau, torch.Tensor):
        tau = torch.as_tensor(np.asarray(tau))
    if tuple(tau.shape) != tuple(primitives.shape) or not tau.is_floating_point():
        raise ValueError("predict_stress must return a floating array of shape (ny, nx, 3)")
    return tau.to(device=W.device, dtype=dtype)


def stress_callback(predict_stress, sgs_weights, W, tau):
    """A jx_stress_fn for the fused stage: W and tau are the device buffers handed to the
    library; exceptions are kept (the C side only sees a nonzero status) and re-raised by
    the caller."""
    errors = []

    def call(user, stage, W_ptr, tau_ptr):
        try:
            tau.copy_(checked_stress(predict_stress, sgs_weights, W, tau.dtype))
            return 0
        except BaseException as e:   # re-raised after the C call returns
            errors.append(e)
            return 1

    return _lib.STRESS_FN(call), errors


def _stress_faces(predict_stress, sgs_weights, W, dtype):
    return stress_face_fluxes(checked_stress(predict_stress, sgs_weights, W, dtype))


def traced_step(U, dt, dx, dy, cs, dfloor, integrator, *, viscosity=0.0, smagorinsky_cs=0.0,
                drag_rate=0.0, raw_force=None, dedt=0.0, trace=False, reconstruct="ppm4",
                rsolver="roe", predict_stress=None, sgs_weights=None):
    """isothermal_step (hydro.py:92-146) as separate launches; with trace=True returns
    (U, stages) with the reference's per-stage dictionaries."""
    schedule = RK_STAGES[integrator]
    RECONSTRUCTIONS[reconstruct]
    RIEMANN_SOLVERS[rsolver]
    U = as_device(U)
    if U.ndim != 3 or U.shape[-1] < 3:
        raise ValueError("state must have shape (ny, nx, nvar>=3)")
    raw = None
    if raw_force is not None:
        raw = as_device(raw_force, dtype=U.dtype, device=U.device)
        if tuple(raw.shape) != (U.shape[0], U.shape[1], 2):
            raise ValueError("raw_force must have shape (ny, nx, 2)")
    U, W = recover_isothermal(U, dfloor)
    U_saved = U
    stages = []
    for stage, coefficients in enumerate(schedule, 1):
        left_x, right_x = reconstruct_faces(W, 1, reconstruct, dfloor)
        left_y, right_y = reconstruct_faces(W, 0, reconstruct, dfloor)
        flux_x = isothermal_flux(left_x, right_x, cs, 1, rsolver)
        flux_y = isothermal_flux(left_y, right_y, cs, 0, rsolver)
        if viscosity > 0.0 or smagorinsky_cs > 0.0:
            flux_x = add_momentum_flux(flux_x, viscous_momentum_flux(W, dx, dy, viscosity,
                                                                     smagorinsky_cs, 1))
            flux_y = add_momentum_flux(flux_y, viscous_momentum_flux(W, dx, dy, viscosity,
                                                                     smagorinsky_cs, 0))
        if predict_stress is not None:
            tau_x, tau_y = _stress_faces(predict_stress, sgs_weights, W, U.dtype)
            flux_x = add_momentum_flux(flux_x, tau_x)
            flux_y = add_momentum_flux(flux_y, tau_y)
        div_flux = flux_divergence(flux_x, flux_y, dx, dy)
        U, U_saved = rk_stage(U, U_saved, div_flux, dt, coefficients)
        if drag_rate > 0.0:
            U = _drag(U, W, coefficients[2], dt, drag_rate)
        kick = raw is not None and stage == len(schedule)
        if kick:
            before_kick = U
            acceleration = normalize_force(U, raw, dt, dedt)
            U = _kick(U, acceleration, dt)
        if trace:
            stages.append(dict(primitives=W, left_x=left_x, right_x=right_x, left_y=left_y,
                               right_y=right_y, flux_x=flux_x, flux_y=flux_y, saved=U_saved,
                               conserved=U))
            if kick:
                stages[-1].update(before_kick=before_kick, acceleration=acceleration)
        U, W = recover_isothermal(U, dfloor)
    if trace:
        return U, tuple(stages)
    return U


def evolve_with_predictor(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, *,
                          viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0, forcing=None,
                          dt_update=1.0, tdriv_start=0.0, dedt=0.0, reconstruct="ppm4",
                          rsolver="roe", predict_stress=None, sgs_weights=None):
    """evolve_isothermal (run.py:20-69) with a learned stress: the loop runs on the host,
    one traced-path step per cycle.  Time and dt are float64, as in the device loop."""
    U = as_device(U)
    coefficients = basis = None
    raw = None
    if forcing is not None:
        # Tables in the state's precision, as the device loop renders them.
        coefficients = as_device(forcing[0], dtype=U.dtype, device=U.device)
        basis = [as_device(b, dtype=U.dtype, device=U.device) for b in forcing[1]]
        raw = torch.zeros((U.shape[0], U.shape[1], 2), dtype=U.dtype, device=U.device)
    time, prev, cycle, update = 0.0, math.inf, 0, -1
    while time < tlim and cycle < nlim:
        _, W = recover_isothermal(U, dfloor)
        dt = isothermal_timestep(W, dx, dy, cs, cfl, prev, time, tlim, viscosity=viscosity,
                                 smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate)
        power = 0.0
        if forcing is not None:
            update, raw = update_force(time, update, raw, coefficients, basis, dt_update,
                                       tdriv_start)
            power = dedt if update >= 0 else 0.0
        U = traced_step(U, dt, dx, dy, cs, dfloor, integrator, viscosity=viscosity,
                        smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate,
                        raw_force=raw if forcing is not None else None, dedt=power,
                        reconstruct=reconstruct, rsolver=rsolver,
                        predict_stress=predict_stress, sgs_weights=sgs_weights)
        time, prev, cycle = time + dt, dt, cycle + 1
    return U, time, prev, cycle
