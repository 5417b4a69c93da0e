"""Device-resident time loop (jathenax/run.py:20-69) and the FP64 box diagnostics.

`evolve_isothermal` keeps the reference signature.  The loop (recover, timestep, forcing
refresh, step) runs on the GPU: each step is a begin-step control kernel, an optional
render, the fused RK stages and, when forced, the kick; the host only polls the control
block between chunks of steps.  `forcing` is ``(coefficients, basis)`` as built by the
reference's CLI (run.py:148-155): coefficients (updates, modes, 2, 2) and the basis
tuple of forcing.forcing_basis.
"""

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, tuning
from .device import as_device, context_for, dtype_id, ptr
from .fluxes import fused_operator_ids
from .hydro import RK_STAGES, run_params


def _evolve_raw(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, **kw):
    """evolve_isothermal with the context's current tiles (used by the tile tuner)."""
    return _evolve(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, apply_table=False, **kw)


def evolve_isothermal(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, *,
                      viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0,
                      forcing=None, dt_update=1.0, tdriv_start=0.0, dedt=0.0,
                      reconstruct="ppm4", rsolver="roe", predict_stress=None, sgs_weights=None):
    """Advance to tlim or nlim cycles; returns (U, time, dt, cycle)."""
    for name, value in (("viscosity", viscosity), ("smagorinsky_cs", smagorinsky_cs),
                        ("drag_rate", drag_rate), ("dedt", dedt)):
        if not math.isfinite(value) or value < 0.0:
            raise ValueError(f"{name} must be finite and nonnegative")
    if forcing is not None and (not 0.0 < dt_update < np.inf or not math.isfinite(tdriv_start)):
        raise ValueError("dt_update must be finite and positive and tdriv_start finite")
    RK_STAGES[integrator]
    if fused_operator_ids(reconstruct, rsolver) is None:
        # A replaced registry operator: the loop on the host over unfused-path steps.
        from .trace import evolve_with_predictor
        return evolve_with_predictor(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim,
                                     viscosity=viscosity, smagorinsky_cs=smagorinsky_cs,
                                     drag_rate=drag_rate, forcing=forcing, dt_update=dt_update,
                                     tdriv_start=tdriv_start, dedt=dedt,
                                     reconstruct=reconstruct, rsolver=rsolver,
                                     predict_stress=predict_stress, sgs_weights=sgs_weights)
    return _evolve(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, viscosity=viscosity,
                   smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate, forcing=forcing,
                   dt_update=dt_update, tdriv_start=tdriv_start, dedt=dedt,
                   reconstruct=reconstruct, rsolver=rsolver, apply_table=True,
                   predict_stress=predict_stress, sgs_weights=sgs_weights)


@dataclass(frozen=True)
class LoopState:
    """Where a run stands between steps: simulation time (float64), the last step's dt
    (+inf before the first), completed steps, and the OU row of the raw force (-1: none)."""
    time: float = 0.0
    dt: float = math.inf
    cycle: int = 0
    update: int = -1


def evolve_resumable(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, *, state=None,
                     viscosity=0.0, smagorinsky_cs=0.0, drag_rate=0.0, forcing=None,
                     dt_update=1.0, tdriv_start=0.0, dedt=0.0, reconstruct="ppm4",
                     rsolver="roe"):
    """evolve_isothermal that starts from `state` (a LoopState; None = a fresh run) and
    returns (U, LoopState).  Resuming from the state and array a run ended with continues
    bit for bit as the uninterrupted run: checkpoint/restart of long forced runs and
    pre-emption in serving (no reference counterpart)."""
    for name, value in (("viscosity", viscosity), ("smagorinsky_cs", smagorinsky_cs),
                        ("drag_rate", drag_rate), ("dedt", dedt)):
        if not math.isfinite(value) or value < 0.0:
            raise ValueError(f"{name} must be finite and nonnegative")
    return _evolve(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, viscosity=viscosity,
                   smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate, forcing=forcing,
                   dt_update=dt_update, tdriv_start=tdriv_start, dedt=dedt,
                   reconstruct=reconstruct, rsolver=rsolver, apply_table=True,
                   start=state or LoopState(), want_state=True)


def _evolve(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, *, viscosity=0.0,
            smagorinsky_cs=0.0, drag_rate=0.0, forcing=None, dt_update=1.0, tdriv_start=0.0,
            dedt=0.0, reconstruct="ppm4", rsolver="roe", apply_table=True,
            predict_stress=None, sgs_weights=None, start=None, want_state=False):
    U = as_device(U)
    p = run_params(U, dx, dy, cs, dfloor, integrator, reconstruct=reconstruct, rsolver=rsolver,
                   viscosity=viscosity, smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate,
                   cfl=cfl, tlim=tlim, nlim=min(int(nlim), 2 ** 31 - 1), dt_update=dt_update,
                   tdriv_start=tdriv_start, dedt=dedt)
    keep = []
    fptr = None
    if forcing is not None:
        coefficients, basis = forcing
        coef = as_device(coefficients, dtype=U.dtype, device=U.device)
        tabs = [as_device(b, dtype=U.dtype, device=U.device) for b in basis]
        if coef.ndim != 4 or tuple(coef.shape[2:]) != (2, 2):
            raise ValueError("forcing coefficients must have shape (updates, modes, 2, 2)")
        nm = coef.shape[1]
        if (tuple(tabs[0].shape) != (nm, U.shape[1]) or tuple(tabs[2].shape) != (nm, U.shape[0])):
            raise ValueError("forcing basis does not match the modes and grid")
        keep = [coef, *tabs]
        f = _lib.Forcing(coefficients=coef.data_ptr(), cos_x=tabs[0].data_ptr(),
                         sin_x=tabs[1].data_ptr(), cos_y=tabs[2].data_ptr(),
                         sin_y=tabs[3].data_ptr(), n_updates=coef.shape[0], n_modes=nm)
        fptr = ctypes.byref(f)
    out = torch.empty_like(U)
    t, dt, cyc = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    ctx = context_for(U)
    if apply_table:
        tuning.apply_single(ctx, p)
    if want_state:
        s0 = _lib.LoopState(time=float(start.time), prev_dt=float(start.dt),
                            cycle=int(start.cycle), update=int(start.update))
        s1 = _lib.LoopState()
        ctx.check(_lib.load().jx_evolve_isothermal_resume(ctx.handle, ctypes.byref(p), ptr(U),
                                                          ptr(out), fptr, ctypes.byref(s0),
                                                          ctypes.byref(s1)), "evolve_isothermal")
        del keep
        return out, LoopState(s1.time, s1.prev_dt, s1.cycle, s1.update)
    if predict_stress is None:
        ctx.check(_lib.load().jx_evolve_isothermal(ctx.handle, ctypes.byref(p), ptr(U), ptr(out),
                                                   fptr, ctypes.byref(t), ctypes.byref(dt),
                                                   ctypes.byref(cyc)), "evolve_isothermal")
    else:
        # The device loop with a stress callback between stages (one step per round trip).
        from .trace import stress_callback
        W = torch.empty_like(U)
        tau = torch.empty((U.shape[0], U.shape[1], 3), dtype=U.dtype, device=U.device)
        fn, errors = stress_callback(predict_stress, sgs_weights, W, tau)
        rc = _lib.load().jx_evolve_isothermal_sgs(ctx.handle, ctypes.byref(p), ptr(U), ptr(out),
                                                  fptr, ptr(W), ptr(tau), fn, None,
                                                  ctypes.byref(t), ctypes.byref(dt),
                                                  ctypes.byref(cyc))
        if errors:
            raise errors[0]
        ctx.check(rc, "evolve_isothermal")
    del keep
    return out, t.value, dt.value, cyc.value


def conserved_sums(U):
    """Float64 sums of each conserved component over the box (run.py:186-187 / (dx dy))."""
    U = as_device(U)
    ny, nx, nv = U.shape
    out = (ctypes.c_double * nv)()
    ctx = context_for(U)
    ctx.check(_lib.load().jx_conserved_sums(ctx.handle, dtype_id(U), ny, nx, nv, ptr(U), out),
              "conserved_sums")
    return np.asarray(out[:], dtype=np.float64)


def check_finished(U, time, cycle, tlim, nlim):
    """The reference CLI's end-of-run check (run.py:178-180)."""
    reached = time == tlim
    finite = bool(torch.isfinite(U).all()) and math.isfinite(time)
    if not finite or (not reached and cycle != nlim):
        raise RuntimeError(f"Periodic run ended at t={time} after {cycle} steps")
