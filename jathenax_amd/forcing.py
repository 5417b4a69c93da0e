"""Continuous Cartesian parabolic OU driving (jathenax/forcing.py).

The seed path (uniform and Gaussian streams, mode list, OU coefficient table) runs on the
host in float64 C++ (csrc/jx_seed.cpp) so that the device receives the same table as the
reference.  Rendering, normalisation and the kick run on the GPU.
"""

import ctypes
from math import pi

import numpy as np
import torch

from . import _lib
from .device import as_device, context_for, dtype_id, ptr

_BATCH = 4096


def ran_st(seed):
    """Shuffled two-stream uniform generator (forcing.py:11-32), as an endless iterator."""
    lib = _lib.load()
    skip = 0
    while True:
        n = skip + _BATCH
        buf = (ctypes.c_double * n)()
        _lib.check(lib.jx_ran_st(int(seed), n, buf), None, "ran_st")
        yield from buf[skip:]
        skip = n


def ran_gaussian_st(seed):
    """Polar Box-Muller normals, second of each pair first (forcing.py:35-45)."""
    lib = _lib.load()
    skip = 0
    while True:
        n = skip + _BATCH
        buf = (ctypes.c_double * n)()
        _lib.check(lib.jx_ran_gaussian_st(int(seed), n, buf), None, "ran_gaussian_st")
        yield from buf[skip:]
        skip = n


def cartesian_modes(nlow, nhigh, kx_bounds, ky_bounds):
    """Signed modes with nlow^2 <= |k|^2 <= nhigh^2, kx outer (forcing.py:48-60)."""
    if not 1 <= nlow < nhigh:
        raise ValueError("parabolic driving requires 1 <= nlow < nhigh")
    lib = _lib.load()
    args = (int(nlow), int(nhigh), int(kx_bounds[0]), int(kx_bounds[1]), int(ky_bounds[0]),
            int(ky_bounds[1]))
    n = lib.jx_cartesian_modes(*args, None, 0)
    if n <= 0:
        raise ValueError("no Cartesian forcing modes in the requested band")
    buf = (ctypes.c_int * (2 * n))()
    lib.jx_cartesian_modes(*args, buf, n)
    return np.frombuffer(buf, dtype=np.int32).astype(np.int64).reshape(n, 2)


def modal_coefficients(modes, lengths, nlow, nhigh, npeak, sol_fraction, rseed,
                       tcorr, dt_update, updates):
    """Float64 OU table (updates, modes, dir, re/im) (forcing.py:63-100)."""
    if not nlow <= npeak <= nhigh or not 0.0 <= sol_fraction <= 1.0:
        raise ValueError("npeak must be in the forcing band and sol_fraction in [0, 1]")
    if not np.isfinite(tcorr) or tcorr < 0.0 or not 0.0 < dt_update < np.inf:
        raise ValueError("tcorr must be finite and nonnegative and dt_update positive")
    modes = np.ascontiguousarray(np.asarray(modes, dtype=np.int32).reshape(-1, 2))
    out = np.empty((int(updates), len(modes), 2, 2))
    _lib.check(_lib.load().jx_modal_coefficients(
        modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(modes), float(lengths[0]),
        float(lengths[1]), int(nlow), int(nhigh), float(npeak), float(sol_fraction),
        int(rseed), float(tcorr), float(dt_update), int(updates),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), None, "modal_coefficients")
    return out


def forcing_basis(modes, x, y, lengths, device="cuda"):
    """cos/sin phase tables (n_modes, nx) and (n_modes, ny) in x's dtype (forcing.py:103-108).

    The tables are evaluated once on the host with the state dtype's arithmetic and
    placed on the device.
    """
    x = np.asarray(x)
    y = np.asarray(y, dtype=x.dtype)
    k = np.asarray(modes) * (2.0 * pi / np.asarray(lengths, dtype=x.dtype))
    px = k[:, 0, None] * x
    py = k[:, 1, None] * y
    dt = torch.float32 if x.dtype == np.float32 else torch.float64
    return tuple(torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
                 for a in (np.cos(px), np.sin(px), np.cos(py), np.sin(py)))


def render_force(amplitudes, basis):
    """Raw acceleration summed over modes in order (forcing.py:111-122)."""
    cx, sx, cy, sy = (as_device(b) for b in basis)
    amp = as_device(amplitudes, dtype=cx.dtype, device=cx.device)
    nm, nx = cx.shape
    ny = cy.shape[1]
    if tuple(amp.shape) != (nm, 2, 2):
        raise ValueError("amplitudes must have shape (n_modes, 2, 2)")
    out = torch.empty((ny, nx, 2), dtype=cx.dtype, device=cx.device)
    ctx = context_for(cx)
    ctx.check(_lib.load().jx_render_force(ctx.handle, dtype_id(cx), ny, nx, nm, ptr(amp),
                                          ptr(cx), ptr(sx), ptr(cy), ptr(sy), ptr(out)),
              "render_force")
    return out


def update_force(time, update, raw_force, coefficients, basis, dt_update, tdriv_start):
    """Re-render when the cycle-start time crosses an OU update (forcing.py:125-134)."""
    requested = int(np.floor((time - tdriv_start) / dt_update)) if time >= tdriv_start else -1
    if requested > update:
        coef = as_device(coefficients)
        if 0 <= requested < coef.shape[0]:
            row = coef[requested]
        else:
            row = torch.full(coef.shape[1:], float("nan"), dtype=coef.dtype, device=coef.device)
        raw_force = render_force(row, basis)
    return requested, raw_force


def normalize_force(U, raw_force, dt, dedt):
    """Zero-net-momentum kick scaled to inject dedt*dt of kinetic energy (forcing.py:137-149)."""
    U = as_device(U)
    raw = as_device(raw_force, dtype=U.dtype, device=U.device)
    ny, nx, nv = U.shape
    out = torch.empty((ny, nx, 2), dtype=U.dtype, device=U.device)
    ctx = context_for(U)
    ctx.check(_lib.load().jx_normalize_force(ctx.handle, dtype_id(U), ny, nx, nv, ptr(U),
                                             ptr(raw), float(dt), float(dedt), ptr(out)),
              "normalize_force")
    return out
