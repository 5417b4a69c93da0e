"""Initial conditions and forcing set-up for the benchmark workloads (host, float64).

Coordinates are the reference's cell centres x_i = (i + 1/2)/nx - 1/2 on [-1/2, 1/2)
(jathenax/run.py:133-135).  Both the reference's problem formulas and the workload
formulas of the benchmark configurations are provided; where they differ the function
name says which is which.
"""

from dataclasses import dataclass

import numpy as np

from . import forcing as _forcing


def cell_centres(n):
    """x_i = (i + 1/2) dx - 1/2 with dx = 1/n, evaluated as the CLI does (run.py:133-135)."""
    return (np.arange(n) + 0.5) * (1.0 / n) - 0.5


def smooth_reference(nx, ny):
    """'smooth' problem of the reference CLI (run.py:158-163)."""
    x, y = cell_centres(nx), cell_centres(ny)
    px, py = np.meshgrid(2.0 * np.pi * x, 2.0 * np.pi * y)
    rho = 1.0 + 0.1 * np.sin(px + py)
    return np.stack((rho, rho * (0.2 * np.cos(py)), rho * (0.2 * np.sin(px))), axis=-1)


def smooth_wave(nx, ny):
    """Smooth periodic flow: rho = 1 + 0.2 sin(2 pi x) cos(2 pi y), vx = 0.1 sin(2 pi y),
    vy = 0.1 cos(2 pi x) (the plain-method and 4096^2 headline workload)."""
    x, y = cell_centres(nx), cell_centres(ny)
    X, Y = np.meshgrid(x, y)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    vx = 0.1 * np.sin(2 * np.pi * Y)
    vy = 0.1 * np.cos(2 * np.pi * X)
    return np.stack((rho, rho * vx, rho * vy), axis=-1)


def kelvin_helmholtz_reference(nx, ny):
    """Reference KH shear layers with tracer, nvar = 4 (initial_conditions.py:6-15)."""
    X, Y = np.meshgrid(cell_centres(nx), cell_centres(ny))
    dist = np.abs(Y) - 0.25
    tr = np.clip(-dist / 0.025, -1.0, 1.0)
    s = 0.5 + tr * (0.75 - 0.25 * tr ** 2)
    rho = np.ones_like(s)
    vx = s - 0.5
    vy = 0.02 * np.sin(4 * np.pi * X) * np.exp(-(dist / 0.05) ** 2)
    return np.stack((rho, rho * vx, rho * vy, rho * s), axis=-1)


def kelvin_helmholtz_tanh(nx, ny, shear=0.5, width=0.01, amp=0.01):
    """KH with tanh interfaces: vx = +shear for |y| < 1/4, -shear outside, vy = amp
    sin(4 pi x) exp(-(|y|-1/4)^2/0.05^2), tracer 1 inside, 0 outside (nvar = 4)."""
    X, Y = np.meshgrid(cell_centres(nx), cell_centres(ny))
    inside = 0.5 * (1.0 + np.tanh((0.25 - np.abs(Y)) / width))
    vx = shear * (2.0 * inside - 1.0)
    vy = amp * np.sin(4 * np.pi * X) * np.exp(-((np.abs(Y) - 0.25) / 0.05) ** 2)
    rho = np.ones_like(X)
    return np.stack((rho, rho * vx, rho * vy, rho * inside), axis=-1)


def at_rest(nx, ny):
    U = np.zeros((ny, nx, 3))
    U[..., 0] = 1.0
    return U


@dataclass
class Driving:
    """Turbulence driving parameters (reference defaults, input.py:58-62)."""
    rseed: int = 1618
    tcorr: float = 0.75
    dt_update: float = 0.0075
    tdriv_start: float = 0.0
    dedt: float = 0.03703705
    nlow: int = 3
    nhigh: int = 5
    npeak: float = 4.0
    sol_fraction: float = 1.0
    kmax: int = 5


def turbulence_forcing(nx, ny, tlim, driving=Driving(), dtype=np.float64, device="cuda"):
    """(coefficients, basis) for evolve_isothermal, built like run.py:145-155."""
    modes = _forcing.cartesian_modes(driving.nlow, driving.nhigh, (-driving.kmax, driving.kmax),
                                     (-driving.kmax, driving.kmax))
    updates = int(np.ceil(max(0.0, tlim - driving.tdriv_start) / driving.dt_update)) + 1
    coef = _forcing.modal_coefficients(modes, (1.0, 1.0), driving.nlow, driving.nhigh,
                                       driving.npeak, driving.sol_fraction, driving.rseed,
                                       driving.tcorr, driving.dt_update, updates)
    x = cell_centres(nx).astype(dtype)
    y = cell_centres(ny).astype(dtype)
    basis = _forcing.forcing_basis(modes, x, y, (1.0, 1.0), device=device)
    return coef.astype(dtype), basis
