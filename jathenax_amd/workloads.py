"""Seeded workload generators for the ensemble and serving configurations.

Each spec is a dict of evolve_isothermal arguments plus the initial state, so the same
instance can be run alone (single-grid path), in an Ensemble, or through the Server.

* ensemble_mix: independent instances with grid sizes log-uniform in 64..512 per axis
  (non-square allowed); a seed-drawn mix of Kelvin-Helmholtz shear layers (shear
  amplitude U(0.3, 0.7), tanh interfaces, passive tracer) and driven turbulence (its own
  forcing seed); reconstruction in {plm, ppm4, wenoz}, Riemann solver in {llf, hlle, roe}
  and RK2/RK3 drawn uniformly.
* serving_requests: Poisson arrivals at a fixed rate; square grids log-uniform in
  128..2048; tlim uniform in [0.1, 2]; Kelvin-Helmholtz or driven turbulence with seeded
  parameters; PPM4 + Roe + RK3.
"""

import numpy as np

from . import forcing as _forcing
from . import problems

# Driven-turbulence physics of the 4096^2 driven workload: solenoidal shell 1 <= |k| <= 3
# peaked at 2, uniform viscosity, Smagorinsky 0.17, linear drag.
TURB = dict(viscosity=1.5e-6, smagorinsky_cs=0.17, drag_rate=0.3333335, dedt=0.03703705,
            nlow=1, nhigh=3, npeak=2.0, kmax=3, tcorr=0.75, sol_fraction=1.0)


def _log_uniform_int(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def turbulence_spec(nx, ny, tlim, seed, dt_update, integrator="rk3", reconstruct="ppm4",
                    rsolver="roe", cfl=0.3, dtype=np.float64, device="cuda"):
    modes = _forcing.cartesian_modes(TURB["nlow"], TURB["nhigh"], (-TURB["kmax"], TURB["kmax"]),
                                     (-TURB["kmax"], TURB["kmax"]))
    updates = int(np.ceil(tlim / dt_update)) + 1
    coef = _forcing.modal_coefficients(modes, (1.0, 1.0), TURB["nlow"], TURB["nhigh"],
                                       TURB["npeak"], TURB["sol_fraction"], int(seed),
                                       TURB["tcorr"], dt_update, updates)
    basis = _forcing.forcing_basis(modes, problems.cell_centres(nx).astype(dtype),
                                   problems.cell_centres(ny).astype(dtype), (1.0, 1.0),
                                   device=device)
    return dict(U=problems.at_rest(nx, ny), dx=1.0 / nx, dy=1.0 / ny, cs=1.0,
                dfloor=1.17549e-38, integrator=integrator, cfl=cfl, tlim=float(tlim),
                nlim=10 ** 9,
                kw=dict(reconstruct=reconstruct, rsolver=rsolver,
                        viscosity=TURB["viscosity"], smagorinsky_cs=TURB["smagorinsky_cs"],
                        drag_rate=TURB["drag_rate"], forcing=(coef.astype(dtype), basis),
                        dt_update=dt_update, tdriv_start=0.0, dedt=TURB["dedt"]),
                kind="turbulence", seed=int(seed))


def kh_spec(nx, ny, tlim, shear, integrator="rk3", reconstruct="ppm4", rsolver="roe",
            cfl=0.4):
    return dict(U=problems.kelvin_helmholtz_tanh(nx, ny, shear=shear), dx=1.0 / nx,
                dy=1.0 / ny, cs=1.0, dfloor=1e-10, integrator=integrator, cfl=cfl,
                tlim=float(tlim), nlim=10 ** 9,
                kw=dict(reconstruct=reconstruct, rsolver=rsolver), kind="kh",
                shear=float(shear))


def ensemble_mix(n=256, seed=20261203, kh_tlim=0.1, turb_tlim=0.05, dt_update=1e-4,
                 dtype=np.float32, device="cuda"):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        nx, ny = _log_uniform_int(rng, 64, 512), _log_uniform_int(rng, 64, 512)
        recon = ("plm", "ppm4", "wenoz")[rng.integers(3)]
        rs = ("llf", "hlle", "roe")[rng.integers(3)]
        integ = ("rk2", "rk3")[rng.integers(2)]
        if rng.random() < 0.5:
            specs.append(kh_spec(nx, ny, kh_tlim, rng.uniform(0.3, 0.7), integ, recon, rs))
        else:
            specs.append(turbulence_spec(nx, ny, turb_tlim, int(rng.integers(1, 2 ** 31 - 1)),
                                         dt_update, integ, recon, rs, dtype=dtype,
                                         device=device))
    return specs


def serving_requests(n=32, rate=20.0, seed=1618, tmin=0.1, tmax=2.0, size_lo=128,
                     size_hi=2048, dt_update=1e-3, dtype=np.float32, device="cuda"):
    """(arrival_offset_seconds, spec) pairs with Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        nside = _log_uniform_int(rng, size_lo, size_hi)
        tlim = rng.uniform(tmin, tmax)
        if rng.random() < 0.5:
            spec = kh_spec(nside, nside, tlim, rng.uniform(0.3, 0.7))
        else:
            spec = turbulence_spec(nside, nside, tlim, int(rng.integers(1, 2 ** 31 - 1)),
                                   dt_update, dtype=dtype, device=device)
        out.append((t, spec))
    return out


def zone_stages(spec, cycles):
    """Zone updates an instance performed: cells x RK stages x cycles."""
    nx, ny = spec["U"].shape[1], spec["U"].shape[0]
    return nx * ny * {"rk2": 2, "rk3": 3, "rk4": 4}[spec["integrator"]] * cycles
