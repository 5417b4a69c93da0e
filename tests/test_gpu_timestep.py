"""Timestep limits (hydro.py:63-89) on shapes that exercise the row-marching kernel's strip,
band and periodic-wrap logic: grids narrower than one 62-column strip, exact multiples,
one column past a strip, single rows, and bands that do not divide the grid."""

import numpy as np
import pytest
import torch

import oracle.jathenax_np as ox

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (3, 5), (7, 130), (67, 61), (5, 62), (9, 63), (130, 124), (201, 125)]


@pytest.fixture(scope="module")
def jx():
    import jathenax_amd
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    return jathenax_amd


def _prims(ny, nx, seed):
    rng = np.random.default_rng(seed)
    W = np.empty((ny, nx, 3))
    W[..., 0] = 0.5 + rng.random((ny, nx))
    W[..., 1:] = rng.normal(0.0, 0.7, (ny, nx, 2))
    return W


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("smag", (0.0, 0.17))
def test_timestep_matches_oracle_fp64(jx, shape, smag):
    ny, nx = shape
    W = _prims(ny, nx, ny * 1000 + nx)
    dx, dy = 1.0 / nx, 1.0 / (ny + 1)
    for nu, drag in ((0.0, 0.0), (1e-3, 0.5)):
        want = ox.isothermal_timestep(W, dx, dy, 1.1, 0.4, np.inf, 0.0, 10.0, viscosity=nu,
                                      smagorinsky_cs=smag, drag_rate=drag)
        got = jx.isothermal_timestep(torch.tensor(W, device="cuda"), dx, dy, 1.1, 0.4, np.inf,
                                     0.0, 10.0, viscosity=nu, smagorinsky_cs=smag,
                                     drag_rate=drag)
        # Same face values and limits as the oracle; the final dt expression may round once
        # differently (as in test_gpu_parity's timestep cases).
        assert abs(got - float(want)) <= 1e-15 * float(want), (shape, smag, nu, got,
                                                                float(want))


@pytest.mark.parametrize("shape", [(3, 5), (67, 61), (201, 125)])
def test_timestep_fp32_within_rounding(jx, shape):
    """FP32 forms the CFL limit from signal rates and the Smagorinsky rate in float32: within
    a few float32 roundings (ulp 6e-8) of the float64 oracle on the same float32 data."""
    ny, nx = shape
    W = _prims(ny, nx, 7).astype(np.float32)
    dx, dy = 1.0 / nx, 1.0 / ny
    for smag in (0.0, 0.17):
        want = ox.isothermal_timestep(W.astype(np.float64), dx, dy, 1.0, 0.4, np.inf, 0.0, 10.0,
                                      smagorinsky_cs=smag)
        got = jx.isothermal_timestep(torch.tensor(W, device="cuda"), dx, dy, 1.0, 0.4, np.inf,
                                     0.0, 10.0, smagorinsky_cs=smag)
        assert abs(got - float(want)) <= 1e-6 * float(want), (shape, smag, got, float(want))


def test_nan_state_gives_nan_dt(jx):
    """A NaN anywhere in the state propagates to dt, as jnp.min does (hydro.py:68)."""
    W = _prims(40, 70, 3)
    W[17, 66, 1] = np.nan
    for dtype in (torch.float64, torch.float32):
        for smag in (0.0, 0.17):
            dt = jx.isothermal_timestep(torch.tensor(W, dtype=dtype, device="cuda"), 0.1, 0.1,
                                        1.0, 0.4, np.inf, 0.0, 10.0, smagorinsky_cs=smag)
            assert np.isnan(dt), (dtype, smag, dt)
