"""The reference's own unit tests (tests/test_jathenax.py:23-242) run through jathenax_amd, in
float64 and in float32.

The reference writes every tolerance as k * finfo(dtype).eps so that its suite also runs in
float32 (JAX_ENABLE_X64=0, README.md:38); the same tolerances here hold the HIP FP32 path,
with its rewritten operator forms (reciprocal-based divisions, the median PPM4 limiter, the
select-free Roe flux with its LLF fallback, rational HLLE, difference-form WENO-Z), to the
reference's own float32 contract.  The names are imported from jathenax_amd as the
reference test imports them from jathenax.  Every call goes through libjathenax_amd.so.
"""

from math import factorial

import numpy as np
import pytest
import torch

from jathenax_amd.fluxes import (RIEMANN_SOLVERS, flux_divergence, isothermal_flux,
                                 stress_face_fluxes, wenoz_face_states)
from jathenax_amd.forcing import normalize_force
from jathenax_amd.hydro import (RK_STAGES, isothermal_step, isothermal_timestep,
                                recover_isothermal, rk_stage)
from jathenax_amd.problems import kelvin_helmholtz_reference
from jathenax_amd.run import evolve_isothermal

pytestmark = pytest.mark.gpu

DTYPES = (torch.float64, torch.float32)


def _dev(a, dtype):
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    return torch.as_tensor(np.asarray(a), dtype=dtype, device="cuda")


def _np(t):
    return t.detach().cpu().numpy()


def _eps(dtype):
    return float(torch.finfo(dtype).eps)


def _kelvin_helmholtz(x, y):
    # initial_conditions.kelvin_helmholtz(x, y) on the test's own coordinates
    assert np.array_equal(x, y)
    return kelvin_helmholtz_reference(len(x), len(y))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(("integrator", "order"), [("rk2", 2), ("rk3", 3), ("rk4", 4)])
def test_rk_stage_schedule(integrator, order, dtype):
    # test_jathenax.py:23-36: dU/dt = U has the stability polynomial of the scheme.
    dt = 0.125
    U = _dev([1.0], dtype)
    U_saved = U
    for coefficients in RK_STAGES[integrator]:
        U, U_saved = rk_stage(U, U_saved, -U, dt, coefficients)
    expected = sum(dt ** n / factorial(n) for n in range(order + 1))
    np.testing.assert_allclose(_np(U), [expected], rtol=32 * _eps(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_rk3_zero_flux_preserves_float32_density(dtype):
    # test_jathenax.py:39-44, and the same through the fused stage: a uniform state has
    # equal fluxes on every face, so RK2 and RK3, whose later rows are blended, return it
    # bit for bit.  RK4's rows are not blended (gam0 + gam1 != 1): the reference's own
    # arithmetic moves the state by an ulp there, and the fused stage moves it alike.
    from oracle import jathenax_np as ox
    rho = _dev(np.linspace(0.75, 1.25, 1024, dtype=np.float32), dtype)
    out = rk_stage(rho, rho, torch.zeros_like(rho), 0.01, RK_STAGES["rk3"][-1])[0]
    np.testing.assert_array_equal(_np(out), _np(rho))
    U = _dev(np.broadcast_to([1.25, 0.0, 0.0], (24, 40, 3)), dtype)
    for integrator in ("rk2", "rk3"):
        np.testing.assert_array_equal(
            _np(isothermal_step(U, 0.01, 1 / 40, 1 / 24, 1.0, 1e-10, integrator)), _np(U))
    got = _np(isothermal_step(U, 0.01, 1 / 40, 1 / 24, 1.0, 1e-10, "rk4"))
    want = ox.isothermal_step(_np(U), 0.01, 1 / 40, 1 / 24, 1.0, 1e-10, "rk4")
    np.testing.assert_allclose(got, want, rtol=0, atol=2 * _eps(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_density_floor_preserves_momentum(dtype):
    # test_jathenax.py:47-50
    U, W = recover_isothermal(_dev([[[0.125, 0.5, -0.25]]], dtype), dfloor=0.25)
    np.testing.assert_array_equal(_np(U), [[[0.25, 0.5, -0.25]]])
    np.testing.assert_array_equal(_np(W), [[[0.25, 2.0, -1.0]]])


@pytest.mark.parametrize("dtype", DTYPES)
def test_isothermal_timestep_limits(dtype):
    # test_jathenax.py:53-58
    W = _dev([[[1.0, 1.0, 2.0], [1.0, 3.0, 4.0]]], dtype)
    assert isothermal_timestep(W, 1.0, 2.0, 1.0, 0.5, np.inf, 0.0, 1.0) == 0.125
    assert isothermal_timestep(W, 1.0, 2.0, 1.0, 0.5, 0.03125, 0.0, 1.0) == 0.0625
    assert isothermal_timestep(W, 1.0, 2.0, 1.0, 0.5, np.inf, 0.9375, 1.0) == 0.0625


@pytest.mark.parametrize("dtype", DTYPES)
def test_viscous_timestep_limits(dtype):
    # test_jathenax.py:61-75
    W = _dev([[[1.0, 0.0, 0.0]] * 2, [[1.0, 1.0, 0.0]] * 2], dtype)
    dx, dy, cfl = 0.25, 0.5, 0.5
    nu, Cs = 0.5, 2.0
    args = (dx, dy, 1.0, cfl, np.inf, 0.0, 1.0)
    assert isothermal_timestep(W, *args, viscosity=nu) == cfl * 0.25 * dx ** 2 / nu
    max_nu_sgs = Cs ** 2 * dx * dy / dy
    np.testing.assert_allclose(isothermal_timestep(W, *args, viscosity=nu, smagorinsky_cs=Cs),
                               cfl * (3 / 16) * dx ** 2 / (nu + 2 * max_nu_sgs))
    W_at_rest = W.clone()
    W_at_rest[..., 1] = 0.0
    assert isothermal_timestep(W_at_rest, *args, smagorinsky_cs=Cs) == cfl * dx


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(("integrator", "order"), [("rk2", 2), ("rk3", 3), ("rk4", 4)])
def test_uniform_linear_drag(integrator, order, dtype):
    # test_jathenax.py:78-89, through the fused stage kernel (drag with the incoming
    # primitives, beta dt per stage).
    initial = _dev(np.broadcast_to([2.0, 0.6, -0.4], (6, 8, 3)), dtype)
    dt, rate, cfl = 0.125, 4.0, 0.4
    _, W = recover_isothermal(initial, dfloor=1e-10)
    bound = isothermal_timestep(W, 1.0, 1.0, 1.0, cfl, np.inf, 0.0, 1.0, drag_rate=rate)
    np.testing.assert_allclose(bound, cfl / rate)
    final = _np(isothermal_step(initial, dt, 1.0, 1.0, 1.0, 1e-10, integrator, drag_rate=rate))
    factor = sum((-rate * dt) ** n / factorial(n) for n in range(order + 1))
    tolerance = 32 * _eps(dtype)
    init = _np(initial)
    np.testing.assert_allclose(final[..., 0], init[..., 0], rtol=tolerance)
    np.testing.assert_allclose(final[..., 1:], init[..., 1:] * factor, rtol=tolerance)


@pytest.mark.parametrize("dtype", DTYPES)
def test_kick_normalization(dtype):
    # test_jathenax.py:92-105
    rho = np.asarray([[[1.0], [2.0]]])
    raw = np.asarray([[[2.0, 1.0], [-1.0, 1.0]]])
    dt, dedt = 0.125, 0.04
    for sign in (-1.0, 1.0):
        momentum = sign * 0.125 * rho * raw
        U = np.concatenate((rho, momentum), axis=-1)
        force = _np(normalize_force(_dev(U, dtype), _dev(raw, dtype), dt, dedt)).astype(np.float64)
        delta = rho * force * dt
        energy = np.mean(0.5 * np.sum(delta * (2.0 * momentum + delta) / rho, axis=-1))
        tolerance = 64 * _eps(dtype)
        np.testing.assert_allclose(delta.sum(axis=(0, 1)), 0.0, atol=tolerance)
        np.testing.assert_allclose(energy, dt * dedt, rtol=tolerance)
    np.testing.assert_array_equal(_np(normalize_force(_dev(U, dtype), _dev(raw, dtype), dt, 0.0)),
                                  0.0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("integrator", RK_STAGES)
def test_periodic_step_conserves_mass_and_momentum(integrator, dtype):
    # test_jathenax.py:108-121
    nx, ny = 8, 6
    x, y = np.meshgrid(2.0 * np.pi * np.arange(nx) / nx, 2.0 * np.pi * np.arange(ny) / ny)
    rho = 1.0 + 0.2 * np.cos(x + y)
    U = _dev(np.stack((rho, 0.2 * rho * np.sin(y), 0.3 * rho * np.cos(x)), axis=-1), dtype)
    final = _np(isothermal_step(U, 0.01, 1 / nx, 1 / ny, 1.0, 1e-10, integrator))
    initial = _np(U)
    assert np.max(np.abs(final - initial)) > 1.0e-4
    np.testing.assert_allclose(final.mean(axis=(0, 1)), initial.mean(axis=(0, 1)), rtol=0,
                               atol=16 * _eps(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("rsolver", RIEMANN_SOLVERS)
def test_passive_scalars_share_mass_flux_without_changing_hydro(rsolver, dtype):
    # test_jathenax.py:124-144: viscosity, Smagorinsky, drag and a kick, one fused step.
    x = (np.arange(12) + 0.5) / 12 - 0.5
    U = _kelvin_helmholtz(x, x)
    np.testing.assert_array_equal(U[..., 0], 1.0)
    assert U[..., 3].min() == 0.0 and U[..., 3].max() == 1.0
    U = _dev(np.concatenate((U, 0.25 * U[..., :1]), axis=-1), dtype)
    force = _dev(np.broadcast_to(np.sin(2 * np.pi * x)[None, :, None], (12, 12, 2)), dtype)

    def step(U):
        return _np(isothermal_step(U, 0.01, 1 / 12, 1 / 12, 1.0, 1e-10, "rk3", viscosity=0.001,
                                   smagorinsky_cs=0.2, drag_rate=0.3, raw_force=force,
                                   dedt=0.04, rsolver=rsolver))

    final = step(U)
    tolerance = 32 * _eps(dtype)
    np.testing.assert_allclose(final[..., :3], step(U[..., :3].contiguous()), rtol=0,
                               atol=tolerance)
    np.testing.assert_allclose(final[..., 4], 0.25 * final[..., 0], rtol=0, atol=tolerance)
    np.testing.assert_allclose(final[..., 3:].mean(axis=(0, 1), dtype=np.float64), [0.5, 0.25],
                               rtol=0, atol=tolerance)
    assert np.max(np.abs(final[..., 3] - _np(U)[..., 3])) > 1e-4


@pytest.mark.parametrize("dtype", DTYPES)
def test_wenoz_float32_flat_stencils_at_jump(dtype):
    # test_jathenax.py:147-151 (a (1, 8) array reconstructed along axis 1)
    W = _dev([[0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0]], dtype)
    left, right = wenoz_face_states(W, 1)
    np.testing.assert_allclose(_np(left), np.roll(_np(W), 1, axis=1), atol=2e-7)
    np.testing.assert_allclose(_np(right), _np(W), atol=2e-7)


@pytest.mark.parametrize("dtype", DTYPES)
def test_learned_stress_couples_conservatively(dtype):
    # test_jathenax.py:154-187
    nx, ny = 8, 5
    dx, dy = 1.0 / nx, 1.0 / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    k = 2.0 * np.pi
    vx, vy = np.meshgrid(np.sin(k * x), np.sin(k * y))
    weights = np.asarray([0.5, 2.0, -1.0, 1.5])
    dt = 0.125
    tau = np.stack((weights[0] * vx, weights[1] * vx + weights[2] * vy, weights[3] * vy), axis=-1)
    flux_x, flux_y = stress_face_fluxes(_dev(tau, dtype))
    result = -dt * _np(flux_divergence(flux_x, flux_y, dx, dy)).astype(np.float64)
    dvx_dx = np.cos(k * x)[None, :] * np.sin(k * dx) / dx
    dvy_dy = np.cos(k * y)[:, None] * np.sin(k * dy) / dy
    expected = -dt * np.stack((0.5 * dvx_dx - dvy_dy, 2.0 * dvx_dx + 1.5 * dvy_dy), axis=-1)
    tolerance = 32 * _eps(dtype)
    np.testing.assert_allclose(result, expected, rtol=tolerance, atol=tolerance)
    np.testing.assert_allclose(result.sum(axis=(0, 1)), 0.0, atol=tolerance * nx * ny)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("integrator", RK_STAGES)
def test_sgs_predictor_uses_each_stage_and_conserves_momentum(integrator, dtype):
    # test_jathenax.py:190-219: the traced stages expose the stress added to each flux.
    x = (np.arange(8) + 0.5) / 8 - 0.5
    U = _dev(_kelvin_helmholtz(x, x), dtype)
    scale = torch.tensor([0.0125, 0.025, -0.0375], dtype=dtype, device="cuda")

    def predict_stress(weights, W):
        assert tuple(W.shape) == (8, 8, 3)
        return weights["scale"] * W

    def step(U, weights):
        return isothermal_step(U, 0.01, 1 / 8, 1 / 8, 1.0, 1e-10, integrator, trace=True,
                               predict_stress=predict_stress, sgs_weights=weights)

    final, stages = step(U, {"scale": scale})
    baseline, _ = step(U, {"scale": torch.zeros_like(scale)})
    final, baseline = _np(final), _np(baseline)
    assert np.max(np.abs(final[..., 1:3] - baseline[..., 1:3])) > 1e-4
    tolerance = 16 * _eps(dtype)
    for stage in stages:
        tau = _np(scale) * _np(stage["primitives"])[..., :3]
        for axis, direction, components in ((1, "x", slice(0, 2)), (0, "y", slice(1, 3))):
            gas = _np(isothermal_flux(stage[f"left_{direction}"], stage[f"right_{direction}"],
                                      1.0, axis))
            stress = tau[..., components]
            expected = 0.5 * (stress + np.roll(stress, 1, axis=axis))
            added = _np(stage[f"flux_{direction}"]) - gas
            np.testing.assert_allclose(added[..., 1:3], expected, rtol=0, atol=tolerance)
            np.testing.assert_allclose(added[..., [0, 3]], 0.0, rtol=0, atol=tolerance)
    np.testing.assert_allclose(final.mean(axis=(0, 1)), _np(U).mean(axis=(0, 1)), rtol=0,
                               atol=tolerance)


@pytest.mark.parametrize("dtype", DTYPES)
def test_zero_sgs_predictor_matches_baseline_evolution(dtype):
    # test_jathenax.py:222-234
    x = (np.arange(8) + 0.5) / 8 - 0.5
    U = _dev(_kelvin_helmholtz(x, x), dtype)
    args = (U, 1 / 8, 1 / 8, 1.0, 1e-10, "rk3", 0.4, 0.1, 10)
    baseline = evolve_isothermal(*args)
    coupled = evolve_isothermal(*args, predict_stress=lambda weights, W: torch.zeros_like(W))
    for a, b in zip(baseline, coupled, strict=True):
        if isinstance(a, torch.Tensor):
            np.testing.assert_array_equal(_np(a), _np(b))
        else:
            assert a == b


def test_sgs_predictor_requires_the_full_grid():
    # test_jathenax.py:237-242
    U = _dev(np.ones((6, 8, 3)), torch.float64)
    with pytest.raises(ValueError, match="predict_stress must return"):
        isothermal_step(U, 0.01, 1 / 8, 1 / 6, 1.0, 1e-10, "rk3",
                        predict_stress=lambda weights, W: W[:1])


# ---- validation/periodic_shear.py:21-93: viscosity on a sinusoidal shear ----------------

SHEAR_N, SHEAR_MODE, SHEAR_AMP = 256, 8, 0.5
SHEAR_NU, SHEAR_CS, SHEAR_TLIM = 1.5e-5, 0.2, 5.0
# FP64: the reference's own tolerances.  FP32: rounding of a 4800-step run, stated here.
SHEAR_RTOL = {torch.float64: dict(discrete=2e-9, smag=1e-2, uniform_atol=2e-11),
              torch.float32: dict(discrete=1e-4, smag=1e-2, uniform_atol=2e-5)}


def _shear():
    h = 1.0 / SHEAR_N
    x = (np.arange(SHEAR_N) + 0.5) * h
    U = np.zeros((SHEAR_N, SHEAR_N, 3))
    U[..., 0] = 1.0
    U[..., 2] = SHEAR_AMP * np.sin(2.0 * np.pi * SHEAR_MODE * x)
    return U, h


def _kinetic_energy(U):
    return 0.5 * np.mean(np.sum(U[..., 1:].astype(np.float64) ** 2, axis=-1) / U[..., 0])


@pytest.mark.parametrize("dtype", DTYPES)
def test_periodic_shear_uniform_viscosity_decay(dtype):
    # periodic_shear.py:44-60: the energy at t=5 against the discrete decay rate
    U0, h = _shear()
    k = 2.0 * np.pi * SHEAR_MODE
    final, time, _, cycles = evolve_isothermal(_dev(U0, dtype), h, h, 1.0, 1e-10, "rk3", 0.4,
                                               SHEAR_TLIM, 10_000, viscosity=SHEAR_NU)
    assert time == SHEAR_TLIM
    final = _np(final)
    tol = SHEAR_RTOL[dtype]
    np.testing.assert_allclose(final[..., 0], 1.0, rtol=0, atol=tol["uniform_atol"])
    np.testing.assert_allclose(final[..., 1], 0.0, rtol=0, atol=tol["uniform_atol"])
    np.testing.assert_allclose(final[..., 2].mean(), 0.0, rtol=0, atol=tol["uniform_atol"])
    E0 = _kinetic_energy(U0)
    k_grid_sq = 4.0 * np.sin(k * h / 2.0) ** 2 / h ** 2
    discrete = E0 * np.exp(-2.0 * SHEAR_NU * k_grid_sq * SHEAR_TLIM)
    np.testing.assert_allclose(_kinetic_energy(final), discrete, rtol=tol["discrete"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_periodic_shear_smagorinsky_initial_rate(dtype):
    # periodic_shear.py:61-67: the first step's fractional energy loss rate
    U0, h = _shear()
    k = 2.0 * np.pi * SHEAR_MODE
    E0 = _kinetic_energy(U0)
    first, _, first_dt, _ = evolve_isothermal(_dev(U0, dtype), h, h, 1.0, 1e-10, "rk3", 0.4,
                                              SHEAR_TLIM, 1, smagorinsky_cs=SHEAR_CS)
    loss_rate = (E0 - _kinetic_energy(_np(first))) / (first_dt * E0)
    exact = 16.0 / (3.0 * np.pi) * (SHEAR_CS * h) ** 2 * SHEAR_AMP * k ** 3
    np.testing.assert_allclose(loss_rate, exact, rtol=SHEAR_RTOL[dtype]["smag"])


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(("nu", "Cs"), [(0.01, 0.0), (0.0, 8.0)])
def test_periodic_shear_diffusive_timestep_limits(dtype, nu, Cs):
    # periodic_shear.py:81-97 (rtol 5e-13 in float64; float32 rounds the limits once)
    U0, h = _shear()
    k = 2.0 * np.pi * SHEAR_MODE
    dt = isothermal_timestep(_dev(U0, dtype), h, h, 1.0, 0.4, np.inf, 0.0, SHEAR_TLIM,
                             viscosity=nu, smagorinsky_cs=Cs)
    if Cs > 0.0:
        max_face_strain = 2.0 * SHEAR_AMP * np.sin(k * h / 2.0) / h
        expected = 0.4 * (3.0 / 16.0) * h ** 2 / (2.0 * (Cs * h) ** 2 * max_face_strain)
    else:
        expected = 0.4 * 0.25 * h ** 2 / nu
    rtol = 5e-13 if dtype == torch.float64 else 8 * _eps(dtype)
    np.testing.assert_allclose(dt, expected, rtol=rtol)
    _, _, second_dt, cycles = evolve_isothermal(_dev(U0, dtype), h, h, 1.0, 1e-10, "rk3", 0.4,
                                                SHEAR_TLIM, 2, viscosity=nu, smagorinsky_cs=Cs)
    assert cycles == 2
    assert 0.0 < second_dt <= 2.0 * dt
