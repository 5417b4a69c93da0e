"""Closed forms and invariants that pin the oracle independently of the golden vectors.

The first group restates the reference's own unit tests (tests/test_jathenax.py) against
the oracle; the second adds mathematical facts (polynomial exactness, consistency of the
Riemann fluxes, mirror/transposition symmetry, the discrete shear-decay solution of
validation/periodic_shear.py) chosen so that a dropped term, a wrong sign or index or a
swapped operand fails one of them.
"""

from math import factorial

import numpy as np
import pytest

from oracle import jathenax_np as ox


@pytest.mark.parametrize(("integrator", "order"), [("rk2", 2), ("rk3", 3), ("rk4", 4)])
def test_rk_stability_polynomial(integrator, order):
    # test_jathenax.py:23-36
    U = Us = np.asarray(1.0)
    dt = 0.125
    for coef in ox.RK_STAGES[integrator]:
        U, Us = ox.rk_stage(U, Us, -U, dt, coef)
    expected = sum(dt ** n / factorial(n) for n in range(order + 1))
    np.testing.assert_allclose(U, expected, rtol=32 * np.finfo(float).eps)


def test_rk3_zero_flux_keeps_float32_bits():
    # test_jathenax.py:39-44
    rho = np.linspace(0.75, 1.25, 1024, dtype=np.float32)
    out = ox.rk_stage(rho, rho, np.zeros_like(rho), 0.01, ox.RK_STAGES["rk3"][-1])[0]
    np.testing.assert_array_equal(out, rho)


def test_density_floor():
    # test_jathenax.py:47-50
    U, W = ox.recover_isothermal(np.asarray([[[0.125, 0.5, -0.25]]]), 0.25)
    np.testing.assert_array_equal(U, [[[0.25, 0.5, -0.25]]])
    np.testing.assert_array_equal(W, [[[0.25, 2.0, -1.0]]])


def test_timestep_limits():
    # test_jathenax.py:53-75
    W = np.asarray([[[1.0, 1.0, 2.0], [1.0, 3.0, 4.0]]])
    assert ox.isothermal_timestep(W, 1.0, 2.0, 1.0, 0.5, np.inf, 0.0, 1.0) == 0.125
    assert ox.isothermal_timestep(W, 1.0, 2.0, 1.0, 0.5, 0.03125, 0.0, 1.0) == 0.0625
    assert ox.isothermal_timestep(W, 1.0, 2.0, 1.0, 0.5, np.inf, 0.9375, 1.0) == 0.0625
    W = np.asarray([[[1.0, 0.0, 0.0]] * 2, [[1.0, 1.0, 0.0]] * 2])
    dx, dy, cfl, nu, Cs = 0.25, 0.5, 0.5, 0.5, 2.0
    args = (dx, dy, 1.0, cfl, np.inf, 0.0, 1.0)
    assert ox.isothermal_timestep(W, *args, viscosity=nu) == cfl * 0.25 * dx ** 2 / nu
    nu_sgs = Cs ** 2 * dx * dy / dy
    np.testing.assert_allclose(ox.isothermal_timestep(W, *args, viscosity=nu, smagorinsky_cs=Cs),
                               cfl * (3 / 16) * dx ** 2 / (nu + 2 * nu_sgs))
    W[..., 1] = 0.0
    assert ox.isothermal_timestep(W, *args, smagorinsky_cs=Cs) == cfl * dx


@pytest.mark.parametrize(("integrator", "order"), [("rk2", 2), ("rk3", 3), ("rk4", 4)])
def test_uniform_linear_drag(integrator, order):
    # test_jathenax.py:78-89
    U0 = np.broadcast_to(np.asarray([2.0, 0.6, -0.4]), (6, 8, 3)).copy()
    dt, rate = 0.125, 4.0
    _, W = ox.recover_isothermal(U0, 1e-10)
    np.testing.assert_allclose(
        ox.isothermal_timestep(W, 1.0, 1.0, 1.0, 0.4, np.inf, 0.0, 1.0, drag_rate=rate), 0.4 / rate)
    final = ox.isothermal_step(U0, dt, 1.0, 1.0, 1.0, 1e-10, integrator, drag_rate=rate)
    factor = sum((-rate * dt) ** n / factorial(n) for n in range(order + 1))
    tol = 32 * np.finfo(float).eps
    np.testing.assert_allclose(final[..., 0], U0[..., 0], rtol=tol)
    np.testing.assert_allclose(final[..., 1:], U0[..., 1:] * factor, rtol=tol)


def test_kick_zero_net_momentum_and_exact_power():
    # test_jathenax.py:92-105
    rho = np.asarray([[[1.0], [2.0]]])
    raw = np.asarray([[[2.0, 1.0], [-1.0, 1.0]]])
    dt, dedt = 0.125, 0.04
    for sign in (-1.0, 1.0):
        mom = sign * 0.125 * rho * raw
        U = np.concatenate((rho, mom), axis=-1)
        f = ox.normalize_force(U, raw, dt, dedt)
        delta = rho * f * dt
        energy = np.mean(0.5 * np.sum(delta * (2.0 * mom + delta) / rho, axis=-1))
        tol = 64 * np.finfo(float).eps
        np.testing.assert_allclose(delta.sum(axis=(0, 1)), 0.0, atol=tol)
        np.testing.assert_allclose(energy, dt * dedt, rtol=tol)
    np.testing.assert_array_equal(ox.normalize_force(U, raw, dt, 0.0), 0.0)


@pytest.mark.parametrize("integrator", ("rk2", "rk3", "rk4"))
def test_step_conserves_mass_and_momentum(integrator):
    # test_jathenax.py:108-121
    nx, ny = 8, 6
    x, y = np.meshgrid(2 * np.pi * np.arange(nx) / nx, 2 * np.pi * np.arange(ny) / ny)
    rho = 1.0 + 0.2 * np.cos(x + y)
    U = np.stack((rho, 0.2 * rho * np.sin(y), 0.3 * rho * np.cos(x)), axis=-1)
    final = ox.isothermal_step(U, 0.01, 1 / nx, 1 / ny, 1.0, 1e-10, integrator)
    assert np.max(np.abs(final - U)) > 1e-4
    np.testing.assert_allclose(final.mean(axis=(0, 1)), U.mean(axis=(0, 1)), rtol=0,
                               atol=16 * np.finfo(float).eps)


def _kh_reference(x, y):
    """Reference KH layers with tracer (initial_conditions.py:6-15), restated."""
    X, Y = np.meshgrid(x, y)
    d = np.abs(Y) - 0.25
    tr = np.clip(-d / 0.025, -1.0, 1.0)
    s = 0.5 + tr * (0.75 - 0.25 * tr ** 2)
    vy = 0.02 * np.sin(4 * np.pi * X) * np.exp(-(d / 0.05) ** 2)
    return np.stack((np.ones_like(s), s - 0.5, vy, s), axis=-1)


@pytest.mark.parametrize("rsolver", ("llf", "hlle", "roe"))
def test_passive_scalars_ride_the_mass_flux(rsolver):
    # test_jathenax.py:124-144
    x = (np.arange(12) + 0.5) / 12 - 0.5
    U = _kh_reference(x, x)
    U = np.concatenate((U, 0.25 * U[..., :1]), axis=-1)
    force = np.broadcast_to(np.sin(2 * np.pi * x)[None, :, None], (12, 12, 2)).copy()

    def step(V):
        return ox.isothermal_step(V, 0.01, 1 / 12, 1 / 12, 1.0, 1e-10, "rk3", viscosity=0.001,
                                  smagorinsky_cs=0.2, drag_rate=0.3, raw_force=force,
                                  dedt=0.04, rsolver=rsolver)
    final = step(U)
    tol = 32 * np.finfo(float).eps
    np.testing.assert_allclose(final[..., :3], step(U[..., :3]), rtol=0, atol=tol)
    np.testing.assert_allclose(final[..., 4], 0.25 * final[..., 0], rtol=0, atol=tol)
    np.testing.assert_allclose(final[..., 3:].mean(axis=(0, 1)), [0.5, 0.25], rtol=0, atol=tol)
    assert np.max(np.abs(final[..., 3] - U[..., 3])) > 1e-4


def test_learned_stress_divergence_of_sines():
    # test_jathenax.py:154-187
    nx, ny = 8, 5
    dx, dy = 1 / nx, 1 / ny
    x, y = (np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy
    k = 2 * np.pi
    vx, vy = np.meshgrid(np.sin(k * x), np.sin(k * y))
    w = [0.5, 2.0, -1.0, 1.5]
    tau = np.stack((w[0] * vx, w[1] * vx + w[2] * vy, w[3] * vy), axis=-1)
    fx, fy = ox.stress_face_fluxes(tau)
    res = -0.125 * ox.flux_divergence(fx, fy, dx, dy)
    dvx = np.cos(k * x)[None, :] * np.sin(k * dx) / dx
    dvy = np.cos(k * y)[:, None] * np.sin(k * dy) / dy
    expect = -0.125 * np.stack((0.5 * dvx - dvy, 2.0 * dvx + 1.5 * dvy), axis=-1)
    tol = 32 * np.finfo(float).eps
    np.testing.assert_allclose(res, expect, rtol=tol, atol=tol)
    with pytest.raises(ValueError):
        ox.stress_face_fluxes(tau[..., :2])


# ---- mathematical pins -------------------------------------------------------------------


def _cell_avg_cubic(n, c):
    """Exact cell averages of a cubic on unit cells 0..n-1 (not periodic: interior only)."""
    xs = np.arange(n + 1, dtype=float)
    prim = c[0] * xs + c[1] * xs ** 2 / 2 + c[2] * xs ** 3 / 3 + c[3] * xs ** 4 / 4
    return np.diff(prim)


@pytest.mark.parametrize("recon", ("ppm4", "ppmx"))
def test_parabolic_edges_exact_for_monotone_cubics(recon):
    # Fourth-order edge interpolation reproduces point values of a cubic from its cell
    # averages; on gently monotone data no limiter branch fires.
    c = (0.3, 0.02, 0.001, 0.0001)
    avg = _cell_avg_cubic(16, c)
    W = np.tile(avg[None, :, None], (1, 1, 3))
    left, right = ox.RECONSTRUCTIONS[recon](W, 1)
    xs = np.arange(16.0)
    exact = c[0] + c[1] * xs + c[2] * xs ** 2 + c[3] * xs ** 3
    np.testing.assert_allclose(right[0, 3:13, 0], exact[3:13], rtol=1e-13)
    np.testing.assert_allclose(left[0, 3:13, 0], exact[3:13], rtol=1e-13)


def test_wenoz_fifth_order_on_smooth_data():
    # WENO-Z edge errors fall at fifth order on a smooth periodic profile.
    errs = []
    for n in (32, 64):
        xe = np.arange(n + 1) / n
        avg = (np.cos(2 * np.pi * xe[:-1]) - np.cos(2 * np.pi * xe[1:])) / (2 * np.pi) * n
        W = np.tile((2.0 + avg)[None, :, None], (1, 1, 3))
        _, right = ox.wenoz_face_states(W, 1)
        errs.append(np.max(np.abs(right[0, :, 0] - (2.0 + np.sin(2 * np.pi * xe[:-1])))))
    assert np.log2(errs[0] / errs[1]) > 4.5


@pytest.mark.parametrize("recon", ("dc", "plm", "ppm4", "ppmx", "wenoz"))
def test_reconstruction_mirror_symmetry(recon):
    # Reversing the axis swaps left and right states at the mirrored faces.
    rng = np.random.default_rng(3)
    W = 1.0 + 0.3 * rng.standard_normal((1, 24, 3))
    L, R = ox.RECONSTRUCTIONS[recon](W, 1)
    Lm, Rm = ox.RECONSTRUCTIONS[recon](W[:, ::-1], 1)
    # Face i (low face of cell i) maps to the face between mirrored cells n-1-i and n-i.
    np.testing.assert_allclose(np.roll(Lm[:, ::-1], 1, axis=1), R, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(np.roll(Rm[:, ::-1], 1, axis=1), L, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("recon", ("dc", "plm", "ppm4", "ppmx", "wenoz"))
def test_reconstruction_preserves_linear_profiles_where_monotone(recon):
    # A linear ramp (away from the periodic seam) is reproduced exactly at interior faces.
    W = np.tile((0.5 + 0.1 * np.arange(20.0))[None, :, None], (1, 1, 3))
    L, R = ox.RECONSTRUCTIONS[recon](W, 1)
    face = 0.5 + 0.1 * (np.arange(20.0) - 0.5)
    if recon == "dc":
        np.testing.assert_allclose(L[0, 4:16, 0], W[0, 3:15, 0])
    else:
        np.testing.assert_allclose(L[0, 4:16, 0], face[4:16], rtol=1e-12)
        np.testing.assert_allclose(R[0, 4:16, 0], face[4:16], rtol=1e-12)


def _phys_flux(rho, vn, vt, cs):
    return np.stack((rho * vn, rho * vn * vn + cs * cs * rho, rho * vn * vt), axis=-1)


@pytest.mark.parametrize("solver", ("llf", "hlle", "roe"))
def test_riemann_consistency(solver):
    rng = np.random.default_rng(5)
    rho = np.exp(rng.uniform(-2, 2, 64))
    vn, vt = rng.uniform(-3, 3, (2, 64))
    f = ox.RIEMANN_SOLVERS[solver](rho, rho, vn, vn, vt, vt, 1.3)
    np.testing.assert_allclose(np.stack(f, axis=-1), _phys_flux(rho, vn, vt, 1.3),
                               rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("solver", ("hlle", "roe"))
def test_riemann_supersonic_upwinding(solver):
    rng = np.random.default_rng(6)
    rl, rr = np.exp(rng.uniform(-1, 1, (2, 32)))
    vt_l, vt_r = rng.uniform(-1, 1, (2, 32))
    f = np.stack(ox.RIEMANN_SOLVERS[solver](rl, rr, np.full(32, 3.0), np.full(32, 3.2),
                                            vt_l, vt_r, 1.0), axis=-1)
    np.testing.assert_allclose(f, _phys_flux(rl, np.full(32, 3.0), vt_l, 1.0), rtol=1e-13)


@pytest.mark.parametrize("recon", ("ppm4", "wenoz"))
@pytest.mark.parametrize("solver", ("llf", "hlle", "roe"))
def test_step_transposition_invariance(recon, solver):
    # Transposing the grid and swapping vx/vy (with dx<->dy) transposes the result exactly.
    rng = np.random.default_rng(8)
    U = np.empty((10, 12, 3))
    U[..., 0] = 1.0 + 0.2 * rng.random((10, 12))
    U[..., 1:] = 0.3 * rng.standard_normal((10, 12, 2))
    a = ox.isothermal_step(U, 0.003, 0.1, 0.07, 1.0, 1e-10, "rk3", viscosity=1e-3,
                           smagorinsky_cs=0.1, reconstruct=recon, rsolver=solver)
    Ut = U.transpose(1, 0, 2)[..., [0, 2, 1]].copy()
    b = ox.isothermal_step(Ut, 0.003, 0.07, 0.1, 1.0, 1e-10, "rk3", viscosity=1e-3,
                           smagorinsky_cs=0.1, reconstruct=recon, rsolver=solver)
    np.testing.assert_allclose(b.transpose(1, 0, 2)[..., [0, 2, 1]], a, rtol=1e-13, atol=1e-15)


def test_uniform_viscosity_shear_decay_is_discrete_laplacian():
    # validation/periodic_shear.py:22-56 on a 32^2 grid: with vx = 0 and vy(x) the hydro
    # fluxes vanish and the y momentum obeys the centred discrete Laplacian, so the energy
    # ratio is the product of the RK3 stability polynomials over the steps actually taken.
    n, mode, amp, nu = 32, 2, 0.5, 2e-3
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    k = 2 * np.pi * mode
    U = np.zeros((n, n, 3))
    U[..., 0] = 1.0
    U[..., 2] = amp * np.sin(k * x)
    e0 = 0.5 * np.mean(U[..., 2] ** 2)
    # The x-face stress on y momentum is u_y + v_x = v_x: coefficient 1 (not 4/3).
    lam = -nu * 4.0 * np.sin(k * h / 2) ** 2 / h ** 2
    t, prev, amp_ratio = 0.0, np.inf, 1.0
    Ucur = U
    for _ in range(20):
        _, W = ox.recover_isothermal(Ucur, 1e-10)
        dt = float(ox.isothermal_timestep(W, h, h, 1.0, 0.4, prev, t, 1.0, viscosity=nu))
        Ucur = ox.isothermal_step(Ucur, dt, h, h, 1.0, 1e-10, "rk3", viscosity=nu)
        z = lam * dt
        amp_ratio *= 1 + z + z * z / 2 + z ** 3 / 6
        t, prev = t + dt, dt
    e = 0.5 * np.mean(Ucur[..., 2] ** 2)
    np.testing.assert_allclose(e / e0, amp_ratio ** 2, rtol=1e-12)
    np.testing.assert_allclose(Ucur[..., 0], 1.0, atol=1e-14)
    np.testing.assert_allclose(Ucur[..., 1], 0.0, atol=1e-14)


def test_smagorinsky_initial_loss_rate():
    # validation/periodic_shear.py:57-66 (continuum mean of |cos|^3 = 4/(3 pi)), 1% on 256^2.
    n, mode, amp, Cs = 256, 8, 0.5, 0.2
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    k = 2 * np.pi * mode
    U = np.zeros((n, n, 3))
    U[..., 0] = 1.0
    U[..., 2] = amp * np.sin(k * x)
    e0 = 0.5 * np.mean(U[..., 2] ** 2)
    _, W = ox.recover_isothermal(U, 1e-10)
    dt = float(ox.isothermal_timestep(W, h, h, 1.0, 0.4, np.inf, 0.0, 5.0, smagorinsky_cs=Cs))
    U1 = ox.isothermal_step(U, dt, h, h, 1.0, 1e-10, "rk3", smagorinsky_cs=Cs)
    rate = (e0 - 0.5 * np.mean(U1[..., 2] ** 2)) / (dt * e0)
    np.testing.assert_allclose(rate, 16 / (3 * np.pi) * (Cs * h) ** 2 * amp * k ** 3, rtol=0.01)


def test_diffusive_timestep_limits():
    # validation/periodic_shear.py:81-93
    n, mode, amp = 64, 8, 0.5
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    k = 2 * np.pi * mode
    W = np.zeros((n, n, 3))
    W[..., 0] = 1.0
    W[..., 2] = amp * np.sin(k * x)
    dt = ox.isothermal_timestep(W, h, h, 1.0, 0.4, np.inf, 0.0, 5.0, viscosity=0.01)
    np.testing.assert_allclose(dt, 0.4 * 0.25 * h ** 2 / 0.01, rtol=5e-13)
    dt = ox.isothermal_timestep(W, h, h, 1.0, 0.4, np.inf, 0.0, 5.0, smagorinsky_cs=8.0)
    smax = 2.0 * amp * np.sin(k * h / 2.0) / h
    nu_max = (8.0 * h) ** 2 * smax
    np.testing.assert_allclose(dt, 0.4 * (3 / 16) * h ** 2 / (2 * nu_max), rtol=5e-13)


def test_ran_st_is_lecuyer_bays_durham():
    # Independent restatement from the published algorithm (L'Ecuyer 1988 combined with a
    # Bays-Durham shuffle, Numerical Recipes 'ran2' constants): long-period stream and
    # uniform moments.
    u = ox.ran_st(1618)
    draws = np.asarray([next(u) for _ in range(20000)])
    assert 0.0 < draws.min() and draws.max() < 1.0
    assert abs(draws.mean() - 0.5) < 0.01 and abs(draws.var() - 1 / 12) < 0.003
    g = ox.ran_gaussian_st(1618)
    gd = np.asarray([next(g) for _ in range(20000)])
    assert abs(gd.mean()) < 0.03 and abs(gd.std() - 1.0) < 0.03
