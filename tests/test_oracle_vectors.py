"""Pin the NumPy oracle against vectors written from the reference's own functions."""

import numpy as np
import pytest

from oracle import jathenax_np as ox

RECONS = ("dc", "plm", "ppm4", "ppmx", "wenoz")
SOLVERS = ("llf", "hlle", "roe")


def _same(a, b):
    np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("axis", (0, 1))
def test_reconstruction_matches_reference(ref_vectors, recon, axis):
    left, right = ox.reconstruct_faces(ref_vectors["recon_W"], axis, recon, 0.9)
    _same(left, ref_vectors[f"recon_{recon}_axis{axis}_left"])
    _same(right, ref_vectors[f"recon_{recon}_axis{axis}_right"])


def test_wenoz_float32_jump_matches_reference(ref_vectors):
    left, right = ox.wenoz_face_states(ref_vectors["wenoz_f32_W"], 1)
    assert left.dtype == np.float32
    _same(left, ref_vectors["wenoz_f32_left"])
    _same(right, ref_vectors["wenoz_f32_right"])


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("axis", (0, 1))
def test_riemann_matches_reference(ref_vectors, solver, axis):
    f = ox.isothermal_flux(ref_vectors["riemann_left"], ref_vectors["riemann_right"], 1.3,
                           axis, solver)
    _same(f, ref_vectors[f"riemann_{solver}_axis{axis}"])


@pytest.mark.parametrize("axis", (0, 1))
def test_viscous_fluxes_match_reference(ref_vectors, axis):
    W = ref_vectors["visc_W"]
    gx, gy = ox.face_velocity_gradients(W, 0.1, 0.07, axis)
    _same(gx, ref_vectors[f"visc_grad_axis{axis}_dvdx"])
    _same(gy, ref_vectors[f"visc_grad_axis{axis}_dvdy"])
    _same(ox.smagorinsky_viscosity(gx, gy, 0.1, 0.07, 0.17), ref_vectors[f"visc_nusgs_axis{axis}"])
    _same(ox.viscous_momentum_flux(W, 0.1, 0.07, 2e-3, 0.17, axis),
          ref_vectors[f"visc_flux_axis{axis}"])
    _same(ox.viscous_momentum_flux(W, 0.1, 0.07, 2e-3, 0.0, axis),
          ref_vectors[f"visc_flux_uniform_axis{axis}"])


def test_stress_faces_match_reference(ref_vectors):
    fx, fy = ox.stress_face_fluxes(ref_vectors["stress_tau"])
    _same(fx, ref_vectors["stress_fx"])
    _same(fy, ref_vectors["stress_fy"])


def test_timestep_matches_reference(ref_vectors):
    W = ref_vectors["visc_W"]
    for nu, cs_smag, drag, prev, t, tlim, expect in ref_vectors["timestep_cases"]:
        dt = ox.isothermal_timestep(W, 0.1, 0.07, 1.1, 0.4, prev, t, tlim,
                                    viscosity=nu, smagorinsky_cs=cs_smag, drag_rate=drag)
        assert float(dt) == expect


@pytest.mark.parametrize("integrator", ("rk2", "rk3", "rk4"))
@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("solver", SOLVERS)
def test_step_matches_reference(ref_vectors, integrator, recon, solver):
    U = ref_vectors["step_U"]
    tag = f"step_{integrator}_{recon}_{solver}"
    trace = integrator == "rk3" and solver == "roe"
    res = ox.isothermal_step(U, 0.004, 0.1, 0.125, 1.0, 1e-3, integrator, trace=trace,
                             reconstruct=recon, rsolver=solver)
    final, stages = res if trace else (res, ())
    _same(final, ref_vectors[tag])
    for s, st in enumerate(stages):
        for key, val in st.items():
            _same(val, ref_vectors[f"{tag}_stage{s}_{key}"])


@pytest.mark.parametrize("integrator", ("rk2", "rk3", "rk4"))
def test_full_physics_step_matches_reference(ref_vectors, integrator):
    U, raw = ref_vectors["step_U"], ref_vectors["step_raw_force"]
    final, stages = ox.isothermal_step(
        U, 0.004, 0.1, 0.125, 1.0, 1e-3, integrator, trace=True, viscosity=2e-3,
        smagorinsky_cs=0.17, drag_rate=0.7, raw_force=raw, dedt=0.05)
    _same(final, ref_vectors[f"stepfull_{integrator}"])
    _same(stages[-1]["before_kick"], ref_vectors[f"stepfull_{integrator}_before_kick"])
    _same(stages[-1]["acceleration"], ref_vectors[f"stepfull_{integrator}_acceleration"])


def test_normalize_matches_reference(ref_vectors):
    U = ref_vectors["step_U"][..., :3]
    for dedt in (0.0, 0.04):
        _same(ox.normalize_force(U, ref_vectors["step_raw_force"], 0.01, dedt),
              ref_vectors[f"normalize_dedt{dedt}"])


def test_seed_streams_match_reference(ref_vectors):
    u = ox.ran_st(1618)
    _same([next(u) for _ in range(256)], ref_vectors["ran_st_1618"])
    u = ox.ran_st(7)
    _same([next(u) for _ in range(64)], ref_vectors["ran_st_7"])
    g = ox.ran_gaussian_st(1618)
    _same([next(g) for _ in range(800)], ref_vectors["gauss_1618"])


@pytest.mark.parametrize("lo,hi,peak,kb", [(3, 5, 4.0, 5), (1, 3, 2.0, 3)])
def test_modal_table_matches_reference(ref_vectors, lo, hi, peak, kb):
    modes = ox.cartesian_modes(lo, hi, (-kb, kb), (-kb, kb))
    _same(modes, ref_vectors[f"modes_{lo}_{hi}"])
    coef = ox.modal_coefficients(modes, (1.0, 1.0), lo, hi, peak, 1.0, 1618, 0.75, 0.0075, 6)
    _same(coef, ref_vectors[f"coef_{lo}_{hi}"])


def test_render_and_basis_match_reference(ref_vectors):
    modes = ox.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = ox.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 0.7, 99, 0.0, 0.01, 3)
    _same(coef, ref_vectors["coef_1_3_sol07_tcorr0"])
    x = (np.arange(16) + 0.5) / 16 - 0.5
    y = (np.arange(12) + 0.5) / 12 - 0.5
    basis = ox.forcing_basis(modes, x, y, (1.0, 1.0))
    for name, b in zip(("cos_x", "sin_x", "cos_y", "sin_y"), basis):
        _same(b, ref_vectors[f"basis_{name}"])
    _same(ox.render_force(coef[1], basis), ref_vectors["render_force"])


@pytest.mark.parametrize("tlim,nlim,tag", [(0.2, 100000, "tlim"), (10.0, 100, "n100")])
def test_smooth_run_matches_reference(ref_vectors, tlim, nlim, tag):
    U, t, dt, cyc = ox.evolve_isothermal(ref_vectors["run_smooth_U0"], 1 / 64, 1 / 48, 1.0,
                                         1e-10, "rk3", 0.4, tlim, nlim)
    _same(U, ref_vectors[f"run_smooth_{tag}_U"])
    _same([float(t), float(dt), float(cyc)], ref_vectors[f"run_smooth_{tag}_tdc"])


@pytest.mark.parametrize("integ,recon,rs", [("rk2", "wenoz", "hlle"), ("rk4", "ppmx", "llf"),
                                            ("rk3", "plm", "roe"), ("rk3", "dc", "hlle")])
def test_kh_viscous_run_matches_reference(ref_vectors, integ, recon, rs):
    U, t, dt, cyc = ox.evolve_isothermal(
        ref_vectors["run_kh_U0"], 1 / 32, 1 / 32, 1.0, 1e-10, integ, 0.4, 0.05, 1000,
        viscosity=1e-3, smagorinsky_cs=0.17, drag_rate=0.5, reconstruct=recon, rsolver=rs)
    _same(U, ref_vectors[f"run_kh_{integ}_{recon}_{rs}_U"])
    _same([float(t), float(dt), float(cyc)], ref_vectors[f"run_kh_{integ}_{recon}_{rs}_tdc"])


def test_driven_run_matches_reference(ref_vectors):
    nx, ny = 24, 20
    xs = (np.arange(nx) + 0.5) / nx - 0.5
    ys = (np.arange(ny) + 0.5) / ny - 0.5
    modes = ox.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = ox.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, 1618, 0.75, 0.004, 40)
    basis = ox.forcing_basis(modes, xs, ys, (1.0, 1.0))
    U0 = np.zeros((ny, nx, 3))
    U0[..., 0] = 1.0
    U, t, dt, cyc = ox.evolve_isothermal(
        U0, 1 / nx, 1 / ny, 1.0, 1.17549e-38, "rk2", 0.3, 0.1, 12, viscosity=1.5e-3,
        smagorinsky_cs=0.17, drag_rate=0.3333335, forcing=(coef, basis), dt_update=0.004,
        tdriv_start=0.0, dedt=0.5, reconstruct="wenoz", rsolver="hlle")
    _same(U, ref_vectors["run_turb_U"])
    _same([float(t), float(dt), float(cyc)], ref_vectors["run_turb_tdc"])
