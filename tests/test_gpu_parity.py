"""Parity of the HIP path with the oracle (float64: relative L-infinity <= 1e-12 unless a
tighter bound is stated; the expression order is the reference's, so most cases agree to
the last bit).  Every call goes through libjathenax_amd.so."""

import numpy as np
import pytest
import torch

from oracle import jathenax_np as ox

pytestmark = pytest.mark.gpu

RECONS = ("dc", "plm", "ppm4", "ppmx", "wenoz")
SOLVERS = ("llf", "hlle", "roe")


def _np(t):
    return t.detach().cpu().numpy()


def _rel_inf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def jx():
    import jathenax_amd
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    return jathenax_amd


@pytest.mark.parametrize("integrator", ("rk2", "rk3", "rk4"))
@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("solver", SOLVERS)
def test_step_every_operator(jx, ref_vectors, integrator, recon, solver):
    U = ref_vectors["step_U"]
    got = jx.isothermal_step(torch.tensor(U, device="cuda"), 0.004, 0.1, 0.125, 1.0, 1e-3,
                             integrator, reconstruct=recon, rsolver=solver)
    assert _rel_inf(_np(got), ref_vectors[f"step_{integrator}_{recon}_{solver}"]) <= 1e-13


@pytest.mark.parametrize("integrator", ("rk2", "rk3", "rk4"))
def test_step_viscous_drag_kick(jx, ref_vectors, integrator):
    U, raw = ref_vectors["step_U"], ref_vectors["step_raw_force"]
    got = jx.isothermal_step(torch.tensor(U, device="cuda"), 0.004, 0.1, 0.125, 1.0, 1e-3,
                             integrator, viscosity=2e-3, smagorinsky_cs=0.17, drag_rate=0.7,
                             raw_force=torch.tensor(raw, device="cuda"), dedt=0.05)
    assert _rel_inf(_np(got), ref_vectors[f"stepfull_{integrator}"]) <= 1e-13


@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("nvar", (3, 5))
def test_step_multiblock_grid(jx, recon, solver, nvar):
    # 300 x 173: several strips in x and y, ragged last tiles, viscosity + Smagorinsky.
    rng = np.random.default_rng(11)
    ny, nx = 173, 300
    U = np.empty((ny, nx, nvar))
    U[..., 0] = 1.0 + 0.3 * rng.random((ny, nx))
    U[..., 1:3] = 0.4 * rng.standard_normal((ny, nx, 2))
    U[..., 3:] = rng.random((ny, nx, nvar - 3)) * U[..., :1]
    kw = dict(viscosity=1e-3, smagorinsky_cs=0.15, drag_rate=0.2, reconstruct=recon,
              rsolver=solver)
    want = ox.isothermal_step(U, 1e-3, 1 / nx, 1 / ny, 1.0, 1e-10, "rk3", **kw)
    got = jx.isothermal_step(torch.tensor(U, device="cuda"), 1e-3, 1 / nx, 1 / ny, 1.0, 1e-10,
                             "rk3", **kw)
    assert _rel_inf(_np(got), want) <= 1e-13


@pytest.mark.parametrize("bx,rows", [(64, 1), (64, 7), (128, 16), (192, 5), (256, 64)])
def test_tile_shapes_are_bitwise_equal(jx, bx, rows):
    from jathenax_amd.device import context_for
    rng = np.random.default_rng(5)
    U = np.empty((97, 131, 4))
    U[..., 0] = 1.0 + 0.3 * rng.random((97, 131))
    U[..., 1:] = 0.3 * rng.standard_normal((97, 131, 3))
    Ud = torch.tensor(U, device="cuda")
    ctx = context_for(Ud)
    kw = dict(viscosity=1e-3, smagorinsky_cs=0.1, reconstruct="wenoz", rsolver="roe")
    ctx.set_tiles(0, 0)
    base = _np(jx.isothermal_step(Ud, 1e-3, 0.01, 0.01, 1.0, 1e-10, "rk3", **kw))
    ctx.set_tiles(bx, rows)
    try:
        got = _np(jx.isothermal_step(Ud, 1e-3, 0.01, 0.01, 1.0, 1e-10, "rk3", **kw))
    finally:
        ctx.set_tiles(0, 0)
    np.testing.assert_array_equal(got, base)


@pytest.mark.parametrize("recon,solver", [("plm", "hlle"), ("ppm4", "roe"), ("ppmx", "llf"),
                                          ("wenoz", "roe")])
def test_fp32_tile_shapes_are_bitwise_equal(jx, recon, solver):
    """FP32 fuses products into adds only within one source expression, so the unrolled row
    loop computes every row with the same operations wherever a tile starts."""
    from jathenax_amd.device import context_for
    rng = np.random.default_rng(5)
    U = np.empty((97, 131, 4))
    U[..., 0] = 1.0 + 0.3 * rng.random((97, 131))
    U[..., 1:] = 0.3 * rng.standard_normal((97, 131, 3))
    Ud = torch.tensor(U, dtype=torch.float32, device="cuda")
    ctx = context_for(Ud)
    kw = dict(viscosity=1e-3, smagorinsky_cs=0.1, reconstruct=recon, rsolver=solver)
    ctx.set_tiles(0, 0)
    base = _np(jx.isothermal_step(Ud, 1e-3, 0.01, 0.01, 1.0, 1e-10, "rk3", **kw))
    try:
        for bx, rows in ((64, 1), (64, 7), (128, 13), (128, 16), (192, 5), (256, 64)):
            ctx.set_tiles(bx, rows)
            got = _np(jx.isothermal_step(Ud, 1e-3, 0.01, 0.01, 1.0, 1e-10, "rk3", **kw))
            np.testing.assert_array_equal(got, base, err_msg=str((bx, rows)))
    finally:
        ctx.set_tiles(0, 0)


@pytest.mark.parametrize("recon,solver", [("ppm4", "roe"), ("plm", "llf"), ("wenoz", "hlle")])
def test_fp32_inviscid_runs_tile_shapes_are_bitwise_equal(jx, recon, solver):
    """The inviscid kernels (the 4096^2 benchmark's path) over several steps: the first-stage
    instantiation, the later stages and the CFL epilogue give the same bits for every tile,
    including bands whose last one ends at row ny (110 = 2 x 55) and a ragged one (28)."""
    from jathenax_amd.device import context_for
    rng = np.random.default_rng(8)
    ny, nx = 110, 131
    U = np.empty((ny, nx, 3))
    U[..., 0] = 1.0 + 0.3 * rng.random((ny, nx))
    U[..., 1:] = 0.3 * rng.standard_normal((ny, nx, 2))
    Ud = torch.tensor(U, dtype=torch.float32, device="cuda")
    ctx = context_for(Ud)
    args = (1 / nx, 1 / ny, 1.0, 1e-6, "rk3", 0.4, 10.0, 4)
    kw = dict(reconstruct=recon, rsolver=solver)
    ctx.set_tiles(0, 0)
    base = jx.evolve_isothermal(Ud, *args, **kw)
    try:
        for bx, rows in ((64, 1), (64, 7), (128, 13), (192, 5), (256, 28), (256, 55), (256, 64)):
            ctx.set_tiles(bx, rows)
            got = jx.evolve_isothermal(Ud, *args, **kw)
            np.testing.assert_array_equal(_np(got[0]), _np(base[0]), err_msg=str((bx, rows)))
            assert got[1:] == base[1:], (bx, rows)
    finally:
        ctx.set_tiles(0, 0)


@pytest.mark.parametrize("bx,rows", [(256, 55), (256, 28), (128, 55)])
def test_fp64_inviscid_step_at_benchmark_tiles(jx, bx, rows):
    """FP64 PPM4/Roe RK3 step against the oracle at the benchmark's tile shapes; the top band
    ends at row ny (110 = 2 x 55), where the saved register is read one iteration ahead."""
    from jathenax_amd.device import context_for
    rng = np.random.default_rng(9)
    ny, nx = 110, 300
    U = np.empty((ny, nx, 3))
    U[..., 0] = 1.0 + 0.3 * rng.random((ny, nx))
    U[..., 1:] = 0.4 * rng.standard_normal((ny, nx, 2))
    want = ox.isothermal_step(U, 1e-3, 1 / nx, 1 / ny, 1.0, 1e-10, "rk3")
    Ud = torch.tensor(U, device="cuda")
    ctx = context_for(Ud)
    ctx.set_tiles(bx, rows)
    try:
        got = jx.isothermal_step(Ud, 1e-3, 1 / nx, 1 / ny, 1.0, 1e-10, "rk3")
    finally:
        ctx.set_tiles(0, 0)
    assert _rel_inf(_np(got), want) <= 1e-13


@pytest.mark.parametrize("tlim,nlim,tag", [(0.2, 100000, "tlim"), (10.0, 100, "n100")])
def test_smooth_run_100_cycles(jx, ref_vectors, tlim, nlim, tag):
    U0 = ref_vectors["run_smooth_U0"]
    U, t, dt, cyc = jx.evolve_isothermal(torch.tensor(U0, device="cuda"), 1 / 64, 1 / 48, 1.0,
                                         1e-10, "rk3", 0.4, tlim, nlim)
    want = ref_vectors[f"run_smooth_{tag}_U"]
    assert _rel_inf(_np(U), want) <= 1e-12
    t_ref, dt_ref, c_ref = ref_vectors[f"run_smooth_{tag}_tdc"]
    assert cyc == int(c_ref)
    assert abs(t - t_ref) <= 1e-12 * abs(t_ref) and abs(dt - dt_ref) <= 1e-12 * dt_ref
    # conservation of mass and momentum to round-off (FP64 box sums)
    s0, s1 = ox.conserved_sums(U0), jx.conserved_sums(U)
    assert np.max(np.abs(s1 - s0)) <= 1e-12 * U0.size


@pytest.mark.parametrize("integ,recon,rs", [("rk2", "wenoz", "hlle"), ("rk4", "ppmx", "llf"),
                                            ("rk3", "plm", "roe"), ("rk3", "dc", "hlle")])
def test_kh_viscous_runs(jx, ref_vectors, integ, recon, rs):
    U, t, dt, cyc = jx.evolve_isothermal(
        torch.tensor(ref_vectors["run_kh_U0"], device="cuda"), 1 / 32, 1 / 32, 1.0, 1e-10, integ,
        0.4, 0.05, 1000, viscosity=1e-3, smagorinsky_cs=0.17, drag_rate=0.5, reconstruct=recon,
        rsolver=rs)
    assert _rel_inf(_np(U), ref_vectors[f"run_kh_{integ}_{recon}_{rs}_U"]) <= 1e-12
    np.testing.assert_allclose([t, dt, cyc], ref_vectors[f"run_kh_{integ}_{recon}_{rs}_tdc"],
                               rtol=1e-12)


def test_driven_run(jx, ref_vectors):
    from jathenax_amd import forcing as jf
    nx, ny = 24, 20
    xs = (np.arange(nx) + 0.5) / nx - 0.5
    ys = (np.arange(ny) + 0.5) / ny - 0.5
    modes = jf.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, 1618, 0.75, 0.004, 40)
    basis = jf.forcing_basis(modes, xs, ys, (1.0, 1.0))
    U0 = np.zeros((ny, nx, 3))
    U0[..., 0] = 1.0
    U, t, dt, cyc = jx.evolve_isothermal(
        torch.tensor(U0, device="cuda"), 1 / nx, 1 / ny, 1.0, 1.17549e-38, "rk2", 0.3, 0.1, 12,
        viscosity=1.5e-3, smagorinsky_cs=0.17, drag_rate=0.3333335, forcing=(coef, basis),
        dt_update=0.004, tdriv_start=0.0, dedt=0.5, reconstruct="wenoz", rsolver="hlle")
    # The kick sums are reduced in a different order than NumPy's: round-off level only.
    assert _rel_inf(_np(U), ref_vectors["run_turb_U"]) <= 1e-12
    np.testing.assert_allclose([t, dt, cyc], ref_vectors["run_turb_tdc"], rtol=1e-12)


def test_standalone_operators(jx, ref_vectors):
    W = ref_vectors["visc_W"]
    for nu, cs_smag, drag, prev, t, tlim, expect in ref_vectors["timestep_cases"]:
        dt = jx.isothermal_timestep(torch.tensor(W, device="cuda"), 0.1, 0.07, 1.1, 0.4, prev, t,
                                    tlim, viscosity=nu, smagorinsky_cs=cs_smag, drag_rate=drag)
        assert abs(dt - expect) <= 1e-15 * expect
    U = ref_vectors["step_U"]
    Uf, Wd = jx.recover_isothermal(torch.tensor(U, device="cuda"), 1e-3)
    Uo, Wo = ox.recover_isothermal(U, 1e-3)
    np.testing.assert_array_equal(_np(Uf), Uo)
    np.testing.assert_array_equal(_np(Wd), Wo)
    tau = ref_vectors["stress_tau"]
    fx, fy = jx.stress_face_fluxes(torch.tensor(tau, device="cuda"))
    np.testing.assert_array_equal(_np(fx), ref_vectors["stress_fx"])
    np.testing.assert_array_equal(_np(fy), ref_vectors["stress_fy"])
    div = jx.flux_divergence(torch.tensor(U, device="cuda"), torch.tensor(U[::-1].copy(),
                                                                          device="cuda"), 0.1, 0.2)
    np.testing.assert_array_equal(_np(div), ox.flux_divergence(U, U[::-1].copy(), 0.1, 0.2))
    for integrator in ("rk3", "rk4"):
        for coef in jx.RK_STAGES[integrator]:
            a, b = (torch.tensor(x, device="cuda") for x in (U, U * 0.9))
            d = torch.tensor(U * 0.1, device="cuda")
            g_u, g_s = jx.rk_stage(a, b, d, 0.01, coef)
            o_u, o_s = ox.rk_stage(U, U * 0.9, U * 0.1, 0.01, coef)
            np.testing.assert_array_equal(_np(g_u), o_u)
            np.testing.assert_array_equal(_np(g_s), o_s)


def test_forcing_operators(jx, ref_vectors):
    from jathenax_amd import forcing as jf
    modes = jf.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 0.7, 99, 0.0, 0.01, 3)
    x = (np.arange(16) + 0.5) / 16 - 0.5
    y = (np.arange(12) + 0.5) / 12 - 0.5
    basis = jf.forcing_basis(modes, x, y, (1.0, 1.0))
    for name, b in zip(("cos_x", "sin_x", "cos_y", "sin_y"), basis):
        np.testing.assert_array_equal(_np(b), ref_vectors[f"basis_{name}"])
    raw = jf.render_force(torch.tensor(coef[1], device="cuda"), basis)
    np.testing.assert_array_equal(_np(raw), ref_vectors["render_force"])
    U = ref_vectors["step_U"][..., :3].copy()
    for dedt in (0.0, 0.04):
        acc = jf.normalize_force(torch.tensor(U, device="cuda"),
                                 torch.tensor(ref_vectors["step_raw_force"], device="cuda"), 0.01,
                                 dedt)
        want = ref_vectors[f"normalize_dedt{dedt}"]
        assert np.max(np.abs(_np(acc) - want)) <= 1e-13 * max(np.max(np.abs(want)), 1.0)


def test_fp32_render_by_mode_groups(jx, ref_vectors):
    """FP32 renders sum modes in groups that share their x basis rows (equal, or with sx
    negated).  The Cartesian band (28 modes, 4 groups) must give the reference render to
    float32 rounding, and so must bases where no rows, or only some, are shared."""
    from jathenax_amd import forcing as jf
    modes = jf.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 0.7, 99, 0.0, 0.01, 3)
    x = ((np.arange(16) + 0.5) / 16 - 0.5).astype(np.float32)
    y = ((np.arange(12) + 0.5) / 12 - 0.5).astype(np.float32)
    basis = jf.forcing_basis(modes, x, y, (1.0, 1.0))
    raw = _np(jf.render_force(torch.tensor(coef[1], dtype=torch.float32, device="cuda"), basis))
    want = ref_vectors["render_force"]
    assert np.max(np.abs(raw - want)) <= 1e-5 * np.max(np.abs(want))

    rng = np.random.default_rng(3)
    M, nx, ny = 9, 300, 37
    cx = rng.uniform(-1, 1, (M, nx)).astype(np.float32)
    sx = rng.uniform(-1, 1, (M, nx)).astype(np.float32)
    cy = rng.uniform(-1, 1, (M, ny)).astype(np.float32)
    sy = rng.uniform(-1, 1, (M, ny)).astype(np.float32)
    amp = rng.uniform(-1, 1, (M, 2, 2)).astype(np.float32)
    for shared in ((), ((0, 3), (0, 5), (2, 8)), ((1, 4),)):
        cxs, sxs = cx.copy(), sx.copy()
        for a, b in shared:            # (0,3), (0,5), (2,8): equal rows; (1,4): sx negated
            cxs[b] = cxs[a]
            sxs[b] = -sxs[a] if (a, b) == (1, 4) else sxs[a]
        want = ox.render_force(amp.astype(np.float64),
                               tuple(t.astype(np.float64) for t in (cxs, sxs, cy, sy)))
        got = _np(jf.render_force(torch.tensor(amp, device="cuda"),
                                  tuple(torch.tensor(t, device="cuda")
                                        for t in (cxs, sxs, cy, sy))))
        # 8 products per mode, each within |amp|: float32 accumulation error bound.
        assert np.max(np.abs(got - want)) <= 8 * M * 2.0 ** -23 * 2.0, shared


@pytest.mark.parametrize("viscosity", (1.5e-3, 0.0))
def test_fp32_driven_run_tracks_fp64(jx, viscosity):
    """Driven runs in FP32 against FP64 from the same state and OU table.  Viscous FP32
    kernels take the kick's second moments from one pass (sums about zero, float64);
    inviscid ones from a second pass about the mean, as FP64 does.  Both must keep the kick
    scale, so the runs agree to FP32 rounding."""
    from jathenax_amd import forcing as jf
    nx, ny = 48, 40
    xs = (np.arange(nx) + 0.5) / nx - 0.5
    ys = (np.arange(ny) + 0.5) / ny - 0.5
    modes = jf.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, 1618, 0.75, 0.004, 40)
    U0 = np.zeros((ny, nx, 3))
    U0[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * xs)[None, :]
    U0[..., 1] = 0.05 * U0[..., 0]    # a mean flow: the mean force and momentum are not zero
    out = {}
    for dt_ in (np.float64, np.float32):
        basis = jf.forcing_basis(modes, xs.astype(dt_), ys.astype(dt_), (1.0, 1.0))
        tdt = torch.float64 if dt_ == np.float64 else torch.float32
        out[dt_] = jx.evolve_isothermal(
            torch.tensor(U0, dtype=tdt, device="cuda"), 1 / nx, 1 / ny, 1.0, 1e-10, "rk2", 0.3,
            0.05, 20, viscosity=viscosity, drag_rate=0.3333335,
            forcing=(coef.astype(dt_), basis), dt_update=0.004, tdriv_start=0.0, dedt=0.5,
            reconstruct="wenoz", rsolver="hlle")
    U64, U32 = _np(out[np.float64][0]), _np(out[np.float32][0]).astype(np.float64)
    assert out[np.float64][3] == out[np.float32][3]
    assert np.max(np.abs(U32 - U64)) <= 2e-5 * np.max(np.abs(U64))


@pytest.mark.parametrize("shape,nlim,tracer", [((48, 40), 24, False), ((300, 173), 6, False),
                                               ((66, 52), 12, True)])
def test_fp32_evaluated_forcing_equals_rendered(jx, shape, nlim, tracer):
    """FP32 forced Smagorinsky runs (the driven-turbulence configuration) evaluate the raw
    force where the kick reads it, from per-row sums of the basis's mode groups; rendering
    the (ny, nx, 2) array every step instead must give the same run bit for bit (with a
    tracer, which the kick carries unchanged)."""
    from jathenax_amd import workloads
    from jathenax_amd.device import context_for
    nx, ny = shape
    s = workloads.turbulence_spec(nx, ny, 0.5, 1618, 0.002, "rk2", "wenoz", "hlle", cfl=0.3,
                                  dtype=np.float32)
    s["kw"]["dedt"] = 0.5
    U = s["U"]
    if tracer:
        ys = (np.arange(ny) + 0.5) / ny
        conc = (np.abs(ys - 0.5) < 0.25).astype(np.float64)[:, None] * np.ones((1, nx))
        U = np.concatenate([U, (U[..., 0] * conc)[..., None]], axis=-1)
    U0 = torch.tensor(U, dtype=torch.float32, device="cuda")
    args = (s["dx"], s["dy"], s["cs"], s["dfloor"], s["integrator"], s["cfl"], s["tlim"], nlim)
    ctx = context_for(U0)
    out = {}
    try:
        for fused in (True, False):
            ctx.set_fused_forcing(fused)
            out[fused] = jx.evolve_isothermal(U0, *args, **s["kw"])
    finally:
        ctx.set_fused_forcing(True)
    assert out[True][3] == out[False][3] == nlim
    assert out[True][1:] == out[False][1:]
    np.testing.assert_array_equal(_np(out[True][0]), _np(out[False][0]))
    assert float(torch.max(torch.abs(out[True][0][..., 1]))) > 0.0   # the force acted


def test_fp32_tracks_fp64(jx):
    from jathenax_amd import problems
    U0 = problems.smooth_wave(128, 128)
    kw = dict(reconstruct="ppm4", rsolver="roe")
    U64, t64, _, c64 = jx.evolve_isothermal(torch.tensor(U0, device="cuda"), 1 / 128, 1 / 128,
                                            1.0, 1e-10, "rk3", 0.4, 0.2, 10000, **kw)
    U32, t32, _, c32 = jx.evolve_isothermal(torch.tensor(U0, dtype=torch.float32, device="cuda"),
                                            1 / 128, 1 / 128, 1.0, 1e-10, "rk3", 0.4, 0.2, 10000,
                                            **kw)
    assert t64 == 0.2 and t32 == 0.2
    l1 = float(np.mean(np.abs(_np(U32)[..., 0] - _np(U64)[..., 0])))
    assert l1 <= 1e-5


def test_zero_step_run_returns_input(jx):
    U0 = np.ones((6, 8, 3))
    U0[0, 0, 0] = -1.0
    U, t, dt, cyc = jx.evolve_isothermal(torch.tensor(U0, device="cuda"), 0.1, 0.1, 1.0, 0.5,
                                         "rk3", 0.4, 1.0, 0)
    np.testing.assert_array_equal(_np(U), U0)
    assert (t, dt, cyc) == (0.0, np.inf, 0)


def test_unknown_options_raise_key_error(jx):
    U = torch.ones((6, 8, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(KeyError):
        jx.isothermal_step(U, 0.01, 0.1, 0.1, 1.0, 1e-10, "rk5")
    with pytest.raises(KeyError):
        jx.isothermal_step(U, 0.01, 0.1, 0.1, 1.0, 1e-10, "rk3", reconstruct="weno7")
    with pytest.raises(KeyError):
        jx.isothermal_step(U, 0.01, 0.1, 0.1, 1.0, 1e-10, "rk3", rsolver="hllc")
