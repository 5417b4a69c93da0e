"""Traced and learned-stress stage paths against the oracle, and against the fused stage.

Float64: relative L-infinity <= 1e-13 against the oracle (the expression order is the
reference's).  An unforced traced step equals the fused step bit for bit; forced steps agree
to 1e-13, because the two paths reduce the kick's box sums in different fixed orders.
Every call goes through libjathenax_amd.so."""

import numpy as np
import pytest
import torch

from oracle import jathenax_np as ox

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def jx():
    import jathenax_amd
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    return jathenax_amd


def _state(ny, nx, nv, seed):
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    U = np.empty((ny, nx, nv))
    U[..., 0] = (1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
                 + 0.02 * rng.random((ny, nx)))
    U[..., 1] = U[..., 0] * (0.4 * np.sin(2 * np.pi * y) + 0.05 * rng.standard_normal((ny, nx)))
    U[..., 2] = U[..., 0] * (0.3 * np.cos(2 * np.pi * x) + 0.05 * rng.standard_normal((ny, nx)))
    for v in range(3, nv):
        U[..., v] = U[..., 0] * rng.random((ny, nx))
    return U


def _stress_np(w, prim):
    r, u, v = prim[..., 0], prim[..., 1], prim[..., 2]
    return w * np.stack((r * u * u, r * u * v, r * v * v), axis=-1)


def _stress_torch(w, prim):
    r, u, v = prim[..., 0], prim[..., 1], prim[..., 2]
    return w * torch.stack((r * u * u, r * u * v, r * v * v), dim=-1)


@pytest.mark.parametrize("recon", ("dc", "plm", "ppm4", "ppmx", "wenoz"))
@pytest.mark.parametrize("axis", (0, 1))
def test_reconstruct_faces(jx, recon, axis):
    from jathenax_amd import fluxes
    U = _state(37, 45, 4, 1)
    _, W = ox.recover_isothermal(U, 1e-6)
    l, r = fluxes.reconstruct_faces(torch.tensor(W, device="cuda"), axis, recon, 1e-6)
    lo, ro = ox.reconstruct_faces(W, axis, recon, 1e-6)
    np.testing.assert_array_equal(_np(l), lo)
    np.testing.assert_array_equal(_np(r), ro)


@pytest.mark.parametrize("rs", ("llf", "hlle", "roe"))
@pytest.mark.parametrize("axis", (0, 1))
def test_isothermal_flux(jx, rs, axis):
    from jathenax_amd import fluxes
    U = _state(33, 40, 5, 2)
    _, W = ox.recover_isothermal(U, 1e-6)
    lo, ro = ox.reconstruct_faces(W, axis, "ppm4", 1e-6)
    got = fluxes.isothermal_flux(torch.tensor(lo, device="cuda"),
                                 torch.tensor(ro, device="cuda"), 1.3, axis, rs)
    assert _rel(_np(got), ox.isothermal_flux(lo, ro, 1.3, axis, rs)) <= 1e-14


@pytest.mark.parametrize("axis", (0, 1))
def test_viscous_momentum_flux(jx, axis):
    from jathenax_amd import fluxes
    U = _state(31, 29, 3, 3)
    _, W = ox.recover_isothermal(U, 1e-6)
    got = fluxes.viscous_momentum_flux(torch.tensor(W, device="cuda"), 0.1, 0.07, 2e-3, 0.17,
                                       axis)
    assert _rel(_np(got), ox.viscous_momentum_flux(W, 0.1, 0.07, 2e-3, 0.17, axis)) <= 1e-14


CASES = [
    ("ppm4", "roe", "rk3", 3, dict(viscosity=2e-3, smagorinsky_cs=0.17, drag_rate=0.7), True),
    ("wenoz", "hlle", "rk2", 4, dict(viscosity=1e-3), False),
    ("ppmx", "llf", "rk4", 5, dict(drag_rate=0.3), True),
    ("plm", "roe", "rk3", 3, {}, False),
]


@pytest.mark.parametrize("recon,rs,integ,nv,phys,forced", CASES)
def test_traced_step_matches_oracle_and_fused(jx, recon, rs, integ, nv, phys, forced):
    U = _state(36, 44, nv, 4)
    raw = np.random.default_rng(5).standard_normal((36, 44, 2)) if forced else None
    kw = dict(reconstruct=recon, rsolver=rs, **phys)
    if forced:
        kw.update(raw_force=raw, dedt=0.05)
    args = (0.004, 1 / 44, 1 / 36, 1.0, 1e-6, integ)
    ref_U, ref_stages = ox.isothermal_step(U, *args, trace=True, **kw)
    gkw = dict(kw)
    if forced:
        gkw["raw_force"] = torch.tensor(raw, device="cuda")
    got_U, got_stages = jx.isothermal_step(torch.tensor(U, device="cuda"), *args, trace=True,
                                           **gkw)
    assert _rel(_np(got_U), ref_U) <= 1e-13
    assert len(got_stages) == len(ref_stages)
    for g, r in zip(got_stages, ref_stages):
        assert set(g) == set(r)
        for key in r:
            assert _rel(_np(g[key]), r[key]) <= 1e-13, key
    fused = jx.isothermal_step(torch.tensor(U, device="cuda"), *args, **gkw)
    if forced:
        # The kick's box sums are reduced in a different (fixed) order on the two paths.
        assert _rel(_np(fused), _np(got_U)) <= 1e-13
    else:
        np.testing.assert_array_equal(_np(fused), _np(got_U))


def test_traced_fp32_tracks_fused_fp32(jx):
    U = _state(64, 48, 4, 6)
    kw = dict(reconstruct="ppm4", rsolver="roe", viscosity=1e-3, drag_rate=0.5)
    args = (0.003, 1 / 48, 1 / 64, 1.0, 1e-6, "rk3")
    U32 = torch.tensor(U, dtype=torch.float32, device="cuda")
    fused = _np(jx.isothermal_step(U32, *args, **kw))
    traced, stages = jx.isothermal_step(U32, *args, trace=True, **kw)
    assert traced.dtype == torch.float32 and stages[0]["flux_x"].dtype == torch.float32
    assert _rel(_np(traced), fused) <= 1e-6


@pytest.mark.parametrize("recon,rs,integ,nv,phys", [
    ("wenoz", "roe", "rk3", 4, dict(viscosity=1e-3)),
    ("ppm4", "hlle", "rk2", 3, {}),
    ("plm", "llf", "rk4", 5, dict(smagorinsky_cs=0.17, drag_rate=0.4)),
])
def test_predict_stress_step_matches_oracle(jx, recon, rs, integ, nv, phys):
    U = _state(40, 32, nv, 7)
    args = (0.004, 1 / 32, 1 / 40, 1.0, 1e-6, integ)
    kw = dict(reconstruct=recon, rsolver=rs, **phys)
    ref = ox.isothermal_step(U, *args, predict_stress=_stress_np, sgs_weights=0.05, **kw)
    Ud = torch.tensor(U, device="cuda")
    fused = jx.isothermal_step(Ud, *args, predict_stress=_stress_torch, sgs_weights=0.05, **kw)
    assert _rel(_np(fused), ref) <= 1e-13
    traced, _ = jx.isothermal_step(Ud, *args, predict_stress=_stress_torch, sgs_weights=0.05,
                                   trace=True, **kw)
    np.testing.assert_array_equal(_np(fused), _np(traced))
    plain = ox.isothermal_step(U, *args, **kw)
    assert _rel(ref, plain) > 1e-8   # the stress is not negligible


def test_predict_stress_contract(jx):
    U = torch.tensor(_state(16, 16, 3, 8), device="cuda")
    bad_shape = lambda w, prim: torch.zeros((16, 16, 2), dtype=prim.dtype, device=prim.device)
    with pytest.raises(ValueError):
        jx.isothermal_step(U, 0.001, 1 / 16, 1 / 16, 1.0, 1e-6, "rk2", predict_stress=bad_shape)
    bad_type = lambda w, prim: torch.zeros((16, 16, 3), dtype=torch.int32, device=prim.device)
    with pytest.raises(ValueError):
        jx.isothermal_step(U, 0.001, 1 / 16, 1 / 16, 1.0, 1e-6, "rk2", predict_stress=bad_type)
    with pytest.raises(ValueError):
        jx.isothermal_step(U, 0.001, 1 / 16, 1 / 16, 1.0, 1e-6, "rk2", predict_stress=bad_shape,
                           trace=True)
    numpy_out = lambda w, prim: np.zeros((16, 16, 3))
    out = jx.isothermal_step(U, 0.001, 1 / 16, 1 / 16, 1.0, 1e-6, "rk2",
                             predict_stress=numpy_out)
    np.testing.assert_array_equal(_np(out), _np(jx.isothermal_step(U, 0.001, 1 / 16, 1 / 16,
                                                                   1.0, 1e-6, "rk2")))


def test_evolve_with_predictor(jx):
    from jathenax_amd import forcing as jf, problems
    nx, ny = 40, 36
    U = _state(ny, nx, 3, 9)
    modes = jf.cartesian_modes(1, 2, (-2, 2), (-2, 2))
    coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 2, 2.0, 1.0, 1618, 0.75, 0.002, 40)
    basis = jf.forcing_basis(modes, problems.cell_centres(nx), problems.cell_centres(ny),
                             (1.0, 1.0))
    kw = dict(viscosity=1e-3, smagorinsky_cs=0.17, drag_rate=0.3, dt_update=0.002,
              tdriv_start=0.0, dedt=0.5, reconstruct="ppm4", rsolver="hlle")
    args = (1 / nx, 1 / ny, 1.0, 1e-6, "rk2", 0.4, 0.05, 12)
    # A zero stress reproduces the fused device loop bit for bit.
    zero = lambda w, prim: torch.zeros_like(prim)
    unforced = {k: v for k, v in kw.items() if k not in ("dt_update", "tdriv_start", "dedt")}
    got = jx.evolve_isothermal(torch.tensor(U, device="cuda"), *args, predict_stress=zero,
                               **unforced)
    fused = jx.evolve_isothermal(torch.tensor(U, device="cuda"), *args, **unforced)
    np.testing.assert_array_equal(_np(got[0]), _np(fused[0]))
    assert got[1:] == fused[1:]
    got = jx.evolve_isothermal(torch.tensor(U, device="cuda"), *args, forcing=(coef, basis),
                               predict_stress=zero, **kw)
    fused = jx.evolve_isothermal(torch.tensor(U, device="cuda"), *args, forcing=(coef, basis),
                                 **kw)
    np.testing.assert_array_equal(_np(got[0]), _np(fused[0]))
    assert got[1:] == fused[1:]
    # A nonzero stress against the oracle's loop.
    coef_np = np.asarray(coef.cpu() if isinstance(coef, torch.Tensor) else coef)
    basis_np = tuple(np.asarray(b.cpu() if isinstance(b, torch.Tensor) else b) for b in basis)
    ref = ox.evolve_isothermal(U, *args, forcing=(coef_np, basis_np), predict_stress=_stress_np,
                               sgs_weights=0.05, **kw)
    got = jx.evolve_isothermal(torch.tensor(U, device="cuda"), *args, forcing=(coef, basis),
                               predict_stress=_stress_torch, sgs_weights=0.05, **kw)
    assert got[3] == ref[3] and abs(got[1] - ref[1]) <= 1e-12 * ref[1]
    assert _rel(_np(got[0]), ref[0]) <= 1e-12
    # The unfused host loop (one traced step per cycle) agrees with the fused loop.
    from jathenax_amd.trace import evolve_with_predictor
    unf = evolve_with_predictor(torch.tensor(U, device="cuda"), *args, forcing=(coef, basis),
                                predict_stress=_stress_torch, sgs_weights=0.05, **kw)
    assert unf[3] == got[3] and _rel(_np(unf[0]), _np(got[0])) <= 1e-12
