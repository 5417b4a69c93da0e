"""The reference's operator surface through jathenax_amd: the registry callables with the
reference signatures (fluxes.py:6-131, 134-224), the gradient and Smagorinsky operators
(fluxes.py:227-245), replaced registry entries, and RK_STAGES read on every call
(hydro.py:20-51, 100).  FP64 results are held to the golden vectors written by the
reference's own functions and to the oracle; the FP32 operator forms to the FP64 vectors by
a per-face error bound.  Every call goes through libjathenax_amd.so."""

import numpy as np
import pytest
import torch

from oracle import jathenax_np as ox

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
RECONS = ("dc", "plm", "ppm4", "ppmx", "wenoz")
SOLVERS = ("llf", "hlle", "roe")


@pytest.fixture(scope="module")
def fx():
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    from jathenax_amd import fluxes
    return fluxes


def _np(t):
    return t.detach().cpu().numpy()


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dtype, device="cuda")


def test_registries_hold_callables_in_reference_order(fx):
    assert list(fx.RECONSTRUCTIONS) == list(RECONS)
    assert list(fx.RIEMANN_SOLVERS) == list(SOLVERS)
    assert fx.RECONSTRUCTIONS["wenoz"] is fx.wenoz_face_states
    assert all(callable(f) for f in (*fx.RECONSTRUCTIONS.values(), *fx.RIEMANN_SOLVERS.values()))
    with pytest.raises(KeyError):
        fx.reconstruct_faces(_dev(np.ones((4, 4, 3))), 1, "weno7", 1e-10)
    with pytest.raises(KeyError):
        fx.isothermal_flux(_dev(np.ones((4, 4, 3))), _dev(np.ones((4, 4, 3))), 1.0, 1, "hllc")


@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("axis", (0, 1))
def test_face_state_functions_match_reference(fx, ref_vectors, recon, axis):
    # RECONSTRUCTIONS[name](W, axis) does not floor (fluxes.py:6-117); reconstruct_faces
    # floors PPMX/WENO-Z face densities (fluxes.py:124-131), here with the golden vectors'
    # dfloor 0.9 so that the floor binds.
    W = ref_vectors["recon_W"]
    left, right = fx.RECONSTRUCTIONS[recon](_dev(W), axis)
    lo, ro = ox.RECONSTRUCTIONS[recon](W, axis)
    np.testing.assert_array_equal(_np(left), lo)
    np.testing.assert_array_equal(_np(right), ro)
    left, right = fx.reconstruct_faces(_dev(W), axis, recon, 0.9)
    np.testing.assert_array_equal(_np(left), ref_vectors[f"recon_{recon}_axis{axis}_left"])
    np.testing.assert_array_equal(_np(right), ref_vectors[f"recon_{recon}_axis{axis}_right"])


@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("axis", (0, 1))
def test_fp32_face_states_track_fp64(fx, ref_vectors, recon, axis):
    # FP32 operator forms (reciprocal divisions, median limiter, WENO-Z difference form)
    # against the float64 reference on the same float32-rounded inputs.  Bound: 16 ulps of
    # the stencil's largest magnitude, except where the FP64 limiter sits within rounding of
    # a branch switch (counted, and required to be rare).
    W32 = ref_vectors["recon_W"].astype(np.float32)
    left, right = fx.RECONSTRUCTIONS[recon](_dev(W32, torch.float32), axis)
    lo, ro = ox.RECONSTRUCTIONS[recon](W32.astype(np.float64), axis)
    scale = np.max(np.abs(W32), axis=axis, keepdims=True).astype(np.float64)
    for got, want in ((_np(left), lo), (_np(right), ro)):
        bad = np.abs(got - want) > 16 * EPS32 * scale
        assert bad.sum() <= 0.01 * bad.size, (recon, axis, int(bad.sum()))


def test_fp32_wenoz_jump_matches_reference_float32(fx, ref_vectors):
    # The reference's float32 WENO-Z on the flat-stencil jump (the vector its own function
    # wrote in float32; test_jathenax.py:147-151 bounds both at 2e-7).
    W = ref_vectors["wenoz_f32_W"]
    left, right = fx.wenoz_face_states(_dev(W, torch.float32), 1)
    np.testing.assert_allclose(_np(left), ref_vectors["wenoz_f32_left"], rtol=0, atol=2e-7)
    np.testing.assert_allclose(_np(right), ref_vectors["wenoz_f32_right"], rtol=0, atol=2e-7)


@pytest.mark.parametrize("rs", SOLVERS)
def test_solver_callables_match_reference(fx, ref_vectors, rs):
    # RIEMANN_SOLVERS[name](rho_l, rho_r, vn_l, vn_r, vt_l, vt_r, cs) in (mass, normal,
    # tangential) order: the golden x-face vectors are exactly this call (axis 1 keeps the
    # component order).
    L, R = ref_vectors["riemann_left"][0], ref_vectors["riemann_right"][0]
    f = fx.RIEMANN_SOLVERS[rs](_dev(L[:, 0]), _dev(R[:, 0]), _dev(L[:, 1]), _dev(R[:, 1]),
                               _dev(L[:, 2]), _dev(R[:, 2]), 1.3)
    want = ref_vectors[f"riemann_{rs}_axis1"][0, :, :3]
    assert tuple(f.shape) == want.shape
    np.testing.assert_allclose(_np(f), want, rtol=1e-14, atol=1e-14 * np.abs(want).max())
    for axis in (0, 1):
        got = fx.isothermal_flux(_dev(ref_vectors["riemann_left"]),
                                 _dev(ref_vectors["riemann_right"]), 1.3, axis, rs)
        want = ref_vectors[f"riemann_{rs}_axis{axis}"]
        np.testing.assert_allclose(_np(got), want, rtol=1e-14, atol=1e-14 * np.abs(want).max())
    if rs == "roe":
        got = fx.roe_isothermal_flux(_dev(ref_vectors["riemann_left"]),
                                     _dev(ref_vectors["riemann_right"]), 1.3, 0)
        np.testing.assert_array_equal(
            _np(got), _np(fx.isothermal_flux(_dev(ref_vectors["riemann_left"]),
                                             _dev(ref_vectors["riemann_right"]), 1.3, 0)))


def _flux_scale(L, R, cs, axis):
    """Per-face magnitude of the terms every solver sums: |F_L| + |F_R| + s (|U_L| + |U_R|)
    over the hydro components, s the larger signal speed."""
    n, t = (1, 2) if axis == 1 else (2, 1)
    rl, rr, ul, ur, tl, tr = L[..., 0], R[..., 0], L[..., n], R[..., n], L[..., t], R[..., t]
    s = np.maximum(np.abs(ul), np.abs(ur)) + cs
    flux = (np.abs(rl * ul) + np.abs(rr * ur) + np.abs(rl * ul * ul) + np.abs(rr * ur * ur)
            + cs * cs * (rl + rr) + np.abs(rl * ul * tl) + np.abs(rr * ur * tr))
    state = rl + rr + np.abs(rl * ul) + np.abs(rr * ur) + np.abs(rl * tl) + np.abs(rr * tr)
    return flux + s * state


@pytest.mark.parametrize("rs", SOLVERS)
@pytest.mark.parametrize("axis", (0, 1))
def test_fp32_solvers_on_supersonic_and_rarefying_faces(fx, ref_vectors, rs, axis):
    # The golden face pairs hold subsonic, supersonic (|v| up to 4 against cs 1.3) and
    # strongly rarefying faces where Roe falls back to LLF.  The FP32 forms (select-free
    # Roe with the fallback, rational HLLE, LLF) against the float64 reference on the same
    # float32-rounded states: every face within 8 float32 ulps of _flux_scale.
    L32 = ref_vectors["riemann_left"].astype(np.float32)
    R32 = ref_vectors["riemann_right"].astype(np.float32)
    L, R = L32.astype(np.float64), R32.astype(np.float64)
    got = _np(fx.isothermal_flux(_dev(L32, torch.float32), _dev(R32, torch.float32), 1.3, axis,
                                 rs)).astype(np.float64)
    want = ox.isothermal_flux(L, R, 1.3, axis, rs)
    scale = _flux_scale(L, R, 1.3, axis)[..., None]
    err = np.abs(got - want) / scale
    assert err.max() <= 8 * EPS32, (rs, axis, err.max() / EPS32)
    # the cases the forms rewrite are all present, and none sits on a branch switch
    n = 1 if axis == 1 else 2
    rl, rr, ul, ur = L[..., 0], R[..., 0], L[..., n], R[..., n]
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    vn = (sl * ul + sr * ur) / (sl + sr)
    am = (rr - rl) * (0.5 + 0.5 * vn / 1.3) - (rr * ur - rl * ul) * 0.5 / 1.3
    rstar = rl + am
    assert (rstar < 0).sum() >= 8 and (np.abs(vn) > 1.3).sum() >= 32
    assert np.min(np.abs(rstar) / (rl + rr)) > 1e-5


@pytest.mark.parametrize("axis", (0, 1))
@pytest.mark.parametrize("dtype", (torch.float64, torch.float32))
def test_face_gradients_and_smagorinsky(fx, ref_vectors, axis, dtype):
    W = ref_vectors["visc_W"]
    gx, gy = fx.face_velocity_gradients(_dev(W, dtype), 0.1, 0.07, axis)
    nu = fx.smagorinsky_viscosity(gx, gy, 0.1, 0.07, 0.17)
    want = (ref_vectors[f"visc_grad_axis{axis}_dvdx"], ref_vectors[f"visc_grad_axis{axis}_dvdy"],
            ref_vectors[f"visc_nusgs_axis{axis}"])
    if dtype == torch.float64:
        for got, w in zip((gx, gy, nu), want):
            np.testing.assert_array_equal(_np(got), w)
        return
    # FP32: reciprocal multiplies instead of divisions; gradients of O(1/dx) differences
    for got, w in zip((gx, gy, nu), want):
        np.testing.assert_allclose(_np(got), w, rtol=0, atol=64 * EPS32 * np.abs(w).max())


# ---- replaced registry entries --------------------------------------------------------


def _state(ny, nx, nv, seed):
    rng = np.random.default_rng(seed)
    U = np.empty((ny, nx, nv))
    U[..., 0] = 1.0 + 0.3 * rng.random((ny, nx))
    U[..., 1:3] = 0.3 * rng.standard_normal((ny, nx, 2)) * U[..., :1]
    U[..., 3:] = rng.random((ny, nx, nv - 3)) * U[..., :1]
    return U


def test_replaced_reconstruction_runs_on_the_unfused_path(fx, monkeypatch):
    import jathenax_amd as jx
    U = _state(24, 20, 4, 3)
    want = _np(jx.isothermal_step(_dev(U), 2e-3, 0.05, 0.04, 1.0, 1e-10, "rk3",
                                  reconstruct="dc"))

    def my_dc(W, axis):   # the caller's own operator, in torch
        return torch.roll(W, 1, dims=axis), W

    monkeypatch.setitem(fx.RECONSTRUCTIONS, "ppm4", my_dc)
    got = jx.isothermal_step(_dev(U), 2e-3, 0.05, 0.04, 1.0, 1e-10, "rk3", reconstruct="ppm4")
    np.testing.assert_array_equal(_np(got), want)
    got, t, dt, cyc = jx.evolve_isothermal(_dev(U), 0.05, 0.04, 1.0, 1e-10, "rk3", 0.4, 1.0, 3,
                                           reconstruct="ppm4")
    want, wt, wdt, wcyc = jx.evolve_isothermal(_dev(U), 0.05, 0.04, 1.0, 1e-10, "rk3", 0.4, 1.0,
                                               3, reconstruct="dc")
    assert (cyc, t, dt) == (wcyc, wt, wdt)
    np.testing.assert_array_equal(_np(got), _np(want))


def test_registry_entries_are_dispatched_by_callable(fx, monkeypatch):
    import jathenax_amd as jx
    U = _state(20, 28, 3, 4)
    hlle = _np(jx.isothermal_step(_dev(U), 2e-3, 0.05, 0.04, 1.0, 1e-10, "rk2", rsolver="hlle"))
    # a built-in operator under another name runs fused as that operator
    monkeypatch.setitem(fx.RIEMANN_SOLVERS, "roe", fx.RIEMANN_SOLVERS["hlle"])
    got = jx.isothermal_step(_dev(U), 2e-3, 0.05, 0.04, 1.0, 1e-10, "rk2", rsolver="roe")
    np.testing.assert_array_equal(_np(got), hlle)
    # the caller's own solver (here wrapping HLLE) runs unfused: FP64 traced == fused
    monkeypatch.setitem(fx.RIEMANN_SOLVERS, "roe",
                        lambda *states: fx.RIEMANN_SOLVERS["hlle"](*states))
    got = jx.isothermal_step(_dev(U), 2e-3, 0.05, 0.04, 1.0, 1e-10, "rk2", rsolver="roe")
    np.testing.assert_array_equal(_np(got), hlle)
    from jathenax_amd.ensemble import Ensemble
    with pytest.raises(ValueError, match="not a built-in operator"):
        Ensemble(torch.float64).add(U, 0.05, 0.04, 1.0, 1e-10, "rk2", 0.4, 0.01, 2,
                                    rsolver="roe")


# ---- RK_STAGES read on every call (hydro.py:100) --------------------------------------


TABLES = {
    "heun": ((0.0, 1.0, 1.0, None), (0.5, 0.5, 0.5, None)),              # rk2 rows, new key
    "euler": ((0.0, 1.0, 1.0, None),),                                   # one stage
    "delta_first": ((0.0, 1.0, 0.6, 0.5), (0.5, 0.25, 0.4, None)),       # delta in row 1
    "blend_first": ((0.5, 0.5, 1.0, None), (0.25, 0.75, 0.25, None),     # blended row 1
                    (2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, None)),
}


@pytest.mark.parametrize("name", sorted(TABLES))
@pytest.mark.parametrize("dtype", (torch.float64, torch.float32))
def test_added_rk_tables_follow_the_reference(name, dtype, monkeypatch):
    import jathenax_amd as jx
    from jathenax_amd import hydro
    monkeypatch.setitem(hydro.RK_STAGES, name, TABLES[name])
    monkeypatch.setitem(ox.RK_STAGES, name, TABLES[name])
    U = _state(26, 30, 4, 5)
    kw = dict(viscosity=1e-3, drag_rate=0.2)
    want = ox.isothermal_step(U, 2e-3, 1 / 30, 1 / 26, 1.0, 1e-10, name, **kw)
    got = _np(jx.isothermal_step(_dev(U, dtype), 2e-3, 1 / 30, 1 / 26, 1.0, 1e-10, name, **kw))
    tol = 1e-13 if dtype == torch.float64 else 64 * EPS32
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))
    # the device loop uses the same rows (the one-stage table writes through a copy)
    got, t, dt, cyc = jx.evolve_isothermal(_dev(U, dtype), 1 / 30, 1 / 26, 1.0, 1e-10, name, 0.3,
                                           1.0, 4, **kw)
    want, wt, wdt, wcyc = ox.evolve_isothermal(U, 1 / 30, 1 / 26, 1.0, 1e-10, name, 0.3, 1.0, 4,
                                               **kw)
    assert cyc == wcyc == 4
    assert np.max(np.abs(_np(got) - want)) <= 4 * tol * np.max(np.abs(want))


def test_edited_rk_row_is_applied(monkeypatch):
    import jathenax_amd as jx
    from jathenax_amd import hydro
    edited = (hydro.RK_STAGES["rk3"][0], (0.3, 0.7, 0.3, None), hydro.RK_STAGES["rk3"][2])
    monkeypatch.setitem(hydro.RK_STAGES, "rk3", edited)
    monkeypatch.setitem(ox.RK_STAGES, "rk3", edited)
    U = _state(18, 22, 3, 6)
    want = ox.isothermal_step(U, 2e-3, 1 / 22, 1 / 18, 1.0, 1e-10, "rk3")
    got = _np(jx.isothermal_step(_dev(U), 2e-3, 1 / 22, 1 / 18, 1.0, 1e-10, "rk3"))
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    from jathenax_amd.ensemble import Ensemble
    with pytest.raises(ValueError, match="built-in"):
        Ensemble(torch.float64).add(U, 1 / 22, 1 / 18, 1.0, 1e-10, "rk3", 0.4, 0.01, 2)
    with pytest.raises(KeyError):
        jx.isothermal_step(_dev(U), 2e-3, 1 / 22, 1 / 18, 1.0, 1e-10, "rk5")


# ---- FP32 against FP64 for every reconstruction x solver pair --------------------------


def _hard_flow(n):
    """Supersonic x drift over a density wave, with a velocity jump of 5 cs at x = 0 (a
    vacuum-forming rarefaction, where Roe falls back to LLF on the first steps) and -5 cs at
    the periodic seam (a shock)."""
    x = (np.arange(n) + 0.5) / n - 0.5
    X, Y = np.meshgrid(x, x)
    rho = 1.0 + 0.5 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    vx = 2.0 + 0.5 * np.cos(2 * np.pi * Y) + 2.5 * np.sign(X)
    vy = 0.8 * np.sin(2 * np.pi * X)
    return np.stack((rho, rho * vx, rho * vy), axis=-1)


# Relative L1 of FP32 against FP64 after 10 RK3 steps of the hard flow, stated bound
# (measured 0.8e-7 .. 4.6e-7 over the 15 pairs, PPMX + HLLE the largest).
PAIR_FP32_REL_L1 = 2e-6


@pytest.mark.parametrize("recon", RECONS)
@pytest.mark.parametrize("rs", SOLVERS)
def test_fp32_every_operator_pair_tracks_fp64(recon, rs):
    import jathenax_amd as jx
    n = 64
    U0 = _hard_flow(n)
    args = (1 / n, 1 / n, 1.0, 1e-10, "rk3", 0.3, 10.0, 10)
    kw = dict(reconstruct=recon, rsolver=rs)
    # the flow exercises the branches the FP32 forms rewrite
    _, W = ox.recover_isothermal(U0, 1e-10)
    L, R = ox.reconstruct_faces(W, 1, recon, 1e-10)
    assert (np.abs(L[..., 1]) > 1.0).mean() >= 0.4       # supersonic x faces
    assert (R[..., 1] - L[..., 1] > 2.0).sum() >= n      # Roe's LLF fallback faces
    U64, t64, dt64, c64 = jx.evolve_isothermal(_dev(U0), *args, **kw)
    Uo = ox.evolve_isothermal(U0, *args, **kw)[0]
    U64 = _np(U64)
    assert np.max(np.abs(U64 - Uo)) <= 1e-12 * np.max(np.abs(Uo))
    U32, t32, dt32, c32 = jx.evolve_isothermal(_dev(U0, torch.float32), *args, **kw)
    assert c32 == c64 == 10
    rel = float(np.sum(np.abs(_np(U32).astype(np.float64) - U64)[..., 0])
                / np.sum(np.abs(U64[..., 0])))
    print(f"FP32 vs FP64 {recon}+{rs}, 10 RK3 steps: relative L1(rho) {rel:.3e}")
    assert rel <= PAIR_FP32_REL_L1
