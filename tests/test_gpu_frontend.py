"""Run front-end on the device: the CLI flow (deck, overrides, run, end-of-run check, .npz
snapshot, drift report) and the device spectra, against the reference's vectors, the oracle
and closed forms."""

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fe():
    from jathenax_amd import frontend
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    return frontend


def _rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_deck_and_cli_overrides(fe, tmp_path, monkeypatch, capsys):
    """test_input.py:31-67 through the front-end's flags: the flags win over the deck, and
    the snapshot holds the run."""
    deck = tmp_path / "input.toml"
    deck.write_text('[problem]\nname = "kh"\n[mesh]\nnx = 8\nny = 10\n'
                    '[time]\nnlim = 2\ntlim = 1\nintegrator = "rk2"\n'
                    '[hydro]\nreconstruct = "plm"\nrsolver = "hlle"\n'
                    '[output]\nfile = "configured.npz"\n')
    output = tmp_path / "overridden.npz"
    monkeypatch.chdir(tmp_path)
    fe.main(["-i", str(deck), "--nx", "12", "--integrator", "rk4", "--rsolver", "llf",
             "--output", str(output)])
    out = capsys.readouterr().out
    assert "RK4" in out and "float32" in out
    assert "Reconstruction plm, Riemann solver llf" in out
    assert "Reached time.nlim=2 before tlim=1" in out
    with np.load(output) as data:
        assert data["problem"].item() == "kh"
        assert data["conserved"].shape == (10, 12, 4)
        assert data["conserved"].dtype == np.float32
        assert data["cycle"] == 2
        assert 0 < data["time"] < 1
    assert not (tmp_path / "configured.npz").exists()


def test_smooth_defaults_match_the_reference_run(fe, ref_vectors, tmp_path):
    """The CLI's 'smooth' problem (64x48, PPM4+Roe, RK3, t = 0.2) in FP64 against the
    reference's own run of it; mass and momenta drift at round-off only."""
    s = fe.with_overrides(fe.defaults("smooth"), output=dict(file=str(tmp_path / "s.npz")))
    r = fe.run_settings(s, dtype=torch.float64)
    assert r["reached_tlim"] and r["time"] == 0.2
    U = r["U"].cpu().numpy()
    assert _rel_inf(U, ref_vectors["run_smooth_tlim_U"]) <= 1e-12
    t_ref, dt_ref, c_ref = ref_vectors["run_smooth_tlim_tdc"]
    assert r["cycle"] == int(c_ref) and abs(r["dt"] - dt_ref) <= 1e-12 * dt_ref
    assert np.max(np.abs(r["drift"])) <= 1e-14
    with np.load(r["output"]) as data:
        np.testing.assert_array_equal(data["conserved"], U)
        np.testing.assert_array_equal(data["initial"], fe.initial_state("smooth", 64, 48))
        np.testing.assert_array_equal(data["x"], (np.arange(64) + 0.5) * (1 / 64) - 0.5)
        assert float(data["time"]) == 0.2 and int(data["cycle"]) == r["cycle"]
    assert r["lines"][-1] == str(r["output"].resolve())


def test_driven_deck_matches_the_oracle(fe, tmp_path):
    """The seed-1618 deck on a 24x20 mesh with a short forcing period, FP64, 10 cycles,
    against the oracle's loop with its own forcing tables."""
    import oracle.jathenax_np as ox
    s = fe.read_input(GOLDEN / "turbulence.toml")
    s = fe.with_overrides(s, mesh=dict(nx=24, ny=20), time=dict(tlim=0.2, nlim=10),
                          turb_driving=dict(dt_update=0.004, dedt=0.5),
                          hydro=dict(viscosity=1.5e-3),
                          output=dict(file=str(tmp_path / "t.npz")))
    r = fe.run_settings(s, dtype=torch.float64)
    assert r["cycle"] == 10 and not r["reached_tlim"]
    d = s["turb_driving"]
    modes = ox.cartesian_modes(d["nlow"], d["nhigh"], (d["min_kx"], d["max_kx"]),
                               (d["min_ky"], d["max_ky"]))
    updates = int(np.ceil(0.2 / d["dt_update"])) + 1
    coef = ox.modal_coefficients(modes, (1.0, 1.0), d["nlow"], d["nhigh"], d["npeak"],
                                 d["sol_fraction"], d["rseed"], d["tcorr"], d["dt_update"],
                                 updates)
    x = (np.arange(24) + 0.5) * (1 / 24) - 0.5
    y = (np.arange(20) + 0.5) * (1 / 20) - 0.5
    U0 = np.zeros((20, 24, 3))
    U0[..., 0] = 1.0
    h = s["hydro"]
    U, t, dt, cyc = ox.evolve_isothermal(
        U0, 1 / 24, 1 / 20, h["iso_sound_speed"], h["dfloor"], "rk3", 0.4, 0.2, 10,
        viscosity=h["viscosity"], drag_rate=s["hydro_srcterms"]["drag_rate"],
        forcing=(coef, ox.forcing_basis(modes, x, y, (1.0, 1.0))), dt_update=d["dt_update"],
        tdriv_start=0.0, dedt=d["dedt"])
    # Kick sums reduce in another order than NumPy's: round-off level only.
    assert _rel_inf(r["U"].cpu().numpy(), U) <= 1e-12
    assert cyc == 10 and abs(r["time"] - t) <= 1e-12 * t
    # Forcing injects momentum only through the zero-mean force: box momenta stay ~0.
    assert abs(r["drift"][0]) <= 1e-14 and np.max(np.abs(r["drift"][1:])) <= 1e-12


def test_device_spectra_closed_forms(fe):
    """Single Fourier modes: vx = A cos(2 pi k y), vy = B sin(2 pi m x), rho = 1. Shell k
    holds A^2/4 of energy and (2 pi k A)^2/4 of enstrophy; shell m likewise for B."""
    from jathenax_amd import spectra
    ny, nx, A, B, k, m = 48, 64, 0.3, 0.2, 3, 7
    x = (np.arange(nx) + 0.5) / nx - 0.5
    y = (np.arange(ny) + 0.5) / ny - 0.5
    U = np.zeros((ny, nx, 3))
    U[..., 0] = 1.0
    U[..., 1] = A * np.cos(2 * np.pi * k * y)[:, None]
    U[..., 2] = B * np.sin(2 * np.pi * m * x)[None, :]
    mode, energy, enstrophy = spectra.shell_spectra(torch.tensor(U, device="cuda"))
    want_e = np.zeros_like(energy)
    want_e[k] += A * A / 4
    want_e[m] += B * B / 4
    want_z = np.zeros_like(enstrophy)
    want_z[k] += (2 * np.pi * k * A) ** 2 / 4
    want_z[m] += (2 * np.pi * m * B) ** 2 / 4
    np.testing.assert_allclose(energy, want_e, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(enstrophy, want_z, rtol=1e-12, atol=1e-12)
    assert mode[-1] == int(np.floor(np.hypot(nx // 2, ny // 2) + 0.5))
    # Centred-difference vorticity of the same field: -dvx/dy + dvy/dx in closed form.
    w = spectra.vorticity(torch.tensor(U, device="cuda"), 1 / nx, 1 / ny).cpu().numpy()
    want_w = (B * np.cos(2 * np.pi * m * x)[None, :] * np.sin(2 * np.pi * m / nx) * nx
              + A * np.sin(2 * np.pi * k * y)[:, None] * np.sin(2 * np.pi * k / ny) * ny)
    np.testing.assert_allclose(w, want_w, rtol=0, atol=1e-12)


def test_device_spectra_match_the_host_restatement(fe):
    """A turbulent-like state: device shells equal the host numpy restatement of
    spectra.py:25-35 (float64, FFT round-off)."""
    from jathenax_amd import accuracy, spectra
    rng = np.random.default_rng(5)
    U = np.empty((40, 56, 3))
    U[..., 0] = 1.0 + 0.2 * rng.random((40, 56))
    U[..., 1:] = U[..., :1] * rng.normal(0, 0.3, (40, 56, 2))
    _, energy, _ = spectra.shell_spectra(torch.tensor(U, device="cuda"))
    np.testing.assert_allclose(energy, accuracy.ke_spectrum(U), rtol=1e-11, atol=1e-17)
    np.testing.assert_allclose(
        spectra.vorticity(torch.tensor(U, device="cuda"), 0.1, 0.2).cpu().numpy(),
        accuracy.vorticity(U, 0.1, 0.2), rtol=1e-13, atol=1e-13)
