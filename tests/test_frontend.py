"""Run front-end on the host: deck reading and validation (jathenax/input.py, run.py:117-131),
the CLI's initial states against the reference's own vectors, the snapshot schema, and the
CLI's argument errors (raised before any device work)."""

import numpy as np
import pytest

from conftest import GOLDEN
from jathenax_amd import frontend, problems


def test_seed1618_deck_is_the_turbulence_defaults():
    """The reference's deck inputs/turbulence.toml equals the defaults but for the output
    file (test_input.py:11-16)."""
    settings = frontend.read_input(GOLDEN / "turbulence.toml")
    want = frontend.defaults("turbulence")
    want["output"]["file"] = "runs/driven_seed1618.npz"
    assert settings == want


@pytest.mark.parametrize(("contents", "message"), [
    ("[hydro]\nviscosty = 0", "unknown input setting hydro.viscosty"),
    ("[mesh]\nnx = true", "mesh.nx expects int"),
    ("[turb_driving]\nrseed = 1", "unsupported input section"),
    ("[problem]\nname = \"sod\"", "problem.name must be"),
    ("problem = 3", "must be a table"),
])
def test_deck_rejects_invalid_settings(tmp_path, contents, message):
    path = tmp_path / "input.toml"
    path.write_text(contents)
    with pytest.raises(ValueError, match=message):
        frontend.read_input(path)


def test_deck_overrides_and_float_settings_take_ints(tmp_path):
    path = tmp_path / "input.toml"
    path.write_text('[problem]\nname = "kh"\n[mesh]\nnx = 8\n[time]\ntlim = 1\n'
                    '[hydro]\nreconstruct = "plm"\n')
    s = frontend.read_input(path)
    assert s["problem"]["name"] == "kh" and s["mesh"] == dict(nx=8, ny=128)
    assert s["time"]["tlim"] == 1 and s["hydro"]["reconstruct"] == "plm"
    # The positional problem wins over the deck's [problem] name.
    assert frontend.read_input(path, "smooth")["problem"]["name"] == "smooth"
    with pytest.raises(ValueError, match="unknown input setting time.dt"):
        frontend.with_overrides(s, time=dict(dt=0.1))


@pytest.mark.parametrize(("section", "key", "value", "message"), [
    ("mesh", "nx", 5, "at least six"),
    ("time", "nlim", 0, "32-bit cycle limit"),
    ("time", "nlim", 2 ** 31, "32-bit cycle limit"),
    ("time", "tlim", float("inf"), "time.tlim must be finite"),
    ("time", "cfl_number", 0.0, "time.cfl_number must be finite"),
    ("hydro", "dfloor", -1.0, "hydro.dfloor must be finite"),
    ("time", "integrator", "rk5", "rk2, rk3, or rk4"),
    ("hydro", "rsolver", "hllc", "invalid hydro.reconstruct"),
    ("sgs", "weights", "w.npz", "SGS weights require a model"),
    ("output", "file", "run.h5", "must end in .npz"),
])
def test_validation_matches_the_cli(section, key, value, message):
    s = frontend.defaults("kh")
    s[section][key] = value
    with pytest.raises(ValueError, match=message):
        frontend.validate(s)


def test_cell_centres_are_the_clis():
    for n in (6, 12, 48, 64, 100, 4096):
        dx = 1.0 / n
        np.testing.assert_array_equal(problems.cell_centres(n), (np.arange(n) + 0.5) * dx - 0.5)


def test_initial_states_match_the_references(ref_vectors):
    """The reference's smooth and KH initial states (written with (i+1/2)/n - 1/2 centres,
    which differ from the CLI's by at most an ulp)."""
    np.testing.assert_allclose(frontend.initial_state("smooth", 64, 48),
                               ref_vectors["run_smooth_U0"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(frontend.initial_state("kh", 32, 32), ref_vectors["run_kh_U0"],
                               rtol=0, atol=1e-15)
    rest = frontend.initial_state("turbulence", 10, 8)
    assert rest.shape == (8, 10, 3)
    np.testing.assert_array_equal(rest[..., 0], 1.0)
    np.testing.assert_array_equal(rest[..., 1:], 0.0)


def test_snapshot_schema(tmp_path):
    U0 = frontend.initial_state("kh", 12, 10)
    x, y = problems.cell_centres(12), problems.cell_centres(10)
    path = frontend.save_snapshot(tmp_path / "sub" / "kh.npz", x, y, U0, U0 * 2, np.float64(0.5),
                                  np.float64(0.01), 7, "kh", "")
    with np.load(path) as data:
        assert sorted(data.files) == sorted(["x", "y", "initial", "conserved", "time", "dt",
                                             "cycle", "problem", "sgs_model"])
        assert data["problem"].item() == "kh" and data["sgs_model"].item() == ""
        assert data["conserved"].shape == (10, 12, 4) and data["cycle"] == 7
        np.testing.assert_array_equal(data["initial"], U0)
        assert float(data["time"]) == 0.5
    with pytest.raises(ValueError, match="must end in .npz"):
        frontend.save_snapshot(tmp_path / "kh.npy", x, y, U0, U0, 0.0, 0.0, 0, "kh", "")


def test_cli_rejects_bad_arguments_before_running(capsys):
    with pytest.raises(SystemExit):
        frontend.main(["kh", "--nx", "4"])
    assert "at least six" in capsys.readouterr().err
    with pytest.raises(SystemExit):
        frontend.main(["kh", "--output", "out.txt"])
    with pytest.raises(SystemExit):
        frontend.main(["sod"])
