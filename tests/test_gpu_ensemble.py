"""Batched ensemble engine: each instance equals its own single-grid run, and an instance's
result does not depend on the rest of the batch (bitwise)."""

import numpy as np
import pytest
import torch

from oracle import jathenax_np as ox

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def jx():
    import jathenax_amd
    assert torch.cuda.is_available()
    return jathenax_amd


def _specs():
    """A mixed set: shapes (square and not, ragged strips), operators, integrators,
    scalars, viscosity/Smagorinsky/drag and one forced instance."""
    from jathenax_amd import forcing as jf, problems
    rng = np.random.default_rng(7)
    out = []
    combos = [((64, 48), "ppm4", "roe", "rk3", 3, {}),
              ((100, 70), "wenoz", "hlle", "rk2", 4, dict(viscosity=1e-3, smagorinsky_cs=0.17)),
              ((128, 128), "plm", "llf", "rk3", 4, dict(drag_rate=0.5)),
              ((58, 131), "ppmx", "roe", "rk4", 3, dict(viscosity=2e-3)),
              ((256, 64), "dc", "hlle", "rk3", 3, {}),
              ((80, 80), "ppm4", "llf", "rk2", 4, dict(smagorinsky_cs=0.2, drag_rate=0.3))]
    for (nx, ny), recon, rs, integ, nv, phys in combos:
        if nv == 4:
            U = problems.kelvin_helmholtz_reference(nx, ny)
        else:
            U = problems.smooth_reference(nx, ny)
        U = U + 0.01 * rng.standard_normal(U.shape) * (np.arange(U.shape[-1]) > 0)
        out.append(dict(U=U, dx=1 / nx, dy=1 / ny, integ=integ, kw=dict(reconstruct=recon,
                                                                        rsolver=rs, **phys)))
    # forced instances of every kick arrangement: with Smagorinsky the final stage writes a
    # scratch register (RK3: B, with a tracer the kick carries; RK4: C, beside S) and the
    # timestep pass kicks it into A; without Smagorinsky the kick is a pass of its own
    modes = jf.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    for (nx, ny), recon, rs, integ, nv, phys, seed in (
            ((66, 52), "ppm4", "roe", "rk3", 4, dict(viscosity=1e-3, smagorinsky_cs=0.17), 11),
            ((60, 70), "plm", "hlle", "rk4", 3, dict(smagorinsky_cs=0.2, drag_rate=0.3), 12),
            ((64, 64), "ppm4", "llf", "rk3", 3, dict(viscosity=1e-3), 13)):
        coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, seed, 0.75, 0.002, 40)
        basis = jf.forcing_basis(modes, problems.cell_centres(nx), problems.cell_centres(ny),
                                 (1.0, 1.0))
        U = problems.kelvin_helmholtz_reference(nx, ny) if nv == 4 else problems.at_rest(nx, ny)
        out.append(dict(U=U, dx=1 / nx, dy=1 / ny, integ=integ,
                        kw=dict(reconstruct=recon, rsolver=rs, forcing=(coef, basis),
                                dt_update=0.002, tdriv_start=0.0, dedt=0.5, **phys)))
    # forced instance
    nx, ny = 72, 60
    coef = jf.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, 1618, 0.75, 0.002, 40)
    basis = jf.forcing_basis(modes, problems.cell_centres(nx), problems.cell_centres(ny),
                             (1.0, 1.0))
    out.append(dict(U=problems.at_rest(nx, ny), dx=1 / nx, dy=1 / ny, integ="rk2",
                    kw=dict(reconstruct="wenoz", rsolver="hlle", viscosity=1.5e-3,
                            smagorinsky_cs=0.17, drag_rate=0.3333335, forcing=(coef, basis),
                            dt_update=0.002, tdriv_start=0.0, dedt=0.5)))
    return out


@pytest.mark.parametrize("graph", (True, False))
def test_instances_match_single_grid_runs_fp64(jx, graph):
    from jathenax_amd.ensemble import Ensemble
    specs = _specs()
    ens = Ensemble(dtype=torch.float64, max_instances=16, pool_bytes=1 << 28, graph=graph)
    ids = []
    for s in specs:
        ids.append(ens.add(torch.tensor(s["U"], device="cuda"), s["dx"], s["dy"], 1.0, 1e-10,
                           s["integ"], 0.4, 0.06, 10 ** 6, **s["kw"]))
    rounds = ens.run_until_done(chunk=12)
    assert rounds > 0
    for iid, s in zip(ids, specs):
        U, t, dt, cyc = ens.result(iid)
        Ur, tr, dtr, cr = jx.evolve_isothermal(torch.tensor(s["U"], device="cuda"), s["dx"],
                                               s["dy"], 1.0, 1e-10, s["integ"], 0.4, 0.06, 10 ** 6,
                                               **s["kw"])
        assert cyc == cr and t == tr, (iid, cyc, cr, t, tr)
        if "forcing" in s["kw"]:
            assert _rel(_np(U), _np(Ur)) <= 1e-12
        else:
            np.testing.assert_array_equal(_np(U), _np(Ur))


def test_instances_match_single_grid_runs_fp32(jx):
    """FP32: both engines evaluate the same operator expressions with products fused only
    within an expression, so unforced instances match their single-grid runs bit for bit.
    Forced ones reduce the kick sums in another order (and the single grid takes the
    viscous kick's second moments from one pass): round-off agreement."""
    from jathenax_amd.ensemble import Ensemble
    specs = _specs()
    ens = Ensemble(dtype=torch.float32, max_instances=16, pool_bytes=1 << 28)
    ids = [ens.add(torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"], s["dy"],
                   1.0, 1e-10, s["integ"], 0.4, 0.06, 10 ** 6, **s["kw"]) for s in specs]
    ens.run_until_done(chunk=12)
    for iid, s in zip(ids, specs):
        U, t, dt, cyc = ens.result(iid)
        Ur, tr, dtr, cr = jx.evolve_isothermal(
            torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"], s["dy"], 1.0,
            1e-10, s["integ"], 0.4, 0.06, 10 ** 6, **s["kw"])
        assert cyc == cr, (iid, cyc, cr)
        if "forcing" in s["kw"]:
            assert _rel(_np(U), _np(Ur)) <= 1e-5
        else:
            assert t == tr
            np.testing.assert_array_equal(_np(U), _np(Ur), err_msg=str(s["kw"]))


def test_instance_alone_equals_instance_in_mix_fp32(jx):
    from jathenax_amd.ensemble import Ensemble
    specs = _specs()
    target = specs[-1]   # the forced one: its kick sums must not see the other instances

    def run(with_others):
        ens = Ensemble(dtype=torch.float32, max_instances=16, pool_bytes=1 << 27)
        if with_others:
            for s in specs[:3]:
                ens.add(torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"],
                        s["dy"], 1.0, 1e-10, s["integ"], 0.4, 0.05, 10 ** 6, **s["kw"])
        iid = ens.add(torch.tensor(target["U"], dtype=torch.float32, device="cuda"),
                      target["dx"], target["dy"], 1.0, 1e-10, target["integ"], 0.4, 0.05, 10 ** 6,
                      **target["kw"])
        if with_others:
            for s in specs[3:-1]:
                ens.add(torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"],
                        s["dy"], 1.0, 1e-10, s["integ"], 0.4, 0.05, 10 ** 6, **s["kw"])
        ens.run_until_done(chunk=12)
        return ens.result(iid)

    a, b = run(False), run(True)
    np.testing.assert_array_equal(_np(a[0]), _np(b[0]))
    assert a[1:] == b[1:]


def test_ensemble_against_oracle(jx, ref_vectors):
    from jathenax_amd.ensemble import Ensemble
    ens = Ensemble(dtype=torch.float64, max_instances=4, pool_bytes=1 << 26)
    U0 = ref_vectors["run_smooth_U0"]
    a = ens.add(torch.tensor(U0, device="cuda"), 1 / 64, 1 / 48, 1.0, 1e-10, "rk3", 0.4, 10.0, 100)
    b = ens.add(torch.tensor(ref_vectors["run_kh_U0"], device="cuda"), 1 / 32, 1 / 32, 1.0, 1e-10,
                "rk2", 0.4, 0.05, 1000, viscosity=1e-3, smagorinsky_cs=0.17, drag_rate=0.5,
                reconstruct="wenoz", rsolver="hlle")
    ens.run_until_done(chunk=12)
    U, t, dt, cyc = ens.result(a)
    assert cyc == 100 and _rel(_np(U), ref_vectors["run_smooth_n100_U"]) <= 1e-12
    U, t, dt, cyc = ens.result(b)
    assert _rel(_np(U), ref_vectors["run_kh_rk2_wenoz_hlle_U"]) <= 1e-12


def test_remove_and_reuse_slots(jx):
    from jathenax_amd import problems
    from jathenax_amd.ensemble import Ensemble
    ens = Ensemble(dtype=torch.float32, max_instances=3, pool_bytes=1 << 24)
    U = torch.tensor(problems.smooth_reference(64, 64), dtype=torch.float32, device="cuda")
    ids = [ens.add(U, 1 / 64, 1 / 64, 1.0, 1e-10, "rk3", 0.4, 0.02, 100) for _ in range(3)]
    with pytest.raises(ValueError):
        ens.add(U, 1 / 64, 1 / 64, 1.0, 1e-10, "rk3", 0.4, 0.02, 100)
    ens.run_until_done(chunk=6)
    ref = ens.result(ids[0])
    ens.remove(ids[1])
    again = ens.add(U, 1 / 64, 1 / 64, 1.0, 1e-10, "rk3", 0.4, 0.02, 100)
    assert again == ids[1]
    ens.run_until_done(chunk=6)
    got = ens.result(again)
    np.testing.assert_array_equal(_np(got[0]), _np(ref[0]))
    with pytest.raises(KeyError):
        ens.add(U, 1 / 64, 1 / 64, 1.0, 1e-10, "rk3", 0.4, 0.02, 100, reconstruct="weno9")


def test_results_do_not_depend_on_tile_height(jx):
    """Forced and Smagorinsky instances (kick sums, CFL and viscous limits) give bitwise the
    same result whatever band height the rounds cut them into."""
    from jathenax_amd.ensemble import Ensemble
    specs = [_specs()[-1], _specs()[1]]

    def run(rows, fill):
        ens = Ensemble(dtype=torch.float32, max_instances=4, pool_bytes=1 << 26)
        ens.set_fill(fill)
        ids = [ens.add(torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"],
                       s["dy"], 1.0, 1e-10, s["integ"], 0.4, 0.03, 10 ** 6, rows=rows, **s["kw"])
               for s in specs]
        ens.run_until_done(chunk=12)
        return [ens.result(i) for i in ids]

    ref = run(8, 2)
    for rows, fill in ((64, 0), (24, 0), (8, 0), (64, 2)):
        got = run(rows, fill)
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(_np(a[0]), _np(b[0]))
            assert a[1:] == b[1:]


def test_pipelined_polling(jx):
    """Every launch snapshots the loop state at its end; poll_launched returns the snapshots
    in launch order while later launches run.  They equal a synchronous run's state after the
    same rounds, and the results of a pipelined run equal a synchronous run's bit for bit."""
    from jathenax_amd.ensemble import Ensemble
    specs = _specs()[:4] + _specs()[-2:]

    def make():
        ens = Ensemble(dtype=torch.float32, max_instances=8, pool_bytes=1 << 27)
        ids = [ens.add(torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"],
                       s["dy"], 1.0, 1e-10, s["integ"], 0.4, 0.04, 10 ** 6, **s["kw"])
               for s in specs]
        return ens, ids

    a, ida = make()
    a.run(6)
    _, t6, c6 = a.poll()
    a.run(6)
    _, t12, c12 = a.poll()
    b, idb = make()
    b.launch(6)
    b.launch(6)
    _, s6, k6, _ = b.poll_launched()
    _, s12, k12, _ = b.poll_launched()
    assert list(k6[idb]) == list(c6[ida]) and list(k12[idb]) == list(c12[ida])
    assert list(s6[idb]) == list(t6[ida]) and list(s12[idb]) == list(t12[ida])
    with pytest.raises(ValueError):
        b.poll_launched()
    # wait() settles every launch, so the snapshot polled next is the next launch's, also
    # when the launch before the wait was never polled.
    b.launch(6)
    b.wait()
    b.launch(6)
    _, s24, k24, _ = b.poll_launched()
    a.run(6)
    a.run(6)
    _, t24, c24 = a.poll()
    assert list(k24[idb]) == list(c24[ida]) and list(s24[idb]) == list(t24[ida])
    a.run_until_done(chunk=6)
    b.launch(6)
    while True:
        b.launch(6)
        done, _, _, active = b.poll_launched()
        if active == 0:
            break
    b.wait()
    for i, j in zip(ida, idb):
        ra, rb = a.result(i), b.result(j)
        np.testing.assert_array_equal(_np(ra[0]), _np(rb[0]))
        assert ra[1:] == rb[1:]


def test_polled_snapshot_of_a_reused_slot(jx):
    """A launch's snapshot taken before an instance was added to a slot shows the slot's
    previous occupant: poll_launched reports such a slot running, at time NaN and cycle -1,
    and the first snapshot taken after the add reports the new instance."""
    from jathenax_amd.ensemble import Ensemble
    s = _specs()[0]
    ens = Ensemble(dtype=torch.float32, max_instances=1, pool_bytes=1 << 26)

    def add():
        return ens.add(torch.tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"],
                       s["dy"], 1.0, 1e-10, s["integ"], 0.4, 0.05, 10 ** 6, **s["kw"])

    a = add()
    ens.launch(6)
    ens.launch(6)
    ens.remove(a)
    b = add()
    assert b == a
    for _ in range(2):   # the two snapshots taken before the add
        done, t, c, _ = ens.poll_launched()
        assert done[b] == 0 and np.isnan(t[b]) and c[b] == -1
    ens.launch(6)
    done, t, c, _ = ens.poll_launched()
    assert done[b] in (0, 1) and np.isfinite(t[b]) and c[b] >= 0
    ens.wait()
