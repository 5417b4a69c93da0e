"""Forced-run resume: a run checkpointed (state + loop state) and resumed continues bit for
bit as the uninterrupted run, on the single-grid path, in the batched engine (mid-step and
end-of-step checkpoints) and under the server's pre-emption."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def jx():
    import jathenax_amd
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    return jathenax_amd


def _driven(nx, ny, dtype, tlim=0.05):
    from jathenax_amd import workloads
    s = workloads.turbulence_spec(nx, ny, tlim, 77, 0.002, "rk3", "wenoz", "hlle", cfl=0.4,
                                  dtype=dtype)
    s["kw"]["dedt"] = 0.5
    return s


@pytest.mark.parametrize("dtype", (torch.float64, torch.float32))
def test_single_grid_resume_is_bitwise(jx, dtype):
    from jathenax_amd.run import LoopState, evolve_resumable
    s = _driven(48, 40, np.float64 if dtype == torch.float64 else np.float32, tlim=0.6)
    U0 = torch.tensor(s["U"], dtype=dtype, device="cuda")
    args = (s["dx"], s["dy"], s["cs"], s["dfloor"], s["integrator"], s["cfl"], s["tlim"])
    full, st_full = evolve_resumable(U0, *args, 40, **s["kw"])
    half, st_half = evolve_resumable(U0, *args, 17, **s["kw"])
    assert st_half.cycle == 17 and st_half.update >= 1
    rest, st_rest = evolve_resumable(half, *args, 40, state=st_half, **s["kw"])
    np.testing.assert_array_equal(_np(rest), _np(full))
    assert st_rest == st_full
    # A fresh resumable run is the plain device loop.
    plain = jx.evolve_isothermal(U0, *args, 40, **s["kw"])
    np.testing.assert_array_equal(_np(plain[0]), _np(full))
    assert (plain[1], plain[2], plain[3]) == (st_full.time, st_full.dt, st_full.cycle)
    # Nothing left to do: the state passes through.
    same, st_same = evolve_resumable(full, *args, 40, state=st_full, **s["kw"])
    np.testing.assert_array_equal(_np(same), _np(full))
    assert st_same == st_full
    assert LoopState() == LoopState(0.0, math.inf, 0, -1)


@pytest.mark.parametrize("rounds", (7, 9, 13))
def test_batch_checkpoint_and_resume(jx, rounds):
    """rk3: 7 and 13 rounds stop mid-step, 9 right after a final stage."""
    from jathenax_amd import problems
    from jathenax_amd.ensemble import Ensemble
    s = _driven(72, 60, np.float64)
    kh = problems.kelvin_helmholtz_tanh(64, 48)
    runs = [(torch.tensor(s["U"], device="cuda"), s["dx"], s["dy"], s["cs"], s["dfloor"],
             "rk3", 0.4, 0.03, 10 ** 6, s["kw"]),
            (torch.tensor(kh, device="cuda"), 1 / 64, 1 / 48, 1.0, 1e-10, "rk2", 0.4, 0.04,
             10 ** 6, dict(reconstruct="ppm4", rsolver="roe", viscosity=1e-3))]

    def full():
        ens = Ensemble(dtype=torch.float64, max_instances=4, pool_bytes=1 << 26)
        ids = [ens.add(U, *a, **kw) for (U, *a, kw) in runs]
        ens.run_until_done(chunk=12)
        return [ens.result(i) for i in ids]

    ref = full()
    ens = Ensemble(dtype=torch.float64, max_instances=4, pool_bytes=1 << 26, graph_rounds=1)
    ids = [ens.add(U, *a, **kw) for (U, *a, kw) in runs]
    ens.run(rounds)
    ckpts = [ens.checkpoint(i) for i in ids]
    for i in ids:
        ens.remove(i)
    ens2 = Ensemble(dtype=torch.float64, max_instances=4, pool_bytes=1 << 26)
    ids2 = []
    for (Uc, st), (_, *a, kw) in zip(ckpts, runs):
        assert st.cycle >= 1
        ids2.append(ens2.add(Uc, *a, state=st, **kw))
    ens2.run_until_done(chunk=12)
    for i, r in zip(ids2, ref):
        got = ens2.result(i)
        np.testing.assert_array_equal(_np(got[0]), _np(r[0]))
        assert got[1:] == r[1:]


def test_server_preemption_keeps_results(jx):
    from jathenax_amd import workloads
    from jathenax_amd.serving import Request, Server
    # The low-priority run lasts ~500 steps (1500 rounds); the urgent one arrives 10 ms in.
    low = workloads.turbulence_spec(96, 96, 2.0, 5, 0.01, "rk3", "ppm4", "roe",
                                    dtype=np.float32)
    low["kw"]["dedt"] = 0.5
    high = workloads.kh_spec(96, 96, 0.02, 0.5)
    probe = Server(pool_bytes=1 << 26, max_instances=4)
    one = max(probe._footprint(low), probe._footprint(high))
    srv = Server(pool_bytes=one + 4096, max_instances=4, rounds_per_poll=2)
    reqs = [Request(rid=0, spec=low, arrival=0.0, priority=0),
            Request(rid=1, spec=high, arrival=0.01, priority=5)]
    out = {r.rid: r for r in srv.serve(reqs)}
    assert all(r.status == "done" for r in out.values())
    assert out[0].preemptions >= 1
    alone = Server(pool_bytes=one + 4096, max_instances=4).serve(
        [Request(rid=0, spec=low, arrival=0.0)])[0]
    np.testing.assert_array_equal(_np(alone.result[0]), _np(out[0].result[0]))
    assert alone.result[1:] == out[0].result[1:]
