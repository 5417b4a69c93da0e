"""Continuous-batching server: results in a mix equal results alone; admission by HBM
footprint; poisoned requests are evicted without disturbing the others."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _requests(n=6, seed=3):
    from jathenax_amd import workloads
    from jathenax_amd.serving import Request
    reqs = workloads.serving_requests(n=n, rate=200.0, seed=seed, tmin=0.01, tmax=0.03,
                                      size_lo=64, size_hi=160, dt_update=1e-3)
    return [Request(rid=k, spec=spec, arrival=t) for k, (t, spec) in enumerate(reqs)]


def test_served_in_mix_equals_served_alone():
    from jathenax_amd.serving import Request, Server
    mix = Server(pool_bytes=1 << 28, max_instances=16).serve(_requests())
    assert all(r.status == "done" for r in mix), [(r.rid, r.status, r.info) for r in mix]
    for r in mix:
        alone = Server(pool_bytes=1 << 28, max_instances=2).serve(
            [Request(rid=r.rid, spec=r.spec, arrival=0.0)])[0]
        assert alone.status == "done"
        np.testing.assert_array_equal(alone.result[0].cpu().numpy(), r.result[0].cpu().numpy())
        assert alone.result[1:] == r.result[1:]


def test_admission_respects_the_hbm_budget():
    from jathenax_amd.serving import Server
    reqs = _requests(n=5, seed=9)
    srv = Server(pool_bytes=1 << 28, max_instances=16)
    biggest = max(srv._footprint(r.spec) for r in reqs)
    srv = Server(pool_bytes=2 * biggest + 4096, max_instances=16)
    peak = []
    orig = srv._admit

    def admit():
        orig()
        peak.append(len(srv.running))
    srv._admit = admit
    out = srv.serve(reqs)
    assert all(r.status == "done" for r in out)
    assert max(peak) <= 2


def test_poisoned_request_is_evicted():
    from jathenax_amd.serving import Server
    reqs = _requests(n=4, seed=5)
    reqs[1].spec = dict(reqs[1].spec)
    U = np.array(reqs[1].spec["U"], copy=True)
    U[3, 3, 1] = np.nan
    reqs[1].spec["U"] = U
    reqs[1].max_rounds = 600
    out = {r.rid: r for r in Server(pool_bytes=1 << 28, max_instances=8).serve(reqs)}
    assert out[1].status == "error"
    assert all(out[k].status == "done" for k in (0, 2, 3))


@pytest.mark.parametrize("threads", (True, False))
def test_large_requests_on_the_single_grid_path(threads):
    """Requests at or above big_zones cells run on the single-grid kernel beside the batch,
    in worker threads or in chunks of the server loop.  Served alone under the same
    threshold a request takes the same path and gives the same result bit for bit; against
    the batched path the result agrees to FP32 round-off (the two kernels are separate code
    that may fuse different products into adds, and forced runs reduce their kick sums in
    another order)."""
    from jathenax_amd.serving import Request, Server
    reqs = _requests(n=6, seed=11)
    sizes = sorted(r.spec["U"].shape[0] * r.spec["U"].shape[1] for r in reqs)
    big = sizes[3]
    out = Server(pool_bytes=1 << 28, max_instances=8, big_zones=big, max_big=1,
                 big_chunk=5, big_threads=threads).serve(reqs)
    assert all(r.status == "done" for r in out), [(r.rid, r.status, r.info) for r in out]
    assert {r.path for r in out} == {"grid", "batch"}
    for r in out:
        alone = Server(pool_bytes=1 << 28, max_instances=2, big_zones=big, big_chunk=7).serve(
            [Request(rid=r.rid, spec=r.spec, arrival=0.0)])[0]
        assert alone.status == "done" and alone.path == r.path
        np.testing.assert_array_equal(alone.result[0].cpu().numpy(), r.result[0].cpu().numpy())
        assert alone.result[1:] == r.result[1:]
        if r.path != "grid":
            continue
        batch = Server(pool_bytes=1 << 28, max_instances=2, big_zones=0).serve(
            [Request(rid=r.rid, spec=r.spec, arrival=0.0)])[0]
        a, b = batch.result[0].cpu().numpy(), r.result[0].cpu().numpy()
        assert batch.result[3] == r.result[3]
        assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))


def test_large_request_continues_after_a_timeout():
    """A timed-out serve() stops a large request's worker between chunks; the request keeps
    its state, the next serve() continues it, and the result equals an uninterrupted run
    bit for bit."""
    from jathenax_amd import workloads
    from jathenax_amd.serving import Request, Server
    # Kelvin-Helmholtz 128^2 to t = 1: hundreds of steps, so the worker stops long before
    # the end.
    spec = workloads.kh_spec(128, 128, 1.0, 0.5)
    big = 128 * 128
    srv = Server(pool_bytes=1 << 28, max_instances=2, big_zones=big, big_chunk=3)
    first = srv.serve([Request(rid=0, spec=spec, arrival=0.0)], timeout=0.0)
    assert first == [] and list(srv.grid_running) == [0]
    (r,) = srv.serve([])
    alone = Server(pool_bytes=1 << 28, max_instances=2, big_zones=big, big_chunk=3).serve(
        [Request(rid=0, spec=spec, arrival=0.0)])[0]
    assert r.status == alone.status == "done" and r.path == alone.path == "grid"
    np.testing.assert_array_equal(r.result[0].cpu().numpy(), alone.result[0].cpu().numpy())
    assert r.result[1:] == alone.result[1:]
