"""Benchmark: zone-updates/s per RK stage of the fused isothermal stage on one MI355X.

Workload (one "step" = one RK3 step = 3 fused stage launches over the grid):
    4096 x 4096, nvar = 3, PPM4 + Roe, RK3, cs = 1, CFL 0.4, dfloor 1e-10, no viscosity,
    drag or forcing; smooth-wave initial state rho = 1 + 0.2 sin(2 pi x) cos(2 pi y),
    vx = 0.1 sin(2 pi y), vy = 0.1 cos(2 pi x); FP32 storage and arithmetic (FP64 time).
The same steps are then timed in FP64 for the FP32/FP64 ratio.  The state is resident in
HBM before the timed region; the time loop runs on the device (evolve_isothermal with
nlim = steps), bracketed by torch.cuda.synchronize() on both sides.

Prints ONE JSON line.  `roofline` describes the fused stage kernel: algorithmic bytes per
launch (stage 1 reads and writes U; stages 2-3 also read the saved register: on average
(8/3)*nvar*w bytes per zone) divided by its average launch time measured with HIP events
on the launch stream.  `cpu_baseline` times the NumPy oracle on one host core.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
METRIC = "zone-updates/s per RK stage, FP32 vs FP64 at 4096²; p50/p99 ensemble latency"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096, help="grid size (n x n)")
    ap.add_argument("--dtype", choices=("f32", "f64"), default="f32")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=4096, help="CPU baseline sample grid size")
    ap.add_argument("--block-x", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="also time plain streaming kernels over arrays of the state's size "
                         "(read-two-write-one add, copy) for the practical HBM rate")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 512 MiB device-to-device copy (PMC byte calibration)")
    ap.add_argument("--tune", action="store_true",
                    help="re-run the tile autotuner for this grid (default: the tiles it "
                         "stored in jathenax_amd/tile_table.json)")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the stored tiles (the default; kept for older command lines)")
    ap.add_argument("--ensemble", type=int, default=256,
                    help="instances in the batched-ensemble leg (0 = skip)")
    ap.add_argument("--serve", type=int, default=100,
                    help="requests in the serving leg (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=3.0, help="Poisson arrival rate, 1/s")
    ap.add_argument("--configs", type=int, default=20,
                    help="steps per workload in the KH 1024^2 / turbulence 4096^2 leg (0 = skip)")
    ap.add_argument("--config-set", choices=("all", "kh", "turbulence"), default="all",
                    help="which workloads the configs leg runs")
    return ap.parse_args()


def gpu_leg(jx, U0, n, steps, warmup, block_x, rows):
    """`warmup` untimed steps, then `steps` timed steps of the device loop (wall clock
    between two synchronisations: `value`), then the same `steps` again with every stage
    launch timed by HIP events on the context stream (the roofline's launch duration).
    The events are taken in their own pass because each one costs the launch stream a few
    microseconds, which would otherwise be charged to the wall-clock figure."""
    import torch
    from jathenax_amd.device import context_for
    ctx = context_for(U0)
    ctx.set_tiles(block_x, rows)
    args = (1.0 / n, 1.0 / n, 1.0, 1e-10, "rk3", 0.4, 1.0e9)
    if warmup:
        jx.evolve_isothermal(U0, *args, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U, t, dt, cyc = jx.evolve_isothermal(U0, *args, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert cyc == steps, (cyc, steps)
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    jx.evolve_isothermal(U0, *args, steps)
    torch.cuda.synchronize()
    wall_events = time.perf_counter() - t0
    launches, kernel_ms = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    from jathenax_amd.hydro import run_params
    used = ctx.tiles_for(run_params(U0, *args[:5]))
    # The final state is checked after both legs: a first torch reduction loads its kernels,
    # which would leave the card idle between the legs.
    return dict(wall=wall, wall_events=wall_events, launches=launches, kernel_ms=kernel_ms,
                tiles=used, U=U)


def ensemble_leg(n_inst):
    """Batched ensemble: n_inst seeded instances (64..512 per axis, KH / driven turbulence,
    mixed operators, RK2/RK3, FP32) in one engine; latency = time from the start of the
    run to each instance's completion, observed between graph replays."""
    import torch
    from jathenax_amd import workloads
    from jathenax_amd.ensemble import Ensemble
    specs = workloads.ensemble_mix(n=n_inst)
    ens = Ensemble(dtype=torch.float32, max_instances=n_inst, pool_bytes=8 << 30)
    ids = [ens.add(torch.as_tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"], s["dy"],
                   s["cs"], s["dfloor"], s["integrator"], s["cfl"], s["tlim"], s["nlim"],
                   **s["kw"]) for s in specs]
    torch.cuda.synchronize()
    finish = {}
    t0 = time.perf_counter()
    # Pipelined: the next 6 rounds are queued before the flags of the previous 6 are read,
    # so the GPU never waits for the host; the timed region ends when the GPU has drained.
    ens.launch(6)
    rounds = 0   # rounds until the last instance was seen done
    while len(finish) < len(ids):
        ens.launch(6)
        done, _, _, _ = ens.poll_launched()
        rounds += 6
        now = time.perf_counter() - t0
        for k, iid in enumerate(ids):
            if k not in finish and done[iid] == 1:
                finish[k] = now
    ens.wait()
    wall = time.perf_counter() - t0
    _, _, cycles = ens.poll()
    work = sum(workloads.zone_stages(s, int(cycles[iid])) for s, iid in zip(specs, ids))
    lat = np.asarray([finish[k] for k in range(len(ids))])
    zones = sum(s["U"].shape[0] * s["U"].shape[1] for s in specs)
    # Every instance against its own single-grid run (after the timed region): unforced runs
    # agree bit for bit (tested); forced runs to round-off, since the kick sums are reduced
    # in another order.
    from jathenax_amd import run
    worst = {"unforced": 0.0, "forced": 0.0}
    same_cycles = bitwise = 0
    for k in range(len(ids)):
        s = specs[k]
        U, t, dt, cyc = ens.result(ids[k])
        Ur, tr, dtr, cr = run.evolve_isothermal(
            torch.as_tensor(s["U"], dtype=torch.float32, device="cuda"), s["dx"], s["dy"],
            s["cs"], s["dfloor"], s["integrator"], s["cfl"], s["tlim"], s["nlim"], **s["kw"])
        key = "unforced" if s["kw"].get("forcing") is None else "forced"
        worst[key] = max(worst[key], float((U - Ur).abs().max() / Ur.abs().max()))
        same_cycles += int(cyc == cr)
        bitwise += int(bool(torch.equal(U, Ur)))
    return {"instances": len(ids), "total_zones": int(zones), "rounds": rounds,
            "wall_s": wall, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "zone_updates_per_s": work / wall,
            "checked_vs_single_grid": {"instances": len(ids),
                                       "same_cycle_count": same_cycles,
                                       "bitwise_equal": bitwise,
                                       "max_rel_diff_unforced": worst["unforced"],
                                       "max_rel_diff_forced": worst["forced"]},
            "workload": "ensemble_mix seed 20261203: sizes log-uniform 64..512 per axis, KH "
                        "(tlim 0.1) / driven turbulence (tlim 0.05), recon plm/ppm4/wenoz, "
                        "rsolver llf/hlle/roe, rk2/rk3, FP32, one batched engine"}


def configs_leg(steps, which="all"):
    """Throughput of the other single-GPU workloads, device-resident loop, `steps` steps each
    after one warm-up step: Kelvin-Helmholtz 1024^2 with tracer (PPM4+Roe, RK3) in FP32 and
    FP64, and driven turbulence 4096^2 FP32 (WENO-Z+HLLE, RK2, CFL 0.3, forcing refreshed
    every step, viscosity, Smagorinsky and drag)."""
    import torch
    from jathenax_amd import problems, run, workloads
    out = {}
    U = problems.kelvin_helmholtz_tanh(1024, 1024)
    kh = (("kh_1024_f32", torch.float32), ("kh_1024_f64", torch.float64))
    for name, tdt in kh if which in ("all", "kh") else ():
        U0 = torch.tensor(U, dtype=tdt, device="cuda")
        args = (1 / 1024, 1 / 1024, 1.0, 1e-10, "rk3", 0.4, 1e9)
        run.evolve_isothermal(U0, *args, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.evolve_isothermal(U0, *args, steps)
        torch.cuda.synchronize()
        w = time.perf_counter() - t0
        out[name] = {"zone_updates_per_s": 1024 * 1024 * 3 * steps / w,
                     "ms_per_step": w / steps * 1e3}
    out["steps"] = steps
    if which == "kh":
        return out
    s = workloads.turbulence_spec(4096, 4096, 0.01, 1618, 5e-5, "rk2", "wenoz", "hlle", cfl=0.3,
                                  dtype=np.float32)
    U0 = torch.tensor(s["U"], dtype=torch.float32, device="cuda")
    args = (s["dx"], s["dy"], s["cs"], s["dfloor"], s["integrator"], s["cfl"], s["tlim"])
    run.evolve_isothermal(U0, *args, 1, **s["kw"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.evolve_isothermal(U0, *args, steps, **s["kw"])
    torch.cuda.synchronize()
    w = time.perf_counter() - t0
    out["turbulence_4096_f32"] = {"zone_updates_per_s": 4096 * 4096 * 2 * steps / w,
                                  "ms_per_step": w / steps * 1e3}
    return out


def serving_leg(n_req, rate, sample=3):
    """Continuous-batching server: Poisson arrivals, grids 128^2..2048^2, tlim U[0.1, 2].
    The `sample` smallest served requests are run again in FP64 from the same initial state
    (the same requests drawn with FP64 forcing tables) on the single-grid path, and their
    FP32 density error is reported (L1 = mean |rho32 - rho64|)."""
    import torch
    from jathenax_amd import run, workloads
    from jathenax_amd.serving import Request, Server
    reqs = [Request(rid=k, spec=spec, arrival=t) for k, (t, spec) in
            enumerate(workloads.serving_requests(n=n_req, rate=rate))]
    zones = {r.rid: r.spec["U"].shape[0] * r.spec["U"].shape[1] for r in reqs}
    keep = set(sorted(zones, key=zones.get)[:sample])
    srv = Server(pool_bytes=64 << 30, max_instances=max(8, n_req), keep_results=True)
    t0 = time.perf_counter()
    out = srv.serve(reqs)
    wall = time.perf_counter() - t0
    work = sum(workloads.zone_stages(r.spec, r.info.get("cycles", 0)) for r in out)
    rep = Server.report(out, wall, work)
    rep["workload"] = (f"{n_req} requests, Poisson rate {rate}/s, square grids log-uniform "
                       "128..2048, tlim U[0.1, 2], KH / driven turbulence, PPM4+Roe RK3, FP32; "
                       "latency = finish - arrival = queue (admitted - arrival) + service")
    got = {r.rid: r for r in out if r.rid in keep and r.status == "done"}
    for r in out:
        if r.rid not in got:
            r.result = None
    specs64 = dict(enumerate(s for _, s in workloads.serving_requests(n=n_req, rate=rate,
                                                                        dtype=np.float64)))
    errs = []
    for rid, r in sorted(got.items()):
        s = specs64[rid]
        U64, t64, _, c64 = run.evolve_isothermal(
            torch.tensor(s["U"], dtype=torch.float64, device="cuda"), s["dx"], s["dy"], s["cs"],
            s["dfloor"], s["integrator"], s["cfl"], s["tlim"], s["nlim"], **s["kw"])
        rho32 = r.result[0][..., 0].double()
        errs.append({"rid": rid, "kind": s["kind"], "n": int(s["U"].shape[1]),
                     "tlim": s["tlim"], "cycles_f32": int(r.result[3]), "cycles_f64": int(c64),
                     "l1_density": float((rho32 - U64[..., 0]).abs().mean())})
        r.result = None
    rep["fp32_vs_fp64"] = errs
    return rep


def cpu_leg(problems, n, kh_steps=5):
    """The NumPy float64 oracle on one host core: one RK3 step of the headline workload at
    n x n, and `kh_steps` steps of Kelvin-Helmholtz 1024^2 with a tracer (nvar 4)."""
    sys.path.insert(0, str(ROOT))
    from oracle import jathenax_np as ox
    U = problems.smooth_wave(n, n)
    t0 = time.perf_counter()
    ox.isothermal_step(U, 1e-4, 1.0 / n, 1.0 / n, 1.0, 1e-10, "rk3")
    sec = time.perf_counter() - t0
    del U
    kh = None
    if kh_steps:
        U = problems.kelvin_helmholtz_tanh(1024, 1024)
        t0 = time.perf_counter()
        ox.evolve_isothermal(U, 1 / 1024, 1 / 1024, 1.0, 1e-10, "rk3", 0.4, 1e9, kh_steps)
        ksec = time.perf_counter() - t0
        kh = {"value": 1024 * 1024 * 3 * kh_steps / ksec, "unit": "zone-updates/s",
              "sample": f"{kh_steps} RK3 steps of KH 1024^2 with tracer, PPM4+Roe ({ksec:.1f} s)"}
    return n * n * 3 / sec, sec, kh


def stream_leg(nelem, reps=50):
    """HBM rate of plain torch streaming kernels over float32 arrays of nelem values: the
    read-two-write-one pattern of RK stages 2-3 (c = a + b) and a copy."""
    import torch
    a = torch.rand(nelem, device="cuda")
    b = torch.rand(nelem, device="cuda")
    c = torch.empty_like(a)
    res = {}
    for name, fn, arrays in (("add_2r1w", lambda: torch.add(a, b, out=c), 3),
                             ("copy_1r1w", lambda: c.copy_(a), 2)):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / reps * 1e-3
        res[name] = {"GB_per_s": arrays * 4 * nelem / sec / 1e9, "us": sec * 1e6}
    return res


def read_traffic(n, dtype, tiles):
    """HBM bytes per stage launch for this grid, precision and tile shape, from the
    committed rocprofv3 --pmc summary; None when that configuration was not measured."""
    f = ROOT / "profiles" / "stage_traffic.json"
    if not f.exists():
        return None
    try:
        data = json.loads(f.read_text())
        return data.get(f"{n}x{n}_{dtype}_{tiles[0]}x{tiles[1]}")
    except (OSError, ValueError):
        return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the single-grid "
              "path would shard by y-slabs with halo exchange, which this build does not run)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import jathenax_amd as jx
    from jathenax_amd import problems
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    n, K = a.n, a.steps
    nvar, stages = 3, 3
    if a.calibrate_copy:
        src = torch.ones(128 * 2 ** 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # The best 4096^2 tile shapes time within a few per cent of each other, so a fresh
    # tuning run picks among them at random; by default the timed leg uses the tiles the
    # tuner stored, whose HBM traffic is measured (roofline.traffic).
    tiles = {}
    if a.tune and not a.no_tune and not (a.block_x or a.rows):
        from jathenax_amd import tuning
        for dt_name, tdt in (("f32", torch.float32), ("f64", torch.float64)):
            U0 = torch.tensor(problems.smooth_wave(n, n), dtype=tdt, device="cuda")
            choice, _ = tuning.tune_single(U0, 1.0 / n, 1.0 / n, 1.0, 1e-10, "rk3", steps=4,
                                           repeats=3)
            tiles[dt_name] = list(choice)
            del U0
        print(f"# bench: tuned tiles {tiles}", file=sys.stderr, flush=True)
    # Both initial states are made (on the host) and put in HBM before any timed work.  The
    # FP64 leg (the reference point of the FP32/FP64 ratio) runs first and the FP32 leg
    # straight after it: a card left idle for a while needs some tens of milliseconds of
    # load before its clocks settle, and the FP32 leg is the headline figure.
    U_host = problems.smooth_wave(n, n)
    U_main = torch.tensor(U_host, dtype=torch.float32 if a.dtype == "f32" else torch.float64,
                          device="cuda")
    U_f64 = None
    if not a.no_fp64 and a.dtype == "f32":
        U_f64 = torch.tensor(U_host, dtype=torch.float64, device="cuda")
    del U_host
    torch.cuda.synchronize()
    f64 = None
    if U_f64 is not None:
        print(f"# bench: fp64 leg, steps={K} warmup={a.warmup}", file=sys.stderr, flush=True)
        f64 = gpu_leg(jx, U_f64, n, K, a.warmup, a.block_x, a.rows)
    print(f"# bench: {n}x{n} {a.dtype} steps={K} warmup={a.warmup}", file=sys.stderr, flush=True)
    main_leg = gpu_leg(jx, U_main, n, K, a.warmup, a.block_x, a.rows)
    zu = n * n * stages * K / main_leg["wall"]
    w = 4 if a.dtype == "f32" else 8
    bytes_per_launch = (8.0 / 3.0) * nvar * w * n * n
    avg_launch_s = main_leg["kernel_ms"] / 1e3 / max(1, main_leg["launches"])
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC, "value": zu, "unit": "zone-updates/s", "n_gpus": 1, "steps": K,
        "warmup": a.warmup, "ms_per_step": main_leg["wall"] / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.dtype, "data": "synthetic",
        "config": {"workload": f"{n}x{n} smooth wave, PPM4+Roe, RK3, nvar=3, CFL 0.4, "
                               f"device-resident time loop", "nx": n, "ny": n,
                   "reconstruct": "ppm4", "rsolver": "roe", "integrator": "rk3",
                   "stages_per_step": stages, "tiles": tiles or "tile table",
                   "tiles_used": list(main_leg["tiles"])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(n, a.dtype, main_leg["tiles"]),
                     "kernel": "jx::stage_kernel (fused RK stage)",
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3, "launches_timed": main_leg["launches"],
                     "ms_per_step_with_launch_events": main_leg["wall_events"] / K * 1e3},
        "finite": bool(torch.isfinite(main_leg["U"]).all()),
    }
    if f64 is not None:
        zu64 = n * n * stages * K / f64["wall"]
        out["fp64"] = {"value": zu64, "ms_per_step": f64["wall"] / K * 1e3,
                       "avg_launch_ms": f64["kernel_ms"] / max(1, f64["launches"]),
                       "fp32_over_fp64": zu / zu64,
                       "fp32_vs_fp64_l1_density_after_steps":
                           float((main_leg["U"][..., 0].double() - f64["U"][..., 0]).abs().mean())}
    del main_leg["U"], U_main
    if f64 is not None:
        del f64["U"], U_f64
    if a.configs > 0:
        print("# bench: other workloads", file=sys.stderr, flush=True)
        out["configs"] = configs_leg(a.configs, a.config_set)
    if a.ensemble > 0:
        print(f"# bench: ensemble leg ({a.ensemble} instances)", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_leg(a.ensemble)
    if a.serve > 0:
        print(f"# bench: serving leg ({a.serve} requests)", file=sys.stderr, flush=True)
        out["serving"] = serving_leg(a.serve, a.serve_rate)
    if a.stream_ceiling:
        print("# bench: streaming ceiling", file=sys.stderr, flush=True)
        out["stream_ceiling"] = stream_leg(n * n * nvar)
    if not a.no_cpu_baseline:
        print("# bench: cpu baseline", file=sys.stderr, flush=True)
        rate, sec, kh = cpu_leg(problems, a.cpu_n)
        out["cpu_baseline"] = {
            "value": rate, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": f"NumPy float64 oracle, one RK3 step of the same workload on a "
                      f"{a.cpu_n}x{a.cpu_n} grid ({sec:.1f} s; host has {os.cpu_count()} cores, "
                      f"NumPy elementwise ops use one)",
            "kh_1024": kh}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
