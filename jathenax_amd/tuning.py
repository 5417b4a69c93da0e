"""Tile-shape autotuner for the fused stage kernel and the batched engine.

Single grid: the stage kernel's tile is (block width in threads, rows per block).  The
default derives rows from occupancy so the grid is one wave of resident blocks; the
tuner times a few candidates with HIP events on the launch stream and keeps the fastest
for that (shape, precision, operators) key.

Batch: each instance's tile height (rows of its 58-column strips) trades the 6-row
warm-up of every tile against load balance.  The tuner times whole ensemble rounds per
grid-size class and keeps the best height per class.

Choices live in a table (jathenax_amd/tile_table.json holds measured defaults); tuning at
run time updates the in-memory table, and save() writes it back.
"""

import json
import time
from pathlib import Path

import numpy as np
import torch

TABLE_PATH = Path(__file__).resolve().parent / "tile_table.json"
SIZE_CLASSES = (96, 192, 384, 768)   # upper bounds of max(nx, ny) per batch class

_table = None


def table():
    global _table
    if _table is None:
        try:
            _table = json.loads(TABLE_PATH.read_text())
        except (OSError, ValueError):
            _table = {}
        _table.setdefault("single", {})
        _table.setdefault("batch_rows", {})
    return _table


def save(path=TABLE_PATH):
    Path(path).write_text(json.dumps(table(), indent=1, sort_keys=True) + "\n")


def single_key(p):
    """Key of a RunParams for the single-grid table."""
    visc = int(p.viscosity > 0.0 or p.smagorinsky_cs > 0.0)
    prec = "f32" if p.dtype == 1 else "f64"
    return f"{p.nx}x{p.ny}x{p.nvar}:{prec}:r{p.reconstruct}s{p.rsolver}v{visc}"


def apply_single(ctx, p):
    """Set the context's tiles from the table (auto when the shape was never tuned),
    unless the caller pinned a tile shape with ctx.set_tiles."""
    if getattr(ctx, "pinned", False):
        return
    bx, rows = table()["single"].get(single_key(p), (0, 0))
    ctx._apply_tiles(bx, rows)


def size_class(nx, ny):
    n = max(nx, ny)
    for k, hi in enumerate(SIZE_CLASSES):
        if n <= hi:
            return k
    return len(SIZE_CLASSES)


def batch_rows(nx, ny, default=32):
    return int(table()["batch_rows"].get(str(size_class(nx, ny)), default))


def _single_candidates(nx, ny):
    # Block counts near a whole number of resident waves matter more than the shape
    # itself (a small second wave of blocks costs a whole block time), hence the dense
    # bands of heights around one full wave of 4096-wide grids at 4 and 5 resident
    # blocks of 256 threads per CU.
    out = [(0, 0)]
    for bx in (256, 128):
        if bx > 64 * ((nx + 6 + 63) // 64) and bx != 128:
            continue
        for rows in (8, 10, 12, 16, 24, 28, 32, 35, 48, 55, 56, 58, 60, 64, 68, 69, 70, 80,
                     96):
            if rows <= ny:
                out.append((bx, rows))
    return out


def tune_single(U, dx, dy, cs, dfloor, integrator, *, steps=3, repeats=2, candidates=None,
                **kw):
    """Time candidate tiles on evolve_isothermal(U, ..., nlim=steps); store and return
    (best (bx, rows), {candidate: seconds per stage launch})."""
    from . import run as _run
    from .device import as_device, context_for
    from .hydro import run_params
    U = as_device(U)
    ctx = context_for(U)
    p = run_params(U, dx, dy, cs, dfloor, integrator, **{k: v for k, v in kw.items()
                                                       if k in ("reconstruct", "rsolver",
                                                                "viscosity", "smagorinsky_cs",
                                                                "drag_rate")})
    key = single_key(p)
    results = {}
    was_pinned = ctx.pinned
    for cand in candidates or _single_candidates(p.nx, p.ny):
        ctx._apply_tiles(*cand)
        # One untimed step first: the first launches after a tile change run cold.
        _run._evolve_raw(U, dx, dy, cs, dfloor, integrator, 0.4, 1e9, 1, **kw)
        best = float("inf")
        for _ in range(repeats):
            torch.cuda.synchronize()
            ctx.set_kernel_timing(True)
            _run._evolve_raw(U, dx, dy, cs, dfloor, integrator, 0.4, 1e9, steps, **kw)
            n, ms = ctx.kernel_timing()
            ctx.set_kernel_timing(False)
            best = min(best, ms / max(n, 1) / 1e3)
        results[cand] = best
    choice = min(results, key=results.get)
    table()["single"][key] = list(choice)
    ctx.pinned = was_pinned
    ctx._apply_tiles(*choice)
    return choice, results


def tune_batch(dtype=torch.float32, candidates=(16, 24, 32, 48, 64, 96), n_inst=24,
               rounds=24, seed=11):
    """Best tile height per size class, from timed rounds of a same-class ensemble."""
    from . import problems
    from .ensemble import Ensemble
    rng = np.random.default_rng(seed)
    bounds = (48,) + SIZE_CLASSES
    chosen = {}
    for k in range(len(SIZE_CLASSES)):
        lo, hi = bounds[k] + 1, bounds[k + 1]
        sizes = [(int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1)))
                 for _ in range(n_inst)]
        timing = {}
        for rows in candidates:
            ens = Ensemble(dtype=dtype, max_instances=n_inst, pool_bytes=1 << 31)
            for nx, ny in sizes:
                ens.add(torch.tensor(problems.kelvin_helmholtz_tanh(nx, ny), dtype=dtype,
                                     device="cuda"), 1 / nx, 1 / ny, 1.0, 1e-10, "rk3", 0.4, 1e9,
                        10 ** 9, reconstruct="ppm4", rsolver="roe", rows=rows)
            ens.run(6)
            t0 = time.perf_counter()
            ens.run(rounds)
            timing[rows] = (time.perf_counter() - t0) / rounds
            del ens
        best = min(timing, key=timing.get)
        chosen[k] = (best, timing)
        table()["batch_rows"][str(k)] = best
    return chosen
