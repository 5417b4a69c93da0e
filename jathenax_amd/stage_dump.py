"""Per-stage dumps in the AthenaK stage-comparison text layout.

The reference compares every RK stage against AthenaK through text files written by its
AthenaK problem generator (validation/athenak_stages.cpp:21-96) and read back by
validation/compare_athenak.py:17-28.  One file per stage, `stage<k>.txt`:

* a header line `ny nx dx dy cs dfloor dt stage_dt` (stage_dt = beta_k dt), the grid size
  as integers and every real at 17 significant digits (round-trip exact for float64);
* then one line per cell, rows j outer and columns i inner, holding nine fields of three
  components (rho, x, y) each followed by a space: primitives, left_x, right_x, left_y,
  right_y, flux_x, flux_y, saved, conserved (the field names of the traced stage
  dictionaries, hydro.py:136-142).

`write_stages` writes the stages that `isothermal_step(..., trace=True)` returns in this
layout, so a run of this library can be compared with AthenaK fixtures by the reference's own
tooling; `read_stage` reads either kind of file back.
"""

from pathlib import Path

import numpy as np

FIELDS = ("primitives", "left_x", "right_x", "left_y", "right_y",
          "flux_x", "flux_y", "saved", "conserved")


def _host64(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def _real(x):
    return format(float(x), ".17g")


def write_stage(path, stage, dx, dy, cs, dfloor, dt, stage_dt):
    """Write one traced stage dictionary to `path` (athenak_stages.cpp:39-96)."""
    arrays = [_host64(stage[name])[..., :3] for name in FIELDS]
    ny, nx = arrays[0].shape[:2]
    for name, a in zip(FIELDS, arrays):
        if a.shape != (ny, nx, 3):
            raise ValueError(f"stage field {name} must have shape (ny, nx, >=3)")
    rows = np.concatenate(arrays, axis=-1).reshape(ny * nx, 3 * len(FIELDS))
    with Path(path).open("w") as out:
        out.write(" ".join((str(int(ny)), str(int(nx)), _real(dx), _real(dy), _real(cs),
                            _real(dfloor), _real(dt), _real(stage_dt))) + "\n")
        for row in rows:
            out.write("".join(_real(v) + " " for v in row) + "\n")


def write_stages(folder, stages, dx, dy, cs, dfloor, dt, integrator):
    """Write stage1.txt ... stageN.txt for the stages of one traced step; returns the paths."""
    from .hydro import RK_STAGES
    rows = RK_STAGES[integrator]
    if len(stages) != len(rows):
        raise ValueError(f"{integrator} has {len(rows)} stages, got {len(stages)}")
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    paths = []
    for k, (stage, coefficients) in enumerate(zip(stages, rows), 1):
        path = folder / f"stage{k}.txt"
        write_stage(path, stage, dx, dy, cs, dfloor, dt, coefficients[2] * dt)
        paths.append(path)
    return paths


def read_stage(path):
    """((dx, dy, cs, dfloor, dt, stage_dt), {field: (ny, nx, 3) float64}), as
    compare_athenak.py:21-28 reads a stage file."""
    with Path(path).open() as file:
        ny, nx, dx, dy, cs, dfloor, dt, stage_dt = map(float, file.readline().split())
        data = np.loadtxt(file, dtype=np.float64).reshape(int(ny), int(nx), len(FIELDS), 3)
    arrays = {name: data[:, :, index, :] for index, name in enumerate(FIELDS)}
    return (dx, dy, cs, dfloor, dt, stage_dt), arrays
