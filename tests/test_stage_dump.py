"""AthenaK stage-dump layout (validation/athenak_stages.cpp:39-96, compare_athenak.py:17-28):
writing and reading back are exact, and the GPU traced stages round-trip bit for bit."""

import numpy as np
import pytest
import torch

from jathenax_amd import stage_dump


def _stage(rng, ny, nx, nv):
    return {name: rng.standard_normal((ny, nx, nv)) * 10.0 ** rng.integers(-12, 12, (ny, nx, nv))
            for name in stage_dump.FIELDS}


def test_write_read_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(4)
    stage = _stage(rng, 5, 7, 4)
    stage["conserved"][0, 0, 0] = 1.0 / 3.0
    stage["saved"][1, 2, 1] = -0.0
    stage_dump.write_stage(tmp_path / "stage1.txt", stage, 0.1, 1 / 7, 1.3, 1e-10, 1e-3, 2.5e-4)
    meta, arrays = stage_dump.read_stage(tmp_path / "stage1.txt")
    assert meta == (0.1, 1 / 7, 1.3, 1e-10, 1e-3, 2.5e-4)
    for name in stage_dump.FIELDS:
        assert arrays[name].shape == (5, 7, 3)
        np.testing.assert_array_equal(arrays[name].view(np.uint64),
                                      stage[name][..., :3].view(np.uint64))


def test_layout_is_the_athenak_one(tmp_path):
    # header "ny nx dx dy cs dfloor dt stage_dt", then per cell 9 fields x 3 components,
    # each value followed by a space, rows j outer, columns i inner
    ny, nx = 2, 3
    stage = {name: np.arange(ny * nx * 3, dtype=np.float64).reshape(ny, nx, 3) + 100 * k
             for k, name in enumerate(stage_dump.FIELDS)}
    stage_dump.write_stage(tmp_path / "s.txt", stage, 0.5, 0.25, 1.0, 1e-10, 0.001, 0.001)
    lines = (tmp_path / "s.txt").read_text().split("\n")
    assert lines[0] == "2 3 0.5 0.25 1 1e-10 0.001 0.001"
    assert len(lines) == 1 + ny * nx + 1 and lines[-1] == ""
    cell = lines[1 + 1 * nx + 2].split(" ")            # j = 1, i = 2
    assert cell[-1] == "" and len(cell) == 27 + 1
    values = [float(v) for v in cell[:-1]]
    for k in range(9):
        assert values[3 * k:3 * k + 3] == [100 * k + (1 * nx + 2) * 3 + c for c in range(3)]


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ("rk2", "rk3", "rk4"))
@pytest.mark.parametrize("dtype", (torch.float64, torch.float32))
def test_traced_stages_round_trip(tmp_path, integrator, dtype):
    import jathenax_amd as jx
    rng = np.random.default_rng(8)
    U = np.empty((12, 10, 4))
    U[..., 0] = 1.0 + 0.3 * rng.random((12, 10))
    U[..., 1:3] = 0.3 * rng.standard_normal((12, 10, 2))
    U[..., 3] = rng.random((12, 10))
    Ud = torch.tensor(U, dtype=dtype, device="cuda")
    dx, dy, cs, dfloor, dt = 0.1, 1 / 12, 1.0, 1e-10, 2e-3
    _, stages = jx.isothermal_step(Ud, dt, dx, dy, cs, dfloor, integrator, trace=True,
                                   viscosity=1e-3)
    paths = stage_dump.write_stages(tmp_path, stages, dx, dy, cs, dfloor, dt, integrator)
    assert [p.name for p in paths] == [f"stage{k}.txt" for k in range(1, len(stages) + 1)]
    for k, (path, stage) in enumerate(zip(paths, stages)):
        meta, arrays = stage_dump.read_stage(path)
        assert meta == (dx, dy, cs, dfloor, dt, jx.RK_STAGES[integrator][k][2] * dt)
        for name in stage_dump.FIELDS:
            want = stage[name][..., :3].detach().cpu().numpy().astype(np.float64)
            np.testing.assert_array_equal(arrays[name], want)
