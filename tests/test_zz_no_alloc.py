"""jx_isothermal_step_noalloc: the step on the library's static scratch (no device
allocation at call time) equals jx_isothermal_step bit for bit, on consecutive calls that
share the scratch (the second with a smaller input), and rejects states beyond the scratch
before doing any work.  Both entry points are called through the C ABI."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _state(ny, nx, nv, seed, dtype):
    rng = np.random.default_rng(seed)
    U = np.empty((ny, nx, nv))
    U[..., 0] = 1.0 + 0.3 * rng.random((ny, nx))
    U[..., 1:3] = 0.3 * rng.standard_normal((ny, nx, 2)) * U[..., :1]
    U[..., 3:] = rng.random((ny, nx, nv - 3)) * U[..., :1]
    return torch.tensor(U, dtype=dtype, device="cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()


def _both(U, integrator, dt, raw=None, dedt=0.0, **kw):
    """(jx_isothermal_step result, jx_isothermal_step_noalloc result) for the same call."""
    from jathenax_amd import _lib
    from jathenax_amd.device import context_for
    from jathenax_amd.hydro import run_params
    ny, nx, _ = U.shape
    p = run_params(U, 1 / nx, 1 / ny, 1.0, 1e-10, integrator, **kw)
    ctx = context_for(U)
    lib = _lib.load()
    ref = torch.empty_like(U)
    out = torch.full_like(U, float("nan"))
    ctx.check(lib.jx_isothermal_step(ctx.handle, ctypes.byref(p), _ptr(U), _ptr(ref), dt,
                                     _ptr(raw), dedt), "step")
    ctx.check(lib.jx_isothermal_step_noalloc(ctx.handle, ctypes.byref(p), _ptr(U), _ptr(out),
                                             dt, _ptr(raw), dedt), "step_noalloc")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_step_equals_step_on_consecutive_calls():
    assert torch.cuda.is_available(), "GPU test needs a visible ROCm device"
    # 1: FP64, rk4 (S register), viscosity, drag and a kick (partial sums)
    U = _state(200, 176, 4, 1, torch.float64)
    raw = torch.tensor(np.random.default_rng(2).standard_normal((200, 176, 2)),
                       device="cuda")
    ref, out = _both(U, "rk4", 1e-3, raw, 0.05, viscosity=1e-3, smagorinsky_cs=0.15,
                     drag_rate=0.3, reconstruct="wenoz", rsolver="hlle")
    np.testing.assert_array_equal(out, ref)
    # 2: a smaller FP32 input on the same scratch, rk3 with the CFL epilogue-free step
    U = _state(40, 56, 3, 3, torch.float32)
    ref, out = _both(U, "rk3", 2e-3)
    np.testing.assert_array_equal(out, ref)
    # 3: FP32 kick on a small grid after the larger FP64 one (stale partial sums unused)
    U = _state(64, 48, 3, 4, torch.float32)
    raw = torch.tensor(np.random.default_rng(5).standard_normal((64, 48, 2)),
                       dtype=torch.float32, device="cuda")
    ref, out = _both(U, "rk2", 1e-3, raw, 0.02, viscosity=1e-3)
    np.testing.assert_array_equal(out, ref)


def test_noalloc_step_allocates_nothing():
    # A state of 24 MiB per register (1024^2 x 3 float64, rk4: 3 registers = 72 MiB) on a
    # context that has never allocated stage registers: the device's free memory does not
    # move by anything near that across the call.
    from jathenax_amd import _lib
    from jathenax_amd.device import Context
    from jathenax_amd.hydro import run_params
    U = _state(1024, 1024, 3, 6, torch.float64)
    out = torch.empty_like(U)
    torch.cuda.synchronize()
    p = run_params(U, 1 / 1024, 1 / 1024, 1.0, 1e-10, "rk4")
    ctx = Context(0, 0)           # a fresh context on the null stream: no buffers yet
    lib = _lib.load()
    free0 = torch.cuda.mem_get_info()[0]
    ctx.check(lib.jx_isothermal_step_noalloc(ctx.handle, ctypes.byref(p), _ptr(U), _ptr(out),
                                             1e-4, _ptr(None), 0.0), "step_noalloc")
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < 8 << 20, (free0 - free1)
    ref = torch.empty_like(U)
    ctx.check(lib.jx_isothermal_step(ctx.handle, ctypes.byref(p), _ptr(U), _ptr(ref), 1e-4,
                                     _ptr(None), 0.0), "step")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())


def test_noalloc_step_rejects_states_beyond_the_scratch():
    from jathenax_amd import _lib
    from jathenax_amd.device import context_for
    from jathenax_amd.hydro import run_params
    lib = _lib.load()
    # 2048^2 x 3 float64 = 96 MiB per register; rk3 needs two
    U = torch.ones((2048, 2048, 3), dtype=torch.float64, device="cuda")
    out = torch.full_like(U, 7.0)
    p = run_params(U, 1 / 2048, 1 / 2048, 1.0, 1e-10, "rk3")
    ctx = context_for(U)
    rc = lib.jx_isothermal_step_noalloc(ctx.handle, ctypes.byref(p), _ptr(U), _ptr(out), 1e-4,
                                        _ptr(None), 0.0)
    assert rc == _lib.JX_EINVAL
    assert b"static scratch" in lib.jx_last_error(ctx.handle)
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())   # nothing ran
    with pytest.raises(ValueError, match="static scratch"):
        ctx.check(rc, "step_noalloc")
