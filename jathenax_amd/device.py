"""Device arrays and contexts: torch (ROCm) tensors carry the device memory and streams."""

import ctypes
import threading

import numpy as np
import torch

from . import _lib

_DTYPES = {torch.float64: 0, torch.float32: 1}
_ctx_lock = threading.Lock()
_contexts = {}


def dtype_id(t):
    try:
        return _DTYPES[t.dtype]
    except KeyError:
        raise ValueError(f"state dtype must be float64 or float32, got {t.dtype}") from None


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("jathenax_amd needs a ROCm GPU (gfx950); none is visible")


def as_device(x, dtype=None, device=None):
    """Return a contiguous CUDA tensor view/copy of x (torch tensor or array-like)."""
    _require_gpu()
    if isinstance(x, torch.Tensor):
        t = x
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        if not t.is_cuda:
            t = t.to(device or "cuda")
        return t.contiguous()
    arr = np.asarray(x)
    if dtype is None:
        dtype = torch.float32 if arr.dtype == np.float32 else torch.float64
    return torch.as_tensor(arr, dtype=dtype, device=device or "cuda").contiguous()


class Context:
    """One C context (device + stream + owned scratch)."""

    def __init__(self, device_index, stream_ptr):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.jx_ctx_create(ctypes.byref(h), int(device_index),
                                     ctypes.c_void_p(stream_ptr)), None, "jx_ctx_create failed")
        self.handle = h
        self.device_index = device_index
        self.pinned = False

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_ctx_destroy(h)
            self.handle = None

    def check(self, rc, what=""):
        _lib.check(rc, self.handle, what)

    def set_fused_forcing(self, on):
        """FP32 forced Smagorinsky runs: evaluate the force where the kick reads it (True,
        the default) or render the raw force array every step (False); same results."""
        self.check(_lib.load().jx_set_fused_forcing(self.handle, int(bool(on))), "forcing")

    def set_kernel_timing(self, on):
        self.check(_lib.load().jx_set_kernel_timing(self.handle, int(bool(on))), "timing")

    def kernel_timing(self):
        """(number of timed stage launches, summed milliseconds) since timing was enabled."""
        n, ms = ctypes.c_int(), ctypes.c_double()
        self.check(_lib.load().jx_kernel_timing(self.handle, ctypes.byref(n), ctypes.byref(ms)),
                   "timing")
        return n.value, ms.value

    def set_tiles(self, block_x=0, rows=0):
        """Pin the stage kernel's tile shape for this context; (0, 0) unpins (tuned/auto)."""
        self.pinned = (int(block_x), int(rows)) != (0, 0)
        self._apply_tiles(block_x, rows)

    def tiles_for(self, p):
        """(block_x, rows) the next step or run with RunParams p will use."""
        bx, rows = ctypes.c_int(), ctypes.c_int()
        self.check(_lib.load().jx_get_tiles(self.handle, ctypes.byref(p), ctypes.byref(bx),
                                            ctypes.byref(rows)), "get_tiles")
        return bx.value, rows.value

    def _apply_tiles(self, block_x, rows):
        self.check(_lib.load().jx_set_tiles(self.handle, int(block_x), int(rows)), "set_tiles")


def context_for(t):
    """Context bound to tensor t's device and the current torch stream there."""
    _require_gpu()
    dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(dev).cuda_stream
    key = (dev, stream)
    with _ctx_lock:
        ctx = _contexts.get(key)
        if ctx is None:
            with torch.cuda.device(dev):
                ctx = Context(dev, stream)
            _contexts[key] = ctx
    return ctx


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()
