"""ctypes binding of libjathenax_amd.so (declarations in include/jathenax_amd.h).

There is no CPU fallback: if the library is missing or no GPU is visible, calls raise.
"""

import ctypes
from ctypes import POINTER, c_double, c_int, c_longlong, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libjathenax_amd.so"

JX_OK, JX_EINVAL, JX_EKEY, JX_EHIP = 0, 1, 2, 3


class RunParams(ctypes.Structure):
    _fields_ = [
        ("ny", c_int), ("nx", c_int), ("nvar", c_int), ("dtype", c_int),
        ("reconstruct", c_int), ("rsolver", c_int), ("integrator", c_int), ("nlim", c_int),
        ("dx", c_double), ("dy", c_double), ("cs", c_double), ("dfloor", c_double),
        ("cfl", c_double), ("tlim", c_double),
        ("viscosity", c_double), ("smagorinsky_cs", c_double), ("drag_rate", c_double),
        ("dt_update", c_double), ("tdriv_start", c_double), ("dedt", c_double),
        ("rk_rows", POINTER(c_double)),
    ]


class Forcing(ctypes.Structure):
    _fields_ = [
        ("coefficients", c_void_p), ("cos_x", c_void_p), ("sin_x", c_void_p),
        ("cos_y", c_void_p), ("sin_y", c_void_p), ("n_updates", c_int), ("n_modes", c_int),
    ]


class LoopState(ctypes.Structure):
    _fields_ = [("time", c_double), ("prev_dt", c_double), ("cycle", c_int), ("update", c_int)]


# jx_stress_fn: int (*)(void* user, int stage, const void* W, void* tau).
STRESS_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_int, c_void_p, c_void_p)

# name -> (restype, argtypes); every symbol the header declares.
SIGNATURES = {
    "jx_ctx_create": (c_int, [POINTER(c_void_p), c_int, c_void_p]),
    "jx_ctx_destroy": (None, [c_void_p]),
    "jx_last_error": (ctypes.c_char_p, [c_void_p]),
    "jx_set_tiles": (c_int, [c_void_p, c_int, c_int]),
    "jx_get_tiles": (c_int, [c_void_p, POINTER(RunParams), POINTER(c_int), POINTER(c_int)]),
    "jx_set_kernel_timing": (c_int, [c_void_p, c_int]),
    "jx_set_fused_forcing": (c_int, [c_void_p, c_int]),
    "jx_kernel_timing": (c_int, [c_void_p, POINTER(c_int), POINTER(c_double)]),
    "jx_isothermal_step": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_void_p, c_double,
                                   c_void_p, c_double]),
    "jx_isothermal_step_noalloc": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_void_p,
                                           c_double, c_void_p, c_double]),
    "jx_isothermal_step_sgs": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_void_p,
                                       c_double, c_void_p, c_double, c_void_p, c_void_p,
                                       STRESS_FN, c_void_p]),
    "jx_evolve_isothermal_sgs": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_void_p,
                                         POINTER(Forcing), c_void_p, c_void_p, STRESS_FN,
                                         c_void_p, POINTER(c_double), POINTER(c_double),
                                         POINTER(c_int)]),
    "jx_evolve_isothermal_resume": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_void_p,
                                            POINTER(Forcing), POINTER(LoopState),
                                            POINTER(LoopState)]),
    "jx_isothermal_timestep": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_double,
                                       c_double, c_double, POINTER(c_double)]),
    "jx_evolve_isothermal": (c_int, [c_void_p, POINTER(RunParams), c_void_p, c_void_p,
                                     POINTER(Forcing), POINTER(c_double), POINTER(c_double),
                                     POINTER(c_int)]),
    "jx_recover_isothermal": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_double,
                                      c_void_p, c_void_p]),
    "jx_rk_stage": (c_int, [c_void_p, c_int, c_longlong, c_void_p, c_void_p, c_void_p, c_double,
                            c_double, c_double, c_double, c_double, c_int, c_void_p, c_void_p]),
    "jx_flux_divergence": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                   c_double, c_double, c_void_p]),
    "jx_stress_face_fluxes": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                      c_void_p]),
    "jx_render_force": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "jx_normalize_force": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                   c_double, c_double, c_void_p]),
    "jx_reconstruct_faces": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int,
                                     c_int, c_double, c_void_p, c_void_p]),
    "jx_isothermal_flux": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                   c_double, c_int, c_int, c_void_p]),
    "jx_viscous_momentum_flux": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                         c_double, c_double, c_double, c_double, c_int,
                                         c_void_p]),
    "jx_face_velocity_gradients": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                           c_double, c_double, c_int, c_void_p, c_void_p]),
    "jx_smagorinsky_viscosity": (c_int, [c_void_p, c_int, c_longlong, c_void_p, c_void_p,
                                         c_double, c_double, c_double, c_void_p]),
    "jx_add_momentum_flux": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_void_p]),
    "jx_apply_drag": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_void_p, c_double,
                              c_double, c_double]),
    "jx_apply_kick": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_void_p,
                              c_double]),
    "jx_conserved_sums": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                  POINTER(c_double)]),
    "jx_ran_st": (c_int, [c_longlong, c_longlong, POINTER(c_double)]),
    "jx_ran_gaussian_st": (c_int, [c_longlong, c_longlong, POINTER(c_double)]),
    "jx_cartesian_modes": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, POINTER(c_int),
                                   c_int]),
    "jx_modal_coefficients": (c_int, [POINTER(c_int), c_int, c_double, c_double, c_int, c_int,
                                      c_double, c_double, c_longlong, c_double, c_double, c_int,
                                      POINTER(c_double)]),
    "jx_batch_create": (c_int, [c_int, c_int, c_int, c_longlong, POINTER(c_void_p)]),
    "jx_batch_destroy": (None, [c_void_p]),
    "jx_batch_last_error": (ctypes.c_char_p, [c_void_p]),
    "jx_batch_footprint": (c_longlong, [c_void_p, POINTER(RunParams), c_int]),
    "jx_batch_pool_free": (c_longlong, [c_void_p]),
    "jx_batch_add": (c_int, [c_void_p, POINTER(RunParams), c_void_p, POINTER(Forcing), c_int,
                             POINTER(c_int)]),
    "jx_batch_add_resume": (c_int, [c_void_p, POINTER(RunParams), c_void_p, POINTER(Forcing),
                                    c_int, POINTER(LoopState), POINTER(c_int)]),
    "jx_batch_checkpoint": (c_int, [c_void_p, c_int, c_void_p, POINTER(LoopState)]),
    "jx_batch_remove": (c_int, [c_void_p, c_int]),
    "jx_batch_run": (c_int, [c_void_p, c_int, POINTER(c_int)]),
    "jx_batch_launch": (c_int, [c_void_p, c_int]),
    "jx_batch_wait": (c_int, [c_void_p, POINTER(c_int)]),
    "jx_batch_poll": (c_int, [c_void_p, POINTER(c_int), POINTER(c_double), POINTER(c_int), c_int]),
    "jx_batch_poll_launched": (c_int, [c_void_p, POINTER(c_int), POINTER(c_double),
                                       POINTER(c_int), c_int, POINTER(c_int)]),
    "jx_batch_result": (c_int, [c_void_p, c_int, c_void_p, POINTER(c_double), POINTER(c_double),
                                POINTER(c_int)]),
    "jx_batch_state": (c_int, [c_void_p, c_int, POINTER(c_void_p)]),
    "jx_batch_set_graph": (c_int, [c_void_p, c_int, c_int]),
    "jx_batch_set_fill": (c_int, [c_void_p, c_int]),
}

_lib = None


def load():
    """Load the library (once) and attach prototypes.  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH.name} is not built; run `python -m jathenax_amd.build` "
                "(the HIP library is required, there is no CPU fallback)")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class JxError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    """Map a C status to the reference's exception types."""
    if rc == JX_OK:
        return
    msg = what
    if ctx is not None:
        raw = load().jx_last_error(ctx)
        msg = (raw.decode() if raw else "") or what
    if rc == JX_EINVAL:
        raise ValueError(msg)
    if rc == JX_EKEY:
        raise KeyError(msg)
    raise JxError(msg or f"jathenax_amd call failed with status {rc}")
