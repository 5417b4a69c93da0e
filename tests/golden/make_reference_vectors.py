"""Write golden input/output vectors by running JathenAX's own functions.

JAX is not installed in this container, so the reference modules are imported over a
small NumPy-backed stand-in for the parts of the ``jax`` API they use (``jax.numpy``
array functions, ``.at[]`` updates, ``lax.fori_loop/cond/while_loop``, ``jit``).  The
reference's arithmetic is elementwise IEEE float64 (X64), which NumPy evaluates without
FMA contraction, so the vectors record the reference algorithm's exact float64 results
under NumPy's reduction order.  XLA's own reduction order and FMA policy are not
reproduced (they cannot be, without JAX).

Run from the repository root, in a container that has the reference checkout:

    python tests/golden/make_reference_vectors.py /root/reference

Output: tests/golden/reference_vectors.npz (inputs and outputs, float64 unless stated).
Only data is written; no reference source is copied.
"""

import functools
import sys
import types
from pathlib import Path

import numpy as np


# --------------------------------------------------------------------------------------
# NumPy stand-in for the jax API surface the reference uses.
# --------------------------------------------------------------------------------------


class _Arr(np.ndarray):
    """ndarray with JAX's functional ``.at[index].set/add/get`` updates."""

    @property
    def at(self):
        return _At(self)


class _At:
    def __init__(self, arr):
        self._arr = arr

    def __getitem__(self, idx):
        return _AtIndex(self._arr, idx)


class _AtIndex:
    def __init__(self, arr, idx):
        self._arr, self._idx = arr, idx

    def set(self, value):
        out = np.array(self._arr, copy=True)
        out[self._idx] = value
        return out.view(_Arr)

    def add(self, value):
        out = np.array(self._arr, copy=True)
        out[self._idx] = out[self._idx] + value
        return out.view(_Arr)

    def get(self, mode=None):
        idx = self._idx
        n = self._arr.shape[0]
        if isinstance(idx, (int, np.integer)) and not 0 <= int(idx) < n:
            return np.full(self._arr.shape[1:], np.nan, self._arr.dtype).view(_Arr)
        return np.asarray(self._arr[idx]).view(_Arr)


def _wrap_out(x):
    if isinstance(x, np.ndarray):
        return x.view(_Arr)
    if isinstance(x, np.generic):
        return np.asarray(x).view(_Arr)
    if isinstance(x, tuple):
        return tuple(_wrap_out(v) for v in x)
    return x


def _wrapped(fn):
    @functools.wraps(fn)
    def inner(*args, **kwargs):
        return _wrap_out(fn(*args, **kwargs))
    return inner


def _asarray(x, dtype=None):
    return np.asarray(x, dtype=dtype).view(_Arr)


def _fori_loop(lo, hi, body, init):
    val = init
    for i in range(int(lo), int(hi)):
        val = body(i, val)
    return val


def _cond(pred, true_fn, false_fn, *ops):
    return true_fn(*ops) if bool(pred) else false_fn(*ops)


def _while_loop(cond_fn, body_fn, state):
    while bool(cond_fn(state)):
        state = body_fn(state)
    return state


def _jit(fn=None, **_):
    if fn is None:
        return lambda f: f
    return fn


def install_numpy_jax():
    jnp = types.ModuleType("jax.numpy")
    for name in ("roll", "where", "fmin", "fmax", "abs", "copysign", "sqrt", "stack",
                 "concatenate", "zeros", "zeros_like", "sum", "mean", "min", "max",
                 "minimum", "maximum", "floor", "cos", "sin", "broadcast_to", "ones",
                 "linspace"):
        setattr(jnp, name, _wrapped(getattr(np, name)))
    jnp.asarray = _asarray
    jnp.array = _asarray
    jnp.finfo = np.finfo
    jnp.issubdtype = np.issubdtype
    jnp.floating = np.floating
    jnp.int32, jnp.float32, jnp.float64 = np.int32, np.float32, np.float64
    jnp.inf, jnp.pi, jnp.nan = np.inf, np.pi, np.nan
    lax = types.ModuleType("jax.lax")
    lax.fori_loop, lax.cond, lax.while_loop = _fori_loop, _cond, _while_loop
    jax = types.ModuleType("jax")
    jax.numpy, jax.lax = jnp, lax
    jax.jit = _jit
    jax.device_put = lambda x, *a, **k: x
    jax.device_get = lambda x: x
    jax.config = types.SimpleNamespace(update=lambda *a, **k: None)
    sys.modules["jax"], sys.modules["jax.numpy"], sys.modules["jax.lax"] = jax, jnp, lax
    # jathenax/input.py imports tomllib (Python >= 3.11); only read_input uses it.
    sys.modules.setdefault("tomllib", types.ModuleType("tomllib"))


# --------------------------------------------------------------------------------------
# Inputs
# --------------------------------------------------------------------------------------


def primitive_field(rng, ny, nx, nvar):
    """Smooth field plus a jump and noise, so every limiter branch is exercised."""
    y, x = np.meshgrid((np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx, indexing="ij")
    W = np.empty((ny, nx, nvar))
    W[..., 0] = 1.0 + 0.3 * np.sin(2 * np.pi * (x + 2 * y)) + 0.5 * (x > 0.5)
    W[..., 1] = 0.8 * np.cos(2 * np.pi * y) + 1.5 * (np.abs(x - 0.3) < 0.1)
    W[..., 2] = 0.6 * np.sin(4 * np.pi * x) - 1.2 * (y > 0.6)
    for s in range(3, nvar):
        W[..., s] = 0.5 + 0.5 * np.tanh((x - 0.4) / 0.05) * (s % 2 * 2 - 1)
    W[..., :] += 0.05 * rng.standard_normal(W.shape)
    W[..., 0] = np.abs(W[..., 0]) + 0.05
    return W


def riemann_states(rng, n):
    """Random face states, including supersonic and strongly rarefying pairs."""
    rho_l = np.exp(rng.uniform(-3, 2, n))
    rho_r = np.exp(rng.uniform(-3, 2, n))
    v = rng.uniform(-4, 4, (4, n))
    v[:, : n // 8] *= 0.01
    rho_r[n // 8: n // 4] = rho_l[n // 8: n // 4]
    v[1, n // 4: n // 2] = -np.abs(v[1, n // 4: n // 2]) - 1.5   # strong rarefaction at
    v[0, n // 4: n // 2] = np.abs(v[0, n // 4: n // 2]) + 1.5    # x faces (right moves +)
    return rho_l, rho_r, v


def main(ref_root):
    install_numpy_jax()
    sys.path.insert(0, str(Path(ref_root).resolve()))
    from jathenax import fluxes, forcing, hydro, run
    from jathenax.initial_conditions import kelvin_helmholtz

    out = {}
    rng = np.random.default_rng(20260101)

    # Reconstructions on a 12x10 grid with two scalars, both axes (fluxes.py:120-131).
    W = primitive_field(rng, 12, 10, 5)
    out["recon_W"] = W
    for name in fluxes.RECONSTRUCTIONS:
        for axis in (0, 1):
            left, right = fluxes.reconstruct_faces(_asarray(W), axis, name, 0.9)
            out[f"recon_{name}_axis{axis}_left"] = np.asarray(left)
            out[f"recon_{name}_axis{axis}_right"] = np.asarray(right)
    W32 = np.asarray([[0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0]], dtype=np.float32)
    left, right = fluxes.wenoz_face_states(_asarray(W32), 1)
    out["wenoz_f32_W"], out["wenoz_f32_left"], out["wenoz_f32_right"] = (
        W32, np.asarray(left), np.asarray(right))

    # Riemann solvers on random face states (fluxes.py:134-219).
    n = 256
    rho_l, rho_r, v = riemann_states(rng, n)
    scal = rng.uniform(0, 1, (2, n))
    left = np.stack((rho_l, v[0], v[2], scal[0]), axis=-1)[None]
    right = np.stack((rho_r, v[1], v[3], scal[1]), axis=-1)[None]
    out["riemann_left"], out["riemann_right"] = left, right
    for name in fluxes.RIEMANN_SOLVERS:
        for axis in (0, 1):
            f = fluxes.isothermal_flux(_asarray(left), _asarray(right), 1.3, axis, name)
            out[f"riemann_{name}_axis{axis}"] = np.asarray(f)

    # Viscous and Smagorinsky face fluxes and gradients (fluxes.py:227-266).
    Wv = primitive_field(rng, 9, 11, 3)
    out["visc_W"] = Wv
    for axis in (0, 1):
        gx, gy = fluxes.face_velocity_gradients(_asarray(Wv), 0.1, 0.07, axis)
        out[f"visc_grad_axis{axis}_dvdx"] = np.asarray(gx)
        out[f"visc_grad_axis{axis}_dvdy"] = np.asarray(gy)
        out[f"visc_nusgs_axis{axis}"] = np.asarray(
            fluxes.smagorinsky_viscosity(gx, gy, 0.1, 0.07, 0.17))
        out[f"visc_flux_axis{axis}"] = np.asarray(
            fluxes.viscous_momentum_flux(_asarray(Wv), 0.1, 0.07, 2e-3, 0.17, axis))
        out[f"visc_flux_uniform_axis{axis}"] = np.asarray(
            fluxes.viscous_momentum_flux(_asarray(Wv), 0.1, 0.07, 2e-3, 0.0, axis))
    tau = rng.standard_normal((9, 11, 3))
    out["stress_tau"] = tau
    sx, sy = fluxes.stress_face_fluxes(_asarray(tau))
    out["stress_fx"], out["stress_fy"] = np.asarray(sx), np.asarray(sy)

    # Timestep limits (hydro.py:63-89).
    cases = []
    for i, (nu, cs_smag, drag, prev, t, tlim) in enumerate((
            (0.0, 0.0, 0.0, np.inf, 0.0, 1.0), (3e-3, 0.0, 0.0, np.inf, 0.0, 1.0),
            (1e-4, 0.2, 0.0, 0.001, 0.0, 1.0), (0.0, 0.3, 5.0, np.inf, 0.0, 1.0),
            (0.0, 0.0, 0.0, np.inf, 0.995, 1.0))):
        dt = hydro.isothermal_timestep(_asarray(Wv), 0.1, 0.07, 1.1, 0.4, prev, t, tlim,
                                       viscosity=nu, smagorinsky_cs=cs_smag, drag_rate=drag)
        cases.append((nu, cs_smag, drag, prev, t, tlim, float(dt)))
    out["timestep_cases"] = np.asarray(cases)

    # Traced single steps: every integrator x reconstruction x solver on a 10x8 grid with
    # a passive scalar, plus the viscous/drag/forced variants (hydro.py:92-146).
    Ws = primitive_field(rng, 8, 10, 4)
    Us = np.concatenate((Ws[..., :1], Ws[..., :1] * Ws[..., 1:]), axis=-1)
    Us[0, 0, :] = (5e-4, 1e-4, -1e-4, 2e-4)    # below the 1e-3 floor: exercises recovery
    out["step_U"] = Us
    keys = ("primitives", "left_x", "right_x", "left_y", "right_y", "flux_x", "flux_y",
            "saved", "conserved")
    for integrator in hydro.RK_STAGES:
        for recon in fluxes.RECONSTRUCTIONS:
            for rs in fluxes.RIEMANN_SOLVERS:
                final, stages = hydro.isothermal_step(
                    _asarray(Us), 0.004, 0.1, 0.125, 1.0, 1e-3, integrator, trace=True,
                    reconstruct=recon, rsolver=rs)
                tag = f"step_{integrator}_{recon}_{rs}"
                out[tag] = np.asarray(final)
                if integrator == "rk3" and rs == "roe":
                    for s, st in enumerate(stages):
                        for k in keys:
                            out[f"{tag}_stage{s}_{k}"] = np.asarray(st[k])
    raw = rng.standard_normal((8, 10, 2))
    out["step_raw_force"] = raw
    for integrator in hydro.RK_STAGES:
        final, stages = hydro.isothermal_step(
            _asarray(Us), 0.004, 0.1, 0.125, 1.0, 1e-3, integrator, trace=True,
            viscosity=2e-3, smagorinsky_cs=0.17, drag_rate=0.7, raw_force=_asarray(raw),
            dedt=0.05, reconstruct="ppm4", rsolver="roe")
        out[f"stepfull_{integrator}"] = np.asarray(final)
        out[f"stepfull_{integrator}_before_kick"] = np.asarray(stages[-1]["before_kick"])
        out[f"stepfull_{integrator}_acceleration"] = np.asarray(stages[-1]["acceleration"])

    # Kick normalisation (forcing.py:137-149).
    Uk = Us[..., :3].copy()
    for dedt in (0.0, 0.04):
        out[f"normalize_dedt{dedt}"] = np.asarray(
            forcing.normalize_force(_asarray(Uk), _asarray(raw), 0.01, dedt))

    # Seed path (forcing.py:11-100).
    uni = forcing.ran_st(1618)
    out["ran_st_1618"] = np.asarray([next(uni) for _ in range(256)])
    uni = forcing.ran_st(7)
    out["ran_st_7"] = np.asarray([next(uni) for _ in range(64)])
    gau = forcing.ran_gaussian_st(1618)
    out["gauss_1618"] = np.asarray([next(gau) for _ in range(800)])
    for lo, hi, peak, kb in ((3, 5, 4.0, 5), (1, 3, 2.0, 3)):
        modes = forcing.cartesian_modes(lo, hi, (-kb, kb), (-kb, kb))
        coef = forcing.modal_coefficients(modes, (1.0, 1.0), lo, hi, peak, 1.0, 1618, 0.75,
                                          0.0075, 6)
        out[f"modes_{lo}_{hi}"] = np.asarray(modes)
        out[f"coef_{lo}_{hi}"] = np.asarray(coef)
    modes = forcing.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = forcing.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 0.7, 99, 0.0, 0.01, 3)
    out["coef_1_3_sol07_tcorr0"] = np.asarray(coef)
    x = (np.arange(16) + 0.5) / 16 - 0.5
    y = (np.arange(12) + 0.5) / 12 - 0.5
    basis = forcing.forcing_basis(_asarray(modes), _asarray(x), _asarray(y), (1.0, 1.0))
    for name, b in zip(("cos_x", "sin_x", "cos_y", "sin_y"), basis):
        out[f"basis_{name}"] = np.asarray(b)
    out["render_force"] = np.asarray(forcing.render_force(_asarray(coef[1]), basis))

    # Whole runs of the reference loop (run.py:20-69).
    # (1) The reference 'smooth' problem, 64x48, PPM4/Roe/RK3, 100 cycles.
    nx, ny = 64, 48
    xs = (np.arange(nx) + 0.5) / nx - 0.5
    ys = (np.arange(ny) + 0.5) / ny - 0.5
    px, py = np.meshgrid(2 * np.pi * xs, 2 * np.pi * ys)
    rho = 1.0 + 0.1 * np.sin(px + py)
    U0 = np.stack((rho, rho * 0.2 * np.cos(py), rho * 0.2 * np.sin(px)), axis=-1)
    out["run_smooth_U0"] = U0
    for tlim, nlim, tag in ((0.2, 100000, "tlim"), (10.0, 100, "n100")):
        U, t, dt, cyc = run.evolve_isothermal(_asarray(U0), 1 / nx, 1 / ny, 1.0, 1e-10, "rk3",
                                              0.4, tlim, nlim)
        out[f"run_smooth_{tag}_U"] = np.asarray(U)
        out[f"run_smooth_{tag}_tdc"] = np.asarray([float(t), float(dt), float(cyc)])
    # (2) KH with tracer, 32x32, every solver menu entry with viscosity/Smagorinsky/drag.
    xk = (np.arange(32) + 0.5) / 32 - 0.5
    Ukh = kelvin_helmholtz(xk, xk)
    out["run_kh_U0"] = Ukh
    for integ, recon, rs in (("rk2", "wenoz", "hlle"), ("rk4", "ppmx", "llf"),
                             ("rk3", "plm", "roe"), ("rk3", "dc", "hlle")):
        U, t, dt, cyc = run.evolve_isothermal(
            _asarray(Ukh), 1 / 32, 1 / 32, 1.0, 1e-10, integ, 0.4, 0.05, 1000,
            viscosity=1e-3, smagorinsky_cs=0.17, drag_rate=0.5, reconstruct=recon, rsolver=rs)
        out[f"run_kh_{integ}_{recon}_{rs}_U"] = np.asarray(U)
        out[f"run_kh_{integ}_{recon}_{rs}_tdc"] = np.asarray([float(t), float(dt), float(cyc)])
    # (3) Driven turbulence, 24x20, seed 1618, shell 1..3, 12 cycles.
    nx, ny = 24, 20
    xs = (np.arange(nx) + 0.5) / nx - 0.5
    ys = (np.arange(ny) + 0.5) / ny - 0.5
    modes = forcing.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    coef = forcing.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, 1618, 0.75, 0.004, 40)
    basis = forcing.forcing_basis(_asarray(modes), _asarray(xs), _asarray(ys), (1.0, 1.0))
    Ut = np.zeros((ny, nx, 3))
    Ut[..., 0] = 1.0
    U, t, dt, cyc = run.evolve_isothermal(
        _asarray(Ut), 1 / nx, 1 / ny, 1.0, 1.17549e-38, "rk2", 0.3, 0.1, 12,
        viscosity=1.5e-3, smagorinsky_cs=0.17, drag_rate=0.3333335,
        forcing=(_asarray(coef), basis), dt_update=0.004, tdriv_start=0.0, dedt=0.5,
        reconstruct="wenoz", rsolver="hlle")
    out["run_turb_U"] = np.asarray(U)
    out["run_turb_tdc"] = np.asarray([float(t), float(dt), float(cyc)])

    dest = Path(__file__).resolve().parent / "reference_vectors.npz"
    np.savez_compressed(dest, **out)
    print(f"wrote {len(out)} arrays to {dest} ({dest.stat().st_size / 1e6:.2f} MB)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
