"""FP32-against-FP64 accuracy of the stage update on the reference's problem setups.

FP32 mode keeps the state and every stage operation in float32 and the time accumulator and
box sums in float64. Each report runs one configuration in both precisions from the same
float64 initial state and compares the end states. The FP64 form is the same kernel in
double; it matches the oracle bit for bit, or to round-off, on the sizes the oracle can run.
Diagnostics follow the reference's reporting scripts:

* density and tracer (rho s / rho) L1 = box average of |a - b| on the unit box;
* vorticity dvy/dx - dvx/dy with centred differences (plot_snapshot.py:28-31);
* kinetic-energy spectrum per shell floor(|k| + 1/2) of sqrt(rho) v (spectra.py:25-35).

    python -m jathenax_amd.accuracy [--smooth N] [--kh N] [--turb N] [--cycles K]

prints one JSON line per configuration.
"""

import argparse
import json
import time

import numpy as np
import torch

from . import problems, workloads
from .run import evolve_isothermal


def l1(a, b):
    return float(np.mean(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


def vorticity(U, dx, dy):
    """dvy/dx - dvx/dy, centred and periodic (plot_snapshot.py:28-31)."""
    U = np.asarray(U, np.float64)
    rho, mx, my = U[..., 0], U[..., 1], U[..., 2]
    vx, vy = mx / rho, my / rho
    dvy_dx = (np.roll(vy, -1, axis=1) - np.roll(vy, 1, axis=1)) / (2 * dx)
    dvx_dy = (np.roll(vx, -1, axis=0) - np.roll(vx, 1, axis=0)) / (2 * dy)
    return dvy_dx - dvx_dy


def ke_spectrum(U):
    """Shell-summed kinetic energy of sqrt(rho) v, shells floor(|k| + 1/2) (spectra.py:25-35)."""
    U = np.asarray(U, np.float64)
    rho, mx, my = U[..., 0], U[..., 1], U[..., 2]
    ny, nx = rho.shape
    kx, ky = np.fft.fftfreq(nx) * nx, np.fft.fftfreq(ny) * ny
    shell = np.floor(np.hypot(kx[None, :], ky[:, None]) + 0.5).astype(int)
    root_rho = np.sqrt(rho)
    ax = np.fft.fft2(root_rho * (mx / rho)) / rho.size
    ay = np.fft.fft2(root_rho * (my / rho)) / rho.size
    return np.bincount(shell.ravel(), (0.5 * (np.abs(ax) ** 2 + np.abs(ay) ** 2)).ravel())


def _pair(U0, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, forcing64=None, forcing32=None,
          **kw):
    """Run FP64 then FP32 from the same initial state; returns both end states and runs."""
    out = {}
    for name, dt_, f in (("f64", torch.float64, forcing64), ("f32", torch.float32, forcing32)):
        U = torch.tensor(U0, dtype=dt_, device="cuda")
        extra = dict(kw)
        if f is not None:
            extra["forcing"] = f
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Ue, t, dt, cyc = evolve_isothermal(U, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim,
                                           **extra)
        torch.cuda.synchronize()
        out[name] = dict(U=Ue.double().cpu().numpy(), time=t, cycles=cyc,
                         wall_s=time.perf_counter() - t0)
    return out


def smooth_report(n=4096, tlim=1.0):
    """The 4096^2 smooth-wave stage workload (PPM4 + Roe, RK3, CFL 0.4) to t = tlim."""
    U0 = problems.smooth_wave(n, n)
    r = _pair(U0, 1.0 / n, 1.0 / n, 1.0, 1e-10, "rk3", 0.4, tlim, 10 ** 9)
    a, b = r["f32"]["U"], r["f64"]["U"]
    return dict(config=f"smooth wave {n}x{n}, PPM4+Roe, RK3, CFL 0.4, t={tlim}",
                cycles=(r["f32"]["cycles"], r["f64"]["cycles"]),
                time=(r["f32"]["time"], r["f64"]["time"]),
                l1_density=l1(a[..., 0], b[..., 0]),
                rel_l1_density=l1(a[..., 0], b[..., 0]) / float(np.mean(np.abs(b[..., 0]))),
                l1_vorticity=l1(vorticity(a, 1 / n, 1 / n), vorticity(b, 1 / n, 1 / n)),
                wall_s=(r["f32"]["wall_s"], r["f64"]["wall_s"]))


def kh_report(n=1024, tlim=3.0):
    """Kelvin-Helmholtz with a tracer (nvar 4), PPM4 + Roe, RK3, CFL 0.4, to t = tlim."""
    U0 = problems.kelvin_helmholtz_tanh(n, n)
    r = _pair(U0, 1.0 / n, 1.0 / n, 1.0, 1e-10, "rk3", 0.4, tlim, 10 ** 9)
    a, b = r["f32"]["U"], r["f64"]["U"]
    return dict(config=f"Kelvin-Helmholtz {n}x{n} + tracer, PPM4+Roe, RK3, CFL 0.4, t={tlim}",
                cycles=(r["f32"]["cycles"], r["f64"]["cycles"]),
                time=(r["f32"]["time"], r["f64"]["time"]),
                l1_density=l1(a[..., 0], b[..., 0]),
                l1_tracer=l1(a[..., 3] / a[..., 0], b[..., 3] / b[..., 0]),
                l1_vorticity=l1(vorticity(a, 1 / n, 1 / n), vorticity(b, 1 / n, 1 / n)),
                wall_s=(r["f32"]["wall_s"], r["f64"]["wall_s"]))


def turbulence_report(n=4096, cycles=200, dt_update=5e-5):
    """Driven turbulence from rest: shell 1..3 solenoidal forcing (seed 1618, refreshed every
    step), viscosity 1.5e-6, Smagorinsky 0.17, drag 0.3333335; WENO-Z + HLLE, RK2, CFL 0.3."""
    specs = {name: workloads.turbulence_spec(n, n, 0.05, 1618, dt_update, "rk2", "wenoz",
                                             "hlle", cfl=0.3, dtype=dt_)
             for name, dt_ in (("f64", np.float64), ("f32", np.float32))}
    s = specs["f64"]
    kw = {k: v for k, v in s["kw"].items() if k != "forcing"}
    r = _pair(s["U"], s["dx"], s["dy"], s["cs"], s["dfloor"], s["integrator"], s["cfl"],
              s["tlim"], cycles, forcing64=s["kw"]["forcing"],
              forcing32=specs["f32"]["kw"]["forcing"], **kw)
    a, b = r["f32"]["U"], r["f64"]["U"]
    w32, w64 = vorticity(a, 1 / n, 1 / n), vorticity(b, 1 / n, 1 / n)
    e32, e64 = ke_spectrum(a), ke_spectrum(b)
    shells = np.arange(1, n // 2)
    ratio = e32[shells] / e64[shells]
    # Shells whose FP64 energy is within 1e-10 of the total carry only round-off (the flow
    # starts at rest and is forced in shells 1..3); their ratio is not meaningful.
    resolved = e64[shells] > 1e-10 * e64.sum()
    return dict(config=f"driven turbulence {n}x{n}, WENO-Z+HLLE, RK2, CFL 0.3, {cycles} cycles",
                cycles=(r["f32"]["cycles"], r["f64"]["cycles"]),
                time=(r["f32"]["time"], r["f64"]["time"]),
                l1_vorticity=l1(w32, w64), mean_abs_vorticity=float(np.mean(np.abs(w64))),
                l1_density=l1(a[..., 0], b[..., 0]),
                ke_shells_1_8_f64=[float(x) for x in e64[1:9]],
                ke_ratio_shells_1_8=[float(x) for x in ratio[:8]],
                max_abs_ke_ratio_minus_1_forced=float(np.max(np.abs(ratio[:3] - 1.0))),
                resolved_shells=int(resolved.sum()),
                max_abs_ke_ratio_minus_1_resolved=float(np.max(np.abs(ratio[resolved] - 1.0))),
                wall_s=(r["f32"]["wall_s"], r["f64"]["wall_s"]))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--smooth", type=int, default=4096)
    ap.add_argument("--smooth-tlim", type=float, default=1.0)
    ap.add_argument("--kh", type=int, default=1024)
    ap.add_argument("--kh-tlim", type=float, default=3.0)
    ap.add_argument("--turb", type=int, default=4096)
    ap.add_argument("--cycles", type=int, default=200)
    a = ap.parse_args()
    if a.smooth:
        print(json.dumps(smooth_report(a.smooth, a.smooth_tlim)), flush=True)
    if a.kh:
        print(json.dumps(kh_report(a.kh, a.kh_tlim)), flush=True)
    if a.turb:
        print(json.dumps(turbulence_report(a.turb, a.cycles)), flush=True)


if __name__ == "__main__":
    main()
