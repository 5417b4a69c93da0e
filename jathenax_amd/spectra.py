"""Shell spectra and vorticity of a state on the device (spectra.py:21-47, plot_snapshot.py:28-31).

The reference computes these with numpy on the host from a saved snapshot. Here the FFTs
run on the GPU in float64 (rocFFT through torch.fft) and only the per-shell sums come back,
so a 4096^2 state is reduced in place without a host copy of the field.

* kinetic energy per shell floor(|k| + 1/2) of sqrt(rho) v, with the FFT normalised by the
  cell count (spectra.py:25-35);
* enstrophy per shell of the spectral vorticity 2 pi i (kx vy^ - ky vx^), with each axis'
  Nyquist wavenumber set to zero (spectra.py:36-41);
* vorticity dvy/dx - dvx/dy by centred periodic differences (plot_snapshot.py:28-31).

    python -m jathenax_amd.spectra snapshot.npz [--output spectra.csv]

writes the reference's CSV (mode,kinetic_energy,enstrophy); plotting is left to the reader.
"""

import argparse
from pathlib import Path

import numpy as np
import torch


def _fields(U):
    U = torch.as_tensor(U)
    if U.device.type != "cuda":
        U = U.to("cuda")
    U = U[..., :3].to(torch.float64)
    rho, mx, my = U.unbind(-1)
    return rho, mx / rho, my / rho


def _wavenumbers(n, device):
    return torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)


def shell_spectra(U, check=True):
    """(mode, energy, enstrophy) as float64 numpy arrays, one entry per shell."""
    rho, vx, vy = _fields(U)
    ny, nx = rho.shape
    dev = rho.device
    kx, ky = _wavenumbers(nx, dev), _wavenumbers(ny, dev)
    shell = torch.floor(torch.hypot(kx[None, :], ky[:, None]) + 0.5).to(torch.int64).ravel()
    nshell = int(shell.max()) + 1
    root = torch.sqrt(rho)
    ax = torch.fft.fft2(root * vx) / rho.numel()
    ay = torch.fft.fft2(root * vy) / rho.numel()
    density = 0.5 * (ax.abs() ** 2 + ay.abs() ** 2)
    energy = torch.bincount(shell, density.ravel(), minlength=nshell)
    if nx % 2 == 0:
        kx[nx // 2] = 0
    if ny % 2 == 0:
        ky[ny // 2] = 0
    omega = 2j * np.pi * (kx[None, :] * torch.fft.fft2(vy) - ky[:, None] * torch.fft.fft2(vx))
    enstrophy = torch.bincount(shell, (0.5 * (omega / rho.numel()).abs() ** 2).ravel(),
                               minlength=nshell)
    if check:
        # Parseval: the shells add up to the box-mean kinetic energy (spectra.py:38).
        total = float(0.5 * torch.mean(rho * (vx * vx + vy * vy)))
        got = float(energy.sum())
        if not abs(got - total) <= 1e-12 * abs(total) + 1e-300:
            raise AssertionError(f"shell energies sum to {got}, box kinetic energy is {total}")
    return (np.arange(nshell), energy.cpu().numpy(), enstrophy.cpu().numpy())


def vorticity(U, dx, dy):
    """dvy/dx - dvx/dy, centred and periodic, float64 on the device."""
    _, vx, vy = _fields(U)
    dvy_dx = (torch.roll(vy, -1, 1) - torch.roll(vy, 1, 1)) / (2 * dx)
    dvx_dy = (torch.roll(vx, -1, 0) - torch.roll(vx, 1, 0)) / (2 * dy)
    return dvy_dx - dvx_dy


def write_csv(path, mode, energy, enstrophy):
    """The reference's spectra CSV (spectra.py:43-46)."""
    np.savetxt(path, np.column_stack((mode, energy, enstrophy)), delimiter=",",
               header="mode,kinetic_energy,enstrophy", comments="",
               fmt=("%d", "%.16e", "%.16e"))


def main(argv=None):
    ap = argparse.ArgumentParser(description="Shell kinetic energy and enstrophy of a snapshot")
    ap.add_argument("snapshot", type=Path)
    ap.add_argument("--output", type=Path, help=".csv path (default: <snapshot>_spectra.csv)")
    a = ap.parse_args(argv)
    out = (a.output or a.snapshot.with_name(a.snapshot.stem + "_spectra")).with_suffix(".csv")
    with np.load(a.snapshot) as snap:
        U = snap["conserved"]
    mode, energy, enstrophy = shell_spectra(torch.as_tensor(U))
    out.parent.mkdir(parents=True, exist_ok=True)
    write_csv(out, mode, energy, enstrophy)
    print(out)


if __name__ == "__main__":
    main()
