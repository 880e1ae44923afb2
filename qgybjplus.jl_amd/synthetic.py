"""Seeded synthetic initial conditions for the benchmark workloads.

The streamfunction has an isotropic spectrum with energy ∝ k⁻³ beyond a peak
at horizontal index wavenumber ``kpeak``, Gaussian random phases, the 2/3 mask
applied, zero mean, a barotropic plus a first-baroclinic-mode part, and is
scaled to a target RMS velocity.  This is the benchmark input only; the
reference's own ``RandomStreamfunction`` generator is a different algorithm.
"""

from __future__ import annotations

import math

import numpy as np


def k3_streamfunction(grid, seed: int = 2026, urms: float = 0.1, kpeak: float = 6.0,
                      baroclinic: bool = True) -> np.ndarray:
    """Physical psi, shape (nz, nx, ny), for a ``RectilinearGrid``."""
    nx, ny, nz = grid.size
    Lz = grid.extent[2]
    rng = np.random.default_rng(seed)
    ix = np.fft.fftfreq(nx) * nx
    iy = np.fft.fftfreq(ny) * ny
    K = np.sqrt(ix[:, None] ** 2 + iy[None, :] ** 2)
    k0 = kpeak * math.sqrt(1.5)
    Ks = np.where(K > 0, K, 1.0)
    amp = np.where(K > 0, Ks ** -1.5 * np.exp(-0.5 * (k0 / Ks) ** 2), 0.0)
    kmax = min(nx, ny) // 3
    jx = np.where(np.arange(nx) <= nx // 2, np.arange(nx), np.arange(nx) - nx)
    jy = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny)
    mask = (jx[:, None] ** 2 + jy[None, :] ** 2) <= kmax ** 2

    def one():
        s = amp * (rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)))
        return np.fft.ifft2(np.where(mask, s, 0)).real

    bt = one()
    out = np.empty((nz, nx, ny))
    bc = one() if baroclinic else None
    for k in range(nz):
        out[k] = bt if bc is None else bt + bc * math.cos(math.pi * (grid.z[k] + Lz) / Lz)
    kx = grid.kx[None, :, None]
    ky = grid.ky[None, None, :]
    sp = np.fft.fft2(out, axes=(1, 2))
    u = np.fft.ifft2(-1j * ky * sp, axes=(1, 2)).real
    v = np.fft.ifft2(1j * kx * sp, axes=(1, 2)).real
    rms = math.sqrt(float(np.mean(u * u + v * v)))
    return out * (urms / rms)


def exponential_N2(N0sq: float = 1e-4, scale: float = 800.0):
    """N²(z) = N0² exp(z / scale)."""
    return lambda z: N0sq * math.exp(z / scale)
