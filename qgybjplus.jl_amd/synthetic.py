"""Seeded synthetic initial conditions for the benchmark workloads.

The streamfunction has an isotropic spectrum with energy ∝ k⁻³ beyond a peak
at horizontal index wavenumber ``kpeak``, Gaussian random phases, the 2/3 mask
applied, zero mean, a barotropic plus a first-baroclinic-mode part, and is
scaled to a target RMS velocity.  This is the benchmark input of the fixed
workloads; the serving workload draws the reference's own
``RandomStreamfunction`` on the device, with the amplitude from
``random_streamfunction_amplitude``.
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
    bc = one() if baroclinic else np.zeros_like(bt)
    c = np.array([math.cos(math.pi * (grid.z[k] + Lz) / Lz) if baroclinic else 0.0 for k in range(nz)])
    out = bt[None, :, :] + c[:, None, None] * bc[None, :, :]
    # RMS velocity of psi_bt + c(z) psi_bc from the two 2-D velocity fields:
    # mean over z of |u_bt + c u_bc|^2 = |u_bt|^2 + 2 <c> u_bt.u_bc + <c^2> |u_bc|^2
    kx = grid.kx[:, None]
    ky = grid.ky[None, :]

    def vel(f):
        sp = np.fft.fft2(f)
        return np.fft.ifft2(-1j * ky * sp).real, np.fft.ifft2(1j * kx * sp).real

    ub, vb = vel(bt)
    uc, vc = vel(bc)
    e = (np.mean(ub * ub + vb * vb) + 2 * np.mean(c) * np.mean(ub * uc + vb * vc) +
         np.mean(c * c) * np.mean(uc * uc + vc * vc))
    return out * (urms / math.sqrt(float(e)))


def random_streamfunction_amplitude(grid, urms: float = 0.1, slope: float = -3.0) -> float:
    """Amplitude of ``RandomStreamfunction`` whose expected RMS velocity
    (mean of u^2 + v^2 over the grid) is `urms`: each kept mode draws
    E|psi-hat|^2 = 2 A |k|^slope (k in index units), so by Parseval
    E mean(u^2 + v^2) = (2 A / (nx ny)^2) sum kh^2 |k|^slope."""
    nx, ny, _ = grid.size
    ix = np.fft.fftfreq(nx) * nx
    iy = np.fft.fftfreq(ny) * ny
    K = np.sqrt(ix[:, None] ** 2 + iy[None, :] ** 2)
    kmax = min(nx, ny) // 3
    jx = np.where(np.arange(nx) <= nx // 2, np.arange(nx), np.arange(nx) - nx)
    jy = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny)
    keep = ((jx[:, None] ** 2 + jy[None, :] ** 2) <= kmax ** 2) & (K > 0)
    kh2 = grid.kx[:, None] ** 2 + grid.ky[None, :] ** 2
    s = float(np.sum(np.where(keep, kh2 * np.where(K > 0, K, 1.0) ** slope, 0.0)))
    return urms * urms * float(nx * ny) ** 2 / (2.0 * s)


def exponential_N2(N0sq: float = 1e-4, scale: float = 800.0):
    """N²(z) = N0² exp(z / scale)."""
    return lambda z: N0sq * math.exp(z / scale)


def wave_modes(grid, n_modes: int = 32, seed: int = 2026, rms: float = 0.1, efold: float = 30.0,
               profile: str = "exponential") -> np.ndarray:
    """Spectral B, shape (nz, nx, ny): a random superposition of `n_modes`
    horizontal Fourier modes (distinct integer wavevectors inside the 2/3
    dealias disk, random phases, equal amplitudes with surface RMS |B| =
    `rms`) times exp(-depth/efold) (or exp(-(depth/efold)^2) for
    profile="gaussian").  Values follow the unnormalised forward FFT
    convention: a mode of physical amplitude a has spectral value nx*ny*a."""
    nx, ny, nz = grid.size
    rng = np.random.default_rng(seed)
    kmax = min(nx, ny) // 3
    modes = set()
    while len(modes) < n_modes:
        i, j = (int(v) for v in rng.integers(-kmax, kmax + 1, size=2))
        if (i, j) != (0, 0) and i * i + j * j <= kmax * kmax:
            modes.add((i, j))
    amp = rms / math.sqrt(n_modes)
    top = np.zeros((nx, ny), dtype=np.complex128)
    for (i, j) in sorted(modes):
        top[i % nx, j % ny] += nx * ny * amp * np.exp(1j * rng.uniform(0.0, 2 * math.pi))
    out = np.empty((nz, nx, ny), dtype=np.complex128)
    for k in range(nz):
        depth = max(0.0, -grid.z[k] - grid.dz / 2)
        out[k] = top * (math.exp(-depth / efold) if profile == "exponential" else math.exp(-(depth / efold) ** 2))
    return out
