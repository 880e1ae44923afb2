"""Builders shared by the tests: one configuration -> a GPU model (through the
C ABI) and the CPU oracle model with identical inputs."""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Callable, Optional, Tuple

import numpy as np

from oracle import qgybj_oracle as O


@dataclass
class Case:
    size: Tuple[int, int, int] = (32, 32, 16)
    extent: Tuple[float, float, float] = (500e3, 500e3, 4000.0)
    centered: bool = True
    f: float = 1e-4
    N2: float = 1e-5                      # constant stratification ...
    N2_func: Optional[Callable] = None    # ... or an analytical N²(z)
    fixed_flow: bool = False
    feedback: bool = True
    formulation: str = "ybj_plus"
    inviscid: bool = False
    linear: bool = False
    no_dispersion: bool = False
    nu_z: float = 0.0
    closure: O.Closure = field(default_factory=O.Closure)
    precision: str = "fp64"


def gpu_model(qg, c: Case, kx=None):
    """kx: wavenumbers to use in place of the grid's (the same array goes to
    oracle_model)."""
    grid = qg.RectilinearGrid(size=c.size, extent=c.extent, centered=c.centered)
    if kx is not None:
        grid.kx = np.array(kx, dtype=float)
        grid.kh2 = grid.kx[:, None] ** 2 + grid.ky[None, :] ** 2
    strat = qg.AnalyticalProfile(c.N2_func) if c.N2_func else qg.ConstantStratification(c.N2)
    if c.closure.kind == "horizontal":
        closure = qg.HorizontalHyperdiffusivity(
            flow=qg.FlowHyperdiffusivity(coefficients=c.closure.flow_coefficients,
                                         orders=c.closure.flow_orders),
            wave=qg.WaveHyperdiffusivity(coefficients=c.closure.wave_coefficients,
                                         orders=c.closure.wave_orders))
    elif c.closure.kind == "flow":
        closure = qg.FlowHyperdiffusivity(coefficients=c.closure.flow_coefficients,
                                          orders=c.closure.flow_orders)
    else:
        closure = qg.WaveHyperdiffusivity(coefficients=c.closure.wave_coefficients,
                                          orders=c.closure.wave_orders)
    return qg.QGYBJModel(
        grid=grid, coriolis=qg.FPlane(c.f), stratification=strat, closure=closure,
        vertical_diffusion=qg.VerticalDiffusivity(c.nu_z),
        flow="fixed" if c.fixed_flow else "evolving",
        feedback="wave_mean" if c.feedback else "none",
        formulation="passive" if c.formulation == "passive" else "ybj_plus",
        inviscid=c.inviscid, linear=c.linear, no_dispersion=c.no_dispersion,
        architecture=qg.GPU(c.precision))


def oracle_model(c: Case, kx=None) -> O.OracleModel:
    g = O.make_grid(c.size, extent=c.extent, centered=c.centered)
    if kx is not None:
        g = replace(g, kx=np.array(kx, dtype=float))
    if c.N2_func:
        N2 = O.stratification_profile(g, c.N2_func)
    else:
        N2 = np.full(c.size[2], O.constant_N2(c.N2))
    opt = O.Options(f=c.f, fixed_flow=c.fixed_flow, feedback=c.feedback,
                    formulation=c.formulation, inviscid=c.inviscid, linear=c.linear,
                    no_dispersion=c.no_dispersion, nu_z=c.nu_z, closure=c.closure)
    return O.OracleModel(g, opt, N2)


def random_spectral(shape, seed, mask=None, scale=1.0, hermitian=False):
    """Seeded complex spectrum; Hermitian (a real field's spectrum) on request."""
    rng = np.random.default_rng(seed)
    a = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * scale
    if hermitian:
        a = np.fft.fft2(np.fft.ifft2(a, axes=(1, 2)).real, axes=(1, 2))
    if mask is not None:
        a = np.where(mask[None], a, 0)
    return a


def k3_streamfunction(grid: O.Grid, seed: int, urms: float = 0.1, kpeak: float = 6.0,
                      baroclinic: bool = True) -> np.ndarray:
    """Physical psi with an isotropic k^-3 spectrum peaked at index wavenumber
    `kpeak`, Gaussian random phases, 2/3-masked, zero mean, scaled to RMS |u|."""
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    rng = np.random.default_rng(seed)
    ix = np.fft.fftfreq(nx) * nx
    iy = np.fft.fftfreq(ny) * ny
    K = np.sqrt(ix[:, None] ** 2 + iy[None, :] ** 2)
    k0 = kpeak * math.sqrt(1.5)
    Ks = np.where(K > 0, K, 1.0)  # the mean (K = 0) gets no amplitude
    amp = np.where(K > 0, Ks ** -1.5 * np.exp(-0.5 * (k0 / Ks) ** 2), 0.0)
    mask = O.dealias_mask(nx, ny)

    def one():
        s = amp * (rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)))
        s = np.where(mask, s, 0)
        return np.fft.ifft2(s).real

    bt = one()
    out = np.empty((nz, nx, ny))
    if baroclinic:
        bc = one()
        for k in range(nz):
            out[k] = bt + bc * math.cos(math.pi * (grid.z[k] + grid.Lz) / grid.Lz)
    else:
        out[:] = bt
    # scale to RMS velocity
    kx = grid.kx[None, :, None]
    ky = grid.ky[None, None, :]
    sp = np.fft.fft2(out, axes=(1, 2))
    u = np.fft.ifft2(-1j * ky * sp, axes=(1, 2)).real
    v = np.fft.ifft2(1j * kx * sp, axes=(1, 2)).real
    rms = math.sqrt(float(np.mean(u * u + v * v)))
    return out * (urms / rms)


def set_both(gm, om: O.OracleModel, psi_phys: np.ndarray, B_phys: np.ndarray, qg, pv_method="qg"):
    """Same physical initial fields on the GPU model and the oracle."""
    qg.set_(gm, psi=qg.FieldArray(psi_phys), pv_method=pv_method)
    qg.set_(gm, waves=qg.FieldArray(B_phys))
    om.set_psi_physical(psi_phys.astype(np.complex128), pv_method)
    om.set_B_physical(B_phys.astype(np.complex128))


def rel(a, b):
    return O.relative_l2(a, b)
