"""Model state files and restart: ``_write_model_state_file!``
(src/core/io.jl:138-231) and ``restore!`` (src/core/io.jl:549-598).

A state file holds the dimensions x, y, z, time; the coordinate variables;
the spectral prognostic fields as ``q_hat_real``, ``q_hat_imag``,
``B_hat_real``, ``B_hat_imag``; optionally the physical psi, A and LA
(real and imaginary parts) and u, v, w; N2 and a_ell; and the attributes
title, model_time, iteration, f0, Lx, Ly, Lz and wave_formulation.  Every
field variable is Float64 over ("x", "y", "z") in the reference's Julia
(column-major) order, which a C-order reader sees as (z, y, x).

The files are written as NetCDF classic (64-bit offset) with scipy's NetCDF
writer, the only NetCDF implementation available here.  NCDatasets.jl reads
them, so the reference's ``restore!`` accepts them; files the reference writes
in NetCDF-4 (HDF5) form cannot be read back here.
"""

from __future__ import annotations

import os
from typing import Optional

import numpy as np
from scipy.io import netcdf_file

PROGNOSTIC = ("q_hat_real", "q_hat_imag", "B_hat_real", "B_hat_imag")


def _to_disk(a_zxy: np.ndarray) -> np.ndarray:
    """(z, x, y) model order -> the (z, y, x) C order of an ("x","y","z") Julia
    variable (``_to_xyz``, src/core/io.jl:11)."""
    return np.ascontiguousarray(np.transpose(a_zxy, (0, 2, 1)), dtype=np.float64)


def _from_disk(a_zyx: np.ndarray) -> np.ndarray:
    """Inverse of _to_disk (``_from_xyz``, src/core/io.jl:12)."""
    return np.ascontiguousarray(np.transpose(a_zyx, (0, 2, 1)))


def _physical(spectral_zxy: np.ndarray) -> np.ndarray:
    """``fft_backward!``: the normalised inverse 2-D transform over (x, y)."""
    return np.fft.ifft2(spectral_zxy, axes=(1, 2))


def write_state_arrays(path: str, *, grid, q_hat: np.ndarray, B_hat: np.ndarray, N2, a_ell, f0: float,
                       time: float = 0.0, iteration: int = 0, wave_formulation: str = "YBJPlus",
                       psi: Optional[np.ndarray] = None, A: Optional[np.ndarray] = None,
                       LA: Optional[np.ndarray] = None, u: Optional[np.ndarray] = None,
                       v: Optional[np.ndarray] = None, w: Optional[np.ndarray] = None) -> str:
    """Write one state file from host arrays in (z, x, y) order (spectral q_hat,
    B_hat; physical psi (real), A, LA (complex), u, v, w)."""
    nx, ny, nz = grid.size
    shape = (nz, nx, ny)
    for name, a in (("q_hat", q_hat), ("B_hat", B_hat)):
        if np.shape(a) != shape:
            raise ValueError(f"{name} must have shape (nz, nx, ny) = {shape}")
    with netcdf_file(path, "w", version=2) as ds:
        ds.createDimension("x", nx)
        ds.createDimension("y", ny)
        ds.createDimension("z", nz)
        ds.createDimension("time", 1)
        for name, vals, dims in (("x", grid.x, ("x",)), ("y", grid.y, ("y",)), ("z", grid.z, ("z",)),
                                 ("time", [time], ("time",))):
            var = ds.createVariable(name, "d", dims)
            var[:] = np.asarray(vals, dtype=np.float64)

        def field(name, values):
            var = ds.createVariable(name, "d", ("z", "y", "x"))
            var[:] = _to_disk(values)

        if psi is not None:
            field("psi", np.real(psi))
        if A is not None:
            field("A_real", np.real(A))
            field("A_imag", np.imag(A))
        if LA is not None:
            field("LA_real", np.real(LA))
            field("LA_imag", np.imag(LA))
        field("q_hat_real", np.real(q_hat))
        field("q_hat_imag", np.imag(q_hat))
        field("B_hat_real", np.real(B_hat))
        field("B_hat_imag", np.imag(B_hat))
        for name, a in (("u", u), ("v", v), ("w", w)):
            if a is not None:
                field(name, np.real(a))
        for name, vals in (("N2", N2), ("a_ell", a_ell)):
            var = ds.createVariable(name, "d", ("z",))
            var[:] = np.asarray(vals, dtype=np.float64)
        ds.title = "QG-YBJ model state"
        ds.model_time = float(time)
        ds.iteration = int(iteration)
        ds.f0 = float(f0)
        Lx, Ly, Lz = grid.extent
        ds.Lx, ds.Ly, ds.Lz = float(Lx), float(Ly), float(Lz)
        ds.wave_formulation = wave_formulation
    return path


def read_prognostic(path: str, size) -> tuple:
    """``_read_restart_on_root`` (src/core/io.jl:549-568): q_hat and B_hat in
    (z, x, y) order, checked against the grid size (nx, ny, nz)."""
    nx, ny, nz = size
    with netcdf_file(path, "r", mmap=False) as ds:
        if not all(name in ds.variables for name in PROGNOSTIC):
            raise ValueError("restart file is missing prognostic fields")
        get = lambda n: np.array(ds.variables[n][:], dtype=np.float64)  # noqa: E731
        q = _from_disk(get("q_hat_real") + 1j * get("q_hat_imag"))
        B = _from_disk(get("B_hat_real") + 1j * get("B_hat_imag"))
    if q.shape != (nz, nx, ny):
        raise ValueError("restart q dimensions do not match model grid")
    if B.shape != q.shape:
        raise ValueError("restart B dimensions do not match restart q")
    return q, B


def write_state(owner, path: str, *, psi: bool = True, waves: bool = True, velocities: bool = False) -> str:
    """``_write_model_state_file!`` for a model or a simulation (its clock gives
    the time and iteration).  The diagnostics are refreshed first, as
    ``_refresh_output_diagnostics!`` (src/core/io.jl:119-131) does: psi from q
    (without the wave feedback), A from B, and u, v, w."""
    from .api import QGYBJModel, Simulation, YBJ
    sim = owner if isinstance(owner, Simulation) else None
    model: QGYBJModel = sim.model if sim is not None else owner
    if model.finalized:
        raise RuntimeError("cannot write the state of a finalized model")
    if psi or velocities:
        model.invert_q_to_psi()
    if waves:
        model.invert_B_to_A()
    if velocities:
        model.compute_velocities(compute_w=True)
    f = model.fields
    grid = model.grid
    out = {}
    if psi:
        out["psi"] = np.real(_physical(f.psi))
    if waves:
        A_hat = f.A
        LA_hat = f.B if isinstance(model.formulation, YBJ) else f.B + 0.25 * grid.kh2[None] * A_hat
        out["A"] = _physical(A_hat)
        out["LA"] = _physical(LA_hat)
    if velocities:
        out.update(u=f.u, v=f.v, w=f.w)
    clock = sim.clock if sim is not None else None
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    return write_state_arrays(path, grid=grid, q_hat=f.q, B_hat=f.B, N2=model.N2, a_ell=model.a_ell,
                              f0=model.coriolis.f, time=clock.time if clock else 0.0,
                              iteration=clock.iteration if clock else 0,
                              wave_formulation=type(model.formulation).__name__, **out)


def restore_(owner, path: str):
    """``restore!(model, path)`` / ``restore!(simulation, path)``: q and B from a
    state file, then psi from q, u and v, and A from B."""
    from .api import Simulation, SimulationState
    if isinstance(owner, Simulation):
        if owner.state in (SimulationState.Running, SimulationState.Failed, SimulationState.Finalized):
            raise RuntimeError(f"cannot modify a {owner.state.name.lower()} simulation")
        restore_(owner.model, path)
        return owner
    model = owner
    if model.finalized:
        raise RuntimeError("cannot restore a finalized model")
    q, B = read_prognostic(path, model.grid.size)
    model.fields.q = q
    model.fields.B = B
    model.invert_q_to_psi()
    model.compute_velocities(compute_w=False)
    model.invert_B_to_A()
    return model
