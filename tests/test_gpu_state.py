"""Checkpoint and restart through a state file (src/core/io.jl:138-231,
549-598): restore! gives back q and B exactly, rebuilds psi, u, v and A, and
the restored model steps bit for bit like the one that wrote the file."""

import math
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from tests.helpers import Case, gpu_model, k3_streamfunction
from oracle import qgybj_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_write_restore_continue(qg, tmp_path, precision):
    print(f"case state file {precision}", flush=True)
    c = Case(size=(32, 32, 8), N2=1e-5, precision=precision)
    grid = O.make_grid(c.size, extent=c.extent, centered=c.centered)
    psi = k3_streamfunction(grid, seed=7, urms=0.3, kpeak=3.0)
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(grid.nz):
        B[k] = 0.1 * math.exp(grid.z[k] / 300.0) * np.exp(2j * math.pi * Y / grid.Ly)
    a = gpu_model(qg, c)
    qg.set_(a, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    sim = qg.Simulation(a, dt=600.0, stop_iteration=7)
    qg.run_(sim)
    path = os.path.join(tmp_path, "state0001.nc")
    qg.write_state(sim, path, psi=True, waves=True, velocities=True)
    with netcdf_file(path, "r", mmap=False) as ds:
        assert ds.iteration == 7 and float(ds.variables["time"][0]) == 7 * 600.0
        psi_hat = a.fields.psi
        phys = np.real(np.fft.ifft2(psi_hat, axes=(1, 2)))
        assert np.allclose(np.array(ds.variables["psi"][:]), np.transpose(phys, (0, 2, 1)), rtol=0, atol=1e-12 * np.abs(phys).max())
        assert np.array_equal(np.array(ds.variables["u"][:]), np.transpose(a.fields.u, (0, 2, 1)))
    b = gpu_model(qg, c)
    qg.restore_(b, path)
    for n in ("q", "B", "psi", "A"):
        assert np.array_equal(getattr(b.fields, n), getattr(a.fields, n)), n
    for n in ("u", "v"):
        assert np.array_equal(getattr(b.fields, n), getattr(a.fields, n)), n
    # restore! rebuilds psi = L^-1 q, u, v from psi and A, C from B
    # (src/core/io.jl:583-587): the oracle's values from the file's q-hat, B-hat
    q_file, B_file = qg.read_prognostic(path, c.size)
    N2 = np.full(c.size[2], O.constant_N2(c.N2))
    a_ell = O.a_ell_from_N2(N2, c.f)
    o_psi = O.invert_q_to_psi(q_file, grid, a_ell)
    o_u, o_v, _ = O.compute_velocities(o_psi, grid, c.f, N2, O.dealias_mask(grid.nx, grid.ny), False)
    o_A, o_C = O.invert_B_to_A(B_file, grid, a_ell)
    tol = 1e-11 if precision == "fp64" else 1e-5
    for n, ref in (("psi", o_psi), ("u", o_u), ("v", o_v), ("A", o_A), ("C", o_C)):
        e = O.relative_l2(getattr(b.fields, n), ref)
        print(f"  restored {n} vs oracle: {e:.2e}", flush=True)
        assert e <= tol, (n, e)
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    qg.step_(a, ts, 5)
    qg.step_(b, ts, 5)
    for n in ("q", "B", "psi", "A", "u", "v", "w"):
        assert np.array_equal(getattr(b.fields, n), getattr(a.fields, n)), n
    a.finalize()
    b.finalize()
