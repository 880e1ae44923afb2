"""State files and restart on the CPU (src/core/io.jl:138-231, 549-598):
the variables, dimensions, layout and attributes of a state file, and
read_prognostic's round trip and checks."""

import os

import numpy as np
import pytest
from scipy.io import netcdf_file


def _grid(qg, size=(8, 6, 4)):
    return qg.RectilinearGrid(size=size, extent=(500e3, 300e3, 4000.0), centered=True)


def test_state_file_layout_and_round_trip(qg, tmp_path):
    g = _grid(qg)
    nx, ny, nz = g.size
    rng = np.random.default_rng(3)
    q = rng.normal(size=(nz, nx, ny)) + 1j * rng.normal(size=(nz, nx, ny))
    B = rng.normal(size=(nz, nx, ny)) + 1j * rng.normal(size=(nz, nx, ny))
    psi = rng.normal(size=(nz, nx, ny))
    path = os.path.join(tmp_path, "state0000.nc")
    qg.state_io.write_state_arrays(path, grid=g, q_hat=q, B_hat=B, N2=np.full(nz, 1e-5),
                                   a_ell=np.full(nz, 1e3), f0=1e-4, time=40.0, iteration=2, psi=psi)
    with netcdf_file(path, "r", mmap=False) as ds:
        assert ds.dimensions["x"] == nx and ds.dimensions["y"] == ny and ds.dimensions["z"] == nz
        assert ds.dimensions["time"] == 1
        # Julia ("x", "y", "z") variables appear to a C-order reader as (z, y, x)
        assert ds.variables["q_hat_real"].dimensions == ("z", "y", "x")
        qr = np.array(ds.variables["q_hat_real"][:])
        assert qr.shape == (nz, ny, nx)
        # element (x=i, y=j, z=k) of the Julia variable is q[k, i, j].real
        assert qr[3, 5, 7] == q[3, 7, 5].real
        assert np.array_equal(np.array(ds.variables["psi"][:]), np.transpose(psi, (0, 2, 1)))
        assert np.array_equal(np.array(ds.variables["x"][:]), g.x)
        assert np.array_equal(np.array(ds.variables["z"][:]), g.z)
        assert float(ds.variables["time"][0]) == 40.0
        assert ds.title == b"QG-YBJ model state" and ds.iteration == 2 and ds.f0 == 1e-4
        assert ds.Lx == 500e3 and ds.wave_formulation == b"YBJPlus"
        assert "A_real" not in ds.variables and "u" not in ds.variables
    q2, B2 = qg.read_prognostic(path, g.size)
    assert np.array_equal(q2, q) and np.array_equal(B2, B)


def test_read_prognostic_checks(qg, tmp_path):
    g = _grid(qg)
    nx, ny, nz = g.size
    path = os.path.join(tmp_path, "s.nc")
    z = np.zeros((nz, nx, ny), complex)
    qg.state_io.write_state_arrays(path, grid=g, q_hat=z, B_hat=z, N2=np.ones(nz), a_ell=np.ones(nz), f0=1e-4)
    with pytest.raises(ValueError, match="do not match model grid"):
        qg.read_prognostic(path, (nx, ny, nz + 1))
    bad = os.path.join(tmp_path, "bad.nc")
    with netcdf_file(bad, "w") as ds:
        ds.createDimension("x", 2)
        ds.createVariable("q_hat_real", "d", ("x",))[:] = 0.0
    with pytest.raises(ValueError, match="missing prognostic fields"):
        qg.read_prognostic(bad, (2, 1, 1))
