"""Device-side initial conditions of set! against the oracle: the seeded
RandomStreamfunction (init_mpi_random_psi!, src/parallel_mpi.jl:1159-1216)
and SurfaceWave (set_surface_waves!, src/simulation.jl:292-329), for a model
and for one member of an ensemble."""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O
from tests.helpers import Case, gpu_model, oracle_model, rel

pytestmark = pytest.mark.gpu

EXP_N2 = lambda z: 1e-4 * math.exp(z / 800.0)  # noqa: E731


@pytest.mark.parametrize("size,slope,seed", [((32, 32, 6), -3.0, 0), ((64, 16, 3), -2.5, 12345),
                                             ((384, 384, 1), -3.0, 2026), ((16, 16, 40), -4.0, 7)])
def test_random_psi_matches_oracle(qg, size, slope, seed):
    """Integer hash exact; log/cos/sin/pow may differ by an ulp from numpy's."""
    print(f"case random psi size={size} slope={slope} seed={seed}", flush=True)
    c = Case(size=size, N2_func=EXP_N2)
    m, o = gpu_model(qg, c), oracle_model(c)
    nx, ny, nz = size
    qg.set_(m, psi=qg.RandomStreamfunction(amplitude=3.0, spectral_slope=slope, seed=seed))
    want = O.random_psi(nx, ny, nz, 3.0, slope, seed)
    got = m.fields.psi
    assert rel(got, want) < 1e-14
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    # set! derives q with add_balanced_component! and u, v (compute_w = false)
    o.fields.psi = want
    o._refresh_flow("qg")
    assert rel(m.fields.q, o.fields.q) < 1e-13
    assert rel(m.fields.u, o.fields.u) < 1e-13
    m.finalize()


def test_random_psi_member_and_errors(qg):
    print("case random psi member", flush=True)
    grid = qg.RectilinearGrid(size=(32, 32, 8), extent=(500e3, 500e3, 4000.0), centered=True)
    strats = [qg.ConstantStratification(1e-5 * (j + 1)) for j in range(3)]
    ens = qg.QGYBJEnsemble(stratifications=strats, grid=grid, coriolis=qg.FPlane(1e-4),
                           feedback=qg.WaveMeanFeedback())
    for j in range(3):
        ens.set_member(j, psi=qg.RandomStreamfunction(amplitude=1.0, seed=100 + j))
    before = ens._h.get_field_member(0, "psi")
    ens.set_member(1, psi=qg.RandomStreamfunction(amplitude=2.0, spectral_slope=-2.0, seed=9))
    assert np.array_equal(ens._h.get_field_member(0, "psi"), before)
    want = O.random_psi(32, 32, 8, 2.0, -2.0, 9)
    assert rel(ens._h.get_field_member(1, "psi"), want) < 1e-14
    with pytest.raises(ValueError):
        ens._h.set_random_psi(1.0, -3.0, -5, member=0)
    with pytest.raises(ValueError):
        qg.RandomStreamfunction(amplitude=-1.0)
    with pytest.raises(ValueError):
        ens._h.set_random_psi(1.0, -3.0, 1, member=3)
    ens.finalize()


@pytest.mark.parametrize("size,profile", [((32, 32, 16), "gaussian"), ((64, 64, 12), "exponential"),
                                          ((384, 384, 3), "gaussian")])
def test_surface_wave_matches_oracle(qg, size, profile):
    """KAT test/test_model_initialization.jl:38-52: the top level of B is the
    amplitude; A follows from B (L+ A = B)."""
    print(f"case surface wave size={size} {profile}", flush=True)
    c = Case(size=size, N2_func=EXP_N2)
    m, o = gpu_model(qg, c), oracle_model(c)
    qg.set_(m, waves=qg.SurfaceWave(0.1, 30.0, profile))
    o.set_surface_waves(0.1, 30.0, profile)
    o.refresh_wave_diagnostics()
    assert rel(m.fields.B, o.fields.B) < 1e-15
    top = np.fft.ifft2(m.fields.B[-1])  # the top cell sits at depth 0: profile 1
    assert np.max(np.abs(top - 0.1)) < 1e-12
    assert rel(m.fields.A, o.fields.A) < 1e-12
    m.finalize()
