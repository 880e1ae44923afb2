"""Energy diagnostics reduced on the device (qg_energies) against the oracle's
restatement of _local_energy_components / _record_energy_diagnostics!
(src/core/io.jl:331-419), and the Simulation's scheduled energy series."""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O
from tests.helpers import Case, gpu_model, oracle_model

pytestmark = pytest.mark.gpu

EXP_N2 = lambda z: 1e-4 * math.exp(z / 800.0)  # noqa: E731


@pytest.mark.parametrize("size,feedback", [((32, 32, 12), True), ((64, 32, 5), False), ((16, 16, 1), True)])
def test_energies_match_oracle(qg, size, feedback):
    print(f"case energies size={size} feedback={feedback}", flush=True)
    c = Case(size=size, N2_func=EXP_N2, feedback=feedback)
    m, o = gpu_model(qg, c), oracle_model(c)
    nx, ny, nz = size
    qg.set_(m, psi=qg.RandomStreamfunction(amplitude=1e6, seed=4), waves=qg.SurfaceWave(0.1, 300.0))
    B = m.fields.B.copy()
    B += O.random_psi(nx, ny, nz, 1e-2, -2.0, 8)  # horizontal structure in B too
    m.fields.B = B
    qg.step_(m, qg.ExponentialRungeKutta2(dt=200.0), 3)
    q, B = m.fields.q, m.fields.B
    want = np.array(O.record_energies(q, B, o.grid, o.a))
    got = m._h.energies()[0]
    assert np.all(want > 0) or nz == 1
    for i, n in enumerate(qg.ENERGY_NAMES):
        assert math.isclose(got[i], want[i], rel_tol=1e-11, abs_tol=1e-300), (n, got[i], want[i])
    # the model's own diagnostics are untouched, and a repeat is identical
    psi_before = m.fields.psi
    assert np.array_equal(m._h.energies()[0], got)
    assert np.array_equal(m.fields.psi, psi_before)
    m.finalize()


def test_member_energies_and_simulation_schedule(qg):
    print("case member energies and schedule", flush=True)
    grid = qg.RectilinearGrid(size=(32, 32, 8), extent=(500e3, 500e3, 4000.0), centered=True)
    strats = [qg.ConstantStratification(1e-5 * (j + 1)) for j in range(3)]
    ens = qg.QGYBJEnsemble(stratifications=strats, grid=grid, coriolis=qg.FPlane(1e-4),
                           feedback=qg.WaveMeanFeedback())
    alone = []
    for j in range(3):
        ens.set_member(j, psi=qg.RandomStreamfunction(amplitude=1e6, seed=20 + j), waves=qg.SurfaceWave(0.1, 200.0))
        mj = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), stratification=strats[j],
                           feedback=qg.WaveMeanFeedback())
        qg.set_(mj, psi=qg.RandomStreamfunction(amplitude=1e6, seed=20 + j), waves=qg.SurfaceWave(0.1, 200.0))
        alone.append(mj)
    allm = ens._h.energies()
    for j in range(3):
        e_alone = alone[j]._h.energies()[0]
        assert np.array_equal(ens._h.energies(j)[0], allm[j])
        assert np.allclose(allm[j], e_alone, rtol=1e-14, atol=0), j
    for mj in alone:
        mj.finalize()
    ens.finalize()
    # a simulation records the series at iteration 0 and every 2 iterations
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), stratification=qg.ConstantStratification(1e-5),
                      feedback=qg.WaveMeanFeedback())
    qg.set_(m, psi=qg.RandomStreamfunction(amplitude=1e6, seed=3), waves=qg.SurfaceWave(0.1, 200.0))
    sim = qg.Simulation(m, dt=100.0, stop_iteration=5, diagnostics=qg.IterationInterval(2))
    qg.run_(sim)
    series = sim.diagnostics_manager.series
    assert series["time"] == [0.0, 200.0, 400.0]
    assert all(len(series[n]) == 3 for n in qg.ENERGY_NAMES)
    assert sim.state is qg.SimulationState.Stopped
    with pytest.raises(RuntimeError):
        qg.run_(sim)
    m.finalize()
