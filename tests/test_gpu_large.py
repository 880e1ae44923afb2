"""Full-size and deep-column cases: the 512x512x256 single large run through
size-independent properties, and column depths that take the other
mirror-tile widths (nz = 256: 4 quarter columns per tile, 512: 2) or the
degenerate ends (nz = 1, 2, 3) against the oracle."""

import math

import numpy as np
import pytest

from tests.helpers import Case, gpu_model, k3_streamfunction, oracle_model, rel, set_both

pytestmark = pytest.mark.gpu


def test_large_run_mixed_against_fp64_and_round_trip(qg):
    from qgybjplus_amd import synthetic
    print("case 512x512x256 mixed vs fp64, 3 steps", flush=True)
    grid = qg.RectilinearGrid(size=(512, 512, 256), extent=(500e3, 500e3, 4000.0), centered=True)
    strat = qg.AnalyticalProfile(synthetic.exponential_N2(1e-4, 800.0))
    psi = synthetic.k3_streamfunction(grid, seed=2026, urms=0.1, kpeak=6.0)
    B = synthetic.wave_modes(grid, 32, seed=2026, rms=0.1, efold=30.0)
    out = {}
    for prec in ("mixed", "fp64"):
        m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), stratification=strat,
                          feedback=qg.WaveMeanFeedback(), architecture=qg.GPU(prec))
        qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B, space="spectral"))
        if prec == "fp64":
            # q -> psi inverts psi -> q at nz = 256 (4 quarter columns per tile)
            psi_hat = m.fields.psi
            m.invert_q_to_psi()
            assert rel(m.fields.psi, psi_hat) < 1e-12
        qg.step_(m, qg.ExponentialRungeKutta2(dt=10.0), 3)
        bad, mx = m._h.check()
        assert not bad and 0 < mx < 1e10
        out[prec] = {n: getattr(m.fields, n) for n in ("q", "B", "psi", "A")}
        m.finalize()
    for n in ("q", "B", "psi", "A"):
        e = rel(out["mixed"][n], out["fp64"][n])
        print(f"  mixed vs fp64 {n}: {e:.2e}", flush=True)
        assert e < 1e-5, (n, e)


@pytest.mark.parametrize("nz", [1, 2, 3, 256, 512])
def test_column_depths_match_oracle(qg, nz):
    print(f"case column depth nz={nz}", flush=True)
    c = Case(size=(16, 16, nz), extent=(500e3, 500e3, 4000.0), N2_func=lambda z: 1e-4 * math.exp(z / 800.0),
             feedback=True)
    m, o = gpu_model(qg, c), oracle_model(c)
    psi = k3_streamfunction(o.grid, seed=3, urms=0.1, kpeak=2.0)
    X, Y = np.meshgrid(o.grid.x, o.grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(nz):
        B[k] = 0.1 * math.exp(o.grid.z[k] / 300.0) * np.exp(1j * 2 * math.pi * X / o.grid.Lx)
    set_both(m, o, psi, B, qg)
    qg.step_(m, qg.ExponentialRungeKutta2(dt=600.0), 2)
    for _ in range(2):
        o.step(600.0)
    # The column operators' condition number grows about as nz^2 (dz = 7.8 m
    # at nz = 512, a = f^2/N^2 spanning two decades), so two correctly
    # rounded solvers drift apart by ~cond * eps: 1e-11 up to nz = 256, 1e-10
    # at 512.
    tol = 1e-11 if nz <= 256 else 1e-10
    for n in ("q", "B", "psi", "A", "C"):
        e = rel(getattr(m.fields, n), getattr(o.fields, n))
        assert e <= tol, (nz, n, e)
    if nz > 2:
        assert rel(m.fields.w, o.fields.w) <= 1e-9
    m.finalize()
