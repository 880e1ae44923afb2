"""The plain-vs-mixed workload at its full size: 256x256x128 over 500 km x
500 km x 4000 m, N²(z) = 1e-4 exp(z/800 m), the seeded k^-3 streamfunction
(peak wavenumber 6, RMS velocity 0.1 m/s), surface wave 0.1 m/s over 50 m,
wave-mean feedback on, YBJ+, dt = 20 s, both models autotuned as the
benchmark runs them.

* FP64 against the oracle after 3 steps: q, B, psi, A within 1e-11 relative
  L2 (the bar of the plain mode), C, u, v likewise and w within 1e-9.
* Mixed against the FP64 GPU path after 1000 steps: q, B, psi, A within
  1e-5 relative L2 (the mixed mode's stated error bar).
"""

import math

import numpy as np
import pytest

import bench
from tests.helpers import Case, oracle_model, rel

pytestmark = pytest.mark.gpu

FIELDS = ("q", "B", "psi", "A")


def _gpu_model(qg, precision):
    grid, m = bench.build_model(qg, precision)
    psi, _ = bench.workload_inputs(grid)
    bench.set_workload(qg, m, psi)
    m.autotune(bench.DT, steps=4)
    return grid, m, psi


def test_config2_fp64_matches_oracle_after_3_steps(qg):
    print("case 256x256x128 fp64 vs oracle, 3 steps", flush=True)
    grid, m, psi = _gpu_model(qg, "fp64")
    print("  tuning:", m._h.tuning(), flush=True)
    c = Case(size=bench.SIZE, extent=bench.EXTENT, f=1e-4, N2_func=lambda z: 1e-4 * math.exp(z / 800.0),
             feedback=True)
    o = oracle_model(c)
    o.set_psi_physical(psi.astype(np.complex128), "qg")
    o.set_surface_waves(0.1, 50.0)
    for n in FIELDS:
        assert rel(getattr(m.fields, n), getattr(o.fields, n)) < 1e-13, ("init", n)
    qg.step_(m, qg.ExponentialRungeKutta2(dt=bench.DT), 3)
    for _ in range(3):
        o.step(bench.DT)
    for n in FIELDS + ("C", "u", "v", "w"):
        e = rel(getattr(m.fields, n), getattr(o.fields, n))
        print(f"  fp64 vs oracle {n}: {e:.2e}", flush=True)
        assert e <= (1e-9 if n == "w" else 1e-11), (n, e)
    m.finalize()


def test_config2_mixed_error_after_1000_steps(qg):
    print("case 256x256x128 mixed vs fp64, 1000 steps", flush=True)
    models = {p: _gpu_model(qg, p)[1] for p in ("fp64", "mixed")}
    ts = qg.ExponentialRungeKutta2(dt=bench.DT)
    for p, m in models.items():
        qg.step_(m, ts, 1000)
        m._h.synchronize()
        bad, mx = m._h.check()
        assert not bad and 0 < mx < 1e10, (p, bad, mx)
    for n in FIELDS + ("C", "u", "v", "w"):
        e = rel(getattr(models["mixed"].fields, n), getattr(models["fp64"].fields, n))
        print(f"  mixed vs fp64 after 1000 steps {n}: {e:.2e}", flush=True)
        # the bar is stated on the prognostic state and its inversions; the
        # derived C, u, v, w are reported under a loose sanity bound
        assert e <= (1e-5 if n in FIELDS else 1e-3), (n, e)
    for m in models.values():
        m.finalize()
