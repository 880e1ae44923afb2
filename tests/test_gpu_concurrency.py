"""The step's independent passes run on a second stream (the q update beside
the B update, A = L+^-1 B beside the q_w -> psi chain, A and C beside the
finalize's velocity and omega passes), as parallel branches of the captured
hipGraphs.  Per-launch profiling runs every pass on the main stream in issue
order, so a profiled model is the serial reference: the concurrent model
(graphs on) must equal it bit for bit after every call."""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O
from tests.helpers import Case, gpu_model, k3_streamfunction

pytestmark = pytest.mark.gpu

FIELDS = ("q", "B", "psi", "A", "C", "u", "v", "w")
EXP_N2 = lambda z: 1e-4 * math.exp(z / 800.0)  # noqa: E731


@pytest.mark.parametrize("variant", [dict(), dict(feedback=False), dict(formulation="passive"),
                                     dict(linear=True), dict(nu_z=0.2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_concurrent_steps_equal_serial_steps(qg, precision, variant):
    c = Case(size=(64, 64, 16), precision=precision, N2_func=EXP_N2, **variant)
    print(f"case concurrency {precision} {variant}", flush=True)
    grid = O.make_grid(c.size, extent=c.extent, centered=c.centered)
    psi = k3_streamfunction(grid, seed=3, urms=0.3, kpeak=3.0)
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(grid.nz):
        B[k] = 0.1 * math.exp(grid.z[k] / 300.0) * np.exp(2j * math.pi * (X / grid.Lx + Y / grid.Ly))
    a, b = gpu_model(qg, c), gpu_model(qg, c)
    for m in (a, b):
        qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    b._h.profile_begin(kernels=True, steps=False)  # serial reference
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    for n in (1, 4, 3, 2):
        qg.step_(a, ts, n)
        qg.step_(b, ts, n)
        for f in FIELDS:
            x, y = getattr(a.fields, f), getattr(b.fields, f)
            assert np.array_equal(x, y), (n, f, float(np.max(np.abs(x - y))))
    ra, rb = a._h.run(600.0, 40), b._h.run(600.0, 40)
    assert ra == rb
    for f in FIELDS:
        assert np.array_equal(getattr(a.fields, f), getattr(b.fields, f)), ("run", f)
    b._h.profile_end()
    a.finalize()
    b.finalize()
