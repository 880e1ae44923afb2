"""Replayed hipGraphs of the step against eager steps: every replay, with eager
work and synchronisation between replays, must give bit-identical state."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE = (32, 32, 12)
EXTENT = (500e3, 500e3, 4000.0)


def _model(qg, precision, members=None):
    grid = qg.RectilinearGrid(size=SIZE, extent=EXTENT)
    kw = dict(grid=grid, coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(),
              architecture=qg.GPU(precision))
    strat = lambda n0: qg.AnalyticalProfile(lambda z: n0 * math.exp(z / 800.0))  # noqa: E731
    if members:
        return qg.QGYBJEnsemble(stratifications=[strat(1e-5 * (i + 1)) for i in range(members)], **kw)
    return qg.QGYBJModel(stratification=strat(2e-5), **kw)


def _init(qg, m, seed):
    grid = m.grid
    nx, ny, nz = grid.size
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    psi = np.empty((nz, nx, ny), complex)
    B = np.empty((nz, nx, ny), complex)
    for k in range(nz):
        psi[k] = 3e3 * np.sin(2 * math.pi * X / grid.extent[0] + rng.uniform(0, 6)) * \
            np.cos(4 * math.pi * Y / grid.extent[1])
        B[k] = 0.1 * math.exp(grid.z[k] / 300.0) * np.exp(2j * math.pi * X / grid.extent[0])
    return psi, B


def _state(m):
    return [getattr(m.fields, n) for n in ("q", "B", "psi", "A", "u", "w")]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager_steps(qg, precision):
    print(f"case graph replay precision={precision}", flush=True)
    g, e = _model(qg, precision), _model(qg, precision)
    e._h.use_graphs(False)
    psi, B = _init(qg, g, 3)
    for m in (g, e):
        qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    # pairs replay the two-step graph and runs of 16 the long graph (17: one
    # long graph and an eager step; 35: two long graphs, a pair and an eager
    # step), odd counts end with an eager step, a new dt rewrites the ETD
    # parameter block in place, set_ invalidates the diagnostics
    plan = [2, 2, 3, 4, 1, 17, ("dt", 300.0), 4, 35, ("set", 5), 6, 2, 16]
    for item in plan:
        if isinstance(item, tuple) and item[0] == "dt":
            ts = qg.ExponentialRungeKutta2(dt=item[1])
            continue
        if isinstance(item, tuple) and item[0] == "set":
            psi2, B2 = _init(qg, g, item[1])
            for m in (g, e):
                qg.set_(m, waves=qg.FieldArray(B2))
            continue
        qg.step_(g, ts, item)
        qg.step_(e, ts, item)
        for a, b in zip(_state(g), _state(e)):
            assert np.array_equal(a, b), item
    g.finalize()
    e.finalize()


def test_graph_replays_match_eager_in_ensemble(qg):
    g, e = _model(qg, "mixed", members=3), _model(qg, "mixed", members=3)
    e._h.use_graphs(False)
    for i in range(3):
        psi, B = _init(qg, g, 10 + i)
        for m in (g, e):
            m.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    for n in (2, 4, 2, 16, 18):
        qg.step_(g, ts, n)
        qg.step_(e, ts, n)
        for i in range(3):
            assert np.array_equal(g.member_fields(i).q, e.member_fields(i).q)
            assert np.array_equal(g.member_fields(i).B, e.member_fields(i).B)
    g.finalize()
    e.finalize()
