"""Steps pruned to the dealiased disk against unpruned steps: the pruned passes
leave out transforms of zeros and values the 2/3 mask (src/physics.jl:28-43)
discards, so every field after any number of steps must be bit for bit the
unpruned one (equal values; a zero may differ in sign).  The initial states
include content outside the disk, which the first (unpruned) step removes."""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O
from tests.helpers import Case, gpu_model, k3_streamfunction, random_spectral

pytestmark = pytest.mark.gpu

ALL = ("q", "B", "psi", "A", "C", "u", "v", "w")
EXP_N2 = lambda z: 1e-4 * math.exp(z / 800.0)  # noqa: E731


def _init(qg, c, seed, outside):
    grid = O.make_grid(c.size, extent=c.extent, centered=c.centered)
    psi = k3_streamfunction(grid, seed=seed, urms=0.3, kpeak=3.0)
    nz, nx, ny = c.size[2], c.size[0], c.size[1]
    mask = None if outside else O.dealias_mask(nx, ny)
    Bh = random_spectral((nz, nx, ny), seed + 1, mask=mask, scale=0.05 / math.sqrt(nx * ny))
    B = np.fft.ifft2(Bh, axes=(1, 2)) * (nx * ny) ** 0.5
    if outside:  # a little of psi outside the disk too
        ph = random_spectral((nz, nx, ny), seed + 2, scale=1.0, hermitian=True)
        psi = psi + 1e-3 * np.abs(psi).max() * np.fft.ifft2(ph, axes=(1, 2)).real / np.abs(ph).max()
    return psi, B


def _pair(qg, c, seed, outside):
    a, b = gpu_model(qg, c), gpu_model(qg, c)
    b._h.set_pruning(False)
    psi, B = _init(qg, c, seed, outside)
    for m in (a, b):
        qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    return a, b


def _same(a, b, label):
    for n in ALL:
        x, y = getattr(a.fields, n), getattr(b.fields, n)
        assert np.all(np.isfinite(x)), (label, n)
        assert np.array_equal(x, y), (label, n, float(np.max(np.abs(x - y))))


CASES = {
    "feedback": dict(),
    "exp_N2": dict(N2_func=EXP_N2),
    "no_feedback": dict(feedback=False),
    "linear": dict(linear=True),
    "passive": dict(formulation="passive"),
    "no_dispersion": dict(no_dispersion=True),
    "inviscid": dict(inviscid=True),
    "nu_z": dict(nu_z=0.3),
}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_pruned_steps_equal_unpruned(qg, name, precision):
    c = Case(size=(32, 32, 12), precision=precision, **CASES[name])
    print(f"case pruning {name} {precision} {c.size}", flush=True)
    a, b = _pair(qg, c, 5, outside=True)
    ts = qg.ExponentialRungeKutta2(dt=900.0)
    for n in (1, 2, 3):
        qg.step_(a, ts, n)
        qg.step_(b, ts, n)
        _same(a, b, f"{name} after +{n}")
    a.finalize()
    b.finalize()


@pytest.mark.parametrize("size", [(8, 8, 4), (48, 32, 6), (96, 96, 5), (256, 256, 8), (384, 384, 3)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pruned_steps_equal_unpruned_sizes(qg, size, precision):
    c = Case(size=size, precision=precision, N2_func=EXP_N2)
    print(f"case pruning size {size} {precision}", flush=True)
    a, b = _pair(qg, c, 11, outside=size[0] <= 96)
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    qg.step_(a, ts, 4)
    qg.step_(b, ts, 4)
    _same(a, b, f"{size}")
    a.finalize()
    b.finalize()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pruned_run_equals_unpruned_run(qg, precision):
    """run!'s device loop (its check fused into the pruned passes): same
    steps, same failure code, same max |ifft2(psi)|, same fields."""
    c = Case(size=(64, 64, 16), precision=precision, N2_func=EXP_N2)
    print(f"case pruning run {precision}", flush=True)
    a, b = _pair(qg, c, 21, outside=True)
    ra = a._h.run(600.0, 70)
    rb = b._h.run(600.0, 70)
    assert ra == rb, (ra, rb)
    _same(a, b, "run")
    a.finalize()
    b.finalize()


def test_pruned_ensemble_equals_unpruned(qg):
    """Stacked members of one grid class, one re-set between steps (its
    content outside the disk forces an unpruned step of the whole ensemble)."""
    print("case pruning ensemble", flush=True)
    grid = qg.RectilinearGrid(size=(32, 32, 8), extent=(500e3, 500e3, 4000.0))
    strat = lambda n0: qg.AnalyticalProfile(lambda z: n0 * math.exp(z / 800.0))  # noqa: E731
    kw = dict(grid=grid, coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(),
              architecture=qg.GPU("mixed"), stratifications=[strat(1e-5 * (i + 1)) for i in range(3)])
    a, b = qg.QGYBJEnsemble(**kw), qg.QGYBJEnsemble(**kw)
    b._h.set_pruning(False)
    c = Case(size=(32, 32, 8))
    for m in (a, b):
        for i in range(3):
            psi, B = _init(qg, c, 30 + i, outside=True)
            m.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    qg.step_(a, ts, 3)
    qg.step_(b, ts, 3)
    psi, B = _init(qg, c, 40, outside=True)
    for m in (a, b):
        m.set_member(1, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    qg.step_(a, ts, 4)
    qg.step_(b, ts, 4)
    for i in range(3):
        fa, fb = a.member_fields(i), b.member_fields(i)
        for n in ("q", "B", "psi", "A", "u", "w"):
            assert np.array_equal(getattr(fa, n), getattr(fb, n)), (i, n)
    a.finalize()
    b.finalize()
