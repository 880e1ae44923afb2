"""qg_run_noalloc, the variant of qg_run that allocates nothing: its chunk
snapshot lives in the library's static device scratch instead of a buffer
the first qg_run call allocates.  Through the C ABI it must give exactly what
qg_run gives (steps taken, failure kind, maximum, every field bit for bit),
on two calls in a row that share the scratch, the second with a smaller
model, and it must refuse a model whose snapshot exceeds the scratch."""

import math

import numpy as np
import pytest

from tests.helpers import Case, gpu_model, k3_streamfunction
from oracle import qgybj_oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("q", "B", "psi", "A", "C", "u", "v", "w")
QG_ERR_ARGUMENT = 1


def _pair(qg, size, precision, urms, seed=5):
    c = Case(size=size, N2=1e-5, precision=precision)
    grid = O.make_grid(c.size, extent=c.extent, centered=c.centered)
    psi = k3_streamfunction(grid, seed=seed, urms=urms, kpeak=3.0)
    X, _ = np.meshgrid(grid.x, grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(grid.nz):
        B[k] = 0.1 * math.exp(grid.z[k] / 300.0) * np.exp(2j * math.pi * X / grid.Lx)
    out = []
    for _ in range(2):
        m = gpu_model(qg, c)
        qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
        out.append(m)
    return out


def _same(a, b):
    for f in FIELDS:
        assert np.array_equal(getattr(b.fields, f), getattr(a.fields, f), equal_nan=True), f


def test_noalloc_run_equals_run_on_two_calls(qg):
    # first call: 64x64x16 FP64, 70 steps (two chunks of 32 and a tail)
    print("case no-alloc run 64x64x16 fp64, 70 steps", flush=True)
    a, b = _pair(qg, (64, 64, 16), "fp64", urms=0.5)
    ref = a._h.run(500.0, 70)
    got = b._h.run_noalloc(500.0, 70)
    assert ref == got and ref[:2] == (70, 0), (ref, got)
    _same(a, b)
    # the same handles go on: the second call from the state the first left
    ref = a._h.run(500.0, 9)
    got = b._h.run_noalloc(500.0, 9)
    assert ref == got, (ref, got)
    _same(a, b)
    for m in (a, b):
        m.finalize()
    # second model, smaller (32x32x8 mixed), a run that blows up: the stop
    # right after the failing step, as qg_run makes it
    print("case no-alloc run 32x32x8 mixed, blow-up", flush=True)
    c, d = _pair(qg, (32, 32, 8), "mixed", urms=5.0)
    ref = c._h.run(1500.0, 140)
    got = d._h.run_noalloc(1500.0, 140)
    assert 2 <= ref[0] < 140 and ref[1] in (1, 2), ref
    assert ref[:2] == got[:2] and (ref[2] == got[2] or (math.isnan(ref[2]) and math.isnan(got[2]))), (ref, got)
    _same(c, d)
    for m in (c, d):
        m.finalize()


def test_noalloc_run_refuses_a_snapshot_beyond_the_scratch(qg):
    """256x256x32 in FP64: 12 * 8 bytes * 2,097,152 points = 201 MB > 128 MiB."""
    print("case no-alloc run size limit", flush=True)
    c = Case(size=(256, 256, 32), N2=1e-5, precision="fp64")
    m = gpu_model(qg, c)
    with pytest.raises(ValueError, match="static scratch"):
        m._h.run_noalloc(20.0, 2)
    # nothing ran: the model still steps with qg_run
    assert m._h.run(20.0, 2)[:2] == (2, 0)
    m.finalize()
