"""The benchmark's 48-member uneven-grid ensemble at full size: 16 members
each of 64²x32, 128²x64 and 256²x64 over 500 km x 500 km x 4000 m, mixed
precision, N² = N0² exp(z/800 m) with per-member N0², the benchmark's seeded
inputs.  After 20 steps, one member of each grid class equals the same model
run alone (each class is one device ensemble: its members are extra levels of
every launch, and a member's columns never mix with another's), and no
member has gone non-finite."""

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

STEPS = 20
FIELDS = ("q", "B", "psi", "A", "C", "u", "v", "w")


def test_full_ensemble_members_equal_alone(qg):
    from qgybjplus_amd import synthetic
    print("case 48-member ensemble, 20 steps, one member per class against alone", flush=True)
    n0 = bench.ensemble_n0()
    ens_list, points = bench.build_ensembles(qg)
    assert points == sum(n * n * nz for n, nz in bench.ENSEMBLE_CLASSES) * bench.ENSEMBLE_MEMBERS
    ts = qg.ExponentialRungeKutta2(dt=bench.DT)
    for ens in ens_list:
        qg.step_(ens, ts, STEPS)
        ens._h.synchronize()
        assert [j for j in range(ens.members) if ens.check_member(j)[0]] == []
    for ci, ens in enumerate(ens_list):
        j = (5 * ci + 3) % bench.ENSEMBLE_MEMBERS
        kw = bench.ensemble_model_kwargs(qg, ci)
        m = qg.QGYBJModel(stratification=qg.AnalyticalProfile(
            synthetic.exponential_N2(float(n0[ci * bench.ENSEMBLE_MEMBERS + j]), 800.0)), **kw)
        qg.set_(m, psi=qg.FieldArray(bench.ensemble_member_psi(kw["grid"], ci, j)),
                waves=qg.SurfaceWave(0.1, 50.0))
        qg.step_(m, ts, STEPS)
        mf = ens.member_fields(j)
        for n in FIELDS:
            ref = getattr(m.fields, n)
            got = getattr(mf, n)
            assert got.shape == ref.shape
            e = float(np.max(np.abs(got - ref))) / max(float(np.max(np.abs(ref))), 1e-300)
            print(f"  class {ens.grid.size} member {j} {n}: {e:.1e}", flush=True)
            assert e <= 1e-14, (ci, j, n, e)
        m.finalize()
    for ens in ens_list:
        ens.finalize()
