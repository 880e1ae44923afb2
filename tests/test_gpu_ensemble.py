"""Ensembles: members of one grid class stacked as extra levels of one device
model.  Each member must give the result of the same model run alone, whatever
the other members hold and whenever they were admitted."""

import math

import numpy as np
import pytest

from tests.helpers import Case, oracle_model, rel

pytestmark = pytest.mark.gpu

SIZE = (32, 32, 12)
EXTENT = (500e3, 500e3, 4000.0)


def _strat(qg, i):
    # per-member N²: exponential profiles with different surface values, and
    # one decreasing upwards (its omega elimination interchanges rows)
    if i == 2:
        return qg.AnalyticalProfile(lambda z: 1e-5 * math.exp(-z / 400.0))
    n0 = (5e-6, 2e-5, 0, 4e-5)[i % 4]
    return qg.AnalyticalProfile(lambda z, n0=n0: n0 * math.exp(z / 800.0))


def _initial(qg, grid, i):
    rng = np.random.default_rng(100 + i)
    nx, ny, nz = grid.size
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    kx, ky = 1 + i % 3, 2 + i % 2
    psi = np.empty((nz, nx, ny), complex)
    B = np.empty((nz, nx, ny), complex)
    for k in range(nz):
        s = math.exp(grid.z[k] / 1500.0)
        psi[k] = 2e3 * s * (np.sin(2 * math.pi * kx * X / grid.extent[0] + rng.uniform(0, 6)) *
                            np.cos(2 * math.pi * ky * Y / grid.extent[1]))
        B[k] = 0.1 * math.exp(grid.z[k] / 200.0) * np.exp(1j * 2 * math.pi * (i + 1) * X / grid.extent[0])
    return psi, B


def _kwargs(qg, precision="fp64"):
    grid = qg.RectilinearGrid(size=SIZE, extent=EXTENT)
    return dict(grid=grid, coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(),
                architecture=qg.GPU(precision))


def _alone(qg, i, nsteps, precision="fp64", dt=600.0):
    kw = _kwargs(qg, precision)
    m = qg.QGYBJModel(stratification=_strat(qg, i), **kw)
    psi, B = _initial(qg, kw["grid"], i)
    qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    qg.step_(m, qg.ExponentialRungeKutta2(dt=dt), nsteps)
    out = {n: getattr(m.fields, n) for n in ("q", "B", "psi", "A", "C", "u", "v", "w")}
    m.finalize()
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_members_equal_alone_runs(qg, precision):
    print(f"case ensemble precision={precision}", flush=True)
    kw = _kwargs(qg, precision)
    M = 4
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, i) for i in range(M)], **kw)
    for i in range(M):
        psi, B = _initial(qg, kw["grid"], i)
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    qg.step_(ens, qg.ExponentialRungeKutta2(dt=600.0), 4)
    for i in range(M):
        ref = _alone(qg, i, 4, precision)
        mf = ens.member_fields(i)
        for n in ("q", "B", "psi", "A", "C", "u", "v", "w"):
            got = getattr(mf, n)
            e = float(np.max(np.abs(got - ref[n]))) / max(float(np.max(np.abs(ref[n]))), 1e-300)
            assert e <= 1e-14, (i, n, e)
    # the whole-stack view is the members in order
    stack = ens.fields.q
    assert stack.shape == (M * SIZE[2], SIZE[0], SIZE[1])
    assert np.array_equal(stack[SIZE[2]:2 * SIZE[2]], ens.member_fields(1).q)
    ens.finalize()


def test_member_matches_oracle(qg):
    """Member 2 (whose omega elimination pivots) against the CPU oracle."""
    kw = _kwargs(qg)
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, i) for i in range(3)], **kw)
    for i in range(3):
        psi, B = _initial(qg, kw["grid"], i)
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    qg.step_(ens, qg.ExponentialRungeKutta2(dt=600.0), 3)
    c = Case(size=SIZE, extent=EXTENT, N2_func=lambda z: 1e-5 * math.exp(-z / 400.0), feedback=True)
    o = oracle_model(c)
    psi, B = _initial(qg, kw["grid"], 2)
    o.set_psi_physical(psi, "qg")
    o.set_B_physical(B)
    for _ in range(3):
        o.step(600.0)
    mf = ens.member_fields(2)
    for n in ("q", "B", "psi", "A"):
        e = rel(getattr(mf, n), getattr(o.fields, n))
        assert e <= 1e-11, (n, e)
    assert rel(mf.w, o.fields.w) <= 1e-9
    ens.finalize()


def test_admission_mid_run_leaves_others_unchanged(qg):
    """Member 1 is admitted after member 0 has taken 3 steps; member 0 after
    6 steps equals a lone 6-step run, member 1 after 3 equals a lone 3-step
    run."""
    kw = _kwargs(qg)
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, 0), _strat(qg, 1)], **kw)
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    psi0, B0 = _initial(qg, kw["grid"], 0)
    ens.set_member(0, psi=qg.FieldArray(psi0), waves=qg.FieldArray(B0))
    qg.step_(ens, ts, 3)
    psi1, B1 = _initial(qg, kw["grid"], 1)
    ens.set_member(1, psi=qg.FieldArray(psi1), waves=qg.FieldArray(B1))
    qg.step_(ens, ts, 3)
    ref0, ref1 = _alone(qg, 0, 6), _alone(qg, 1, 3)
    for n in ("q", "B", "psi", "A"):
        a0 = getattr(ens.member_fields(0), n)
        a1 = getattr(ens.member_fields(1), n)
        assert np.max(np.abs(a0 - ref0[n])) <= 1e-14 * np.max(np.abs(ref0[n])), n
        assert np.max(np.abs(a1 - ref1[n])) <= 1e-14 * np.max(np.abs(ref1[n])), n
    ens.finalize()


def test_member_check_isolates_nonfinite_member(qg):
    kw = _kwargs(qg)
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, 0), _strat(qg, 1)], **kw)
    for i in range(2):
        psi, B = _initial(qg, kw["grid"], i)
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    bad = ens.member_fields(1).q
    bad[3, 2, 1] = complex(float("nan"), 0.0)
    ens.member_fields(1).q = bad
    assert ens.check_member(0)[0] is False
    assert ens.check_member(1)[0] is True
    assert ens.fields.q.shape[0] == 2 * SIZE[2]
    ens.finalize()


def test_active_members_and_member_moves(qg):
    """Only members [0, active) step; a member moved to another slot goes on
    exactly as it would have in its own slot."""
    kw = _kwargs(qg)
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, i) for i in range(3)], **kw)
    for i in range(3):
        psi, B = _initial(qg, kw["grid"], i)
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    frozen = ens.member_fields(2).q
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    ens.set_active_members(2)
    qg.step_(ens, ts, 2)
    assert np.array_equal(ens.member_fields(2).q, frozen)
    # member 2 (still at step 0) replaces member 0; both remaining runs go on
    ens.move_member(2, 0)
    qg.step_(ens, ts, 3)
    ref2, ref1 = _alone(qg, 2, 3), _alone(qg, 1, 5)
    for n in ("q", "B", "psi", "A"):
        a, b = getattr(ens.member_fields(0), n), ref2[n]
        assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b)), n
        a, b = getattr(ens.member_fields(1), n), ref1[n]
        assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b)), n
    ens.set_active_members(0)
    qg.step_(ens, ts, 2)  # a no-op
    ens.finalize()


def test_stratification_change_under_captured_graphs(qg):
    """Graphs of the steady-state step are captured while no member's omega
    elimination interchanges rows; then member 1 is given the profile whose
    elimination does (and a fresh start).  The graphs must be re-captured with
    the pivoted-column kernel: member 1 then equals the same run alone, w
    included, and member 0 goes on as its own lone run."""
    kw = _kwargs(qg)
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, 0), _strat(qg, 1)], **kw)
    for i in range(2):
        psi, B = _initial(qg, kw["grid"], i)
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    qg.step_(ens, ts, 5)  # one eager step, then graph-replayed pairs
    ens.set_member_stratification(1, _strat(qg, 2))
    psi2, B2 = _initial(qg, kw["grid"], 2)
    ens.set_member(1, psi=qg.FieldArray(psi2), waves=qg.FieldArray(B2))
    qg.step_(ens, ts, 5)
    ref2, ref0 = _alone(qg, 2, 5), _alone(qg, 0, 10)
    for n in ("q", "B", "psi", "A", "C", "u", "v", "w"):
        a, b = getattr(ens.member_fields(1), n), ref2[n]
        assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b)), ("member 1", n)
        a, b = getattr(ens.member_fields(0), n), ref0[n]
        assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b)), ("member 0", n)
    ens.finalize()


def test_failed_stratification_change_keeps_the_member(qg):
    """A profile whose Thomas elimination is singular is refused with the
    reference's error, and the member keeps its previous profile and tables:
    its next steps equal those of a lone run that never saw the refusal.  On
    a 2e7 m box the longest wave has kh^2 dz^2 ~ 1.1e-8: with a = f^2/N^2 =
    1e-12 the first pivots -(a + kh^2 dz^2 [/4]) fall below sqrt(eps) ~
    1.5e-8, while with a = 1e-3 the smallest pivot of either operator is
    3.3e-8."""
    grid = qg.RectilinearGrid(size=SIZE, extent=(2e7, 2e7, 4000.0))
    kw = dict(grid=grid, coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(), architecture=qg.GPU("fp64"))
    strat = qg.ConstantStratification(1e-5)
    ens = qg.QGYBJEnsemble(stratifications=[strat, strat], **kw)
    psi, B = _initial(qg, grid, 1)
    for i in range(2):
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    with pytest.raises(qg.QgError) as ei:
        ens.set_member_stratification(1, qg.ConstantStratification(1e4))
    assert ei.value.code == 3 and str(ei.value).startswith("Singular matrix in Thomas solver")
    assert np.array_equal(ens.N2[SIZE[2]:], ens.N2[:SIZE[2]])
    qg.step_(ens, qg.ExponentialRungeKutta2(dt=600.0), 3)
    m = qg.QGYBJModel(stratification=strat, **kw)
    qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    qg.step_(m, qg.ExponentialRungeKutta2(dt=600.0), 3)
    for n in ("q", "B", "psi", "A", "w"):
        ref = getattr(m.fields, n)
        for i in range(2):
            got = getattr(ens.member_fields(i), n)
            assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref)), (i, n)
    m.finalize()
    ens.finalize()


def test_in_step_tuning_leaves_members_unchanged(qg):
    """The tuner's in-step refinement replays whole ensemble steps per
    candidate: every member's fields are put back, and the tuned ensemble's
    members step as lone runs do (FP64: tile variants are bitwise, column
    shapes agree to a few ulp)."""
    print("case ensemble in-step tuning", flush=True)
    kw = _kwargs(qg)
    M = 3
    ens = qg.QGYBJEnsemble(stratifications=[_strat(qg, i) for i in range(M)], **kw)
    for i in range(M):
        psi, B = _initial(qg, kw["grid"], i)
        ens.set_member(i, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
    names = ("q", "B", "psi", "A", "C", "u", "v", "w")
    before = [{n: getattr(ens.member_fields(i), n).copy() for n in names} for i in range(M)]
    ens.autotune(600.0, steps=2, force=True, in_step=True)
    for i in range(M):
        mf = ens.member_fields(i)
        for n in names:
            assert np.array_equal(getattr(mf, n), before[i][n]), (i, n)
    qg.step_(ens, qg.ExponentialRungeKutta2(dt=600.0), 4)
    for i in range(M):
        ref = _alone(qg, i, 4)
        mf = ens.member_fields(i)
        for n in ("q", "B", "psi", "A", "C", "w"):
            e = rel(getattr(mf, n), ref[n])
            assert e <= 1e-12, (i, n, e)
    ens.finalize()
