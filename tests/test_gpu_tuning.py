"""The autotuner's search space and its effect on results: every tile
variant of the FFT passes gives bitwise the result of the base variant, every
column-solve launch shape agrees with the default to round-off, and
qg_autotune leaves the model's state where it was."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXTENT = (500e3, 500e3, 4000.0)
FIELDS = ("q", "B", "psi", "A", "C", "u", "v", "w")


def _model(qg, size, precision):
    grid = qg.RectilinearGrid(size=size, extent=EXTENT)
    return qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(),
                         stratification=qg.AnalyticalProfile(lambda z: 1e-4 * math.exp(z / 800.0)),
                         architecture=qg.GPU(precision))


def _init(qg, m, seed=4):
    g = m.grid
    nx, ny, nz = g.size
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    psi = np.empty((nz, nx, ny), complex)
    B = np.empty((nz, nx, ny), complex)
    for k in range(nz):
        psi[k] = 2e3 * np.sin(2 * math.pi * X / g.extent[0] + rng.uniform(0, 6)) * \
            np.cos(4 * math.pi * Y / g.extent[1] + rng.uniform(0, 6))
        B[k] = 0.1 * math.exp(g.z[k] / 200.0) * np.exp(2j * math.pi * (X / g.extent[0] + 2 * Y / g.extent[1]))
    qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))


def _run(qg, m, choices, steps=4, dt=300.0):
    _init(qg, m)
    m.set_tuning(choices)
    qg.step_(m, qg.ExponentialRungeKutta2(dt=dt), steps)
    return {n: getattr(m.fields, n) for n in FIELDS}


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


FFT = ("ya_qw", "xa_qw", "ya2_qeff", "yb_adv", "x_c2r", "xb_q", "xb_B", "yb2_q", "yb2_B", "yf_omega", "xf",
       "y_r2c", "y_herm")
COLUMN = ("thomas_psi", "thomas_A", "thomas_AC", "omega", "thomas_pair")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("size", [(64, 32, 6), (16, 256, 4)])
def test_fft_tile_variants_are_bitwise_identical(qg, precision, size):
    """(16, 256, 4): 256-point columns, where the half tiles of the B update
    copy their base field into LDS while the transforms run."""
    print(f"case tile variants {precision} {size}", flush=True)
    m = _model(qg, size, precision)
    ref = _run(qg, m, {})
    for v in (0, 1, 2):
        out = _run(qg, m, {p: v for p in FFT})
        for n in FIELDS:
            assert np.array_equal(out[n], ref[n]), (precision, v, n)
    m.finalize()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_alternative_fft_plan_agrees_to_round_off(qg, precision, tol):
    """Variant 3 runs the alternative FFT plan (radix 16 x 16 at N = 256): a
    different association of the same transform, equal to round-off."""
    print(f"case alternative plan {precision}", flush=True)
    m = _model(qg, (256, 32, 4), precision)
    ref = _run(qg, m, {})
    out = _run(qg, m, {p: 3 for p in FFT})
    for n in FIELDS:
        assert _rel(out[n], ref[n]) <= tol, (precision, n, _rel(out[n], ref[n]))
    m.finalize()


@pytest.mark.parametrize("nz", [5, 12, 40, 128])
def test_column_shapes_agree_with_default(qg, nz):
    print(f"case column shapes nz={nz}", flush=True)
    m = _model(qg, (16, 16, nz), "fp64")
    ref = _run(qg, m, {})
    tried = 0
    for shape in range(8):  # every compiled shape (kScanShapes)
        try:
            out = _run(qg, m, {p: shape for p in COLUMN})
        except ValueError:  # the shape does not fit this depth
            continue
        tried += 1
        for n in FIELDS:
            e = _rel(out[n], ref[n])
            assert e <= 1e-12, (nz, shape, n, e)
    assert tried >= 3
    m.finalize()


def test_autotune_restores_state_and_caches(qg):
    print("case autotune", flush=True)
    m = _model(qg, (64, 64, 16), "mixed")
    _init(qg, m)
    every = ("q", "B", "psi", "A", "C", "u", "v", "w")
    before = {n: m._h.get_field(n) for n in every}
    choice = m.autotune(dt=300.0, steps=2, force=True)
    assert set(choice) == set(FFT + COLUMN)
    assert all(-1 <= choice[p] <= 3 for p in FFT)
    for n in every:  # the diagnostics too, not only the prognostic state
        assert np.array_equal(m._h.get_field(n), before[n]), n
    # stepping after tuning matches an untuned model to round-off: the
    # column shapes differ in the last FP64 bits, which the FP32 field
    # storage of mixed mode rounds to at most an FP32 ulp here and there
    u = _model(qg, (64, 64, 16), "mixed")
    _init(qg, u)
    ts = qg.ExponentialRungeKutta2(dt=300.0)
    qg.step_(m, ts, 6)
    qg.step_(u, ts, 6)
    for n in ("q", "B", "psi", "A"):
        assert _rel(m._h.get_field(n), u._h.get_field(n)) <= 1e-6, n
    # a second model of the same grid class takes the cached choice
    assert u.autotune(dt=300.0) == choice
    with pytest.raises(ValueError):
        u.set_tuning({"ya_qw": 7})
    with pytest.raises(ValueError):
        u.set_tuning({"thomas_psi": 99})
    m.finalize()
    u.finalize()


def test_autotune_in_step_restores_state(qg):
    """The in-step refinement replays whole steps per candidate: the state
    is still put back, the choice is valid, is cached apart from the
    per-launch choice, and the tuned model steps as an untuned one."""
    print("case autotune in_step", flush=True)
    m = _model(qg, (64, 64, 16), "mixed")
    _init(qg, m)
    every = ("q", "B", "psi", "A", "C", "u", "v", "w")
    before = {n: m._h.get_field(n) for n in every}
    choice = m.autotune(dt=300.0, steps=2, force=True, in_step=True)
    assert set(choice) == set(FFT + COLUMN)
    assert all(-1 <= choice[p] <= 3 for p in FFT)
    for n in every:
        assert np.array_equal(m._h.get_field(n), before[n]), n
    u = _model(qg, (64, 64, 16), "mixed")
    _init(qg, u)
    assert u.autotune(dt=300.0, in_step=True) == choice  # from the cache
    u.set_tuning({p: -1 for p in FFT + COLUMN})
    ts = qg.ExponentialRungeKutta2(dt=300.0)
    qg.step_(m, ts, 6)
    qg.step_(u, ts, 6)
    for n in ("q", "B", "psi", "A"):
        assert _rel(m._h.get_field(n), u._h.get_field(n)) <= 1e-6, n
    m.finalize()
    u.finalize()


def test_autotune_on_base_only_lengths(qg):
    """96 x 80 has no tile variants compiled: the tuner keeps every FFT pass
    at the base variant, still tunes the column shapes, refuses other FFT
    variants, and the tuned model steps as the untuned one to round-off."""
    print("case autotune 96x80", flush=True)
    m = _model(qg, (96, 80, 16), "mixed")
    _init(qg, m)
    choice = m.autotune(dt=300.0, steps=2, force=True)
    assert all(choice[p] == -1 for p in FFT), choice
    with pytest.raises(ValueError):
        m.set_tuning({"ya_qw": 3})
    u = _model(qg, (96, 80, 16), "mixed")
    _init(qg, u)
    ts = qg.ExponentialRungeKutta2(dt=300.0)
    qg.step_(m, ts, 4)
    qg.step_(u, ts, 4)
    for n in ("q", "B", "psi", "A"):
        assert _rel(m._h.get_field(n), u._h.get_field(n)) <= 1e-6, n
    m.finalize()
    u.finalize()
