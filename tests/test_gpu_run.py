"""run!'s stepping loop with the termination check made on the device
(qg_run; src/simulation.jl:621-625 and _check_termination_conditions!,
src/simulation.jl:25-44).  Against plain stepping with a host check after
every step: the same fields bit for bit, the same maximum (the failing
step's, else the last step's), and a stop right after the first failing
step with the reference's failure kind."""

import math

import numpy as np
import pytest

from tests.helpers import Case, gpu_model, k3_streamfunction
from oracle import qgybj_oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("q", "B", "psi", "A", "C", "u", "v", "w")


def _pair(qg, precision, urms, seed=5):
    c = Case(size=(32, 32, 8), N2=1e-5, precision=precision)
    grid = O.make_grid(c.size, extent=c.extent, centered=c.centered)
    psi = k3_streamfunction(grid, seed=seed, urms=urms, kpeak=3.0)
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(grid.nz):
        B[k] = 0.1 * math.exp(grid.z[k] / 300.0) * np.exp(2j * math.pi * X / grid.Lx)
    out = []
    for _ in range(2):
        m = gpu_model(qg, c)
        qg.set_(m, psi=qg.FieldArray(psi), waves=qg.FieldArray(B))
        out.append(m)
    return out


def _stepwise(qg, m, dt, n):
    """step! then the check, one step at a time: (steps, failure, max |ifft
    psi| of the failing step, else of the last)."""
    ts = qg.ExponentialRungeKutta2(dt=dt)
    v = 0.0
    for i in range(1, n + 1):
        qg.step_(m, ts, 1)
        bad, v = m._h.check()
        if bad or not v <= 1e10:
            return i, (1 if bad else 2), v
    return n, 0, v


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_equals_stepping_with_checks(qg, precision):
    print(f"case run without failure {precision}", flush=True)
    a, b = _pair(qg, precision, urms=0.5)
    n = 70  # two chunks of 32 steps and a tail, graph-replayed pairs and single steps
    ref = _stepwise(qg, a, 500.0, n)
    got = b._h.run(500.0, n)
    assert ref[:2] == (n, 0) and got[:2] == (n, 0)
    assert got[2] == ref[2] > 0
    for f in FIELDS:
        assert np.array_equal(getattr(b.fields, f), getattr(a.fields, f)), f
    a.finalize()
    b.finalize()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_replays_chunks_whose_bound_is_too_high(qg, precision):
    """qg_run passes a stretch of steps on a bound of max |ifft(psi)| (the
    largest over the levels of sum |psi| over the spectrum / (nx ny)).
    Here max |ifft(psi)| is about 5e9 while the bound is about 2.4e10: every
    stretch is stepped again with the exact check, nothing fails, and the
    result is that of step-then-check, bit for bit."""
    print(f"case run with the bound above the limit {precision}", flush=True)
    a, b = _pair(qg, precision, urms=1e5)
    nx, ny, _ = a.grid.size
    bound = float(np.max(np.sum(np.abs(a.fields.psi), axis=(1, 2)))) / (nx * ny)
    exact = a._h.check()[1]
    assert bound > 1.5e10 and 1e9 < exact < 0.9e10, (bound, exact)
    n = 40  # one chunk of 32 steps and a tail
    ref = _stepwise(qg, a, 1e-3, n)
    got = b._h.run(1e-3, n)
    assert ref[:2] == (n, 0) and got[:2] == (n, 0)
    assert got[2] == ref[2] > 1e9
    for f in FIELDS:
        assert np.array_equal(getattr(b.fields, f), getattr(a.fields, f)), f
    a.finalize()
    b.finalize()


@pytest.mark.parametrize("precision,urms,dt", [("fp64", 5.0, 1500.0), ("mixed", 5.0, 1500.0),
                                               ("fp64", 2.0, 2000.0)])
def test_run_stops_right_after_the_failing_step(qg, precision, urms, dt):
    print(f"case run blow-up {precision} urms={urms} dt={dt}", flush=True)
    a, b = _pair(qg, precision, urms=urms)
    n = 140
    k, kind, mx = _stepwise(qg, a, dt, n)
    assert 2 <= k < n and kind in (1, 2), (k, kind)
    done, failure, m = b._h.run(dt, n)
    assert (done, failure) == (k, kind)
    assert m == mx or (math.isnan(m) and math.isnan(mx))
    for f in FIELDS:
        assert np.array_equal(getattr(b.fields, f), getattr(a.fields, f), equal_nan=True), f
    a.finalize()
    b.finalize()


def test_run_loop_raises_the_reference_errors(qg):
    print("case run! blow-up error", flush=True)
    a, b = _pair(qg, "fp64", urms=5.0)
    sim = qg.Simulation(b, dt=1500.0, stop_iteration=140)
    with pytest.raises(FloatingPointError, match="blowing up|non-finite"):
        qg.run_(sim)
    assert sim.state is qg.SimulationState.Failed
    k, _, _ = _stepwise(qg, a, 1500.0, 140)
    assert sim.clock.iteration == k
    a.finalize()
    b.finalize()


def test_run_loop_stops_at_time_and_records_diagnostics(qg):
    """Stretches of device steps end where the host loop would act: the stop
    time (with its tolerance) and every due diagnostics record."""
    print("case run! stop time and diagnostics", flush=True)
    a, b = _pair(qg, "fp64", urms=0.5)
    sim = qg.Simulation(b, dt=500.0, stop_time=500.0 * 45, diagnostics=qg.IterationInterval(10))
    qg.run_(sim)
    assert sim.clock.iteration == 45
    assert sim.state is qg.SimulationState.Stopped
    ts = qg.ExponentialRungeKutta2(dt=500.0)
    series = []
    for i in range(0, 46):
        if i % 10 == 0:
            series.append(a.energies()["wave_KE"])
        if i < 45:
            qg.step_(a, ts, 1)
    assert sim.diagnostics_manager.series["wave_KE"] == series
    for f in ("q", "B", "psi"):
        assert np.array_equal(getattr(b.fields, f), getattr(a.fields, f)), f
    a.finalize()
    b.finalize()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_progress_maxima_match_the_fields(qg, precision, tol):
    """_progress_maxima (src/simulation.jl:46-81): max |ifft2(B + kh2 A / 4)|
    and max hypot(u, v), against numpy on the downloaded fields."""
    print(f"case progress maxima {precision}", flush=True)
    a, _ = _pair(qg, precision, urms=0.5)
    qg.step_(a, qg.ExponentialRungeKutta2(dt=500.0), 3)
    m = a.progress_maxima()
    kh2 = a.grid.kh2[None]
    la = np.fft.ifft2(a.fields.B + 0.25 * kh2 * a.fields.A, axes=(1, 2))
    wave = float(np.max(np.abs(la)))
    flow = float(np.max(np.hypot(a.fields.u, a.fields.v)))
    assert abs(m["wave_speed"] - wave) <= tol * wave, (m, wave)
    assert abs(m["flow_speed"] - flow) <= tol * flow, (m, flow)
    a.finalize()


def test_progress_lines_do_not_change_the_run(qg, capsys):
    print("case run! progress lines", flush=True)
    a, b = _pair(qg, "fp64", urms=0.5)
    qg.run_(qg.Simulation(a, dt=500.0, stop_iteration=23))
    qg.run_(qg.Simulation(b, dt=500.0, stop_iteration=23), progress=True, diagnostics_interval=5)
    lines = [ln for ln in capsys.readouterr().out.splitlines() if ln.startswith("iteration=")]
    assert [int(ln.split("|")[0].split("=")[1]) for ln in lines] == [5, 10, 15, 20]
    assert "max_wave_speed=" in lines[0] and "max_flow_speed=" in lines[0]
    for f in FIELDS:
        assert np.array_equal(getattr(b.fields, f), getattr(a.fields, f)), f
    a.finalize()
    b.finalize()
