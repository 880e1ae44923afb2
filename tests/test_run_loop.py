"""The host side of run! (src/simulation.jl:570-662) without a GPU: the
stretches of steps handed to the device end exactly where the reference's
per-iteration loop would stop, record diagnostics or print progress."""

import pytest


class _Model:
    finalized = False


def _brute(sim, mgr, report):
    """Iterations until the per-iteration loop stops or acts on the host
    (a progress report every `report` iterations, 0: none)."""
    from copy import deepcopy
    s = deepcopy(sim)
    n = 0
    while True:
        if s.stop_iteration is not None and s.clock.iteration >= s.stop_iteration:
            return n
        if not s._before_stop_time():
            return n
        n += 1
        s.clock.iteration += 1
        s.clock.time += s.timestepper.dt
        if (report and s.clock.iteration % report == 0) or (mgr is not None and mgr.due(s.clock)):
            return n


@pytest.mark.parametrize("dt,stop_iteration,stop_time,schedule,report,start", [
    (20.0, 100, None, None, 0, (0.0, 0)),
    (20.0, 100, None, None, 1, (0.0, 0)),
    (20.0, 100, None, None, 10, (0.0, 0)),                    # progress every 10 iterations
    (20.0, 95, None, ("it", 7), 10, (0.0, 3)),
    (20.0, None, 2000.0, None, 0, (0.0, 0)),                  # stop time on the grid of steps (tolerance)
    (0.1, None, 1.0, None, 0, (0.0, 0)),                      # 0.1 is inexact: the tolerance decides
    (0.3, None, 1.0, None, 0, (0.0, 0)),
    (20.0, 1000, None, ("it", 7), 0, (0.0, 0)),
    (20.0, 1000, None, ("it", 7), 0, (140.0, 7)),             # resumed on a due iteration
    (0.1, None, 50.0, ("time", 0.7), 0, (0.0, 0)),
    (3.0, None, 1e5, ("time", 100.0), 0, (33.0, 11)),
])
def test_stretch_ends_where_the_loop_acts(qg, dt, stop_iteration, stop_time, schedule, report, start):
    sim = qg.Simulation(_Model(), dt=dt, stop_iteration=stop_iteration, stop_time=stop_time)
    sim.clock.time, sim.clock.iteration = start
    mgr = None
    if schedule is not None:
        kind, v = schedule
        sched = qg.IterationInterval(v) if kind == "it" else qg.TimeInterval(v)
        mgr = qg.EnergyDiagnostics(sched, start[0])
        mgr.last_iteration = start[1]
    # walk the whole run stretch by stretch, as run! does
    total = 0
    for _ in range(10000):
        n = sim._steps_to_next_event(mgr, report)
        assert n == _brute(sim, mgr, report)
        if n == 0:
            break
        for _ in range(n):
            sim.clock.iteration += 1
            sim.clock.time += dt
        total += n
        if mgr is not None and mgr.due(sim.clock):
            mgr.last_iteration = sim.clock.iteration
            if mgr.next_time is not None:
                while mgr.next_time <= sim.clock.time + 1e-9:
                    mgr.next_time += mgr.schedule.interval
    assert total > 0


def test_run_options_are_checked(qg):
    with pytest.raises(ValueError, match="diagnostics_interval must be positive"):
        qg.Simulation(_Model(), dt=1.0, stop_iteration=3, diagnostics_interval=0)
    sim = qg.Simulation(_Model(), dt=1.0, stop_iteration=3)
    assert sim.diagnostics_interval == 10 and sim.verbose
