"""The serving loop: runs admitted into shared ensembles as slots and memory
free up give, each, the result of the same run alone."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _trace(qg, n=6):
    return qg.serving.poisson_trace(n, rate=50.0, seed=11, sizes=(32, 64), depths=(16, 32), steps=(20, 60))


def test_runs_in_the_mix_equal_runs_alone(qg):
    trace = _trace(qg)
    print("case serving mix:", [(r.nx, r.nz, r.steps, r.stratification) for r in trace], flush=True)
    server = qg.serving.Server(precision="mixed", slots=2, chunk=16, keep_fields=True)
    results, summary = server.run(trace, realtime=False)
    assert summary["requests"] == len(trace) and not summary["nonfinite_runs"]
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0
    for req, res in zip(sorted(trace, key=lambda r: r.rid), results):
        alone = qg.serving.run_alone(req, precision="mixed")
        for n in ("q", "B", "psi"):
            a, b = res.fields[n], alone[n]
            assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b)), (req.rid, n)


def test_runs_in_the_mix_equal_runs_alone_tuned(qg):
    """As above with the autotuner on, as the benchmark's serving leg runs:
    each pool tunes its grid class once, and a lone run of the same class
    takes the cached choice, so the results agree bit for bit again."""
    trace = _trace(qg, 4)
    server = qg.serving.Server(precision="mixed", slots=2, chunk=16, keep_fields=True, autotune=True)
    results, summary = server.run(trace, realtime=False)
    assert summary["requests"] == len(trace) and not summary["nonfinite_runs"]
    for req, res in zip(sorted(trace, key=lambda r: r.rid), results):
        alone = qg.serving.run_alone(req, precision="mixed", autotune=True)
        for n in ("q", "B", "psi"):
            a, b = res.fields[n], alone[n]
            assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b)), (req.rid, n)


def test_admission_waits_for_memory(qg):
    """A budget that holds one pool at a time: the second grid class is
    admitted only after the first pool has drained and been released."""
    trace = _trace(qg, 4)
    for i, r in enumerate(trace):  # two grid classes, two runs each
        r.nx, r.nz = (32, 16) if i % 2 == 0 else (64, 16)
    one = max(qg.serving._lib.memory_required(nx=n, ny=n, nz=16, precision=qg.serving._lib.PRECISION_MIXED,
                                              members=2) for n in (32, 64))
    server = qg.serving.Server(precision="mixed", slots=2, chunk=16, memory_budget=one)
    results, summary = server.run(trace, realtime=False)
    assert summary["requests"] == 4 and not summary["nonfinite_runs"]
    by_class = {}
    for req, res in zip(sorted(trace, key=lambda r: r.rid), results):
        by_class.setdefault(req.nx, []).append(res)
    first, second = sorted(by_class.values(), key=lambda rs: min(r.admitted for r in rs))
    assert min(r.admitted for r in second) >= max(r.completed for r in first) - 1e-6
