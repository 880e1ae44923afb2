"""Continuous-batching serving loop for QG-YBJ+ runs on one GPU.

Runs arrive over time.  Runs of one grid class (nx, ny, nz) share a pool: a
device ensemble (``QGYBJEnsemble``) whose passes cover all of its slots in
one launch, so the work of every resident run overlaps on the device.  A run
is admitted into a free slot as soon as one frees up; a pool is created for a
new grid class when the free HBM (less a reserve) holds it.  Each scheduler
round advances every pool that has resident runs by the largest number of
steps no resident run overshoots; pools step concurrently from worker threads
(each pool has its own HIP stream).  A run's result is read back and its
slot freed when its steps are done.  Every run's result is the one the same
model gives alone: members of an ensemble do not interact.

The reference has no serving layer; this module is the host-side scheduler
over the library's ensemble entry points (qg_set_field_member,
qg_set_member_stratification, qg_check_member, qg_memory_required,
qg_device_memory).
"""

from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from . import api as qg
from . import synthetic

EXTENT = (500e3, 500e3, 4000.0)


@dataclass
class RunRequest:
    """One run: grid class, step count and its seeded initial state."""
    rid: int
    nx: int
    nz: int
    steps: int
    seed: int
    wave_amplitude: float
    stratification: str  # "constant" (N² = 1e-5) or "exponential" (1e-4 exp(z/800))
    arrival: float = 0.0  # seconds after the start of the trace

    @property
    def grid_class(self) -> Tuple[int, int, int]:
        return (self.nx, self.nx, self.nz)

    @property
    def gridpoints(self) -> int:
        return self.nx * self.nx * self.nz

    def strat(self):
        if self.stratification == "constant":
            return qg.ConstantStratification(1e-5)
        return qg.AnalyticalProfile(synthetic.exponential_N2(1e-4, 800.0))

    def initial_state(self, grid) -> Tuple["qg.RandomStreamfunction", "qg.SurfaceWave"]:
        """The reference's seeded k^-3 streamfunction (RandomStreamfunction,
        generated on the device) scaled to an expected RMS velocity of
        0.1 m/s, and a surface wave over a 50 m mixed layer."""
        amp = synthetic.random_streamfunction_amplitude(grid, urms=0.1, slope=-3.0)
        return (qg.RandomStreamfunction(amplitude=amp, spectral_slope=-3.0, seed=self.seed),
                qg.SurfaceWave(self.wave_amplitude, 50.0))


@dataclass
class RunResult:
    rid: int
    arrival: float
    admitted: float
    completed: float
    steps: int
    gridpoints: int
    nonfinite: bool
    fields: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


def poisson_trace(n: int, rate: float, seed: int = 7, sizes=(64, 128, 256, 384), depths=(32, 64, 128),
                  steps=(100, 2000)) -> List[RunRequest]:
    """n requests with exponential inter-arrival times (rate per second),
    nx = ny uniform over `sizes`, nz uniform over `depths`, steps
    log-uniform over `steps`, wave amplitude U[0.05, 0.2], constant or
    exponential N² with equal odds, and a per-request seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        nsteps = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))))
        out.append(RunRequest(rid=i, nx=int(rng.choice(sizes)), nz=int(rng.choice(depths)), steps=nsteps,
                              seed=int(seed * 100003 + i), wave_amplitude=float(rng.uniform(0.05, 0.2)),
                              stratification="constant" if rng.random() < 0.5 else "exponential",
                              arrival=t))
    return out


def expected_work(sizes=(64, 128, 256, 384), depths=(32, 64, 128), steps=(100, 2000)) -> float:
    """Mean gridpoint-steps of one request of `poisson_trace`."""
    mean_pts = np.mean([s * s for s in sizes]) * np.mean(depths)
    mean_steps = (steps[1] - steps[0]) / math.log(steps[1] / steps[0])
    return float(mean_pts * mean_steps)


class Pool:
    """The ensemble of one grid class and the runs in its slots.  Resident
    runs occupy the leading slots [0, n) and only those are stepped: a run
    that finishes is replaced by the last resident one (qg_move_member)."""

    def __init__(self, grid_class, slots: int, precision: str, dt: float, autotune: bool = False):
        nx, ny, nz = grid_class
        self.grid = qg.RectilinearGrid(size=(nx, ny, nz), extent=EXTENT, centered=True)
        self.slots = slots
        self.ens = qg.QGYBJEnsemble(stratifications=[qg.ConstantStratification(1e-5)] * slots, grid=self.grid,
                                    coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(),
                                    formulation=qg.YBJPlus(), architecture=qg.GPU(precision))
        if autotune:  # timed on the zero state of every slot; cached per grid class
            self.ens.autotune(dt, steps=2)
        self.ens.set_active_members(0)
        self.ts = qg.ExponentialRungeKutta2(dt=dt)
        self.runs: List[RunRequest] = []
        self.left: List[int] = []
        self.admitted: List[float] = []

    def has_room(self) -> bool:
        return len(self.runs) < self.slots

    def admit(self, req: RunRequest, now: float):
        slot = len(self.runs)
        self.ens.set_member_stratification(slot, req.strat())
        psi, wave = req.initial_state(self.grid)
        self.ens.set_member(slot, psi=psi, waves=wave)
        self.runs.append(req)
        self.left.append(req.steps)
        self.admitted.append(now)
        self.ens.set_active_members(len(self.runs))

    def active(self) -> bool:
        return bool(self.runs)

    def chunk(self, cap: int) -> int:
        return min(self.left + [cap])

    def advance(self, n: int):
        qg.step_(self.ens, self.ts, n)
        self.left = [x - n for x in self.left]

    def finish(self, now_fn, keep_fields: bool, arrival_of) -> List[RunResult]:
        done = []
        for i in reversed(range(len(self.runs))):
            if self.left[i] != 0:
                continue
            r = self.runs[i]
            bad, _ = self.ens.check_member(i)
            mf = self.ens.member_fields(i)
            fields = {n: getattr(mf, n) for n in ("q", "B", "psi")} if keep_fields else {}
            done.append(RunResult(r.rid, arrival_of(r), self.admitted[i], now_fn(), r.steps, r.gridpoints, bad,
                                  fields))
            last = len(self.runs) - 1
            if i != last:
                self.ens.move_member(last, i)
                self.runs[i], self.left[i], self.admitted[i] = self.runs[last], self.left[last], self.admitted[last]
            self.runs.pop()
            self.left.pop()
            self.admitted.pop()
            self.ens.set_active_members(len(self.runs))
        return done

    def close(self):
        self.ens.finalize()


class Server:
    """Admission by free HBM, continuous batching per grid class."""

    def __init__(self, *, precision: str = "mixed", dt: float = 20.0, slots: int = 4, chunk: int = 50,
                 reserve_bytes: int = 8 << 30, memory_budget: Optional[int] = None, keep_fields: bool = False,
                 workers: int = 4, autotune: bool = False):
        self.autotune = autotune  # tune each new pool's tile variants and launch shapes
        self.precision = precision
        self.dt = dt
        self.slots = slots
        self.chunk_cap = chunk
        self.reserve = reserve_bytes
        self.budget = memory_budget  # cap on bytes held by pools (None: free HBM decides)
        self.keep_fields = keep_fields
        self.pools: Dict[Tuple[int, int, int], Pool] = {}
        self.pool_bytes: Dict[Tuple[int, int, int], int] = {}
        self.workers = workers

    def _prec_id(self):
        return _lib.PRECISION_MIXED if self.precision == "mixed" else _lib.PRECISION_FP64

    def _can_hold(self, grid_class) -> bool:
        nx, ny, nz = grid_class
        need = _lib.memory_required(nx=nx, ny=ny, nz=nz, precision=self._prec_id(), members=self.slots)
        if self.budget is not None:
            return sum(self.pool_bytes.values()) + need <= self.budget
        free, _ = _lib.device_memory()
        return free - need >= self.reserve

    def _release_idle_pool(self) -> bool:
        for gc, p in list(self.pools.items()):
            if not p.active():
                p.close()
                del self.pools[gc]
                del self.pool_bytes[gc]
                return True
        return False

    def _try_admit(self, req: RunRequest, now: float) -> bool:
        gc = req.grid_class
        pool = self.pools.get(gc)
        if pool is None:
            while not self._can_hold(gc):
                if not self._release_idle_pool():
                    return False
            nx, ny, nz = gc
            pool = Pool(gc, self.slots, self.precision, self.dt, autotune=self.autotune)
            self.pools[gc] = pool
            self.pool_bytes[gc] = _lib.memory_required(nx=nx, ny=ny, nz=nz, precision=self._prec_id(),
                                                       members=self.slots)
        if not pool.has_room():
            return False
        pool.admit(req, now)
        return True

    def run(self, trace: List[RunRequest], realtime: bool = True) -> Tuple[List[RunResult], dict]:
        """Serve `trace`.  With realtime, requests become visible at their
        arrival offsets from the start and latency counts from the arrival;
        otherwise all are queued at once and latency counts from the start."""
        queue = sorted(trace, key=lambda r: r.arrival)
        pending: List[RunRequest] = []
        results: List[RunResult] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        nxt = 0
        busy = 0.0
        lock = threading.Lock()
        with ThreadPoolExecutor(max_workers=self.workers) as ex:
            while len(results) < len(trace):
                t = now()
                while nxt < len(queue) and (not realtime or queue[nxt].arrival <= t):
                    pending.append(queue[nxt])
                    nxt += 1
                still = []
                for req in pending:
                    if not self._try_admit(req, now()):
                        still.append(req)
                pending = still
                active = [p for p in self.pools.values() if p.active()]
                if not active:
                    if nxt < len(queue):
                        time.sleep(max(0.0, min(queue[nxt].arrival - now(), 0.05)))
                        continue
                    if pending:
                        raise RuntimeError("a request cannot be admitted even with every pool idle")
                    break
                r0 = now()
                jobs = [(p, p.chunk(self.chunk_cap)) for p in active]
                list(ex.map(lambda pj: pj[0].advance(pj[1]), jobs))
                busy += now() - r0
                for p in active:
                    done = p.finish(now, self.keep_fields, (lambda r: r.arrival) if realtime else (lambda r: 0.0))
                    with lock:
                        results.extend(done)
        makespan = now()
        for p in self.pools.values():
            p.close()
        self.pools.clear()
        self.pool_bytes.clear()
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        work = sum(r.gridpoints * r.steps for r in results)
        summary = {
            "requests": len(results),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(np.mean(lat)),
            "gridpoint_steps": int(work),
            "makespan_s": makespan,
            "value": work / makespan / 1e6,
            "unit": "Mgridpt-steps/s",
            "device_busy_fraction": busy / makespan if makespan > 0 else 0.0,
            "nonfinite_runs": [r.rid for r in results if r.nonfinite],
        }
        return sorted(results, key=lambda r: r.rid), summary


def run_alone(req: RunRequest, precision: str = "mixed", dt: float = 20.0,
              autotune: bool = False) -> Dict[str, np.ndarray]:
    """The same run on a model of its own: the reference result for the
    alone-versus-in-the-mix check.  With autotune the model takes the tuning
    of its grid class (cached in the process by the pool that tuned it)."""
    grid = qg.RectilinearGrid(size=(req.nx, req.nx, req.nz), extent=EXTENT, centered=True)
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), stratification=req.strat(),
                      feedback=qg.WaveMeanFeedback(), formulation=qg.YBJPlus(), architecture=qg.GPU(precision))
    if autotune:
        m.autotune(dt, steps=2)
    psi, wave = req.initial_state(grid)
    qg.set_(m, psi=psi, waves=wave)
    qg.step_(m, qg.ExponentialRungeKutta2(dt=dt), req.steps)
    out = {n: getattr(m.fields, n) for n in ("q", "B", "psi")}
    m.finalize()
    return out
