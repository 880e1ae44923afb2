"""Benchmark: ETD-RK2 steps of the QG-YBJ+ model at 256x256x128 on one MI355X.

Workload (the plain-vs-mixed configuration): 256x256x128 over 500 km x 500 km
x 4000 m, f = 1e-4, N²(z) = 1e-4 exp(z/800 m), seeded k^-3 streamfunction
peaked at wavenumber 6 with RMS velocity 0.1 m/s, surface wave 0.1 m/s over a
50 m mixed layer, wave-mean feedback on, YBJ+, default closure, dt = 20 s.
One "step" is one ETD-RK2 step of the whole grid; inputs are resident in HBM
when the timed region starts.  The headline `value` is the mixed-precision
throughput; the FP64 throughput and the mixed-vs-FP64 relative L2 error per
field after at least 1000 steps (untimed steps are added when warmup+steps is
fewer) are reported beside it.

Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mgridpt-steps/s at 256²×128, FP64 vs mixed at stated err; p50/p99 latency"
PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
SIZE = (256, 256, 128)
EXTENT = (500e3, 500e3, 4000.0)
DT = 20.0


# Kernel behind each of the library's per-launch names: (kernel, tuned pass,
# launch kind[, trailing template arguments]).  Several names share a kernel (the two velocity passes); the
# stage and final launches of one kernel are told apart by launch order in the
# rocprofv3 summary.
LAUNCH_KERNEL = {
    "ya_T1": ("k_ya_qw", "ya_qw", None), "ya_qw": ("k_ya_qw", "ya_qw", None),
    "xa_qw": ("k_xa_qw", "xa_qw", None), "ya2_qeff": ("k_ya2_qeff", "ya2_qeff", None),
    "yb_uv": ("k_yb_adv", "yb_adv", None, 3), "yb_adv": ("k_yb_adv", "yb_adv", None, 12),
    "yb_adv_uv": ("k_yb_adv", "yb_adv", None, 15),
    "x_c2r_uv": ("k_x_c2r", "x_c2r", None), "x_c2r_w": ("k_x_c2r", "x_c2r", None),
    "xb_q": ("k_xb_q", "xb_q", None, 0), "xb_q_uv": ("k_xb_q", "xb_q", None, 1), "xb_B": ("k_xb_B", "xb_B", None),
    "xb_all": ("k_xb_all", "xb_B", None, 0), "xb_all_uv": ("k_xb_all", "xb_B", None, 1),
    "yb2_q_stage": ("k_yb2_q", "yb2_q", "stage", 1), "yb2_q_final": ("k_yb2_q", "yb2_q", "final", 1),
    "yb2_B_stage": ("k_yb2_B", "yb2_B", "stage"), "yb2_B_final": ("k_yb2_B", "yb2_B", "final"),
    "yf_omega": ("k_yf_omega", "yf_omega", None, 1, 0), "xf": ("k_xf", "xf", None, 1),
    "y_r2c_omega": ("k_y_r2c_mask", "y_r2c", None), "y_herm_w": ("k_y_herm", "y_herm", None),
    "thomas_psi": ("k_colscan<0", "thomas_psi", None), "thomas_A": ("k_colscan<1", "thomas_A", None),
    "thomas_AC": ("k_colscan<2", "thomas_AC", None), "omega_gtsv": ("k_colscan<3", "omega", None),
}
# Column-solve launch shapes (levels per partition, quarter columns per
# block) in the library's shape-index order, and the default index for a
# column depth above 8.
SCAN_SHAPES = ((8, 8), (16, 8), (32, 8), (16, 4), (32, 4), (64, 4), (16, 16), (32, 16))
SCAN_DEFAULT = 1
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_256x256x128_mixed.json")


def kernel_symbol(name, choice):
    """The demangled symbol rocprofv3 reports for launch `name` of the FP32
    mixed-precision model at 256x256 under tuning `choice` (pass -> variant
    or shape index, -1 = default), with its launch kind."""
    base, tpass, kind, *extra = LAUNCH_KERNEL[name]
    v = int((choice or {}).get(tpass, -1))
    if base.startswith("k_colscan"):
        m, qc = SCAN_SHAPES[v if v >= 0 else SCAN_DEFAULT]
        return f"{base}, {m}, {qc}, float>", kind
    tail = "".join(f", {e}" for e in extra)  # compile-time output set / input mode
    return f"{base}<float, {SIZE[0]}, {v if v >= 0 else 1}{tail}>", kind


def measured_traffic(name, choice):
    """HBM bytes per launch of `name` from the committed rocprofv3 FETCH_SIZE +
    WRITE_SIZE summary (already scaled by the copy-kernel calibration), or
    None when that summary has no entry for the kernel."""
    if name not in LAUNCH_KERNEL or not os.path.exists(TRAFFIC_FILE):
        return None
    sym, kind = kernel_symbol(name, choice)
    with open(TRAFFIC_FILE) as fh:
        entry = json.load(fh).get("kernels", {}).get(sym)
    if entry is None:
        return None
    if kind is not None and kind in entry.get("by_launch", {}):
        entry = entry["by_launch"][kind]
    return entry["hbm_bytes_per_launch"]


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def build_model(qg, precision):
    from qgybjplus_amd import synthetic  # noqa: F401  (registered by load_package)
    grid = qg.RectilinearGrid(size=SIZE, extent=EXTENT, centered=True)
    model = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4),
                          stratification=qg.AnalyticalProfile(synthetic.exponential_N2(1e-4, 800.0)),
                          flow=qg.EvolvingFlow(), feedback=qg.WaveMeanFeedback(),
                          formulation=qg.YBJPlus(), architecture=qg.GPU(precision))
    return grid, model


def workload_inputs(grid):
    """The workload's initial fields: physical psi (the seeded k^-3 field) and
    the physical surface wave 0.1 m/s exp(-(depth/50 m)^2) that
    SurfaceWave(0.1, 50.0) writes."""
    from qgybjplus_amd import synthetic
    psi = synthetic.k3_streamfunction(grid, seed=2026, urms=0.1, kpeak=6.0)
    nx, ny, nz = grid.size
    B = np.empty((nz, nx, ny), dtype=np.complex128)
    for k in range(nz):
        depth = max(0.0, -grid.z[k] - grid.dz / 2)
        B[k] = 0.1 * math.exp(-(depth ** 2) / 50.0 ** 2)
    return psi, B


def set_workload(qg, model, psi):
    """set!(model, psi = psi, pv_method = :qg) and the surface wave."""
    qg.set_(model, psi=qg.FieldArray(psi), pv_method="qg")
    qg.set_(model, waves=qg.SurfaceWave(0.1, 50.0))


def timed_run(qg, model, steps, warmup):
    """The timed region: `steps` steps as the library runs them (steady-state
    pairs replay a captured hipGraph), with one HIP event per step interval."""
    h = model._h
    ts = qg.ExponentialRungeKutta2(dt=DT)
    if warmup:
        qg.step_(model, ts, warmup)
    h.synchronize()
    h.profile_begin(kernels=False, steps=True)
    t0 = time.perf_counter()
    qg.step_(model, ts, steps)
    h.synchronize()
    t1 = time.perf_counter()
    prof = h.profile_end()
    return t1 - t0, prof


def kernel_run(qg, model, steps):
    """`steps` more steps run eagerly with HIP events around every kernel
    launch on the library's stream: per-kernel durations and algorithmic
    bytes for the roofline."""
    h = model._h
    h.profile_begin(kernels=True, steps=False)
    qg.step_(model, qg.ExponentialRungeKutta2(dt=DT), steps)
    h.synchronize()
    return h.profile_end()


def run_loop_leg(qg, model, steps):
    """run!(simulation) for `steps` iterations of `model`: the termination
    check (non-finite q, B, psi; max |ifft psi| <= 1e10) after every step, as
    the reference's run! makes it, here on the device."""
    log(f"run! loop over {steps} iterations")
    qg.run_(qg.Simulation(model, dt=DT, stop_iteration=4))  # captures the checked step graphs
    sim = qg.Simulation(model, dt=DT, stop_iteration=steps)
    model._h.synchronize()
    t0 = time.perf_counter()
    qg.run_(sim)
    model._h.synchronize()
    t = time.perf_counter() - t0
    npts = SIZE[0] * SIZE[1] * SIZE[2]
    return {"iterations": steps, "value": npts * steps / t / 1e6, "unit": "Mgridpt-steps/s",
            "ms_per_step": 1e3 * t / steps, "check": "every step, on the device"}


ENSEMBLE_CLASSES = ((64, 32), (128, 64), (256, 64))  # (nx = ny, nz), 16 members each
ENSEMBLE_MEMBERS = 16


def ensemble_n0(seed=2026):
    """N0² of the 48 members, class by class."""
    rng = np.random.default_rng(seed)
    return rng.uniform(5e-6, 5e-5, size=len(ENSEMBLE_CLASSES) * ENSEMBLE_MEMBERS)


def ensemble_member_psi(grid, ci, j, seed=2026):
    """Member j of grid class ci: the class's seeded k^-3 field plus 5 % of a
    member-seeded one."""
    from qgybjplus_amd import synthetic
    base = synthetic.k3_streamfunction(grid, seed=seed, urms=0.1, kpeak=6.0)
    noise = synthetic.k3_streamfunction(grid, seed=seed + 1 + ci * ENSEMBLE_MEMBERS + j, urms=0.1, kpeak=6.0)
    return base + 0.05 * noise


def ensemble_model_kwargs(qg, ci):
    """Grid and physics of class ci, for an ensemble or a member run alone."""
    n, nz = ENSEMBLE_CLASSES[ci]
    grid = qg.RectilinearGrid(size=(n, n, nz), extent=EXTENT, centered=True)
    return dict(grid=grid, coriolis=qg.FPlane(1e-4), feedback=qg.WaveMeanFeedback(), formulation=qg.YBJPlus(),
                architecture=qg.GPU("mixed"))


def build_ensembles(qg, seed=2026):
    """The three device ensembles of the batched-ensemble workload with their
    initial states set; returns them and the total grid points."""
    from qgybjplus_amd import synthetic
    n0 = ensemble_n0(seed)
    ens_list, points = [], 0
    for ci, (n, nz) in enumerate(ENSEMBLE_CLASSES):
        kw = ensemble_model_kwargs(qg, ci)
        strats = [qg.AnalyticalProfile(synthetic.exponential_N2(float(n0[ci * ENSEMBLE_MEMBERS + j]), 800.0))
                  for j in range(ENSEMBLE_MEMBERS)]
        ens = qg.QGYBJEnsemble(stratifications=strats, **kw)
        for j in range(ENSEMBLE_MEMBERS):
            ens.set_member(j, psi=qg.FieldArray(ensemble_member_psi(kw["grid"], ci, j, seed)),
                           waves=qg.SurfaceWave(0.1, 50.0))
        ens_list.append(ens)
        points += n * n * nz * ENSEMBLE_MEMBERS
    return ens_list, points


def ensemble_leg(qg, steps, warmup=2, seed=2026, autotune=True):
    """The batched-ensemble workload: 48 members of uneven grids, 16 each of
    64²x32, 128²x64 and 256²x64 over 500 km x 500 km x 4000 m, N² = N0²
    exp(z/800) with N0² ~ U[5e-6, 5e-5], psi = the seeded k^-3 field plus 5 %
    of a member-seeded one, surface wave 0.1 m/s over 50 m, feedback on,
    mixed precision, dt = 20 s.  Each grid class is one device ensemble whose
    passes cover its 16 members in one launch.  The classes are timed one
    after the other (per class) and then stepping concurrently, each on its
    own HIP stream from its own host thread, as the serving loop runs its
    pools; `value` is the concurrent throughput over all 48 members."""
    ens_list, points = build_ensembles(qg, seed)
    ts = qg.ExponentialRungeKutta2(dt=DT)
    for ens in ens_list:
        if autotune:
            ens.autotune(DT, steps=4)
        qg.step_(ens, ts, warmup)
    per_class = []
    t0 = time.perf_counter()
    for ens in ens_list:
        c0 = time.perf_counter()
        qg.step_(ens, ts, steps)
        ens._h.synchronize()
        el = time.perf_counter() - c0
        nx, ny, nz = ens.grid.size
        per_class.append({"grid": [nx, ny, nz], "members": ens.members, "seconds": el,
                          "Mgridpt_steps_per_s": nx * ny * nz * ens.members * steps / el / 1e6})
    sequential = time.perf_counter() - t0
    errors = []

    def advance(ens):
        try:
            qg.step_(ens, ts, steps)
            ens._h.synchronize()
        except Exception as e:  # reported after the join
            errors.append(e)

    threads = [threading.Thread(target=advance, args=(ens,)) for ens in ens_list]
    t0 = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    total = time.perf_counter() - t0
    if errors:
        raise errors[0]
    bad = [j for ens in ens_list for j in range(ens.members) if ens.check_member(j)[0]]
    for ens in ens_list:
        ens.finalize()
    return {"workload": "48-member ensemble: 16 x 64^2x32, 16 x 128^2x64, 16 x 256^2x64, mixed precision, "
                        "N0^2 ~ U[5e-6, 5e-5], perturbed k^-3 psi, dt = 20 s",
            "members": len(ENSEMBLE_CLASSES) * ENSEMBLE_MEMBERS, "gridpoints": points, "steps": steps,
            "warmup": warmup, "value": points * steps / total / 1e6, "unit": "Mgridpt-steps/s",
            "seconds": total, "classes": "concurrent, one HIP stream each",
            "sequential": {"value": points * steps / sequential / 1e6, "seconds": sequential},
            "per_class": per_class, "nonfinite_members": bad}


C3_SIZE = (512, 512, 256)


def large_leg(qg, steps, kernel_steps=20, seed=2026, autotune=True):
    """The single large run: 512x512x256 over 500 km x 500 km x 4000 m, the
    256²x128 case's stratification and seeded k^-3 psi, B = 32 random
    horizontal modes with a 30 m surface e-folding scale, feedback on, mixed
    precision, dt = 10 s.  Throughput over `steps` graph-replayed steps, then
    per-kernel HIP events over `kernel_steps` eager steps for the roofline of
    each kernel at a size well past the 256 MB infinity cache."""
    from qgybjplus_amd import synthetic
    grid = qg.RectilinearGrid(size=C3_SIZE, extent=EXTENT, centered=True)
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4),
                      stratification=qg.AnalyticalProfile(synthetic.exponential_N2(1e-4, 800.0)),
                      feedback=qg.WaveMeanFeedback(), formulation=qg.YBJPlus(), architecture=qg.GPU("mixed"))
    qg.set_(m, psi=qg.FieldArray(synthetic.k3_streamfunction(grid, seed=seed, urms=0.1, kpeak=6.0)),
            waves=qg.FieldArray(synthetic.wave_modes(grid, 32, seed=seed, rms=0.1, efold=30.0), space="spectral"))
    ts = qg.ExponentialRungeKutta2(dt=10.0)
    choice = m.autotune(10.0, steps=3) if autotune else None
    qg.step_(m, ts, 2)
    m._h.synchronize()
    t0 = time.perf_counter()
    qg.step_(m, ts, steps)
    m._h.synchronize()
    el = time.perf_counter() - t0
    m._h.profile_begin(kernels=True, steps=False)
    qg.step_(m, ts, kernel_steps)
    m._h.synchronize()
    prof = m._h.profile_end()
    bad, mx = m._h.check()
    m.finalize()
    npts = C3_SIZE[0] * C3_SIZE[1] * C3_SIZE[2]
    kern = {k: {"launches": v["count"], "avg_us": 1e3 * v["ms"] / v["count"],
                "GBps": v["bytes"] / (v["ms"] / 1e3) / 1e9, "frac_of_peak": v["bytes"] / (v["ms"] / 1e3) / PEAK_HBM}
            for k, v in prof["kernels"].items()}
    return {"workload": "512x512x256, 500 km x 500 km x 4000 m, N2=1e-4*exp(z/800), seeded k^-3 psi, "
                        "B = 32 random modes x exp(-depth/30 m), feedback on, YBJ+, mixed precision, dt=10 s",
            "steps": steps, "value": npts * steps / el / 1e6, "unit": "Mgridpt-steps/s",
            "ms_per_step": 1e3 * el / steps, "kernel_profile_steps": kernel_steps, "kernels": kern,
            "autotune": choice,
            "nonfinite": bool(bad), "max_abs_psi": mx}


def serving_leg(qg, n_requests, capacity_mpts, seed=7, utilization=0.8, autotune=True):
    """The serving workload: Poisson arrivals of QG-YBJ+ runs with grids
    uniform over {64,128,256,384}² x {32,64,128}, steps log-uniform over
    [100, 2000], seeded k^-3 psi, wave amplitude U[0.05, 0.2], constant or
    exponential N², mixed precision, dt = 20 s.  The arrival rate is set for
    `utilization` of the measured single-model throughput.  Runs share
    per-grid-class device ensembles (continuous batching); reports p50/p99
    completion latency (arrival to result on the host) and throughput."""
    from qgybjplus_amd import serving
    rate = utilization * capacity_mpts * 1e6 / serving.expected_work()
    trace = serving.poisson_trace(n_requests, rate=rate, seed=seed)
    server = serving.Server(precision="mixed", slots=4, chunk=50, autotune=autotune)
    _, summary = server.run(trace, realtime=True)
    summary.update({"workload": f"{n_requests} Poisson-arrival runs, grids {{64,128,256,384}}^2 x {{32,64,128}}, "
                                "steps log-uniform [100, 2000], mixed precision, dt = 20 s",
                    "arrival_rate_per_s": rate, "target_utilization": utilization,
                    "latency": "arrival to result on the host"})
    return summary


def host_cores():
    """Host threads this process may use (at most 16, the share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(psi_phys, B_phys, budget_s=25.0, workers=1):
    """The numpy oracle (an op-for-op restatement of the reference's CPU step)
    on the same workload over a bounded number of steps.  workers = 1 is the
    reference's own configuration (FFTW single-threaded, src/parallel_mpi.jl:50);
    with more workers the 2-D transforms run on a thread pool (scipy.fft) while
    the column solves and pointwise work stay single-threaded numpy."""
    from oracle import qgybj_oracle as O
    from qgybjplus_amd import synthetic
    g = O.make_grid(SIZE, extent=EXTENT, centered=True)
    N2 = O.stratification_profile(g, synthetic.exponential_N2(1e-4, 800.0))
    m = O.OracleModel(g, O.Options(f=1e-4, feedback=True), N2)
    m.set_psi_physical(psi_phys.astype(np.complex128), "qg")
    m.set_B_physical(B_phys)
    prev = O.set_fft_workers(workers)
    try:
        steps = 0
        t0 = time.perf_counter()
        while True:
            m.step(DT)
            steps += 1
            el = time.perf_counter() - t0
            if el > budget_s or el / steps * (steps + 1) > budget_s * 1.5:
                break
    finally:
        O.set_fft_workers(prev)
    npts = SIZE[0] * SIZE[1] * SIZE[2]
    how = ("single thread, numpy pocketfft standing in for single-threaded FFTW" if workers == 1 else
           f"{workers} threads in the 2-D transforms (scipy.fft pocketfft), column solves and pointwise "
           "work single-threaded numpy")
    return {"value": npts * steps / el / 1e6, "unit": "Mgridpt-steps/s", "cores": workers, "kind": "port",
            "sample": f"{steps} ETD-RK2 step(s) of the 256x256x128 workload by the numpy oracle ({how}), "
                      f"{el:.1f} s"}


ERROR_STEPS = 1000  # the steps the mixed-precision error bar is stated after


def hbm_ceiling(nbytes, nbuf=16, reps=64):
    """The streaming rate this GPU reaches on arrays of one field's size:
    back-to-back device-to-device copies (the runtime's copy kernel) of
    `nbytes` among `nbuf` buffers, each copy reading the buffer written
    nbuf/2 copies earlier (nbuf/2 * nbytes past it, well beyond the 256 MB
    infinity cache) and writing another, as the step's kernels read arrays
    written several launches before.  Through the HIP runtime directly; bytes
    read plus bytes written per second."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]

    def ok(err, what):
        if err != 0:
            raise RuntimeError(f"{what}: hip error {err}")

    bufs = []
    try:
        for _ in range(nbuf):
            p = C.c_void_p()
            ok(hip.hipMalloc(C.byref(p), nbytes), "hipMalloc")
            bufs.append(p)
            ok(hip.hipMemset(p, 0, nbytes), "hipMemset")
        ev = [C.c_void_p(), C.c_void_p()]
        for e in ev:
            ok(hip.hipEventCreate(C.byref(e)), "hipEventCreate")

        def copies(n):
            for i in range(n):
                k = i % nbuf
                ok(hip.hipMemcpyAsync(bufs[(k + nbuf // 2) % nbuf], bufs[k], nbytes, 3, None), "hipMemcpyAsync")

        copies(nbuf)
        ok(hip.hipEventRecord(ev[0], None), "hipEventRecord")
        copies(reps)
        ok(hip.hipEventRecord(ev[1], None), "hipEventRecord")
        ok(hip.hipEventSynchronize(ev[1]), "hipEventSynchronize")
        ms = C.c_float()
        ok(hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1]), "hipEventElapsedTime")
        for e in ev:
            hip.hipEventDestroy(e)
    finally:
        for p in bufs:
            hip.hipFree(p)
    rate = 2.0 * nbytes * reps / (ms.value / 1e3)
    return {"copy_bytes": nbytes, "GBps": rate / 1e9, "frac_of_peak": rate / PEAK_HBM,
            "method": f"{reps} device-to-device copies of {nbytes >> 20} MiB among {nbuf} buffers, each reading "
                      f"the buffer written {nbuf // 2} copies before (runtime copy kernel, HIP events)"}


def per_kernel_table(kern):
    tot = sum(x["ms"] for x in kern.values())
    return {k: {"launches": v["count"], "avg_us": 1e3 * v["ms"] / v["count"],
                "GBps": v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] > 0 else None,
                "frac_of_peak": v["bytes"] / (v["ms"] / 1e3) / PEAK_HBM if v["ms"] > 0 else None,
                "share": v["ms"] / tot, "bytes_per_launch": v["bytes"] / v["count"]}
            for k, v in kern.items()}


def fp64_only(qg, args):
    """The FP64 model alone (its profiler run): autotuned, `steps` timed
    steps, then per-kernel events over min(steps, 200) eager steps."""
    grid, m = build_model(qg, "fp64")
    psi, _ = workload_inputs(grid)
    set_workload(qg, m, psi)
    choice = m.autotune(DT, steps=16, in_step=True) if args.autotune else None
    log("timing the FP64 model")
    t, _ = timed_run(qg, m, args.steps, args.warmup)
    kernel_steps = min(args.steps, 200)
    prof = kernel_run(qg, m, kernel_steps)
    m.finalize()
    npts = SIZE[0] * SIZE[1] * SIZE[2]
    return {"metric": METRIC, "value": npts * args.steps / t / 1e6, "unit": "Mgridpt-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t / args.steps,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "the 256x256x128 workload, FP64 model only", "grid": list(SIZE), "dt": DT,
                       "precision": "fp64"},
            "kernels_fp64": per_kernel_table(prof["kernels"]), "kernel_profile_steps": kernel_steps,
            "autotune": {"fp64": {"choice": choice}} if choice is not None else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="steps of the 48-member batched-ensemble leg (0 skips it)")
    ap.add_argument("--large-steps", type=int, default=500,
                    help="steps of the 512x512x256 leg (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests of the serving leg (0 skips it)")
    ap.add_argument("--run-steps", type=int, default=200,
                    help="iterations of run!(simulation) (every step checked) on the mixed model")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="run every pass at its default tile variant / launch shape")
    ap.add_argument("--modes", choices=("both", "mixed", "fp64"), default="both",
                    help="'mixed' or 'fp64' times only that model (for profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the whole grid is resident on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        evict = torch.ones(2 * n // 4, dtype=torch.float32, device="cuda")  # pushes src out of L2/MALL
        torch.cuda.synchronize()
        dst.copy_(src)  # runtime blit kernel
        evict.fill_(2.0)
        torch.neg(src, out=dst)  # elementwise kernel: reads n bytes, writes n bytes
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n}), flush=True)
        return

    from __graft_entry__ import load_package
    qg = load_package()

    if args.modes == "fp64":
        print(json.dumps(fp64_only(qg, args)), flush=True)
        return

    grid, m32 = build_model(qg, "mixed")
    m64 = build_model(qg, "fp64")[1] if args.modes == "both" else None
    psi, B = workload_inputs(grid)
    tuning = {}
    for name, m in (("mixed", m32), ("fp64", m64)):
        if m is None:
            continue
        set_workload(qg, m, psi)
        if args.autotune:
            t0 = time.perf_counter()
            tuning[name] = {"choice": m.autotune(DT, steps=16, in_step=True), "seconds": time.perf_counter() - t0}
            log(f"autotuned the {name} model in {tuning[name]['seconds']:.2f} s")

    npts = SIZE[0] * SIZE[1] * SIZE[2]
    log("timing the mixed-precision model")
    t32, s32 = timed_run(qg, m32, args.steps, args.warmup)
    v32 = npts * args.steps / t32 / 1e6
    errs = {}
    t64 = v64 = None
    err_steps = args.steps + args.warmup
    if m64 is not None:
        log("timing the FP64 model")
        t64, _ = timed_run(qg, m64, args.steps, args.warmup)
        v64 = npts * args.steps / t64 / 1e6
        # untimed: both models on to the step count the error bar is stated at
        more = max(0, ERROR_STEPS - err_steps)
        if more:
            log(f"{more} untimed steps of both models for the error after {ERROR_STEPS} steps")
            for m in (m32, m64):
                qg.step_(m, qg.ExponentialRungeKutta2(dt=DT), more)
            err_steps += more
        for n in ("q", "B", "psi", "A", "C", "u", "v", "w"):
            a, b = getattr(m32.fields, n), getattr(m64.fields, n)
            errs[n] = float(np.linalg.norm(a - b) / np.linalg.norm(b))

    # per-kernel roofline pass, after the error comparison so both models
    # have taken the same steps there
    kernel_steps = min(args.steps, 200)
    log(f"per-kernel events over {kernel_steps} more mixed-precision steps")
    p32 = kernel_run(qg, m32, kernel_steps)
    p64 = kernel_run(qg, m64, kernel_steps) if m64 is not None else None
    kern = p32["kernels"]
    dom = max(kern, key=lambda k: kern[k]["ms"])
    avg_s = kern[dom]["ms"] / kern[dom]["count"] / 1e3
    bytes_per_launch = kern[dom]["bytes"] / kern[dom]["count"]
    achieved = bytes_per_launch / avg_s
    steps_ms = np.array(s32["steps_ms"]) if s32["steps_ms"] else np.array([t32 / args.steps * 1e3])
    traffic = measured_traffic(dom, tuning.get("mixed", {}).get("choice"))

    per_kernel = per_kernel_table(kern)
    # bytes per step of each mode: FP32 storage can at most halve them, and
    # every kernel is HBM-bound, so their ratio caps the mixed speedup
    bytes32 = sum(v["bytes"] for v in kern.values()) / kernel_steps
    bytes64 = sum(v["bytes"] for v in p64["kernels"].values()) / kernel_steps if p64 else None
    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "Mgridpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage of all fields incl. state; f64 column recurrences and ETD coefficients",
        "data": "synthetic",
        "config": {"workload": "QG-YBJ+ ETD-RK2 step, 256x256x128, 500km x 500km x 4000m, "
                               "N2=1e-4*exp(z/800), seeded k^-3 psi (k_peak=6, u_rms=0.1 m/s), "
                               "surface wave 0.1 m/s over 50 m, feedback on, YBJ+, dt=20 s",
                   "grid": list(SIZE), "dt": DT, "precision": "mixed"},
        "fp64": ({"value": v64, "ms_per_step": 1e3 * t64 / args.steps, "dtype": "f64"}
                 if v64 is not None else None),
        "mixed_speedup_vs_fp64": v32 / v64 if v64 else None,
        "mixed_speedup_cap": ({"value": bytes64 / bytes32, "bytes_per_step_fp64": bytes64,
                               "bytes_per_step_mixed": bytes32,
                               "why": "every kernel is HBM-bound: the speedup is capped by the ratio of "
                                      "algorithmic bytes per step (FP32 storage at most halves them; the "
                                      "FP64 column tables and ETD coefficients stay), below the 3x target"}
                              if bytes64 else None),
        "mixed_rel_l2_vs_fp64": {"after_steps": err_steps, **errs} if errs else None,
        "step_latency_ms": {"p50": float(np.percentile(steps_ms, 50)),
                            "p99": float(np.percentile(steps_ms, 99))},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_s * 1e6,
                     "kernel_profile_steps": kernel_steps},
        "kernels_mixed": per_kernel,
        "kernels_fp64": per_kernel_table(p64["kernels"]) if p64 else None,
        "autotune": tuning or None,
    }
    out["run_loop"] = run_loop_leg(qg, m32, args.run_steps) if args.run_steps > 0 else None
    # the streaming rate at one complex FP32 field's size: what "% of 8 TB/s"
    # can reach on this GPU at this array size
    out["hbm_ceiling"] = hbm_ceiling(npts * 8)
    m32.finalize()
    if m64 is not None:
        m64.finalize()
    if args.large_steps > 0:
        log("512x512x256 leg")
    out["large"] = large_leg(qg, args.large_steps, autotune=args.autotune) if args.large_steps > 0 else None
    if args.ensemble_steps > 0:
        log("batched-ensemble leg")
    out["ensemble"] = (ensemble_leg(qg, args.ensemble_steps, autotune=args.autotune)
                       if args.ensemble_steps > 0 else None)
    if args.serve_requests > 0:
        log(f"serving leg ({args.serve_requests} requests)")
    out["serving"] = (serving_leg(qg, args.serve_requests, v32, autotune=args.autotune)
                      if args.serve_requests > 0 else None)
    if not args.no_cpu:
        log("CPU baseline (numpy oracle, one core)")
        out["cpu_baseline"] = cpu_baseline(psi, B, args.cpu_budget)
        cores = host_cores()
        if cores > 1:
            log(f"CPU baseline (numpy oracle, {cores} threads in the transforms)")
            out["cpu_baseline_all_cores"] = cpu_baseline(psi, B, args.cpu_budget, workers=cores)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
