"""Benchmark: ETD-RK2 steps of the QG-YBJ+ model at 256x256x128 on one MI355X.

Workload (the plain-vs-mixed configuration): 256x256x128 over 500 km x 500 km
x 4000 m, f = 1e-4, N²(z) = 1e-4 exp(z/800 m), seeded k^-3 streamfunction
peaked at wavenumber 6 with RMS velocity 0.1 m/s, surface wave 0.1 m/s over a
50 m mixed layer, wave-mean feedback on, YBJ+, default closure, dt = 20 s.
One "step" is one ETD-RK2 step of the whole grid; inputs are resident in HBM
when the timed region starts.  The headline `value` is the mixed-precision
throughput; the FP64 throughput and the mixed-vs-FP64 relative L2 error per
field after warmup+steps steps are reported beside it.

Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mgridpt-steps/s at 256²×128, FP64 vs mixed at stated err; p50/p99 latency"
PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
SIZE = (256, 256, 128)
EXTENT = (500e3, 500e3, 4000.0)
DT = 20.0


def build_model(qg, precision):
    from qgybjplus_amd import synthetic  # noqa: F401  (registered by load_package)
    grid = qg.RectilinearGrid(size=SIZE, extent=EXTENT, centered=True)
    model = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4),
                          stratification=qg.AnalyticalProfile(synthetic.exponential_N2(1e-4, 800.0)),
                          flow=qg.EvolvingFlow(), feedback=qg.WaveMeanFeedback(),
                          formulation=qg.YBJPlus(), architecture=qg.GPU(precision))
    return grid, model


def timed_run(qg, model, steps, warmup):
    h = model._h
    ts = qg.ExponentialRungeKutta2(dt=DT)
    if warmup:
        qg.step_(model, ts, warmup)
    h.synchronize()
    h.profile_begin(kernels=True, steps=True)
    t0 = time.perf_counter()
    qg.step_(model, ts, steps)
    h.synchronize()
    t1 = time.perf_counter()
    prof = h.profile_end()
    return t1 - t0, prof


def cpu_baseline(psi_phys, B_phys, budget_s=25.0):
    """The numpy oracle (an op-for-op restatement of the reference's CPU step)
    on the same workload, one host core, over a bounded number of steps."""
    from oracle import qgybj_oracle as O
    from qgybjplus_amd import synthetic
    g = O.make_grid(SIZE, extent=EXTENT, centered=True)
    N2 = O.stratification_profile(g, synthetic.exponential_N2(1e-4, 800.0))
    m = O.OracleModel(g, O.Options(f=1e-4, feedback=True), N2)
    m.set_psi_physical(psi_phys.astype(np.complex128), "qg")
    m.set_B_physical(B_phys)
    steps = 0
    t0 = time.perf_counter()
    while True:
        m.step(DT)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or el / steps * (steps + 1) > budget_s * 1.5:
            break
    npts = SIZE[0] * SIZE[1] * SIZE[2]
    return {"value": npts * steps / el / 1e6, "unit": "Mgridpt-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} ETD-RK2 step(s) of the 256x256x128 workload by the numpy oracle "
                      f"(single thread), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the whole grid is resident on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    qg = load_package()
    from qgybjplus_amd import synthetic

    grid, m32 = build_model(qg, "mixed")
    _, m64 = build_model(qg, "fp64")
    psi = synthetic.k3_streamfunction(grid, seed=2026, urms=0.1, kpeak=6.0)
    nz = SIZE[2]
    B = np.empty((nz, SIZE[0], SIZE[1]), dtype=np.complex128)
    for k in range(nz):
        depth = max(0.0, -grid.z[k] - grid.dz / 2)
        B[k] = 0.1 * math.exp(-(depth ** 2) / 50.0 ** 2)
    for m in (m32, m64):
        qg.set_(m, psi=qg.FieldArray(psi), pv_method="qg")
        qg.set_(m, waves=qg.SurfaceWave(0.1, 50.0))

    npts = SIZE[0] * SIZE[1] * SIZE[2]
    t32, p32 = timed_run(qg, m32, args.steps, args.warmup)
    t64, p64 = timed_run(qg, m64, args.steps, args.warmup)
    v32 = npts * args.steps / t32 / 1e6
    v64 = npts * args.steps / t64 / 1e6

    errs = {}
    for n in ("q", "B", "psi", "A"):
        a, b = getattr(m32.fields, n), getattr(m64.fields, n)
        errs[n] = float(np.linalg.norm(a - b) / np.linalg.norm(b))

    kern = p32["kernels"]
    dom = max(kern, key=lambda k: kern[k]["ms"])
    avg_s = kern[dom]["ms"] / kern[dom]["count"] / 1e3
    bytes_per_launch = kern[dom]["bytes"] / kern[dom]["count"]
    achieved = bytes_per_launch / avg_s
    steps_ms = np.array(p32["steps_ms"]) if p32["steps_ms"] else np.array([t32 / args.steps * 1e3])
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_256x256x128_mixed.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(dom)

    per_kernel = {k: {"launches": v["count"], "avg_us": 1e3 * v["ms"] / v["count"],
                      "GBps": v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] > 0 else None,
                      "share": v["ms"] / sum(x["ms"] for x in kern.values())}
                  for k, v in kern.items()}
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
        "dtype": "mixed: f32 transforms/products, f64 state/column solves/ETD coefficients",
        "data": "synthetic",
        "config": {"workload": "QG-YBJ+ ETD-RK2 step, 256x256x128, 500km x 500km x 4000m, "
                               "N2=1e-4*exp(z/800), seeded k^-3 psi (k_peak=6, u_rms=0.1 m/s), "
                               "surface wave 0.1 m/s over 50 m, feedback on, YBJ+, dt=20 s",
                   "grid": list(SIZE), "dt": DT, "precision": "mixed"},
        "fp64": {"value": v64, "ms_per_step": 1e3 * t64 / args.steps, "dtype": "f64"},
        "mixed_speedup_vs_fp64": v32 / v64,
        "mixed_rel_l2_vs_fp64": {"after_steps": args.steps + args.warmup, **errs},
        "step_latency_ms": {"p50": float(np.percentile(steps_ms, 50)),
                            "p99": float(np.percentile(steps_ms, 99))},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_s * 1e6},
        "kernels_mixed": per_kernel,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(psi, B, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    m32.finalize()
    m64.finalize()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
