"""GPU parity of the full ETD-RK2 step (step!/_advance_etdrk2!) against the
CPU oracle and the reference's exact-damping known-answer tests."""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O
from tests.helpers import Case, gpu_model, k3_streamfunction, oracle_model, rel, set_both

pytestmark = pytest.mark.gpu

EXP_N2 = lambda z: 1e-4 * math.exp(z / 800.0)  # noqa: E731
FIELDS = ("q", "B", "psi", "A")


def test_etd_exact_damping_kat(qg):
    """test/test_model_etdrk2.jl:5-41: passive, linear, fixed flow, no
    dispersion; B[2,2,1] damped by exp(-0.1*0.3*kh2) per step, q unchanged."""
    grid = qg.RectilinearGrid(size=(8, 8, 4), extent=(2 * math.pi, 2 * math.pi, 1.0))
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1.0), stratification=qg.ConstantStratification(1.0),
                      closure=qg.HorizontalHyperdiffusivity(flow=qg.FlowHyperdiffusivity(0),
                                                            wave=qg.WaveHyperdiffusivity(0.3, order=2)),
                      flow=qg.FixedFlow(), formulation=qg.PassiveWave(), linear=qg.LinearDynamics(),
                      no_dispersion=qg.NoDispersion())
    q = np.zeros((4, 8, 8), complex)
    B = np.zeros((4, 8, 8), complex)
    q[1, 1, 0] = 0.5 - 0.1j
    B[1, 1, 0] = 1.2 - 0.7j
    m.fields.q = q
    m.fields.B = B
    ts = qg.ExponentialRungeKutta2(dt=0.1)
    qg.step_(m, ts)
    damping = math.exp(-0.1 * 0.3 * grid.kh2[1, 0])
    assert m.fields.q[1, 1, 0] == q[1, 1, 0]
    assert abs(m.fields.B[1, 1, 0] - damping * B[1, 1, 0]) <= 1e-14 * abs(damping * B[1, 1, 0])
    qg.step_(m, ts)
    assert abs(m.fields.B[1, 1, 0] - damping ** 2 * B[1, 1, 0]) <= 2e-14 * abs(B[1, 1, 0])
    m.finalize()


def test_three_step_multimode_damping_kat(qg):
    """test/test_mpi_stepping_regression.jl:20-25,93-112: 8^3, dt = 1e-3, three
    steps, every wave mode damped by exp(-3 dt 0.3 kh2) to 1e-13."""
    grid = qg.RectilinearGrid(size=(8, 8, 8), extent=(2 * math.pi, 2 * math.pi, 1.0))
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1.0), stratification=qg.ConstantStratification(1.0),
                      closure=qg.HorizontalHyperdiffusivity(flow=qg.FlowHyperdiffusivity(0),
                                                            wave=qg.WaveHyperdiffusivity(0.3, order=2)),
                      flow=qg.FixedFlow(), formulation=qg.PassiveWave(), linear=qg.LinearDynamics(),
                      no_dispersion=qg.NoDispersion())
    q = np.zeros((8, 8, 8), complex)
    B = np.zeros((8, 8, 8), complex)
    psi = np.zeros((8, 8, 8), complex)
    q[1, 1, 0] = 0.5 - 0.1j
    q[1, 7, 0] = np.conj(q[1, 1, 0])
    psi[1, 1, 0] = -0.2 + 0.05j
    psi[1, 7, 0] = np.conj(psi[1, 1, 0])
    B[2, 1, 1] = 1.2 - 0.7j
    B[2, 2, 0] = 0.6 + 0.2j
    B[5, 2, 0] = -0.4 + 0.9j
    m.fields.q, m.fields.B, m.fields.psi = q, B, psi
    qg.step_(m, qg.ExponentialRungeKutta2(dt=1e-3), 3)
    expected = np.exp(-3 * 1e-3 * 0.3 * grid.kh2)[None] * B
    assert np.max(np.abs(m.fields.B - expected)) < 1e-13
    m.finalize()


def test_linear_dispersion_closed_form(qg):
    """Constant a, no flow, one vertical cosine mode: B is an eigenvector of
    the YBJ+ operator, rhs_B = i*sigma*B, so each ETD-RK2 step multiplies B by
    1 + i*sigma*h - (sigma*h)^2/2 (h = dt, no damping)."""
    nz = 8
    grid = qg.RectilinearGrid(size=(16, 16, nz), extent=(1e5, 1e5, 1000.0))
    f, N2 = 1e-4, 1e-5
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(f), stratification=qg.ConstantStratification(N2),
                      closure=qg.HorizontalHyperdiffusivity(flow=qg.FlowHyperdiffusivity(0),
                                                            wave=qg.WaveHyperdiffusivity(0)),
                      flow=qg.FixedFlow(), feedback="none")
    a = f * f / (math.sqrt(N2) ** 2)
    dz = grid.dz_solver
    i, j, mode = 2, 1, 3
    kh2 = grid.kh2[i, j]
    zc = np.arange(nz)
    B = np.zeros((nz, 16, 16), complex)
    B[:, i, j] = np.cos(mode * math.pi * (zc + 0.5) / nz) * (0.3 + 0.1j)
    m.fields.B = B
    mu = -(4 * a * math.sin(mode * math.pi / (2 * nz)) ** 2 + kh2 * dz * dz / 4) / dz ** 2
    sigma = (f / 2) * kh2 / mu
    h = 50.0
    qg.step_(m, qg.ExponentialRungeKutta2(dt=h))
    factor = 1 + 1j * sigma * h - (sigma * h) ** 2 / 2
    got = m.fields.B[:, i, j]
    assert np.max(np.abs(got - factor * B[:, i, j])) < 1e-12 * np.max(np.abs(B))
    m.finalize()


def _c1_case():
    return Case(size=(64, 64, 32), extent=(500e3, 500e3, 4000.0), centered=True, f=1e-4, N2=1e-5,
                feedback=True)


def test_config1_plumbing_100_steps(qg):
    """64x64x32, f-plane, constant N2 = 1e-5, psi = 1e3 sin cos, surface wave
    0.1 m/s over 30 m, feedback on, YBJ+, dt = 20 s, 100 ETD-RK2 steps:
    psi, q, B, A within 1e-11 relative L2 of the oracle."""
    c = _c1_case()
    m, o = gpu_model(qg, c), oracle_model(c)
    psi_fn = lambda x, y, z: 1e3 * np.sin(2 * math.pi * x / 500e3) * np.cos(2 * math.pi * y / 500e3)  # noqa: E731
    qg.set_(m, psi=psi_fn, waves=qg.SurfaceWave(0.1, 30.0))
    o.set_psi_function(lambda x, y, z: 1e3 * math.sin(2 * math.pi * x / 500e3) * math.cos(2 * math.pi * y / 500e3))
    o.set_surface_waves(0.1, 30.0)
    for name in FIELDS:
        assert rel(getattr(m.fields, name), getattr(o.fields, name)) < 1e-13, name
    sim = qg.Simulation(m, dt=20.0, stop_iteration=100)
    qg.run_(sim)
    for _ in range(100):
        o.step(20.0)
    errs = {n: rel(getattr(m.fields, n), getattr(o.fields, n)) for n in FIELDS + ("u", "v")}
    print("config-1 relative L2 after 100 steps:", errs, flush=True)
    for n in FIELDS:
        assert errs[n] <= 1e-11, (n, errs[n])
    m.finalize()


@pytest.mark.parametrize("variant", ["feedback", "nofeedback", "fixed", "passive", "linear",
                                     "nodisp", "inviscid", "nu_z", "flow_closure", "nonsquare",
                                     "omega_pivot"])
def test_step_variants_match_oracle(qg, variant):
    base = dict(size=(32, 32, 12), extent=(500e3, 500e3, 4000.0), N2_func=EXP_N2, feedback=True)
    if variant == "nofeedback":
        base["feedback"] = False
    elif variant == "fixed":
        base["fixed_flow"] = True
    elif variant == "passive":
        base["formulation"] = "passive"
    elif variant == "linear":
        base["linear"] = True
    elif variant == "nodisp":
        base["no_dispersion"] = True
    elif variant == "inviscid":
        base["inviscid"] = True
    elif variant == "nu_z":
        base["nu_z"] = 1e-3
    elif variant == "flow_closure":
        base["closure"] = O.Closure(flow_coefficients=(1e3,), flow_orders=(2,), kind="flow")
    elif variant == "nonsquare":
        base["size"] = (64, 32, 8)
        base["extent"] = (600e3, 300e3, 3000.0)
    elif variant == "omega_pivot":
        # N² falling 2.3x per level towards the surface: dgtsv interchanges rows
        base["N2_func"] = lambda z: 1e-5 * math.exp(-z / 400.0)
    c = Case(**base)
    print(f"case step variant={variant}", flush=True)
    m, o = gpu_model(qg, c), oracle_model(c)
    psi = k3_streamfunction(o.grid, seed=11, urms=0.1, kpeak=3.0)
    X, Y = np.meshgrid(o.grid.x, o.grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(o.grid.nz):
        B[k] = 0.1 * math.exp(o.grid.z[k] / 200.0) * np.exp(1j * 2 * math.pi * X / o.grid.Lx)
    set_both(m, o, psi, B, qg)
    for n in FIELDS:
        assert rel(getattr(m.fields, n), getattr(o.fields, n)) < 1e-12, ("init", n)
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    qg.step_(m, ts, 5)
    for _ in range(5):
        o.step(600.0)
    for n in FIELDS + ("C", "u", "v", "w"):
        e = rel(getattr(m.fields, n), getattr(o.fields, n))
        print(f"  {variant} {n}: {e:.2e}", flush=True)
        tol = 1e-9 if n == "w" else 1e-11
        assert e <= tol, (variant, n, e)
    m.finalize()


def test_mixed_precision_error_is_stated(qg):
    c = Case(size=(64, 64, 16), extent=(500e3, 500e3, 4000.0), N2_func=EXP_N2, feedback=True)
    cm = Case(**{**c.__dict__, "precision": "mixed"})
    m64, m32 = gpu_model(qg, c), gpu_model(qg, cm)
    o = oracle_model(c)
    psi = k3_streamfunction(o.grid, seed=5, urms=0.1)
    X, Y = np.meshgrid(o.grid.x, o.grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(o.grid.nz):
        B[k] = 0.1 * math.exp(o.grid.z[k] / 50.0)
    for m in (m64, m32):
        qg.set_(m, psi=qg.FieldArray(psi), pv_method="qg")
        qg.set_(m, waves=qg.FieldArray(B))
    qg.step_(m64, qg.ExponentialRungeKutta2(dt=20.0), 50)
    qg.step_(m32, qg.ExponentialRungeKutta2(dt=20.0), 50)
    for n in FIELDS:
        e = rel(getattr(m32.fields, n), getattr(m64.fields, n))
        print(f"mixed vs fp64 after 50 steps {n}: {e:.2e}", flush=True)
        assert e < 1e-4, (n, e)
    m64.finalize()
    m32.finalize()


@pytest.mark.parametrize("size", [(384, 384, 8), (96, 96, 12), (160, 80, 6), (1024, 40, 3)])
def test_other_plans_step_matches_oracle(qg, size):
    """The whole step on grids whose transforms use the radix-3 plans (384 =
    8*6*8, the serving workload's largest grid; 96 = 4*6*4) and the radix-5
    ones (160 = 4*10*4, 80 = 4*5*4, 40 = 2*10*2) and 1024 = 16*4*16, 2 steps
    against the oracle."""
    c = Case(size=size, extent=(500e3, 500e3, 4000.0), N2_func=EXP_N2, feedback=True)
    print(f"case step {size}", flush=True)
    m, o = gpu_model(qg, c), oracle_model(c)
    psi = k3_streamfunction(o.grid, seed=21, urms=0.1, kpeak=6.0)
    X, Y = np.meshgrid(o.grid.x, o.grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(o.grid.nz):
        B[k] = 0.1 * math.exp(o.grid.z[k] / 300.0) * np.exp(1j * 2 * math.pi * (3 * X / o.grid.Lx + Y / o.grid.Ly))
    set_both(m, o, psi, B, qg)
    qg.step_(m, qg.ExponentialRungeKutta2(dt=60.0), 2)
    for _ in range(2):
        o.step(60.0)
    for n in FIELDS + ("C", "u", "v", "w"):
        e = rel(getattr(m.fields, n), getattr(o.fields, n))
        print(f"  {size} {n}: {e:.2e}", flush=True)
        assert e <= (1e-9 if n == "w" else 1e-11), (n, e)
    m.finalize()


@pytest.mark.parametrize("size", [(4, 4, 2), (8, 6, 4), (2, 2, 2), (3, 1, 2), (64, 48, 24), (5, 9, 3),
                                  (25, 27, 4), (36, 50, 6), (45, 54, 3), (75, 72, 4), (100, 90, 3),
                                  (125, 150, 2), (200, 216, 2), (15, 30, 3), (60, 45, 3), (120, 135, 2),
                                  (81, 90, 2), (162, 243, 2)])
def test_more_sizes_step_matches_oracle(qg, size):
    """Grids of the reference's own tests ((4, 4, 2), (8, 6, 4), (2, 2, 2),
    (3, 1, 2), (64, 48, 24): test/test_core_components.jl and others), odd
    lengths and the other radix-2-3-5 plans: 3 steps against the oracle."""
    c = Case(size=size, extent=(500e3, 500e3, 4000.0), N2_func=EXP_N2, feedback=True)
    print(f"case step {size}", flush=True)
    m, o = gpu_model(qg, c), oracle_model(c)
    nx, ny, nz = size
    rng = np.random.default_rng(sum(size))
    # random smooth-ish fields with content at every wavenumber (inside and
    # outside the 2/3 disk)
    psi = 1e3 * rng.normal(size=(nz, nx, ny))
    X, Y = np.meshgrid(o.grid.x, o.grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(nz):
        B[k] = 0.1 * math.exp(o.grid.z[k] / 300.0) * np.exp(1j * 2 * math.pi * (X / o.grid.Lx + Y / o.grid.Ly))
        B[k] += 0.01 * (rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)))
    set_both(m, o, psi, B, qg)
    qg.step_(m, qg.ExponentialRungeKutta2(dt=60.0), 3)
    for _ in range(3):
        o.step(60.0)
    for n in FIELDS + ("C", "u", "v", "w"):
        a, b = getattr(m.fields, n), getattr(o.fields, n)
        if n == "w":  # where w is round-off (a single wavenumber shell), against the flow's scale
            d = float(np.linalg.norm(np.ravel(a - b)))
            e = d / max(float(np.linalg.norm(np.ravel(b))), 1e-6 * float(np.linalg.norm(o.fields.u)), 1e-300)
        else:
            e = rel(a, b)
        print(f"  {size} {n}: {e:.2e}", flush=True)
        assert e <= (1e-9 if n == "w" else 1e-11), (n, e)
    m.finalize()


def test_asymmetric_wavenumbers_step_matches_oracle(qg):
    """Wavenumbers need not be sign-symmetric: here one kx is off by a few
    ulp, so kx^2 differs from its mirror's.  With a hyperdiffusion strong
    enough that every ETD coefficient differs from its undamped value (the
    coefficients are computed per point from its kh2), 3 steps match the
    oracle given the same kx."""
    c = Case(size=(32, 32, 12), extent=(500e3, 500e3, 4000.0), N2_func=EXP_N2, feedback=True,
             closure=O.Closure(flow_coefficients=(5e3,), flow_orders=(2,), wave_coefficients=(5e3,),
                               wave_orders=(2,)))
    kx = O.make_grid(c.size, extent=c.extent, centered=True).kx.copy()
    kx[-1] = kx[-1] * (1 + 2.0 ** -40)
    assert kx[-1] ** 2 != kx[1] ** 2
    print("case step with asymmetric kx", flush=True)
    m, o = gpu_model(qg, c, kx=kx), oracle_model(c, kx=kx)
    psi = k3_streamfunction(o.grid, seed=5, urms=0.1, kpeak=4.0)
    X, Y = np.meshgrid(o.grid.x, o.grid.y, indexing="ij")
    B = np.empty(psi.shape, complex)
    for k in range(o.grid.nz):
        B[k] = 0.1 * math.exp(o.grid.z[k] / 200.0) * np.exp(1j * 2 * math.pi * (5 * X / o.grid.Lx + 2 * Y / o.grid.Ly))
    set_both(m, o, psi, B, qg)
    ts = qg.ExponentialRungeKutta2(dt=600.0)
    qg.step_(m, ts, 3)
    for _ in range(3):
        o.step(600.0)
    for n in FIELDS + ("C", "u", "v", "w"):
        e = rel(getattr(m.fields, n), getattr(o.fields, n))
        print(f"  asymmetric kx {n}: {e:.2e}", flush=True)
        assert e <= (1e-9 if n == "w" else 1e-11), (n, e)
    m.finalize()
