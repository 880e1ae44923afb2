"""Pins of the CPU oracle (no GPU): the reference's own known-answer tests,
library cross-checks and closed forms chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails one of them."""

import math

import numpy as np
import pytest
from scipy.linalg import solve_banded

from oracle import qgybj_oracle as O


def test_grid_kat():
    """test/test_core_components.jl:4-52."""
    g = O.make_grid((8, 6, 4), x=(-4.0, 4.0), y=(-3.0, 3.0), z=(-2.0, 0.0))
    assert (g.Lx, g.Ly, g.Lz) == (8.0, 6.0, 2.0)
    assert (g.x0, g.y0) == (-4.0, -3.0)
    assert np.array_equal(g.x, np.arange(-4.0, 4.0, 1.0))
    assert np.array_equal(g.y, np.arange(-3.0, 3.0, 1.0))
    assert np.array_equal(g.z, [-1.75, -1.25, -0.75, -0.25])
    assert np.array_equal(g.z_faces, np.arange(-2.0, 0.01, 0.5))
    assert (g.dx, g.dy, g.dz) == (1.0, 1.0, 0.5)
    assert g.kx[0] == 0 and g.ky[0] == 0
    assert g.kh2.shape == (8, 6)
    assert g.kh2[1, 1] == g.kx[1] ** 2 + g.ky[1] ** 2
    c = O.make_grid((4, 4, 2), extent=(8, 4, 2), centered=True)
    assert (c.x0, c.y0) == (-4.0, -2.0)
    for bad in (dict(size=(0, 8, 4), extent=(1, 1, 1)), dict(size=(8, 8, 4), extent=(1, -1, 1)),
                dict(size=(8, 8, 4), extent=(1, 1, 1), centered=True, x=(-0.5, 0.5))):
        with pytest.raises(ValueError):
            O.make_grid(**bad)


def test_wavenumber_ordering():
    """FFTW order: 0, 1, ..., n/2-1, -n/2, ..., -1 times 2*pi/L (src/core/grid.jl:122-125)."""
    k = O.wavenumbers(8, 2 * math.pi)
    assert np.array_equal(k, [0, 1, 2, 3, -4, -3, -2, -1])
    k = O.wavenumbers(5, 2 * math.pi)
    assert np.array_equal(k, [0, 1, 2, -2, -1])


def test_int_factor_kat():
    """test/test_core_components.jl:100-110."""
    kx, ky, dt = 3.0, 4.0, 2.0
    closure = O.Closure(flow_coefficients=(0.0,), flow_orders=(4,),
                        wave_coefficients=(1e5,), wave_orders=(4,))
    assert O.int_factor(kx, ky, dt, closure) == 0
    assert O.int_factor(kx, ky, dt, closure, waves=True) == dt * 1e5 * (kx ** 2 + ky ** 2) ** 2
    flow = O.Closure(flow_coefficients=(0.25,), flow_orders=(4,), kind="flow")
    assert O.int_factor(kx, ky, dt, flow) == dt * 0.25 * (kx ** 2 + ky ** 2) ** 2
    wave = O.Closure(wave_coefficients=(1e5,), wave_orders=(4,), kind="wave")
    assert O.int_factor(kx, ky, dt, wave) == 0
    assert O.int_factor(kx, ky, dt, closure, waves=True, inviscid=True) == 0


def test_dealias_mask_radius():
    L = O.dealias_mask(64, 64)
    assert L[21, 0] and not L[22, 0] and L[0, 64 - 21] and not L[0, 64 - 22]
    assert L[14, 14] and not L[15, 15]  # 14²+14² <= 21² < 15²+15²
    assert np.array_equal(L, L[(-np.arange(64)) % 64][:, (-np.arange(64)) % 64])  # symmetric


def test_etd_coefficients_closed_forms_and_series_continuity():
    h = 2.5
    for x in (0.3, 1.0, 4.0):
        E, a1, a2 = O.etd_coefficients(x, h)
        assert E == pytest.approx(math.exp(-x), rel=1e-15)
        assert a1 == pytest.approx(h * (1 - math.exp(-x)) / x, rel=1e-14)
        assert a2 == pytest.approx(h * (math.exp(-x) - 1 + x) / x ** 2, rel=1e-12)
    lo = O.etd_coefficients(1e-6 * (1 - 1e-9), h)
    hi = O.etd_coefficients(1e-6 * (1 + 1e-9), h)
    assert abs(lo[1] - hi[1]) < 1e-13 * h and abs(lo[2] - hi[2]) < 1e-9 * h
    E, a1, a2 = O.etd_coefficients(0.0, h)
    assert (E, a1, a2) == (1.0, h, h / 2)


def test_thomas_against_solve_banded():
    rng = np.random.default_rng(0)
    n, m = 17, 9
    dl = rng.uniform(0.1, 1, (n, m))
    du = rng.uniform(0.1, 1, (n, m))
    d = -(dl + du + rng.uniform(0.01, 1, (n, m)))
    b = rng.normal(size=(n, m))
    x = O.thomas(dl, d, du, b)
    for c in range(m):
        ab = np.zeros((3, n))
        ab[0, 1:] = du[:-1, c]
        ab[1] = d[:, c]
        ab[2, :-1] = dl[1:, c]
        assert np.allclose(x[:, c], solve_banded((1, 1), ab, b[:, c]), rtol=1e-12, atol=0)


def test_thomas_singular_pivot_raises():
    d = np.array([1e-9, -2.0, -2.0])
    with pytest.raises(O.ThomasSingularError):
        O.thomas(np.ones(3), d, np.ones(3), np.ones(3))


@pytest.mark.parametrize("pivot", [False, True])
def test_gtsv_against_dense_solve(pivot):
    rng = np.random.default_rng(1 + pivot)
    n, m = 9, 6
    dl = rng.normal(size=(n - 1, m))
    du = rng.normal(size=(n - 1, m))
    d = rng.normal(size=(n, m)) * (0.1 if pivot else 10.0)
    b = rng.normal(size=(n, m))
    x = O.gtsv(dl, d, du, b)
    for c in range(m):
        A = np.diag(d[:, c]) + np.diag(dl[:, c], -1) + np.diag(du[:, c], 1)
        assert np.allclose(x[:, c], np.linalg.solve(A, b[:, c]), rtol=1e-9, atol=1e-11)


def _exp_grid(nz=10, nx=16):
    g = O.make_grid((nx, nx, nz), extent=(5e5, 5e5, 4000.0))
    N2 = O.stratification_profile(g, lambda z: 1e-4 * math.exp(z / 800))
    return g, O.a_ell_from_N2(N2, 1e-4), N2


def test_stratification_sampled_on_bottom_faces():
    """N² is evaluated at z[k] - dz/2 (src/stratification.jl:308-318)."""
    g, a, N2 = _exp_grid()
    for k in range(g.nz):
        assert N2[k] == 1e-4 * math.exp((g.z[k] - g.dz / 2) / 800)
        assert a[k] == (1e-4 * 1e-4) / N2[k]
    assert O.constant_N2(1e-5) == math.sqrt(1e-5) ** 2


def test_q_to_psi_inverts_compute_q_from_psi():
    g, a, _ = _exp_grid()
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(g.nz, g.nx, g.ny)) + 1j * rng.normal(size=(g.nz, g.nx, g.ny))
    psi[:, 0, 0] = 0
    q = O.compute_q_from_psi(psi, g, a)
    back = O.invert_q_to_psi(q, g, a)
    assert O.relative_l2(back, psi) < 1e-10
    # barotropic column: q = -kh2 psi exactly for a z-independent psi
    bt = np.broadcast_to(psi[0], psi.shape).copy()
    assert np.allclose(O.compute_q_from_psi(bt, g, a), -g.kh2 * bt, rtol=1e-12, atol=0)


def test_B_to_A_inverts_Lplus_and_fixes_the_gauge():
    g, a, _ = _exp_grid()
    rng = np.random.default_rng(4)
    A0 = rng.normal(size=(g.nz, g.nx, g.ny)) + 1j * rng.normal(size=(g.nz, g.nx, g.ny))
    A0[:, 0, 0] -= A0[:, 0, 0].mean()
    B = O.apply_Lplus(A0, g, a)
    A, C = O.invert_B_to_A(B, g, a)
    assert O.relative_l2(A, A0) < 1e-9
    assert np.allclose(C[:-1], (A[1:] - A[:-1]) / g.dz_solver) and np.all(C[-1] == 0)
    assert abs(A[:, 0, 0].sum()) < 1e-12 * np.abs(A[:, 0, 0]).max()


def test_qw_closed_form_kat():
    """test/test_wave_feedback.jl:34-66."""
    f = 2e-4
    g = O.make_grid((16, 16, 4), x=(-math.pi, math.pi), y=(-math.pi, math.pi), z=(-1.0, 0.0))
    L = O.dealias_mask(16, 16)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    B = O.fft_forward(np.broadcast_to(np.sin(X) + 1j * np.sin(Y), (4, 16, 16)))
    qw = O.compute_qw_complex(B, g, f, L)
    phys = O.fft_backward(qw)
    exp = (-np.cos(X) * np.cos(Y) + 0.5 * (np.cos(2 * X) + np.cos(2 * Y))) / f
    assert np.allclose(phys.real, np.broadcast_to(exp, phys.shape), rtol=5e-13, atol=5e-11)
    assert np.max(np.abs(phys.imag)) < 5e-11
    # q_w ∝ 1/f (test/test_mpi_stepping_regression.jl:122-151)
    assert np.max(np.abs(O.compute_qw_complex(B, g, 2 * f, L) - 0.5 * qw)) < 1e-13 * np.abs(qw).max()
    # replace_q: q = 0 -> q_eff = -q_w (test/test_wave_feedback.jl:91-96)
    q_eff = np.where(L[None], 0 - qw, 0)
    assert np.max(np.abs(q_eff + qw)) == 0


def test_advection_single_mode_closed_form():
    """u = U cos(ly) (from psi = -U sin(ly)/l), chi = cos(kx):
    div(u chi) = -k U cos(ly) sin(kx) for the B (complex) and q (real) paths."""
    n = 16
    g = O.make_grid((n, n, 2), extent=(2 * math.pi, 2 * math.pi, 1.0))
    L = O.dealias_mask(n, n)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    U, k, l = 0.7, 2, 3
    u = np.broadcast_to(U * np.cos(l * Y), (2, n, n))
    v = np.zeros_like(u)
    chi = O.fft_forward(np.broadcast_to(np.cos(k * X), (2, n, n)))
    for use_real in (True, False):
        nk = O.convol_advect(u, v, chi, g, L, use_real)
        phys = O.fft_backward(nk).real
        assert np.allclose(phys, np.broadcast_to(-k * U * np.cos(l * Y) * np.sin(k * X), phys.shape),
                           atol=1e-13)


def test_refraction_single_mode_closed_form():
    """psi = sin(kx) -> zeta = -k² sin(kx); B = e^{ily}: r = zeta*B."""
    n = 16
    g = O.make_grid((n, n, 2), extent=(2 * math.pi, 2 * math.pi, 1.0))
    L = O.dealias_mask(n, n)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    psi = O.fft_forward(np.broadcast_to(np.sin(2 * X), (2, n, n)))
    B = O.fft_forward(np.broadcast_to(np.exp(1j * 3 * Y), (2, n, n)))
    r = O.fft_backward(O.refraction_B(B, psi, g, L))
    assert np.allclose(r, np.broadcast_to(-4 * np.sin(2 * X) * np.exp(3j * Y), r.shape), atol=1e-13)


def test_omega_rhs_closed_form():
    """psi = Z1(z) sin x + Z2(z) cos(2y)/2.  With a the forward difference of
    Z over dz (zero at the top level) and b its two-level average (Z itself
    at the top): psi_z = a1 sin x + a2 cos(2y)/2 and lap psi_bar =
    -b1 sin x - 2 b2 cos 2y, so 2 J(psi_z, lap psi_bar) =
    2 (4 a1 b2 - a2 b1) cos x sin 2y.  A flipped or swapped Jacobian term, or
    the average replaced by the level's own psi, fails this (the two
    horizontal scales differ, so J(psi_z, lap psi_z) does not vanish)."""
    n, nz = 16, 6
    g = O.make_grid((n, n, nz), extent=(2 * math.pi, 2 * math.pi, 1.0))
    L = O.dealias_mask(n, n)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    Z1 = np.cos(math.pi * (g.z + 1.0))
    Z2 = np.sin(2 * math.pi * g.z) + 0.3
    psi = O.fft_forward(Z1[:, None, None] * np.sin(X)[None] + 0.5 * Z2[:, None, None] * np.cos(2 * Y)[None])
    got = O.fft_backward(O.omega_eqn_rhs(psi, g, L)).real

    def diff_avg(Z):
        a, b = np.zeros(nz), Z.copy()
        a[: nz - 1] = (Z[1:] - Z[:-1]) / g.dz_solver
        b[: nz - 1] = 0.5 * (Z[1:] + Z[:-1])
        return a, b

    (a1, b1), (a2, b2) = diff_avg(Z1), diff_avg(Z2)
    want = 2 * (4 * a1 * b2 - a2 * b1)[:, None, None] * (np.cos(X) * np.sin(2 * Y))[None]
    assert np.abs(a1 * a2).max() > 0.1  # the average matters
    assert np.allclose(got, want, atol=1e-11)


@pytest.mark.parametrize("m", [1, 2, 3])
def test_solve_omega_discrete_sine_modes(m):
    """Constant N2, rhs = R sin(theta k) cos(x) on the interior levels with
    theta = m pi/(nz-1): the second difference maps sin(theta k) to
    (2 cos theta - 2) sin(theta k) with w = 0 at both ends, so
    w_k = (f/N2) R sin(theta k) cos(x) / ((f^2/N2)(2 cos theta - 2)/dz^2 - kh2),
    kh2 = 1 for cos(x)."""
    n, nz = 8, 10
    g = O.make_grid((n, n, nz), extent=(2 * math.pi, 2 * math.pi, 900.0))
    X, _ = np.meshgrid(g.x, g.y, indexing="ij")
    f, N2c, R = 1e-4, 2.5e-5, 3e-11
    theta = m * math.pi / (nz - 1)
    k = np.arange(nz)
    prof = R * np.sin(theta * k)
    prof[0] = prof[-1] = 0.0
    rhsk = O.fft_forward(prof[:, None, None] * np.cos(X)[None])
    w = O.fft_backward(O.solve_omega(rhsk, g, f, np.full(nz, N2c))).real
    dz = g.dz_solver
    lam = (f * f / N2c) * (2 * math.cos(theta) - 2) / (dz * dz) - 1.0
    want = (f / N2c) * prof[:, None, None] * np.cos(X)[None] / lam
    want[0] = want[-1] = 0.0
    assert np.allclose(w, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


def test_omega_rhs_single_mode_and_w_dense_solve():
    """psi = cos(z-mode) * sin(kx): J(psi_z, lap psi_bar) vanishes for one
    horizontal mode, so the omega RHS is zero; a two-mode psi gives a
    nonzero RHS and w solves the interior system (dense brute force)."""
    n, nz = 8, 8
    g = O.make_grid((n, n, nz), extent=(2 * math.pi, 2 * math.pi, 1.0))
    L = O.dealias_mask(n, n)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    zs = np.cos(math.pi * (g.z + 1.0))[:, None, None]
    one = O.fft_forward(zs * np.sin(X)[None])
    assert np.max(np.abs(O.omega_eqn_rhs(one, g, L))) < 1e-12
    two = O.fft_forward(zs * np.sin(X)[None] + np.cos(Y)[None] * 0.5)
    rhs = O.omega_eqn_rhs(two, g, L)
    assert np.max(np.abs(rhs)) > 1e-3
    f = 1.0
    N2 = np.linspace(1.0, 3.0, nz)
    w = O.compute_vertical_velocity(two, g, f, N2, L)
    wk = O.fft_forward(w)
    dz = g.dz_solver
    for (i, j) in [(1, 0), (0, 1), (1, 1), (2, 1)]:
        kh2 = g.kh2[i, j]
        m = nz - 2
        A = np.zeros((m, m))
        for iz in range(m):
            c = (f * f / N2[iz + 1]) / (dz * dz)
            A[iz, iz] = -2 * c - kh2
            if iz > 0:
                A[iz, iz - 1] = c
            if iz < m - 1:
                A[iz, iz + 1] = c
        b = np.array([(f / N2[iz + 1]) * rhs[iz + 1, i, j] for iz in range(m)])
        ref = np.linalg.solve(A, b)
        # w is real: its spectrum at (i,j) is the Hermitian part of the solve
        sol = wk[1:nz - 1, i, j]
        assert np.allclose(sol, ref, atol=1e-12 * (1 + np.abs(ref).max())) or \
            np.allclose(sol, 0.5 * (ref + np.conj(np.linalg.solve(A, np.array(
                [(f / N2[iz + 1]) * rhs[iz + 1, -i, -j] for iz in range(m)])))), atol=1e-12)


def _damping_model(nz):
    g = O.make_grid((8, 8, nz), extent=(2 * math.pi, 2 * math.pi, 1.0))
    opt = O.Options(f=1.0, fixed_flow=True, feedback=False, formulation="passive", linear=True,
                    no_dispersion=True,
                    closure=O.Closure(flow_coefficients=(0.0,), flow_orders=(4,),
                                      wave_coefficients=(0.3,), wave_orders=(2,)))
    return g, O.OracleModel(g, opt, np.full(nz, 1.0))


def test_etd_exact_damping_kat():
    """test/test_model_etdrk2.jl:5-41."""
    g, m = _damping_model(4)
    m.fields.q[1, 1, 0] = 0.5 - 0.1j
    m.fields.B[1, 1, 0] = 1.2 - 0.7j
    q0, B0 = m.fields.q[1, 1, 0], m.fields.B[1, 1, 0]
    m.step(0.1)
    damping = math.exp(-0.1 * 0.3 * g.kh2[1, 0])
    assert m.fields.q[1, 1, 0] == q0
    assert m.fields.B[1, 1, 0] == pytest.approx(damping * B0, rel=1e-14)
    m.step(0.1)
    assert m.fields.B[1, 1, 0] == pytest.approx(damping ** 2 * B0, rel=2e-14)


def test_three_step_damping_kat():
    """test/test_mpi_stepping_regression.jl:93-112."""
    g, m = _damping_model(8)
    B = np.zeros((8, 8, 8), complex)
    B[2, 1, 1], B[2, 2, 0], B[5, 2, 0] = 1.2 - 0.7j, 0.6 + 0.2j, -0.4 + 0.9j
    m.fields.B = B.copy()
    for _ in range(3):
        m.step(1e-3)
    assert np.max(np.abs(m.fields.B - np.exp(-3e-3 * 0.3 * g.kh2)[None] * B)) < 1e-13


def test_linear_dispersion_closed_form():
    nz = 8
    g = O.make_grid((16, 16, nz), extent=(1e5, 1e5, 1000.0))
    f, N2 = 1e-4, 1e-5
    opt = O.Options(f=f, fixed_flow=True, feedback=False,
                    closure=O.Closure((0.0,), (4,), (0.0,), (4,)))
    m = O.OracleModel(g, opt, np.full(nz, O.constant_N2(N2)))
    a = m.a[0]
    dz = g.dz_solver
    i, j, mode = 2, 1, 3
    B = np.zeros((nz, 16, 16), complex)
    B[:, i, j] = np.cos(mode * math.pi * (np.arange(nz) + 0.5) / nz) * (0.3 + 0.1j)
    m.fields.B = B.copy()
    mu = -(4 * a * math.sin(mode * math.pi / (2 * nz)) ** 2 + g.kh2[i, j] * dz * dz / 4) / dz ** 2
    s = (f / 2) * g.kh2[i, j] / mu * 50.0
    m.step(50.0)
    assert np.max(np.abs(m.fields.B[:, i, j] - (1 + 1j * s - s * s / 2) * B[:, i, j])) < 1e-15


def test_initialization_kats():
    """test/test_model_initialization.jl:25-52: barotropic q = -kh² psi,
    SurfaceWave top value = amplitude, A idempotent."""
    g = O.make_grid((8, 8, 4), x=(-math.pi, math.pi), y=(-math.pi, math.pi), z=(-1.0, 0.0))
    opt = O.Options(f=1.0, closure=O.Closure((0.0,), (4,), (0.0,), (4,)))
    m = O.OracleModel(g, opt, np.full(4, 1.0))
    m.set_psi_function(lambda x, y, z: math.sin(x) * math.cos(y), "barotropic")
    assert np.allclose(m.fields.q, -g.kh2[None] * m.fields.psi)
    m.set_surface_waves(0.2, 0.1, "gaussian")
    Bp = O.fft_backward(m.fields.B)
    assert np.max(np.abs(Bp[-1] - 0.2)) < 1e-12
    A0 = m.fields.A.copy()
    m.refresh_wave_diagnostics()
    assert np.allclose(m.fields.A, A0)


def test_config1_short_run_is_finite_and_checked():
    g = O.make_grid((32, 32, 16), extent=(500e3, 500e3, 4000.0), centered=True)
    m = O.OracleModel(g, O.Options(f=1e-4, feedback=True), np.full(16, O.constant_N2(1e-5)))
    m.set_psi_function(lambda x, y, z: 1e3 * math.sin(2 * math.pi * x / 500e3) * math.cos(2 * math.pi * y / 500e3))
    m.set_surface_waves(0.1, 30.0)
    for _ in range(3):
        m.step(20.0)
    bad, mx = m.check()
    assert not bad and 0 < mx < 1e10


def test_mix64_is_splitmix64():
    """The reference's _mix64 (src/parallel_mpi.jl:1117-1122) is SplitMix64's
    output function: from state 0 it gives SplitMix64's published sequence."""
    gamma = 0x9E3779B97F4A7C15
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    got = [O.mix64((n * gamma) & ((1 << 64) - 1)) for n in range(4)]
    assert got == expect
    # _hash_u01 maps into (0, 1] with 53-bit resolution
    u = O.hash_u01(O.mix64(7), (np.arange(1, 2001), 3, 5), 0x1)
    assert np.all(u > 0) and np.all(u <= 1) and np.all(u * 2.0 ** 53 == np.round(u * 2.0 ** 53))
    assert abs(np.mean(u) - 0.5) < 0.03


@pytest.mark.parametrize("size", [(16, 16, 3), (24, 12, 2), (32, 32, 1)])
def test_random_psi_structure(size):
    """init_mpi_random_psi! (src/parallel_mpi.jl:1159-1216): Hermitian (its
    inverse transform is real), masked, zero mean, deterministic per seed."""
    nx, ny, nz = size
    p = O.random_psi(nx, ny, nz, 2.0, -3.0, 11)
    phys = np.fft.ifft2(p, axes=(1, 2))
    assert np.max(np.abs(phys.imag)) <= 1e-15 * np.max(np.abs(phys.real))
    assert np.all(p[:, ~O.dealias_mask(nx, ny)] == 0)
    assert np.all(p[:, 0, 0] == 0)
    assert np.array_equal(p, O.random_psi(nx, ny, nz, 2.0, -3.0, 11))
    assert not np.array_equal(p, O.random_psi(nx, ny, nz, 2.0, -3.0, 12))
    # levels draw independently
    if nz > 1:
        assert not np.allclose(p[0], p[1])
    with pytest.raises(ValueError):
        O.random_psi(nx, ny, nz, 1.0, -3.0, -1)


def test_random_psi_energy_law():
    """Mean |psi-hat|^2 over the modes of one shell is 2*amplitude*k^slope
    (E[randn^2] = 1): a dropped factor 2 or a wrong exponent fails."""
    nx = ny = 64
    p = O.random_psi(nx, ny, 16, 3.0, -2.5, 5)
    i = np.fft.fftfreq(nx) * nx
    K = np.sqrt(i[:, None] ** 2 + i[None, :] ** 2)
    for lo, hi in ((1.0, 8.0), (8.0, 14.0), (14.0, 21.0)):
        band = (K >= lo) & (K < hi) & O.dealias_mask(nx, ny)
        vals = np.abs(p[:, band]) ** 2 / (2 * 3.0 * K[band] ** -2.5)
        # about half the modes are independent draws: sd of the mean < 0.04
        assert abs(np.mean(vals) - 1.0) < 0.12, (lo, hi, np.mean(vals))


def _energy_case(seed=3, size=(16, 16, 6)):
    nx, ny, nz = size
    g = O.make_grid(size, extent=(500e3, 500e3, 4000.0), centered=True)
    N2 = O.stratification_profile(g, lambda z: 1e-4 * math.exp(z / 800.0))
    a = O.a_ell_from_N2(N2, 1e-4)
    rng = np.random.default_rng(seed)
    mask = O.dealias_mask(nx, ny)

    def field(scale):
        f = np.fft.fft2(rng.normal(size=(nz, nx, ny)), axes=(1, 2)) * scale
        f = np.where(mask[None], f, 0)
        f[:, 0, 0] = 0  # zero horizontal mean: Parseval without the kh = 0 weight
        return f
    return g, a, field(1e3), field(0.1)


def test_energy_components_parseval():
    """Each energy equals the physical-space mean of its density (Parseval,
    numpy's ifft convention), and the wave KE is reached through the L+
    relation LA = B + kh^2 A / 4 from B and A instead of C."""
    g, a, psi, B = _energy_case()
    A, C = O.invert_B_to_A(B, g, a)
    wke, wpe, wce, mke, mpe = O.energy_components(psi, A, C, g, a)
    nz = psi.shape[0]
    ifft = lambda f: np.fft.ifft2(f, axes=(-2, -1))  # noqa: E731
    kx, ky = g.kx[:, None], g.ky[None, :]
    u, v = ifft(-1j * ky * psi), ifft(1j * kx * psi)
    assert math.isclose(mke, 0.5 * np.mean(np.abs(u) ** 2 + np.abs(v) ** 2), rel_tol=1e-12)
    pz = ifft(np.diff(psi, axis=0) / g.dz)
    assert math.isclose(mpe, 0.5 / nz * sum(a[k] * np.mean(np.abs(pz[k]) ** 2) for k in range(nz - 1)),
                        rel_tol=1e-12)
    cx, cy = ifft(1j * kx * C), ifft(1j * ky * C)
    assert math.isclose(wpe, 0.5 / nz * sum(np.mean(0.5 / a[k] * (np.abs(cx[k]) ** 2 + np.abs(cy[k]) ** 2))
                                            for k in range(nz)), rel_tol=1e-12)
    lapA = ifft(-g.kh2 * A)
    assert math.isclose(wce, 0.5 / nz * sum(np.mean(np.abs(lapA[k]) ** 2) / (8 * a[k] ** 2) for k in range(nz)),
                        rel_tol=1e-12)
    LA = B + g.kh2 * A / 4
    assert math.isclose(wke, 0.5 * np.mean(np.abs(ifft(LA)) ** 2), rel_tol=1e-9)


def test_record_energies_uses_plain_q():
    """_record_energy_diagnostics! inverts q itself (no wave feedback)."""
    g, a, psi, B = _energy_case(5)
    q = O.compute_q_from_psi(psi, g, a)
    e = O.record_energies(q, B, g, a)
    A, C = O.invert_B_to_A(B, g, a)
    e2 = O.energy_components(psi, A, C, g, a)
    assert np.allclose(e, e2, rtol=1e-10, atol=0)
