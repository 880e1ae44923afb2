"""GPU parity of the transforms and the model-level operators, through the C
ABI, against the CPU oracle and the reference's known-answer tests."""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O
from tests.helpers import Case, gpu_model, oracle_model, random_spectral, rel

pytestmark = pytest.mark.gpu

EXP_N2 = lambda z: 1e-4 * math.exp(z / 800.0)  # noqa: E731


@pytest.mark.parametrize("size", [(8, 8, 4), (32, 16, 3), (64, 64, 5), (128, 128, 2),
                                  (256, 256, 2), (384, 384, 1), (512, 512, 1),
                                  # the lengths compiled for the base tile variant only,
                                  # radix-5 plans among them
                                  (10, 12, 3), (20, 24, 2), (40, 48, 2), (80, 96, 2), (160, 192, 1),
                                  (320, 640, 1), (768, 1024, 1),
                                  # single-DFT lengths 1-6 (the reference's own test grids),
                                  # odd lengths and the other 2-3-5 plans
                                  (1, 2, 2), (3, 4, 2), (5, 6, 2), (9, 18, 2), (25, 27, 2), (36, 45, 2),
                                  (50, 54, 2), (72, 75, 2), (90, 100, 2), (108, 125, 2), (144, 150, 1),
                                  (180, 200, 1), (216, 250, 1), (288, 300, 1), (360, 400, 1), (432, 500, 1),
                                  (576, 600, 1), (720, 800, 1), (864, 1000, 1),
                                  # plans with a radix-15 or radix-30 stage
                                  (15, 30, 2), (60, 120, 2), (135, 225, 1), (240, 375, 1),
                                  (480, 540, 1), (750, 900, 1), (960, 15, 1),
                                  # plans with a radix-9 or radix-25 stage
                                  (81, 162, 2), (243, 270, 1), (324, 405, 1), (450, 486, 1),
                                  (625, 648, 1), (675, 729, 1), (810, 972, 1)])
def test_fft_matches_dft_and_round_trips(qg, size):
    """KAT test/test_mpi_extension.jl:62-90 (round trip < 1e-12) and the
    transform contract of src/transforms.jl:19-23 against numpy's DFT."""
    print(f"case fft size={size}", flush=True)
    m = gpu_model(qg, Case(size=size, feedback=False))
    nx, ny, nz = size
    rng = np.random.default_rng(1)
    a = rng.normal(size=(nz, nx, ny)) + 1j * rng.normal(size=(nz, nx, ny))
    fw = m.fft_forward(a)
    assert rel(fw, np.fft.fft2(a, axes=(1, 2))) < 1e-14
    bw = m.fft_backward(a)
    assert rel(bw, np.fft.ifft2(a, axes=(1, 2))) < 1e-14
    rt = m.fft_backward(fw)
    assert np.max(np.abs(rt - a)) < 1e-12
    m.finalize()


@pytest.mark.parametrize("n2", ["const", "exp"])
def test_invert_q_to_psi(qg, n2):
    c = Case(size=(32, 32, 12), N2_func=EXP_N2 if n2 == "exp" else None)
    m, o = gpu_model(qg, c), oracle_model(c)
    q = random_spectral((12, 32, 32), 3, scale=1e-5)
    q[:, 0, 0] = 0.0
    m.fields.q = q
    m.invert_q_to_psi()
    psi = m.fields.psi
    ref = O.invert_q_to_psi(q, o.grid, o.a)
    assert rel(psi, ref) < 1e-14
    # the inversion inverts compute_q_from_psi (src/initialization.jl:597-644)
    back = O.compute_q_from_psi(psi, o.grid, o.a)
    assert rel(back, q) < 1e-9
    m.finalize()


def test_invert_B_to_A_with_gauge(qg):
    c = Case(size=(16, 16, 10), N2_func=EXP_N2)
    m, o = gpu_model(qg, c), oracle_model(c)
    B = random_spectral((10, 16, 16), 4, scale=0.1)
    m.fields.B = B
    m.invert_B_to_A()
    A, C = m.fields.A, m.fields.C
    Ar, Cr = O.invert_B_to_A(B, o.grid, o.a)
    assert rel(A, Ar) < 1e-14
    assert rel(C, Cr) < 1e-14
    # kh = 0 column: vertical mean removed (src/elliptic.jl:930-945)
    assert abs(np.sum(A[:, 0, 0])) < 1e-12 * np.max(np.abs(A[:, 0, 0]))
    # L+ A = B on every kh != 0 column
    LA = O.apply_Lplus(A, o.grid, o.a)
    kh2 = o.grid.kh2
    assert rel(LA[:, kh2 != 0], B[:, kh2 != 0]) < 1e-9
    m.finalize()


def test_q_from_psi_methods(qg):
    c = Case(size=(16, 16, 6), N2_func=EXP_N2)
    m, o = gpu_model(qg, c), oracle_model(c)
    psi = random_spectral((6, 16, 16), 5)
    m.fields.psi = psi
    m._h.compute_q_from_psi(barotropic=False)
    assert rel(m.fields.q, O.compute_q_from_psi(psi, o.grid, o.a)) < 1e-15
    m._h.compute_q_from_psi(barotropic=True)
    assert rel(m.fields.q, O.compute_barotropic_q(psi, o.grid)) < 1e-15
    m.finalize()


def test_qw_closed_form(qg):
    """KAT test/test_wave_feedback.jl:34-66: B = sin x + i sin y on 16x16x4
    over [-pi,pi)^2, f = 2e-4 -> q_w = (-cos x cos y + (cos 2x + cos 2y)/2)/f."""
    f = 2e-4
    grid = qg.RectilinearGrid(size=(16, 16, 4), x=(-math.pi, math.pi), y=(-math.pi, math.pi),
                              z=(-1.0, 0.0))
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(f), stratification=qg.ConstantStratification(1e-4),
                      closure=qg.HorizontalHyperdiffusivity(
                          flow=qg.FlowHyperdiffusivity(0), wave=qg.WaveHyperdiffusivity(0)),
                      flow="fixed", feedback="none", no_dispersion=True)
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    Bp = np.broadcast_to(np.sin(X) + 1j * np.sin(Y), (4, 16, 16))
    qg.set_(m, waves=qg.FieldArray(np.array(Bp)))
    qw = m.compute_qw()
    phys = m.fft_backward(qw)
    expected = (-np.cos(X) * np.cos(Y) + 0.5 * (np.cos(2 * X) + np.cos(2 * Y))) / f
    assert np.allclose(phys.real, np.broadcast_to(expected, phys.shape), rtol=5e-13, atol=5e-11)
    assert np.max(np.abs(phys.imag)) < 5e-11
    m.finalize()


def test_qw_random_and_inverse_f_scaling(qg):
    """q_w against the oracle, and q_w(f=2) = q_w(f=1)/2
    (test/test_mpi_stepping_regression.jl:122-151)."""
    c = Case(size=(64, 32, 4), extent=(2 * math.pi, 2 * math.pi, 1.0), f=1.0)
    m, o = gpu_model(qg, c), oracle_model(c)
    B = random_spectral((4, 64, 32), 7, mask=o.L, scale=0.3)
    m.fields.B = B
    qw = m.compute_qw()
    ref = O.compute_qw_complex(B, o.grid, 1.0, o.L)
    assert rel(qw, ref) < 1e-13
    c2 = Case(size=(64, 32, 4), extent=(2 * math.pi, 2 * math.pi, 1.0), f=2.0)
    m2 = gpu_model(qg, c2)
    m2.fields.B = B
    qw2 = m2.compute_qw()
    assert np.max(np.abs(qw2 - 0.5 * qw)) < 1e-13 * np.max(np.abs(qw))
    m.finalize()
    m2.finalize()


def test_velocities_and_omega_w(qg):
    c = Case(size=(32, 32, 10), N2_func=EXP_N2)
    m, o = gpu_model(qg, c), oracle_model(c)
    psi = random_spectral((10, 32, 32), 8, scale=1e3, hermitian=True)
    m.fields.psi = psi
    m.compute_velocities(compute_w=True)
    u, v, w = O.compute_velocities(psi, o.grid, c.f, o.N2, o.L, True)
    assert rel(m.fields.u, u) < 1e-13
    assert rel(m.fields.v, v) < 1e-13
    assert rel(m.fields.w, w) < 1e-12
    m.compute_velocities(compute_w=False)
    assert np.all(m.fields.w == 0)
    m.finalize()


def test_check_reports_max_psi_and_nonfinite(qg):
    """_check_termination_conditions! (src/simulation.jl:25-44) on the
    spectrum of a real streamfunction (Hermitian, as every psi the model
    holds): max |ifft2(psi)| and the non-finite test."""
    c = Case(size=(16, 16, 4))
    m = gpu_model(qg, c)
    psi = random_spectral((4, 16, 16), 9, scale=10.0, hermitian=True)
    m.fields.psi = psi
    bad, mx = m._h.check()
    assert not bad
    assert abs(mx - np.max(np.abs(np.fft.ifft2(psi, axes=(1, 2))))) < 1e-12 * mx
    B = np.zeros((4, 16, 16), complex)
    B[1, 2, 3] = np.nan
    m.fields.B = B
    bad, _ = m._h.check()
    assert bad
    m.finalize()
