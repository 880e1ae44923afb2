"""CPU oracle: a numpy FP64 restatement of one QG-YBJ+ ETD-RK2 step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP library.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, and there only to compute the expected
answer or to time the CPU baseline.  The product path never calls it.

It restates, operation for operation, the reference's serial ("direct") CPU
path of subhk/QGYBJplus.jl v1.0.27 (Julia).  Every function cites the
reference ``file:line`` it follows.  Arrays use the reference's index order
``[k, i, j]`` = (z, x, y); numpy's C order is only a storage detail here.

Third-party arithmetic.  The reference gets its FFTs from FFTW (FFTW.jl "1",
FFTW 3.3.x, via PencilFFTs "0.14-0.16" on a 1x1 topology) and its omega-
equation solve from LAPACK ``dgtsv`` (OpenBLAS through LinearAlgebra); neither
is vendored under the reference and no Manifest pins a version
(``Project.toml:19-27``).  Here:

* the 2-D DFT is numpy's pocketfft (``numpy.fft.fft2``/``ifft2``): forward
  unnormalised with sign exp(-i k.x), inverse normalised by 1/(nx*ny), the
  contract of ``src/transforms.jl:19-23`` and ``src/parallel_mpi.jl:735-797``;
* ``dgtsv`` is restated from LAPACK's published algorithm (Gaussian
  elimination with partial pivoting for tridiagonal systems) in ``gtsv``.

Pins (what checks each function; see ``tests/test_oracle.py``):

* ``make_grid``/``wavenumbers``: reference KAT ``test/test_core_components.jl:4-52``.
* ``int_factor``: reference KAT ``test/test_core_components.jl:100-110``.
* ``etd_coefficients``: closed forms and series continuity at |x| = 1e-6.
* ``thomas``: ``scipy.linalg.solve_banded`` on random diagonally dominant systems.
* ``gtsv``: dense ``numpy.linalg.solve`` (pivoting and non-pivoting systems).
* ``invert_q_to_psi``: round trip through ``compute_q_from_psi``
  (``src/initialization.jl:597-644``, the forward operator of the inversion).
* ``invert_B_to_A``: the L+ forward stencil and the kh=0 gauge/mean property.
* ``compute_qw_complex``: reference KAT ``test/test_wave_feedback.jl:34-66``
  (closed form) and the 1/f scaling ``test/test_mpi_stepping_regression.jl:122-151``.
* ``convol_advect``/``refraction_B``: single-Fourier-mode closed forms.
* ``advance_etdrk2``: reference KATs ``test/test_model_etdrk2.jl:5-41`` and
  ``test/test_mpi_stepping_regression.jl:93-112`` (exact damping), and the
  closed-form linear dispersion of one vertical cosine mode.
* ``omega_eqn_rhs``: single-mode closed form; ``compute_vertical_velocity``:
  dense brute-force solve of the interior system.
* ``set_surface_waves``/``compute_barotropic_q``: reference KAT
  ``test/test_model_initialization.jl:25-52``.
* ``mix64``/``hash_u01``: the published SplitMix64 output sequence for seed 0
  (the reference's ``_mix64`` is SplitMix64's finalizer with its increment);
  ``random_psi``: Hermitian symmetry (real inverse transform), the 2/3 mask,
  the zero mean, the k^slope energy law of its construction, and seed
  determinism (the reference pins determinism only,
  ``test/test_model_initialization.jl:93-98``).  Its ``log``/``cos``/``pow``
  are the platform's, so exact values can differ from Julia's by an ulp.
* ``energy_components``: Parseval against physical-space means of the
  velocity, shear, |grad C|^2 and |lap A|^2, and the wave KE through the
  L+ relation LA = B + kh^2 A / 4 (reference KAT ``test/test_simulation_lifecycle.jl:108-114``),
  which reaches it from B and A instead of C.

* ``omega_eqn_rhs``: the closed form 2 J(psi_z, lap psi_bar) = a_k cos x sin y
  of psi = Z(z) sin x + cos(y)/2 under the reference's forward difference and
  two-level average (a_k = (Z_{k+1} - Z_k)/dz, zero at the top).
* ``solve_omega``: the discrete sine modes sin(m pi k/(nz-1)) of the interior
  second difference, whose eigenvalues 2 cos(theta) - 2 give w in closed
  form for constant N2, and a dense solve for varying N2.

Signs of the step's assembly (``tests/test_oracle_pins.py``):

* ``OracleModel._rhs``: rhs_B = -div(u B) + i (f/2) kh^2 A - (i/2) zeta B
  term by term on three distinct Fourier modes; rhs_q = -J(psi, q) on a
  two-mode barotropic flow; + nu_z d_zz q on Neumann cosine eigenmodes (end
  rows included).
* ``OracleModel.step`` (q): one step of pure vertical diffusion multiplies
  an eigenmode by 1 + h mu + (h mu)^2 / 2.
* ``OracleModel._diagnose_flow``: q = 0 with the closed-form q_w gives
  psi = L^-1(-q_w) and restores q exactly
  (``test/test_wave_feedback.jl:91-96`` run through the function).
* ``compute_velocities``: u = -psi_y, v = psi_x on psi = sin 2x + sin 3y.
* Each of these pins, and the older ones above, is checked against copies of
  this module with one planted one-token mistake (a flipped sign, a dropped
  term, a wrong end row); every mutant fails at least one of them.

Parity unpinned: no reference test checks the omega ``w`` values themselves
(the closed forms above pin the equation as ``src/operators.jl:365-493``
writes it), and true reference trajectories of the full nonlinear step (the
Julia reference cannot run in this environment).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence, Tuple

import numpy as np

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


class ThomasSingularError(RuntimeError):
    """Raised where ``thomas_solve!`` errors (``src/elliptic.jl:1177-1193``)."""


# ----------------------------------------------------------------------------
# Geometry  (src/core/grid.jl)
# ----------------------------------------------------------------------------

def wavenumbers(n: int, length: float) -> np.ndarray:
    """FFTW-ordered wavenumbers, ``src/core/grid.jl:122-125``."""
    scale = (2.0 * math.pi) / length
    m = np.arange(n)
    idx = np.where(m <= (n + 1) // 2 - 1, m, m - n).astype(np.float64)
    return scale * idx


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    Lx: float
    Ly: float
    Lz: float
    x0: float
    y0: float
    z_bottom: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    z_faces: np.ndarray
    dx: float
    dy: float
    dz: float
    kx: np.ndarray
    ky: np.ndarray

    @property
    def kh2(self) -> np.ndarray:
        """``kx_i^2 + ky_j^2`` as ``src/core/grid.jl:196-198``; shape (nx, ny)."""
        return self.kx[:, None] ** 2 + self.ky[None, :] ** 2

    @property
    def dz_solver(self) -> float:
        """``G.z[2]-G.z[1]`` used by every vertical operator (e.g. ``src/elliptic.jl:231``)."""
        return float(self.z[1] - self.z[0]) if self.nz > 1 else 1.0


def make_grid(size: Tuple[int, int, int], extent=None, x=None, y=None, z=None,
              centered: bool = False) -> Grid:
    """``RectilinearGrid`` constructor, ``src/core/grid.jl:127-204``."""
    nx, ny, nz = (int(v) for v in size)
    if min(nx, ny, nz) <= 0:
        raise ValueError("all grid dimensions must be positive")
    if centered and (x is not None or y is not None):
        raise ValueError("centered=true cannot be combined with explicit x or y ranges")
    if extent is None:
        if x is None or y is None or z is None:
            raise ValueError("provide extent=(Lx, Ly, Lz) or explicit x, y, z bounds")
        Lx, Ly, Lz = float(x[1]) - float(x[0]), float(y[1]) - float(y[0]), float(z[1]) - float(z[0])
    else:
        Lx, Ly, Lz = (float(v) for v in extent)
        if not all(math.isfinite(v) and v > 0 for v in (Lx, Ly, Lz)):
            raise ValueError("all grid extents must be finite and positive")
    if min(Lx, Ly, Lz) <= 0:
        raise ValueError("upper bound must exceed lower bound")
    x0 = float(x[0]) if x is not None else (-Lx / 2 if centered else 0.0)
    y0 = float(y[0]) if y is not None else (-Ly / 2 if centered else 0.0)
    z_bottom = float(z[0]) if z is not None else -Lz
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    xs = np.array([x0 + i * dx for i in range(nx)])
    ys = np.array([y0 + j * dy for j in range(ny)])
    zf = np.array([z_bottom + k * dz for k in range(nz + 1)])
    zc = np.array([(zf[k] + zf[k + 1]) / 2 for k in range(nz)])
    return Grid(nx, ny, nz, Lx, Ly, Lz, x0, y0, z_bottom, xs, ys, zc, zf, dx, dy, dz,
                wavenumbers(nx, Lx), wavenumbers(ny, Ly))


# ----------------------------------------------------------------------------
# Physics helpers  (src/physics.jl, src/stratification.jl)
# ----------------------------------------------------------------------------

def dealias_mask(nx: int, ny: int) -> np.ndarray:
    """Radial 2/3 mask, ``src/physics.jl:28-43``; shape (nx, ny), True = keep."""
    kmax = min(nx, ny) // 3
    i = np.arange(nx)
    j = np.arange(ny)
    ix = np.where(i <= nx // 2, i, i - nx)
    jy = np.where(j <= ny // 2, j, j - ny)
    return (ix[:, None] ** 2 + jy[None, :] ** 2) <= kmax ** 2


def constant_N2(N2: float) -> float:
    """``ConstantStratification`` evaluated through ``ConstantN(sqrt(N²))``:
    ``src/core/runtime.jl:29-30`` + ``src/stratification.jl:201-203``."""
    n0 = math.sqrt(float(N2))
    return n0 * n0


def stratification_profile(grid: Grid, N2_of_z: Callable[[float], float]) -> np.ndarray:
    """N² at the bottom face of each cell, ``src/stratification.jl:308-318``
    (``AnalyticalProfile`` returning N², clamped at 0 per ``:263-270``)."""
    out = np.empty(grid.nz)
    for k in range(grid.nz):
        zu = grid.z[k] - grid.dz / 2
        out[k] = max(float(N2_of_z(zu)), 0.0)
    return out


def a_ell_from_N2(N2: np.ndarray, f: float) -> np.ndarray:
    """``a = f²/max(N², √eps)``, ``src/physics.jl:2-25``."""
    N2 = np.asarray(N2, dtype=np.float64)
    if N2.size == 0:
        raise ValueError("the stratification profile cannot be empty")
    if not np.all(np.isfinite(N2)) or not np.all(N2 > 0):
        raise ValueError("N² must be finite and positive")
    return (f * f) / np.maximum(N2, SQRT_EPS)


# ----------------------------------------------------------------------------
# Closure and ETD coefficients  (src/nonlinear.jl, src/timestep.jl)
# ----------------------------------------------------------------------------

@dataclass
class Closure:
    """``HorizontalHyperdiffusivity`` (flow + wave terms), ``src/core/components.jl:105-208``.

    ``kind`` is "horizontal", "flow" (FlowHyperdiffusivity only) or "wave".
    """
    flow_coefficients: Tuple[float, ...] = (0.01, 10.0)
    flow_orders: Tuple[int, ...] = (4, 12)
    wave_coefficients: Tuple[float, ...] = (0.0, 10.0)
    wave_orders: Tuple[int, ...] = (4, 12)
    kind: str = "horizontal"


def _component_int_factor(kh2, dt, coefficients, orders):
    """``src/nonlinear.jl:91-97``: ``Δt·Σ νᵢ·kh²^(orderᵢ÷2)``."""
    total = 0.0
    for c, o in zip(coefficients, orders):
        total = total + c * np.power(kh2, o // 2)
    return dt * total


def int_factor(kx, ky, dt, closure: Closure, waves=False, inviscid=False):
    """``int_factor``, ``src/nonlinear.jl:1160-1183``."""
    kx = np.asarray(kx, dtype=np.float64)
    ky = np.asarray(ky, dtype=np.float64)
    kh2 = kx ** 2 + ky ** 2
    zero = np.zeros_like(kh2)
    if inviscid:
        return zero
    if closure.kind == "horizontal":
        if waves:
            return _component_int_factor(kh2, dt, closure.wave_coefficients, closure.wave_orders)
        return _component_int_factor(kh2, dt, closure.flow_coefficients, closure.flow_orders)
    if closure.kind == "flow":
        return zero if waves else _component_int_factor(kh2, dt, closure.flow_coefficients,
                                                        closure.flow_orders)
    if closure.kind == "wave":
        return _component_int_factor(kh2, dt, closure.wave_coefficients,
                                     closure.wave_orders) if waves else zero
    raise ValueError(closure.kind)


def etd_coefficients(x, h):
    """``E = e^{-x}``, ``hφ1``, ``hφ2`` with the series branch for |x| < 1e-6,
    ``src/timestep.jl:177-190``."""
    x = np.asarray(x, dtype=np.float64)
    E = np.exp(-x)
    x2 = x * x
    s1 = h * (1 - x / 2 + x2 / 6 - x2 * x / 24 + x2 * x2 / 120)
    s2 = h * (1 / 2 - x / 6 + x2 / 24 - x2 * x / 120 + x2 * x2 / 720)
    with np.errstate(divide="ignore", invalid="ignore"):
        em1 = np.expm1(-x)
        d1 = h * (-em1) / x
        d2 = h * (em1 + x) / (x * x)
    small = np.abs(x) < 1e-6
    return E, np.where(small, s1, d1), np.where(small, s2, d2)


# ----------------------------------------------------------------------------
# Transforms  (src/transforms.jl:19-23, src/parallel_mpi.jl:735-797)
# ----------------------------------------------------------------------------

# Threads of the 2-D transforms: 1 (numpy.fft, the reference's single-threaded
# FFTW configuration, src/parallel_mpi.jl:50) unless set_fft_workers raises it
# (scipy.fft's pocketfft with a thread pool over the z-planes).
_FFT_WORKERS = 1


def set_fft_workers(n: int) -> int:
    """Set the transform thread count; returns the previous one."""
    global _FFT_WORKERS
    prev, _FFT_WORKERS = _FFT_WORKERS, max(1, int(n))
    return prev


def fft_forward(a: np.ndarray) -> np.ndarray:
    """Per-z-plane 2-D C2C forward DFT, unnormalised."""
    a = np.asarray(a, dtype=np.complex128)
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.fft2(a, axes=(1, 2), workers=_FFT_WORKERS)
    return np.fft.fft2(a, axes=(1, 2))


def fft_backward(a: np.ndarray) -> np.ndarray:
    """Per-z-plane 2-D inverse DFT, normalised by 1/(nx*ny)."""
    a = np.asarray(a, dtype=np.complex128)
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.ifft2(a, axes=(1, 2), workers=_FFT_WORKERS)
    return np.fft.ifft2(a, axes=(1, 2))


# ----------------------------------------------------------------------------
# Tridiagonal solvers  (src/elliptic.jl:1168-1202; LAPACK dgtsv)
# ----------------------------------------------------------------------------

def thomas(dl, d, du, b):
    """Column-batched ``thomas_solve!`` (``src/elliptic.jl:1168-1202``).

    ``dl, d, du, b`` have shape (n, ...); every trailing index is an
    independent system solved with the reference's exact operation order.
    """
    n = d.shape[0]
    c = np.array(du, dtype=np.float64, copy=True)
    dd = np.array(d, dtype=np.float64, copy=True)
    x = np.array(b, dtype=np.float64, copy=True)
    if np.any(np.abs(dd[0]) < SQRT_EPS):
        raise ThomasSingularError("Singular matrix in Thomas solver: d[1] ≈ 0")
    c[0] = c[0] / dd[0]
    x[0] = x[0] / dd[0]
    for i in range(1, n):
        denom = dd[i] - dl[i] * c[i - 1]
        if np.any(np.abs(denom) < SQRT_EPS):
            raise ThomasSingularError(f"Singular matrix in Thomas solver at i={i + 1}")
        c[i] = c[i] / denom
        x[i] = (x[i] - dl[i] * x[i - 1]) / denom
    for i in range(n - 2, -1, -1):
        x[i] = x[i] - c[i] * x[i + 1]
    return x


def gtsv(dl, d, du, b):
    """Column-batched LAPACK ``dgtsv`` (one right-hand side): Gaussian
    elimination with partial pivoting, restated from LAPACK's published
    algorithm.  ``dl``/``du`` have length n-1, ``d``/``b`` length n along
    axis 0.  Called at ``src/operators.jl:448,459``."""
    dl = np.array(dl, dtype=np.float64, copy=True)
    d = np.array(d, dtype=np.float64, copy=True)
    du = np.array(du, dtype=np.float64, copy=True)
    b = np.array(b, dtype=np.float64, copy=True)
    n = d.shape[0]
    du2 = np.zeros_like(d)  # second superdiagonal fill-in (LAPACK reuses dl)
    for i in range(n - 1):
        di, di1 = np.copy(d[i]), np.copy(d[i + 1])
        dli, dui = np.copy(dl[i]), np.copy(du[i])
        bi, bi1 = np.copy(b[i]), np.copy(b[i + 1])
        nopiv = np.abs(di) >= np.abs(dli)
        with np.errstate(divide="ignore", invalid="ignore"):
            f_np = dli / di
            f_pv = di / dli
        # no interchange
        d1_np = di1 - f_np * dui
        b1_np = bi1 - f_np * bi
        # interchange rows i and i+1
        d0_pv = dli
        d1_pv = dui - f_pv * di1
        if i < n - 2:
            fill_pv = np.copy(du[i + 1])
            du1_pv = -f_pv * fill_pv
        else:
            fill_pv = np.zeros_like(di)
            du1_pv = None
        du0_pv = di1
        b0_pv = bi1
        b1_pv = bi - f_pv * bi1
        if np.any(nopiv & (di == 0)):
            raise RuntimeError("gtsv: exactly singular")
        d[i] = np.where(nopiv, di, d0_pv)
        d[i + 1] = np.where(nopiv, d1_np, d1_pv)
        du2[i] = np.where(nopiv, 0.0, fill_pv)
        if i < n - 2:
            du[i + 1] = np.where(nopiv, du[i + 1], du1_pv)
        du[i] = np.where(nopiv, dui, du0_pv)
        b[i] = np.where(nopiv, bi, b0_pv)
        b[i + 1] = np.where(nopiv, b1_np, b1_pv)
    if np.any(d[n - 1] == 0):
        raise RuntimeError("gtsv: exactly singular")
    b[n - 1] = b[n - 1] / d[n - 1]
    if n > 1:
        b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2]
    for i in range(n - 3, -1, -1):
        b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / d[i]
    return b


# ----------------------------------------------------------------------------
# Elliptic inversions  (src/elliptic.jl)
# ----------------------------------------------------------------------------

def _column_coeffs(a, kh2, dz2, kfac):
    """(dl, d, du) for the Neumann operator with ``-kfac·kh²Δz²`` on the
    diagonal; shapes (nz, nx, ny)."""
    nz = a.shape[0]
    shape = (nz,) + kh2.shape
    dl = np.zeros(shape)
    d = np.zeros(shape)
    du = np.zeros(shape)
    t = kfac(kh2 * dz2)
    d[0] = -(a[0] + t)
    du[0] = a[0]
    for k in range(1, nz - 1):
        dl[k] = a[k - 1]
        d[k] = -(a[k] + a[k - 1] + t)
        du[k] = a[k]
    dl[nz - 1] = a[nz - 2]
    d[nz - 1] = -(a[nz - 2] + t)
    return dl, d, du


def invert_q_to_psi(q: np.ndarray, grid: Grid, a: np.ndarray) -> np.ndarray:
    """``_invert_q_to_psi_direct!``, ``src/elliptic.jl:211-314``."""
    nz = grid.nz
    kh2 = grid.kh2
    psi = np.zeros_like(q, dtype=np.complex128)
    nonzero = kh2 != 0
    if nz == 1:
        with np.errstate(divide="ignore", invalid="ignore"):
            psi[0] = np.where(nonzero, -q[0] / kh2, 0.0)
        return psi
    dz = grid.dz_solver
    dz2 = dz ** 2
    kh2_s = np.where(nonzero, kh2, 1.0)  # kh=0 column handled below (ψ=0)
    dl, d, du = _column_coeffs(a, kh2_s, dz2, lambda t: t)
    sr = thomas(dl, d, du, dz2 * q.real)
    si = thomas(dl, d, du, dz2 * q.imag)
    psi = sr + 1j * si
    psi[:, ~nonzero] = 0.0
    return psi


def invert_B_to_A(B: np.ndarray, grid: Grid, a: np.ndarray):
    """``_invert_B_to_A_direct!``, ``src/elliptic.jl:854-990``; returns (A, C)."""
    nz = grid.nz
    kh2 = grid.kh2
    nonzero = kh2 != 0
    A = np.zeros_like(B, dtype=np.complex128)
    C = np.zeros_like(B, dtype=np.complex128)
    if nz == 1:
        with np.errstate(divide="ignore", invalid="ignore"):
            A[0] = np.where(nonzero, -4 * B[0] / kh2, 0.0)
        return A, C
    dz = grid.dz_solver
    dz2 = dz ** 2
    kh2_s = np.where(nonzero, kh2, 1.0)
    dl, d, du = _column_coeffs(a, kh2_s, dz2, lambda t: t / 4)
    sr = thomas(dl, d, du, dz2 * B.real)
    si = thomas(dl, d, du, dz2 * B.imag)
    A = sr + 1j * si
    # kh = 0 columns: gauge A[1]=0, then remove the vertical mean (:892-945)
    for (i, j) in zip(*np.nonzero(~nonzero)):
        dl0 = np.zeros(nz)
        d0 = np.zeros(nz)
        du0 = np.zeros(nz)
        d0[0] = -a[0]
        du0[0] = a[0]
        for k in range(1, nz - 1):
            dl0[k] = a[k - 1]
            d0[k] = -(a[k] + a[k - 1])
            du0[k] = a[k]
        dl0[nz - 1] = a[nz - 2]
        d0[nz - 1] = -a[nz - 2]
        rr = dz2 * B[:, i, j].real
        ri = dz2 * B[:, i, j].imag
        d0[0] = 1.0
        du0[0] = 0.0
        rr[0] = 0.0
        ri[0] = 0.0
        xr = thomas(dl0, d0, du0, rr)
        xi = thomas(dl0, d0, du0, ri)
        mean = 0.0 + 0.0j
        for k in range(nz):
            mean = mean + (xr[k] + 1j * xi[k])
        mean = mean / nz
        A[:, i, j] = (xr + 1j * xi) - mean
    C[: nz - 1] = (A[1:] - A[: nz - 1]) / dz
    C[nz - 1] = 0.0
    return A, C


def compute_q_from_psi(psi: np.ndarray, grid: Grid, a: np.ndarray) -> np.ndarray:
    """``compute_q_from_psi!``, ``src/initialization.jl:597-644``."""
    nz = grid.nz
    kh2 = grid.kh2
    q = np.zeros_like(psi, dtype=np.complex128)
    if nz == 1:
        q[0] = -kh2 * psi[0]
        return q
    dz = grid.dz_solver
    dz2 = dz ** 2
    for k in range(1, nz - 1):
        vert = a[k] * psi[k + 1] - (a[k] + a[k - 1]) * psi[k] + a[k - 1] * psi[k - 1]
        q[k] = -kh2 * psi[k] + vert / dz2
    vert = a[0] * (psi[1] - psi[0])
    q[0] = -kh2 * psi[0] + vert / dz2
    vert = a[nz - 2] * (psi[nz - 2] - psi[nz - 1])
    q[nz - 1] = -kh2 * psi[nz - 1] + vert / dz2
    return q


def compute_barotropic_q(psi: np.ndarray, grid: Grid) -> np.ndarray:
    """``compute_barotropic_q_from_psi!``, ``src/initialization.jl:653-668``."""
    return -grid.kh2 * psi


def apply_Lplus(A: np.ndarray, grid: Grid, a: np.ndarray) -> np.ndarray:
    """Forward YBJ+ operator L+A whose inverse is ``invert_B_to_A``
    (the stencil of ``src/elliptic.jl:957-969`` applied to A, divided by Δz²)."""
    nz = grid.nz
    kh2 = grid.kh2
    dz2 = grid.dz_solver ** 2
    out = np.zeros_like(A)
    for k in range(nz):
        acc = -(kh2 * dz2) / 4 * A[k]
        if k > 0:
            acc = acc + a[k - 1] * (A[k - 1] - A[k])
        if k < nz - 1:
            acc = acc + a[k] * (A[k + 1] - A[k])
        out[k] = acc / dz2
    return out


# ----------------------------------------------------------------------------
# Nonlinear terms  (src/nonlinear.jl)
# ----------------------------------------------------------------------------

def prefilter(a: np.ndarray, L: np.ndarray) -> np.ndarray:
    """``_prefilter_spectral!``, ``src/nonlinear.jl:64-78``."""
    return np.where(L[None], a, 0.0 + 0.0j)


def compute_qw_complex(B: np.ndarray, grid: Grid, f: float, L: np.ndarray) -> np.ndarray:
    """``compute_qw_complex!``, ``src/nonlinear.jl:859-944``."""
    if not (math.isfinite(f) and f != 0):
        raise ValueError("wave feedback requires a finite, nonzero Coriolis frequency")
    inv_f = 1.0 / float(f)
    kx = grid.kx[None, :, None]
    ky = grid.ky[None, None, :]
    Bf = prefilter(B, L)
    Bx = (1j * kx) * Bf
    By = (1j * ky) * Bf
    Br = fft_backward(Bf)
    Bxr = fft_backward(Bx)
    Byr = fft_backward(By)
    J = np.conj(Bxr) * Byr - np.conj(Byr) * Bxr
    qwr = (0.5j * J).real.astype(np.complex128)
    mag2 = (np.conj(Br) * Br).real.astype(np.complex128)
    temp = fft_forward(mag2)
    qw = fft_forward(qwr)
    kh2 = grid.kh2[None]
    return np.where(L[None], inv_f * (qw - 0.25 * kh2 * temp), 0.0 + 0.0j)


def convol_advect(u, v, chi, grid: Grid, L, use_real: bool):
    """``_convol_advect!``, ``src/nonlinear.jl:414-470`` (q: use_real, B: complex)."""
    chir = fft_backward(prefilter(chi, L))
    val = chir.real if use_real else chir
    uk = fft_forward((u * val).astype(np.complex128))
    vk = fft_forward((v * val).astype(np.complex128))
    kx = grid.kx[None, :, None]
    ky = grid.ky[None, None, :]
    return np.where(L[None], (1j * kx) * uk + (1j * ky) * vk, 0.0 + 0.0j)


def refraction_B(B, psi, grid: Grid, L):
    """``refraction_waqg_B!``, ``src/nonlinear.jl:621-676``."""
    kh2 = grid.kh2[None]
    zeta = np.where(L[None], -kh2 * psi, 0.0 + 0.0j)
    zr = fft_backward(zeta)
    Br = fft_backward(prefilter(B, L))
    rB = fft_forward(zr.real * Br)
    return np.where(L[None], rB, 0.0 + 0.0j)


def dissipation_q_nv(q, nu_z, grid: Grid):
    """``_dissipation_q_nv_direct!``, ``src/nonlinear.jl:1010-1044``."""
    nz = grid.nz
    if nz <= 1:
        return np.zeros_like(q)
    dz = grid.z[1] - grid.z[0]
    inv = 1 / (dz * dz)
    out = np.empty_like(q)
    out[0] = nu_z * (q[1] - q[0]) * inv
    out[nz - 1] = nu_z * (q[nz - 2] - q[nz - 1]) * inv
    for k in range(1, nz - 1):
        out[k] = nu_z * (q[k + 1] - 2 * q[k] + q[k - 1]) * inv
    return out


# ----------------------------------------------------------------------------
# Velocities and the omega equation  (src/operators.jl, src/diagnostics.jl)
# ----------------------------------------------------------------------------

def omega_eqn_rhs(psi, grid: Grid, L):
    """``_omega_eqn_rhs_direct!``, ``src/diagnostics.jl:149-229``."""
    nz = grid.nz
    dz = grid.dz_solver
    kx = grid.kx[None, :, None]
    ky = grid.ky[None, None, :]
    kh2 = grid.kh2[None]
    psiz = np.zeros_like(psi)
    psiz[: nz - 1] = (psi[1:] - psi[: nz - 1]) / dz
    psiavg = psi.copy()
    psiavg[: nz - 1] = 0.5 * (psi[1:] + psi[: nz - 1])
    bx = (1j * kx) * psiz
    by = (1j * ky) * psiz
    xx = (((-1j) * kx) * kh2) * psiavg
    xy = (((-1j) * ky) * kh2) * psiavg
    bxr = fft_backward(bx)
    byr = fft_backward(by)
    xxr = fft_backward(xx)
    xyr = fft_backward(xy)
    rr = 2.0 * (bxr.real * xyr.real - byr.real * xxr.real)
    rhs = fft_forward(rr)
    return np.where(L[None], rhs, 0.0 + 0.0j)


def compute_vertical_velocity(psi, grid: Grid, f: float, N2: np.ndarray, L):
    """``_compute_vertical_velocity_direct!``, ``src/operators.jl:365-493``."""
    return fft_backward(solve_omega(omega_eqn_rhs(psi, grid, L), grid, f, N2)).real


def solve_omega(rhsk, grid: Grid, f: float, N2: np.ndarray):
    """The interior solve of ``_compute_vertical_velocity_direct!``
    (``src/operators.jl:365-493``): per horizontal mode with kh2 > 0,
    (f^2/N2_k)(w_{k+1} - 2 w_k + w_{k-1})/dz^2 - kh2 w_k = (f/N2_k) rhs_k on
    the levels 1 .. nz-2, w = 0 on the first and last level; returns w-hat."""
    nz = grid.nz
    wk = np.zeros_like(rhsk)
    kh2 = grid.kh2
    pos = kh2 > 0
    n = nz - 2
    if nz > 2 and np.any(pos):
        dz = grid.dz_solver
        f2 = f ** 2
        cz = np.array([(f2 / N2[k]) / (dz * dz) for k in range(1, nz - 1)])
        ii, jj = np.nonzero(pos)
        m = ii.size
        d = np.empty((n, m))
        for iz in range(n):
            d[iz] = -2 * cz[iz] - kh2[ii, jj]
        dl = np.empty((max(n - 1, 0), m))
        du = np.empty((max(n - 1, 0), m))
        for iz in range(n - 1):
            dl[iz] = cz[iz + 1]
            du[iz] = cz[iz]
        rhs = np.empty((n, m), dtype=np.complex128)
        for iz in range(n):
            k = iz + 1
            rhs[iz] = (f / N2[k]) * rhsk[k, ii, jj]
        sr = gtsv(dl, d, du, rhs.real)
        si = gtsv(dl, d, du, rhs.imag)
        wk[1: nz - 1, ii, jj] = sr + 1j * si
    return wk


def compute_velocities(psi, grid: Grid, f: float, N2: np.ndarray, L, compute_w: bool):
    """``compute_velocities!``, ``src/operators.jl:195-274`` (QG omega w)."""
    kx = grid.kx[None, :, None]
    ky = grid.ky[None, None, :]
    uk = -(1j * ky) * psi
    vk = (1j * kx) * psi
    u = fft_backward(uk).real
    v = fft_backward(vk).real
    if compute_w:
        w = compute_vertical_velocity(psi, grid, f, N2, L)
    else:
        w = np.zeros_like(u)
    return u, v, w


# ----------------------------------------------------------------------------
# The model and the step  (src/timestep.jl, src/core/runtime.jl)
# ----------------------------------------------------------------------------

@dataclass
class Options:
    """The typed flags ``ETDModelOptions`` reads, ``src/timestep.jl:22-57``."""
    f: float
    fixed_flow: bool = False
    feedback: bool = False          # WaveMeanFeedback (vs NoFeedback / NoWaveFeedback)
    formulation: str = "ybj_plus"   # "ybj_plus" | "passive"  ("ybj" is out of scope)
    inviscid: bool = False
    linear: bool = False
    no_dispersion: bool = False
    nu_z: float = 0.0
    closure: Closure = field(default_factory=Closure)

    @property
    def passive(self):
        return self.formulation == "passive"

    @property
    def feedback_enabled(self):
        return (not self.fixed_flow) and self.feedback


@dataclass
class Fields:
    """``ModelFields`` (``src/core/fields.jl:11-24``), indexed [k, i, j]."""
    q: np.ndarray
    B: np.ndarray
    psi: np.ndarray
    A: np.ndarray
    C: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray

    @classmethod
    def zeros(cls, grid: Grid):
        s = (grid.nz, grid.nx, grid.ny)
        c = lambda: np.zeros(s, dtype=np.complex128)
        r = lambda: np.zeros(s, dtype=np.float64)
        return cls(c(), c(), c(), c(), c(), r(), r(), r())

    def copy(self):
        return Fields(*(getattr(self, n).copy() for n in
                        ("q", "B", "psi", "A", "C", "u", "v", "w")))


class OracleModel:
    """``QGYBJModel`` + ``step!`` on the CPU (``src/core/model.jl:79-121``,
    ``src/core/runtime.jl:34-70,239-263``)."""

    def __init__(self, grid: Grid, options: Options, N2: np.ndarray):
        if options.formulation not in ("ybj_plus", "passive"):
            raise ValueError("only the YBJ+ and passive formulations are restated")
        self.grid = grid
        self.opt = options
        self.N2 = np.asarray(N2, dtype=np.float64)
        self.a = a_ell_from_N2(self.N2, options.f)
        self.L = dealias_mask(grid.nx, grid.ny)
        self.fields = Fields.zeros(grid)

    # -- set! (src/simulation.jl:230-329,461-503) ---------------------------
    def set_psi_physical(self, psi_phys: np.ndarray, pv_method: str = "qg"):
        """``set_mean_flow!`` from physical values, then the PV and u,v refresh."""
        f = self.fields
        f.psi = fft_forward(psi_phys.astype(np.complex128))
        self._refresh_flow(pv_method)

    def set_psi_function(self, fn, pv_method: str = "qg"):
        g = self.grid
        phys = np.empty((g.nz, g.nx, g.ny), dtype=np.complex128)
        for k in range(g.nz):
            for j in range(g.ny):
                for i in range(g.nx):
                    phys[k, i, j] = fn(g.x[i], g.y[j], g.z[k])
        self.set_psi_physical(phys, pv_method)

    def _refresh_flow(self, pv_method):
        f = self.fields
        if pv_method in ("qg", "balanced"):
            f.q = compute_q_from_psi(f.psi, self.grid, self.a)
        elif pv_method in ("barotropic", "asselin"):
            f.q = compute_barotropic_q(f.psi, self.grid)
        elif pv_method != "none":
            raise ValueError(pv_method)
        f.u, f.v, f.w = compute_velocities(f.psi, self.grid, self.opt.f, self.N2, self.L, False)

    def set_surface_waves(self, amplitude: float, scale: float, profile: str = "gaussian"):
        """``set_surface_waves!``, ``src/simulation.jl:292-329``."""
        g = self.grid
        phys = np.empty((g.nz, g.nx, g.ny), dtype=np.complex128)
        for k in range(g.nz):
            depth = max(0.0, -g.z[k] - g.dz / 2)
            prof = math.exp(-(depth ** 2) / scale ** 2) if profile == "gaussian" \
                else math.exp(-depth / scale)
            phys[k] = complex(amplitude * prof)
        self.set_B_physical(phys)

    def set_B_physical(self, B_phys: np.ndarray):
        self.fields.B = fft_forward(B_phys.astype(np.complex128))
        self.refresh_wave_diagnostics()

    def refresh_wave_diagnostics(self):
        """``_refresh_wave_diagnostics!`` for YBJ+ (``src/simulation.jl:442-458``)."""
        self.fields.A, self.fields.C = invert_B_to_A(self.fields.B, self.grid, self.a)

    # -- ETD-RK2 (src/timestep.jl:192-415) -----------------------------------
    def _diagnose_flow(self, S: Fields, compute_w: bool):
        o = self.opt
        if not o.fixed_flow:
            q_base = None
            if o.feedback_enabled:
                q_base = S.q.copy()
                qw = compute_qw_complex(S.B, self.grid, o.f, self.L)
                S.q = np.where(self.L[None], q_base - qw, 0.0 + 0.0j)
            S.psi = invert_q_to_psi(S.q, self.grid, self.a)
            if q_base is not None:
                S.q = q_base
        S.u, S.v, S.w = compute_velocities(S.psi, self.grid, o.f, self.N2, self.L, compute_w)

    def _rhs(self, S: Fields):
        o = self.opt
        g = self.grid
        L = self.L
        self._diagnose_flow(S, compute_w=False)
        nq = convol_advect(S.u, S.v, S.q, g, L, use_real=True)
        nB = convol_advect(S.u, S.v, S.B, g, L, use_real=False)
        rB = refraction_B(S.B, S.psi, g, L)
        if o.linear:
            nq = np.zeros_like(nq)
            nB = np.zeros_like(nB)
        if o.passive:
            rB = np.zeros_like(rB)
        if o.passive or o.no_dispersion:
            S.A = np.zeros_like(S.A)
            S.C = np.zeros_like(S.C)
        else:
            S.A, S.C = invert_B_to_A(S.B, g, self.a)
        dq = dissipation_q_nv(S.q, o.nu_z, g)
        if o.inviscid:
            dq = np.zeros_like(dq)
        kh2 = g.kh2[None]
        alpha = o.f / 2
        rq = np.zeros_like(S.q) if o.fixed_flow else -nq + dq
        rb = -nB + 1j * alpha * kh2 * S.A - 0.5j * rB
        rq = np.where(L[None], rq, 0.0 + 0.0j)
        rb = np.where(L[None], rb, 0.0 + 0.0j)
        return rq, rb

    def _finalize(self, S: Fields):
        self._diagnose_flow(S, compute_w=True)
        o = self.opt
        if o.passive or o.no_dispersion:
            S.A = np.zeros_like(S.A)
            S.C = np.zeros_like(S.C)
        else:
            S.A, S.C = invert_B_to_A(S.B, self.grid, self.a)

    def etd_tables(self, dt):
        g = self.grid
        KX, KY = np.meshgrid(g.kx, g.ky, indexing="ij")
        o = self.opt
        lq = int_factor(KX, KY, dt, o.closure, waves=False, inviscid=o.inviscid)
        lB = int_factor(KX, KY, dt, o.closure, waves=True, inviscid=o.inviscid)
        return etd_coefficients(lq, dt), etd_coefficients(lB, dt)

    def step(self, dt: float):
        """``step!`` → ``_advance_etdrk2!`` (``src/timestep.jl:335-415``)."""
        if not (math.isfinite(dt) and dt > 0):
            raise ValueError("Δt must be finite and positive")
        o = self.opt
        L = self.L[None]
        Sn = self.fields
        (Eq, h1q, h2q), (EB, h1B, h2B) = self.etd_tables(dt)
        Eq, h1q, h2q, EB, h1B, h2B = (t[None] for t in (Eq, h1q, h2q, EB, h1B, h2B))
        rq0, rb0 = self._rhs(Sn)
        St = Sn.copy()
        qs = Sn.q if o.fixed_flow else Eq * Sn.q + h1q * rq0
        St.q = np.where(L, qs, 0.0 + 0.0j)
        St.B = np.where(L, EB * Sn.B + h1B * rb0, 0.0 + 0.0j)
        rq1, rb1 = self._rhs(St)
        Sp = Sn.copy()
        qn = Sn.q if o.fixed_flow else Eq * Sn.q + h1q * rq0 + h2q * (rq1 - rq0)
        Sp.q = np.where(L, qn, 0.0 + 0.0j)
        Sp.B = np.where(L, EB * Sn.B + h1B * rb0 + h2B * (rb1 - rb0), 0.0 + 0.0j)
        self._finalize(Sp)
        self.fields = Sp
        return self

    def check(self):
        """``_check_termination_conditions!`` (``src/simulation.jl:25-44``):
        returns (any non-finite in q,B,ψ, max|ifft ψ|)."""
        f = self.fields
        bad = not (np.all(np.isfinite(f.q)) and np.all(np.isfinite(f.B))
                   and np.all(np.isfinite(f.psi)))
        return bad, float(np.max(np.abs(fft_backward(f.psi))))


def relative_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a-b||₂ / ||b||₂ (0 when both vanish)."""
    nb = float(np.linalg.norm(np.ravel(b)))
    nd = float(np.linalg.norm(np.ravel(a) - np.ravel(b)))
    if nb == 0.0:
        return nd
    return nd / nb


# ---------------------------------------------------------------------------
# Seeded random streamfunction (src/parallel_mpi.jl:1115-1216)
# ---------------------------------------------------------------------------

_M64 = (1 << 64) - 1


def mix64(x: int) -> int:
    """_mix64 (src/parallel_mpi.jl:1117-1122) on Python integers mod 2^64."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def _mix64_np(x: np.ndarray) -> np.ndarray:
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def hash_u01(seed: int, vals, tag: int):
    """_hash_u01 (src/parallel_mpi.jl:1124-1131): vals are integer arrays (or
    ints) of the 1-based indices; returns Float64((x >> 11) + 1) * 2^-53."""
    with np.errstate(over="ignore"):
        x = np.full(np.broadcast(*[np.asarray(v) for v in vals]).shape, seed, dtype=np.uint64)
        for v in vals:
            x = _mix64_np(x ^ np.asarray(v, dtype=np.int64).astype(np.uint64))
        x = _mix64_np(x ^ np.uint64(tag))
    return ((x >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53


def hash_randn(seed: int, vals, tag: int):
    """_hash_randn (src/parallel_mpi.jl:1133-1137): Box-Muller of two hashes."""
    u1 = hash_u01(seed, vals, tag)
    u2 = hash_u01(seed, vals, tag + 1)
    return np.sqrt(-2 * np.log(u1)) * np.cos(2 * math.pi * u2)


def random_psi(nx: int, ny: int, nz: int, amplitude: float, slope: float, seed: int) -> np.ndarray:
    """init_mpi_random_psi! (src/parallel_mpi.jl:1159-1216) with seed_offset 0:
    psi-hat [k, i, j], energy amplitude*|k|^slope per mode, Hermitian by
    construction (the canonical member of a conjugate pair draws), then the
    2/3 mask (apply_dealiasing_mask!)."""
    if seed < 0:
        raise ValueError("seed must be non-negative")
    base = mix64(seed & _M64)
    ig = np.arange(1, nx + 1)[:, None] * np.ones((1, ny), dtype=np.int64)
    jg = np.ones((nx, 1), dtype=np.int64) * np.arange(1, ny + 1)[None, :]
    kx = np.where(ig <= nx // 2, ig - 1, ig - 1 - nx)
    ky = np.where(jg <= ny // 2, jg - 1, jg - 1 - ny)
    k_total = np.sqrt((kx * kx + ky * ky).astype(np.float64))
    with np.errstate(divide="ignore"):
        energy = amplitude * np.where(k_total > 0, k_total, 1.0) ** slope
    ic = np.where(ig == 1, 1, nx - ig + 2)
    jc = np.where(jg == 1, 1, ny - jg + 2)
    selfconj = (ig == ic) & (jg == jc)
    canonical = (ig < ic) | ((ig == ic) & (jg <= jc))
    ican = np.where(canonical, ig, ic)
    jcan = np.where(canonical, jg, jc)
    mask = dealias_mask(nx, ny)
    out = np.zeros((nz, nx, ny), dtype=np.complex128)
    for k in range(nz):
        kg = k + 1
        amp_self = np.sqrt(2 * energy) * hash_randn(base, (kg, ig, jg), 0x10)
        amp = np.sqrt(2 * energy) * hash_randn(base, (kg, ican, jcan), 0x20)
        phase = 2 * math.pi * hash_u01(base, (kg, ican, jcan), 0x30)
        val = amp * np.cos(phase) + 1j * (amp * np.sin(phase))
        val = np.where(canonical, val, np.conj(val))
        v = np.where(selfconj, amp_self + 0j, val)
        v[(kx == 0) & (ky == 0)] = 0
        out[k] = np.where(mask, v, 0)
    return out


# ---------------------------------------------------------------------------
# Energy diagnostics (src/core/io.jl:331-419)
# ---------------------------------------------------------------------------

def energy_components(psi, A, C, grid: Grid, a: np.ndarray):
    """_local_energy_components (src/core/io.jl:331-395) on one model, in the
    reference's summation order (j, i outer; k inner).  Returns
    (wave_KE, wave_PE, wave_CE, mean_flow_KE, mean_flow_PE)."""
    nz, nx, ny = psi.shape
    dz = abs(grid.z[1] - grid.z[0]) if nz > 1 else 1.0
    mask = dealias_mask(nx, ny)
    kh2 = grid.kh2
    w = np.ones((nx, ny))
    w[0, 0] = 0.5
    abs2 = lambda v: v.real * v.real + v.imag * v.imag  # noqa: E731
    mke = np.zeros((nx, ny)); mpe = np.zeros((nx, ny))
    wke = np.zeros((nx, ny)); wpe = np.zeros((nx, ny)); wce = np.zeros((nx, ny))
    for k in range(nz):
        ak = a[k]
        mke += w * kh2 * abs2(psi[k])
        if k < nz - 1:
            mpe += w * ak * abs2((psi[k + 1] - psi[k]) / dz)
        if nz == 1:
            LA = np.zeros_like(C[k])
        elif k == 0:
            LA = ak * C[k] / dz
        elif k == nz - 1:
            LA = -a[k - 1] * C[k - 1] / dz
        else:
            LA = (ak * C[k] - a[k - 1] * C[k - 1]) / dz
        wke += w * abs2(LA)
        wpe += w * (0.5 / ak) * kh2 * abs2(C[k])
        wce += w * (1 / (8 * ak ** 2)) * kh2 ** 2 * abs2(A[k])
    norm = 0.5 / ((nx * ny) ** 2 * nz)
    return tuple(float(np.sum(np.where(mask, t, 0.0))) * norm for t in (wke, wpe, wce, mke, mpe))


def record_energies(q, B, grid: Grid, a: np.ndarray):
    """_record_energy_diagnostics! (src/core/io.jl:398-419): psi from q
    without the wave feedback, A and C from B, then the five energies."""
    psi = invert_q_to_psi(q, grid, a)
    A, C = invert_B_to_A(B, grid, a)
    return energy_components(psi, A, C, grid, a)
