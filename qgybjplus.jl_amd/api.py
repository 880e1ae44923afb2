"""Host-side mirror of the reference's public model API on the GPU library.

Names, argument meaning and error behaviour follow subhk/QGYBJplus.jl:
``RectilinearGrid`` (src/core/grid.jl), the typed components
(src/core/components.jl), ``QGYBJModel`` (src/core/model.jl), ``set!``
(src/simulation.jl:461-503), ``ExponentialRungeKutta2``/``step!``
(src/timestep.jl:60-70, src/core/runtime.jl:239-263), ``Simulation``/``run!``
(src/simulation.jl:204-662).  Julia's bang functions are spelled ``set_``,
``step_`` and ``run_``; ``ArgumentError`` is ``ValueError``.

The model state lives in GPU memory; every numerical operation runs in
``libqgybj.so``.  Only user callbacks (an analytical psi(x, y, z), a surface
wave profile) are evaluated on the host, as in the reference.
"""

from __future__ import annotations

import enum
import logging
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


# ---------------------------------------------------------------------------
# Geometry (src/core/grid.jl)
# ---------------------------------------------------------------------------

def _spectral_wavenumbers(n: int, length: float) -> np.ndarray:
    """FFTW ordering, src/core/grid.jl:122-125."""
    scale = (2.0 * math.pi) / length
    m = np.arange(n)
    return scale * np.where(m <= (n + 1) // 2 - 1, m, m - n).astype(np.float64)


class RectilinearGrid:
    """``RectilinearGrid(; size, extent, x, y, z, centered)``, src/core/grid.jl:127-204."""

    def __init__(self, *, size, extent=None, x=None, y=None, z=None, centered=False):
        if not (isinstance(size, tuple) and len(size) == 3):
            raise ValueError("size must be a three-element tuple")
        if not all(isinstance(v, (int, np.integer)) for v in size):
            raise ValueError("grid dimensions must be integers")
        nx, ny, nz = (int(v) for v in size)
        if min(nx, ny, nz) <= 0:
            raise ValueError("all grid dimensions must be positive")
        if centered and (x is not None or y is not None):
            raise ValueError("centered=true cannot be combined with explicit x or y ranges")

        def bounds(b, name):
            if b is None:
                return None
            if not (isinstance(b, tuple) and len(b) == 2):
                raise ValueError(f"{name} must be a two-element tuple")
            lo, hi = float(b[0]), float(b[1])
            if not (math.isfinite(lo) and math.isfinite(hi)):
                raise ValueError(f"{name} bounds must be finite")
            if not hi > lo:
                raise ValueError(f"{name} upper bound must exceed its lower bound")
            return lo, hi

        xb, yb, zb = bounds(x, "x"), bounds(y, "y"), bounds(z, "z")
        if extent is None:
            if xb is None or yb is None or zb is None:
                raise ValueError("provide extent=(Lx, Ly, Lz) or explicit x, y and z bounds")
            Lx, Ly, Lz = xb[1] - xb[0], yb[1] - yb[0], zb[1] - zb[0]
        else:
            if not (isinstance(extent, tuple) and len(extent) == 3):
                raise ValueError("extent must be a three-element tuple")
            Lx, Ly, Lz = (float(v) for v in extent)
            if not all(math.isfinite(v) for v in (Lx, Ly, Lz)):
                raise ValueError("all grid extents must be finite")
            if not all(v > 0 for v in (Lx, Ly, Lz)):
                raise ValueError("all grid extents must be positive")
            for b, L, name in ((xb, Lx, "x"), (yb, Ly, "y"), (zb, Lz, "z")):
                if b is not None and not math.isclose(b[1] - b[0], L, rel_tol=10 * 2.2e-16,
                                                      abs_tol=10 * 2.2e-16 * max(abs(L), 1.0)):
                    raise ValueError(f"{name} range length does not match extent")
        x0 = xb[0] if xb is not None else (-Lx / 2 if centered else 0.0)
        y0 = yb[0] if yb is not None else (-Ly / 2 if centered else 0.0)
        zbot = zb[0] if zb is not None else -Lz
        self.size = (nx, ny, nz)
        self.extent = (Lx, Ly, Lz)
        self.origin = (x0, y0)
        self.dx, self.dy, self.dz = Lx / nx, Ly / ny, Lz / nz
        self.x = np.array([x0 + i * self.dx for i in range(nx)])
        self.y = np.array([y0 + j * self.dy for j in range(ny)])
        self.z_faces = np.array([zbot + k * self.dz for k in range(nz + 1)])
        self.z = np.array([(self.z_faces[k] + self.z_faces[k + 1]) / 2 for k in range(nz)])
        self.x_faces = np.array([x0 + i * self.dx for i in range(nx + 1)])
        self.y_faces = np.array([y0 + j * self.dy for j in range(ny + 1)])
        self.kx = _spectral_wavenumbers(nx, Lx)
        self.ky = _spectral_wavenumbers(ny, Ly)
        self.kh2 = self.kx[:, None] ** 2 + self.ky[None, :] ** 2

    @property
    def dz_solver(self) -> float:
        """Spacing the vertical operators use (``G.z[2]-G.z[1]``)."""
        return float(self.z[1] - self.z[0]) if self.size[2] > 1 else 1.0

    def __repr__(self):
        return f"RectilinearGrid(size={self.size}, extent={self.extent}, origin={self.origin})"


def dealias_mask(grid: RectilinearGrid) -> np.ndarray:
    """Radial 2/3 mask (src/physics.jl:28-43); shape (nx, ny)."""
    nx, ny = grid.size[:2]
    kmax = min(nx, ny) // 3
    i, j = np.arange(nx), np.arange(ny)
    ix = np.where(i <= nx // 2, i, i - nx)
    jy = np.where(j <= ny // 2, j, j - ny)
    return (ix[:, None] ** 2 + jy[None, :] ** 2) <= kmax ** 2


# ---------------------------------------------------------------------------
# Components (src/core/components.jl, src/stratification.jl)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class FPlane:
    f: float

    def __post_init__(self):
        v = float(self.f)
        if not math.isfinite(v):
            raise ValueError(f"f must be finite (got {self.f})")
        if v == 0:
            raise ValueError("f must be non-zero")
        object.__setattr__(self, "f", v)


@dataclass(frozen=True)
class ConstantStratification:
    N2: float

    def __post_init__(self):
        v = float(self.N2)
        if not math.isfinite(v):
            raise ValueError(f"N² must be finite (got {self.N2})")
        if not v > 0:
            raise ValueError(f"N² must be positive (got {self.N2})")
        object.__setattr__(self, "N2", v)

    def evaluate(self, z: float) -> float:
        """Through ``ConstantN(sqrt(N²))`` (src/core/runtime.jl:29-30)."""
        n0 = math.sqrt(self.N2)
        return n0 * n0


@dataclass(frozen=True)
class AnalyticalProfile:
    """``AnalyticalProfile(func; returns=:N²)`` (src/stratification.jl:95-111,263-270)."""
    func: Callable[[float], float]
    returns: str = "N2"

    def __post_init__(self):
        if self.returns not in ("N", "N2"):
            raise ValueError("returns must be :N or :N²")

    def evaluate(self, z: float) -> float:
        val = float(self.func(z))
        return max(val, 0.0) if self.returns == "N2" else max(val * val, 0.0)


class FlowEvolution: ...
class FixedFlow(FlowEvolution): ...
class EvolvingFlow(FlowEvolution): ...
class FeedbackMode: ...
class NoFeedback(FeedbackMode): ...
class WaveMeanFeedback(FeedbackMode): ...
class NoWaveFeedback(FeedbackMode): ...
class WaveFormulation: ...
class YBJPlus(WaveFormulation): ...
class YBJ(WaveFormulation): ...
class PassiveWave(WaveFormulation): ...
class DissipationMode: ...
class Dissipative(DissipationMode): ...
class Inviscid(DissipationMode): ...
class DynamicsMode: ...
class NonlinearDynamics(DynamicsMode): ...
class LinearDynamics(DynamicsMode): ...
class DispersionMode: ...
class Dispersive(DispersionMode): ...
class NoDispersion(DispersionMode): ...


def _terms(field_name, coefficients, orders):
    if len(coefficients) == 0:
        raise ValueError(f"{field_name} hyperdiffusivity requires at least one term")
    if len(coefficients) != len(orders):
        raise ValueError(f"{field_name} hyperdiffusivity coefficients and orders must have equal lengths")
    vals = tuple(float(c) for c in coefficients)
    if not all(math.isfinite(v) and v >= 0 for v in vals):
        raise ValueError(f"{field_name} hyperdiffusivity coefficients must be finite and non-negative")
    ords = tuple(int(o) for o in orders)
    if not all(o > 0 and o % 2 == 0 for o in ords):
        raise ValueError(f"{field_name} hyperdiffusivity orders must be positive and even")
    return vals, ords


class FlowHyperdiffusivity:
    """Balanced-flow damping only (src/core/components.jl:105-160)."""

    def __init__(self, coefficient=None, *, order: int = 4, coefficients=None, orders=None):
        if coefficients is None:
            if coefficient is None:
                raise ValueError("coefficient is required")
            coefficients, orders = (coefficient,), (order,)
        self.coefficients, self.orders = _terms("flow", coefficients, orders)


class WaveHyperdiffusivity:
    """Wave damping only (src/core/components.jl:162-192)."""

    def __init__(self, coefficient=None, *, order: int = 4, coefficients=None, orders=None):
        if coefficients is None:
            if coefficient is None:
                raise ValueError("coefficient is required")
            coefficients, orders = (coefficient,), (order,)
        self.coefficients, self.orders = _terms("wave", coefficients, orders)


class HorizontalHyperdiffusivity:
    """Flow + wave damping; default 4th+12th order (src/core/components.jl:194-208)."""

    def __init__(self, *, flow: Optional[FlowHyperdiffusivity] = None,
                 wave: Optional[WaveHyperdiffusivity] = None):
        self.flow = flow if flow is not None else FlowHyperdiffusivity(
            coefficients=(0.01, 10.0), orders=(4, 12))
        self.wave = wave if wave is not None else WaveHyperdiffusivity(
            coefficients=(0.0, 10.0), orders=(4, 12))


@dataclass(frozen=True)
class VerticalDiffusivity:
    coefficient: float = 0.0

    def __post_init__(self):
        v = float(self.coefficient)
        if not math.isfinite(v):
            raise ValueError("vertical diffusivity must be finite")
        if v < 0:
            raise ValueError("vertical diffusivity must be non-negative")
        object.__setattr__(self, "coefficient", v)


@dataclass(frozen=True)
class SurfaceWave:
    """Horizontally uniform surface wave (src/core/components.jl:211-215,274-283)."""
    amplitude: float
    scale: float
    profile: str = "gaussian"

    def __post_init__(self):
        if not (math.isfinite(self.amplitude) and math.isfinite(self.scale)):
            raise ValueError("wave amplitude and scale must be finite")
        if not self.scale > 0:
            raise ValueError(f"wave scale must be positive (got {self.scale})")
        if self.profile not in ("gaussian", "exponential"):
            raise ValueError("profile must be :gaussian or :exponential")


@dataclass(frozen=True)
class RandomStreamfunction:
    """Deterministic random streamfunction spectrum (src/core/components.jl:
    218-222,264-272), generated on the device by the reference's counter hash
    (init_mpi_random_psi!, src/parallel_mpi.jl:1159-1216)."""
    amplitude: float = 1.0
    spectral_slope: float = -3.0
    seed: int = 0

    def __post_init__(self):
        if not (math.isfinite(self.amplitude) and math.isfinite(self.spectral_slope)):
            raise ValueError("random streamfunction parameters must be finite")
        if self.amplitude < 0:
            raise ValueError("amplitude must be non-negative")
        object.__setattr__(self, "seed", int(self.seed))


@dataclass(frozen=True)
class FieldArray:
    """A 3-D field for ``set_`` (src/core/components.jl:225-241)."""
    values: np.ndarray
    space: str = "physical"
    layout: str = "zxy"

    def __post_init__(self):
        if np.ndim(self.values) != 3:
            raise ValueError("field arrays must be three-dimensional")
        if self.space not in ("physical", "spectral"):
            raise ValueError("field space must be :physical or :spectral")
        if self.layout not in ("zxy", "xyz"):
            raise ValueError("field layout must be :zxy or :xyz")


@dataclass(frozen=True)
class GPU:
    """Architecture selecting this library: one MI355X holds the whole grid.

    ``precision`` is "fp64" (plain) or "mixed" (FP32 transforms and products,
    FP64 state, column solves and ETD coefficients)."""
    precision: str = "fp64"

    def __post_init__(self):
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")


def _component(value, kinds, sym_map, bool_map, what):
    if isinstance(value, kinds):
        return value
    if isinstance(value, bool) and bool_map is not None:
        return bool_map[value]()
    if isinstance(value, str) and value in sym_map:
        return sym_map[value]()
    raise ValueError(f"{what} must be one of {sorted(sym_map)} or a component instance")


# ---------------------------------------------------------------------------
# Model (src/core/model.jl, src/core/runtime.jl)
# ---------------------------------------------------------------------------

class _Fields:
    """Accessor mirroring ``model.fields`` (ModelFields, src/core/fields.jl).

    Reading an attribute downloads the field in (z, x, y) order; assigning one
    uploads it (spectral values for q, B, psi, A, C)."""

    def __init__(self, model):
        object.__setattr__(self, "_m", model)

    def __getattr__(self, name):
        if name in _lib.FIELD_IDS:
            return self._m._h.get_field(name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _lib.FIELD_IDS:
            self._m._h.set_field(name, value, _lib.SPACE_SPECTRAL)
        else:
            raise AttributeError(name)


class QGYBJModel:
    """``QGYBJModel(; grid, coriolis, stratification, closure, ...)`` on one GPU."""

    def __init__(self, *, grid: RectilinearGrid, coriolis=None, stratification=None,
                 closure=None, vertical_diffusion=None, flow="evolving", feedback="none",
                 formulation=None, ybj_plus: bool = True, inviscid=False, linear=False,
                 no_dispersion=False, architecture: GPU = GPU(), verbose: bool = False):
        self._configure(grid=grid, coriolis=coriolis, closure=closure, vertical_diffusion=vertical_diffusion,
                        flow=flow, feedback=feedback, formulation=formulation, ybj_plus=ybj_plus,
                        inviscid=inviscid, linear=linear, no_dispersion=no_dispersion,
                        architecture=architecture)
        self.stratification = self._stratification(stratification)
        self.N2, self.a_ell = self._profiles(self.stratification)
        self._h = _lib.Handle(**self._handle_args(self.N2, self.a_ell, 1))
        self.fields = _Fields(self)
        self.finalized = False

    def _configure(self, *, grid, coriolis, closure, vertical_diffusion, flow, feedback, formulation,
                   ybj_plus, inviscid, linear, no_dispersion, architecture):
        if not isinstance(grid, RectilinearGrid):
            raise ValueError("grid must be a RectilinearGrid")
        self.grid = grid
        self.coriolis = coriolis if isinstance(coriolis, FPlane) else FPlane(1e-4 if coriolis is None else coriolis)
        self.closure = closure if closure is not None else HorizontalHyperdiffusivity()
        vd = vertical_diffusion if vertical_diffusion is not None else VerticalDiffusivity()
        self.vertical_diffusion = vd if isinstance(vd, VerticalDiffusivity) else VerticalDiffusivity(vd)
        self.flow = _component(flow, FlowEvolution, {"fixed": FixedFlow, "evolving": EvolvingFlow},
                               {True: FixedFlow, False: EvolvingFlow}, "flow")
        self.feedback = _component(feedback, FeedbackMode,
                                   {"none": NoFeedback, "off": NoFeedback, "wave_mean": WaveMeanFeedback,
                                    "on": WaveMeanFeedback, "no_wave_feedback": NoWaveFeedback},
                                   {True: WaveMeanFeedback, False: NoFeedback}, "feedback")
        if formulation is None:
            self.formulation = YBJPlus() if ybj_plus else YBJ()
        else:
            self.formulation = _component(formulation, WaveFormulation,
                                          {"ybj_plus": YBJPlus, "ybj": YBJ, "passive": PassiveWave},
                                          None, "formulation")
        self.dissipation = _component(inviscid, DissipationMode, {}, {True: Inviscid, False: Dissipative},
                                      "inviscid")
        self.dynamics = _component(linear, DynamicsMode, {}, {True: LinearDynamics, False: NonlinearDynamics},
                                   "linear")
        self.dispersion = _component(no_dispersion, DispersionMode, {},
                                     {True: NoDispersion, False: Dispersive}, "no_dispersion")
        self.architecture = architecture
        self.dealias_mask = dealias_mask(grid)

    @staticmethod
    def _stratification(stratification):
        if stratification is None:
            return ConstantStratification(1e-5)
        if isinstance(stratification, (int, float)):
            return ConstantStratification(stratification)
        return stratification

    def _profiles(self, stratification):
        grid = self.grid
        nz = grid.size[2]
        # N² on the bottom face of each cell (src/stratification.jl:308-318)
        N2 = np.array([float(stratification.evaluate(grid.z[k] - grid.dz / 2)) for k in range(nz)])
        if not np.all(np.isfinite(N2)) or not np.all(N2 > 0):
            raise ValueError("N² must be finite and positive at every level")
        f = self.coriolis.f
        return N2, (f * f) / np.maximum(N2, SQRT_EPS)  # src/physics.jl:2-25

    def _handle_args(self, N2, a_ell, members):
        nx, ny, nz = self.grid.size
        cl = self.closure
        if isinstance(cl, HorizontalHyperdiffusivity):
            flow_terms = list(zip(cl.flow.coefficients, cl.flow.orders))
            wave_terms = list(zip(cl.wave.coefficients, cl.wave.orders))
        elif isinstance(cl, FlowHyperdiffusivity):
            flow_terms, wave_terms = list(zip(cl.coefficients, cl.orders)), []
        elif isinstance(cl, WaveHyperdiffusivity):
            flow_terms, wave_terms = [], list(zip(cl.coefficients, cl.orders))
        else:
            raise ValueError("closure must be a hyperdiffusivity component")
        form = {YBJPlus: _lib.YBJ_PLUS, PassiveWave: _lib.PASSIVE, YBJ: _lib.YBJ}[type(self.formulation)]
        return dict(
            nx=nx, ny=ny, nz=nz, f=self.coriolis.f, dz=self.grid.dz_solver, kx=self.grid.kx, ky=self.grid.ky,
            N2=N2, a_ell=a_ell, flow_terms=flow_terms, wave_terms=wave_terms,
            nu_z=self.vertical_diffusion.coefficient,
            fixed_flow=isinstance(self.flow, FixedFlow),
            wave_feedback=isinstance(self.feedback, WaveMeanFeedback),
            formulation=form, inviscid=isinstance(self.dissipation, Inviscid),
            linear=isinstance(self.dynamics, LinearDynamics),
            no_dispersion=isinstance(self.dispersion, NoDispersion),
            precision=_lib.PRECISION_MIXED if self.architecture.precision == "mixed" else _lib.PRECISION_FP64,
            members=members)

    # model-level operators (src/core/runtime.jl:169-206)
    def invert_q_to_psi(self):
        self._h.invert_q_to_psi()
        return self

    def invert_B_to_A(self):
        self._h.invert_B_to_A()
        return self

    def compute_velocities(self, compute_w: bool = True):
        self._h.compute_velocities(compute_w)
        return self

    def fft_forward(self, physical) -> np.ndarray:
        return self._h.fft(physical, -1)

    def fft_backward(self, spectral) -> np.ndarray:
        return self._h.fft(spectral, +1)

    def compute_qw(self) -> np.ndarray:
        return self._h.compute_qw()

    def progress_maxima(self) -> dict:
        """``_progress_maxima`` (src/simulation.jl:46-81) on the device: the
        largest physical wave velocity |LA| and flow speed hypot(u, v)."""
        w, f = self._h.progress_maxima()
        return {"wave_speed": w, "flow_speed": f}

    def energies(self) -> dict:
        """The five energies of _local_energy_components (src/core/io.jl:331-395),
        reduced on the device from psi = L^-1 q and A, C from B."""
        return dict(zip(ENERGY_NAMES, (float(v) for v in self._h.energies(-1)[0])))

    def autotune(self, dt: float, steps: int = 3, force: bool = False, in_step: bool = False) -> dict:
        """Choose the FFT-pass tile variants and column-solve launch shapes by
        timing `steps` steps of size `dt` per candidate on this model's state
        (which is restored).  The choice is cached per grid class in the
        process; `force` re-times.  `in_step`: then time the candidates of the
        passes that take most of the step as whole graph-replayed steps and
        keep the fastest (about three seconds more at 256x256x128).  Returns
        the choice per pass."""
        return self._h.autotune(dt, steps, force, in_step)

    def tuning(self) -> dict:
        return self._h.tuning()

    def set_tuning(self, choices: dict):
        self._h.set_tuning(choices)
        return self

    def finalize(self):
        if not self.finalized:
            self._h.close()
            self.finalized = True
        return self

    def __repr__(self):
        return f"QGYBJModel(grid={self.grid.size}, architecture={self.architecture})"


def finalize_model_(model: QGYBJModel):
    return model.finalize()


class _MemberFields:
    """``fields`` of one ensemble member: (z, x, y) arrays of its nz levels."""

    def __init__(self, ens, member):
        object.__setattr__(self, "_e", ens)
        object.__setattr__(self, "_m", member)

    def __getattr__(self, name):
        if name in _lib.FIELD_IDS:
            return self._e._h.get_field_member(self._m, name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _lib.FIELD_IDS:
            self._e._h.set_field_member(self._m, name, value, _lib.SPACE_SPECTRAL)
        else:
            raise AttributeError(name)


class QGYBJEnsemble(QGYBJModel):
    """Ensemble of models of one grid class on one GPU.

    Every setting is shared except the stratification: ``stratifications``
    gives one per member.  The members are stacked as extra levels of one
    device model, so every pass of the ETD-RK2 step covers all of them in one
    launch and ``step_`` advances all members together.  Member m is read and
    set through ``member_fields(m)`` and ``set_member``; its results are those
    of the same model run alone."""

    def __init__(self, *, stratifications, **kwargs):
        strat = list(stratifications)
        if not strat:
            raise ValueError("an ensemble needs at least one member")
        kwargs.pop("stratification", None)
        self._configure(**{k: kwargs.get(k, d) for k, d in (
            ("grid", None), ("coriolis", None), ("closure", None), ("vertical_diffusion", None),
            ("flow", "evolving"), ("feedback", "none"), ("formulation", None), ("ybj_plus", True),
            ("inviscid", False), ("linear", False), ("no_dispersion", False), ("architecture", GPU()))})
        self.stratifications = [self._stratification(x) for x in strat]
        prof = [self._profiles(x) for x in self.stratifications]
        self.members = len(prof)
        self.N2 = np.concatenate([p[0] for p in prof])
        self.a_ell = np.concatenate([p[1] for p in prof])
        self._h = _lib.Handle(**self._handle_args(self.N2, self.a_ell, self.members))
        self.fields = _Fields(self)
        self.finalized = False

    def member_fields(self, member: int):
        if not 0 <= member < self.members:
            raise IndexError(f"member {member} out of range")
        return _MemberFields(self, member)

    def set_member(self, member: int, *, psi=None, pv_method: str = "qg", waves=None):
        """``set!`` restricted to one member (the others are not touched)."""
        if self.finalized:
            raise RuntimeError("cannot modify a finalized model")
        h = self._h
        if psi is not None:
            if isinstance(psi, RandomStreamfunction):
                h.set_random_psi(psi.amplitude, psi.spectral_slope, psi.seed, member=member)
            elif callable(psi):
                h.set_field_member(member, "psi", _evaluate_on_nodes(self.grid, psi), _lib.SPACE_PHYSICAL)
            else:
                vals, space = _field_values_zxy(self, psi)
                h.set_field_member(member, "psi", vals,
                                   _lib.SPACE_PHYSICAL if space == "physical" else _lib.SPACE_SPECTRAL)
            if pv_method in ("qg", "balanced", "barotropic", "asselin"):
                h.compute_q_from_psi_member(member, barotropic=pv_method in ("barotropic", "asselin"))
            elif pv_method != "none":
                raise ValueError("pv_method must be :qg, :barotropic, or :none")
        if waves is not None:
            if isinstance(waves, SurfaceWave):
                h.set_surface_wave(_surface_wave_profile(self.grid, waves), member=member)
            else:
                vals, space = _field_values_zxy(self, waves)
                h.set_field_member(member, "B", vals,
                                   _lib.SPACE_PHYSICAL if space == "physical" else _lib.SPACE_SPECTRAL)
            h.invert_B_to_A_member(member)
        return self

    def check_member(self, member: int):
        return self._h.check_member(member)

    def energies_member(self, member: int) -> dict:
        return dict(zip(ENERGY_NAMES, (float(v) for v in self._h.energies(member)[0])))

    def set_active_members(self, active: int):
        """Step only members [0, active) from now on."""
        self._h.set_active_members(active)
        return self

    def move_member(self, src: int, dst: int):
        """Copy member src (state, stratification and column eliminations)
        over member dst."""
        self._h.move_member(src, dst)
        self.stratifications[dst] = self.stratifications[src]
        nz = self.grid.size[2]
        self.N2[dst * nz:(dst + 1) * nz] = self.N2[src * nz:(src + 1) * nz]
        self.a_ell[dst * nz:(dst + 1) * nz] = self.a_ell[src * nz:(src + 1) * nz]
        return self

    def set_member_stratification(self, member: int, stratification):
        """Give slot `member` another N²(z) (its column eliminations are
        recorded again); the other members are not touched."""
        strat = self._stratification(stratification)
        N2, a = self._profiles(strat)
        self._h.set_member_stratification(member, N2, a)
        nz = self.grid.size[2]
        self.N2[member * nz:(member + 1) * nz] = N2
        self.a_ell[member * nz:(member + 1) * nz] = a
        self.stratifications[member] = strat
        return self

    def __repr__(self):
        return f"QGYBJEnsemble(grid={self.grid.size}, members={self.members}, architecture={self.architecture})"


# ---------------------------------------------------------------------------
# set! (src/simulation.jl:230-329, 405-503)
# ---------------------------------------------------------------------------

def _evaluate_on_nodes(grid, fn) -> np.ndarray:
    nx, ny, nz = grid.size
    X, Y = np.meshgrid(grid.x, grid.y, indexing="ij")
    out = np.empty((nz, nx, ny), dtype=np.complex128)
    for k in range(nz):
        try:
            vals = fn(X, Y, grid.z[k])
            vals = np.broadcast_to(np.asarray(vals, dtype=np.complex128), (nx, ny))
        except Exception:
            vals = np.array([[fn(grid.x[i], grid.y[j], grid.z[k]) for j in range(ny)]
                             for i in range(nx)], dtype=np.complex128)
        out[k] = vals
    return out


def _field_values_zxy(model, source) -> Tuple[np.ndarray, str]:
    if isinstance(source, FieldArray):
        vals = np.asarray(source.values)
        if source.layout == "xyz":
            vals = np.transpose(vals, (2, 0, 1))
        space = source.space
    else:
        vals, space = np.asarray(source), "physical"
    nx, ny, nz = model.grid.size
    if vals.shape != (nz, nx, ny):
        raise ValueError(f"initial field has size {vals.shape}; expected {(nz, nx, ny)} in z-x-y layout")
    return vals.astype(np.complex128), space


def _surface_wave_profile(g, wave) -> np.ndarray:
    """The amplitude times the vertical profile of set_surface_waves!
    (src/simulation.jl:317-324) at each level's depth max(0, -z - dz/2); the
    horizontally uniform field and its transform are formed on the device."""
    nz = g.size[2]
    prof = np.empty(nz)
    for k in range(nz):
        depth = max(0.0, -g.z[k] - g.dz / 2)
        p = math.exp(-(depth ** 2) / wave.scale ** 2) if wave.profile == "gaussian" \
            else math.exp(-depth / wave.scale)
        prof[k] = wave.amplitude * p
    return prof


def set_(model, *, psi=None, mean_flow=None, pv_method: str = "qg", waves=None, B=None,
         verbose: bool = False):
    """``set!(model; ψ, pv_method, waves)`` (src/simulation.jl:461-503)."""
    if isinstance(model, Simulation):
        model._assert_mutable()
        model = model.model
    if model.finalized:
        raise RuntimeError("cannot modify a finalized model")
    h = model._h
    flow = mean_flow if mean_flow is not None else psi
    wave = waves if waves is not None else B
    if flow is not None:
        if isinstance(flow, RandomStreamfunction):
            h.set_random_psi(flow.amplitude, flow.spectral_slope, flow.seed)
        elif callable(flow):
            h.set_field("psi", _evaluate_on_nodes(model.grid, flow), _lib.SPACE_PHYSICAL)
        elif isinstance(flow, (FieldArray, np.ndarray)):
            vals, space = _field_values_zxy(model, flow)
            h.set_field("psi", vals, _lib.SPACE_PHYSICAL if space == "physical" else _lib.SPACE_SPECTRAL)
        else:
            raise ValueError("mean_flow/ψ must be a function, RandomStreamfunction, array or FieldArray")
        if pv_method in ("qg", "balanced"):
            h.compute_q_from_psi(barotropic=False)
        elif pv_method in ("barotropic", "asselin"):
            h.compute_q_from_psi(barotropic=True)
        elif pv_method != "none":
            raise ValueError("pv_method must be :qg, :barotropic, or :none")
        h.compute_velocities(compute_w=False)
    if wave is not None:
        if isinstance(wave, SurfaceWave):
            members = getattr(model, "members", 1)
            h.set_surface_wave(np.tile(_surface_wave_profile(model.grid, wave), members))
        elif isinstance(wave, (FieldArray, np.ndarray)):
            vals, space = _field_values_zxy(model, wave)
            h.set_field("B", vals, _lib.SPACE_PHYSICAL if space == "physical" else _lib.SPACE_SPECTRAL)
        else:
            raise ValueError("waves/B must be a SurfaceWave, array or FieldArray")
        if isinstance(model.formulation, YBJ):
            raise NotImplementedError("the normal YBJ formulation is not part of this library")
        h.invert_B_to_A()
    return model


# ---------------------------------------------------------------------------
# Time stepping (src/timestep.jl:60-70, src/core/runtime.jl:239-263)
# ---------------------------------------------------------------------------

class ExponentialRungeKutta2:
    def __init__(self, *, dt: float):
        v = float(dt)
        if not math.isfinite(v):
            raise ValueError("Δt must be finite")
        if not v > 0:
            raise ValueError("Δt must be positive")
        self.dt = v
        self.workspace = None


def step_(model: QGYBJModel, timestepper: ExponentialRungeKutta2, nsteps: int = 1):
    """``step!(model, timestepper)``; the workspace lives on the device."""
    if model.finalized:
        raise RuntimeError("cannot use operators after model finalization")
    model._h.step(timestepper.dt, nsteps)
    timestepper.workspace = "device"
    return model


class SimulationState(enum.Enum):
    """``SimulationState`` (src/core/simulation.jl:9-16)."""
    Ready = 0
    Running = 1
    Stopped = 2
    Failed = 3
    Finalized = 4


@dataclass(frozen=True)
class IterationInterval:
    """Every `interval` iterations (src/core/components.jl:297-305)."""
    interval: int

    def __post_init__(self):
        if not int(self.interval) > 0:
            raise ValueError(f"interval must be positive (got {self.interval})")
        object.__setattr__(self, "interval", int(self.interval))


@dataclass(frozen=True)
class TimeInterval:
    """Every `interval` seconds of model time (src/core/components.jl:285-296)."""
    interval: float

    def __post_init__(self):
        v = float(self.interval)
        if not math.isfinite(v):
            raise ValueError(f"interval must be finite (got {self.interval})")
        if not v > 0:
            raise ValueError(f"interval must be positive (got {self.interval})")
        object.__setattr__(self, "interval", v)


ENERGY_NAMES = ("wave_KE", "wave_PE", "wave_CE", "mean_flow_KE", "mean_flow_PE")


class EnergyDiagnostics:
    """``EnergyDiagnosticsManager`` (src/core/io.jl:243-326,398-431) kept in
    memory: the time series of the five energies, recorded on the device by
    ``qg_energies`` when the schedule is due (the NetCDF writer is not part of
    this library)."""

    def __init__(self, schedule, start_time: float = 0.0):
        if not isinstance(schedule, (IterationInterval, TimeInterval)):
            raise ValueError("simulation diagnostics must be an IterationInterval, TimeInterval, or False")
        self.schedule = schedule
        self.last_iteration = None
        self.next_time = start_time + schedule.interval if isinstance(schedule, TimeInterval) else None
        self.series = {"time": [], **{n: [] for n in ENERGY_NAMES}}

    def due(self, clock, initial: bool = False) -> bool:
        """``_diagnostics_due`` (src/core/io.jl:299-314)."""
        if self.last_iteration == clock.iteration:
            return False
        if initial:
            return True
        if isinstance(self.schedule, IterationInterval):
            return clock.iteration % self.schedule.interval == 0
        tol = 8 * np.finfo(float).eps * max(abs(clock.time), abs(self.schedule.interval), 1.0)
        return self.next_time is not None and clock.time >= self.next_time - tol

    def record(self, model, clock):
        """``_record_energy_diagnostics!`` (src/core/io.jl:398-431)."""
        e = model.energies()
        self.series["time"].append(clock.time)
        for n in ENERGY_NAMES:
            self.series[n].append(e[n])
        self.last_iteration = clock.iteration
        if isinstance(self.schedule, TimeInterval):  # _advance_diagnostics_deadline!
            if self.next_time is None:
                self.next_time = clock.time + self.schedule.interval
            tol = 8 * np.finfo(float).eps * max(abs(clock.time), abs(self.schedule.interval), 1.0)
            while self.next_time <= clock.time + tol:
                self.next_time += self.schedule.interval


@dataclass
class Clock:
    time: float = 0.0
    iteration: int = 0


class Simulation:
    """``Simulation(model; Δt, stop_time, stop_iteration)`` (src/simulation.jl:204-227)."""

    def __init__(self, model: QGYBJModel, *, dt: float = 1.0, stop_time=None, stop_iteration=None,
                 check_every: int = 1, diagnostics=False, diagnostics_interval: int = 10, verbose: bool = True):
        if model.finalized:
            raise RuntimeError("cannot create a simulation from a finalized model")
        if not int(diagnostics_interval) > 0:
            raise ValueError("diagnostics_interval must be positive")
        self.model = model
        # run options (SimulationRunOptions, src/core/simulation.jl:19-33)
        self.diagnostics_interval = int(diagnostics_interval)
        self.verbose = bool(verbose)
        self.diagnostics = diagnostics  # False, IterationInterval or TimeInterval
        self.diagnostics_manager: Optional[EnergyDiagnostics] = None
        self.timestepper = ExponentialRungeKutta2(dt=dt)
        self.clock = Clock()
        self.stop_time = stop_time
        self.stop_iteration = stop_iteration
        self.check_every = int(check_every)
        self.state = SimulationState.Ready

    def _assert_mutable(self):
        if self.state in (SimulationState.Running, SimulationState.Failed, SimulationState.Finalized):
            raise RuntimeError(f"cannot modify a {self.state.name.lower()} simulation")

    def _before_stop_time(self, t=None) -> bool:
        """``_before_stop_time`` (src/simulation.jl:15-23) at model time t
        (default: the clock's)."""
        if self.stop_time is None:
            return True
        t = self.clock.time if t is None else t
        tol = 8 * np.finfo(float).eps * max(abs(t), abs(self.stop_time), abs(self.timestepper.dt), 1.0)
        return t < self.stop_time - tol

    def _steps_to_next_event(self, mgr, report: int) -> int:
        """Iterations the run! loop takes from the current clock before it
        stops (stop iteration or stop time) or has to do something on the host
        after a step (a due diagnostics record, or a progress report every
        `report` iterations; 0: none): the clock is advanced on a copy exactly
        as the loop advances it."""
        c = Clock(self.clock.time, self.clock.iteration)
        dt = self.timestepper.dt
        n = 0
        while True:
            if self.stop_iteration is not None and c.iteration >= self.stop_iteration:
                break
            if not self._before_stop_time(c.time):
                break
            n += 1
            c.iteration += 1
            c.time += dt
            if (report and c.iteration % int(report) == 0) or (mgr is not None and mgr.due(c)) \
                    or n >= self._max_stretch:
                break
        return n

    _max_stretch = 1 << 16  # iterations handed to the device at once


def _raise_termination(nonfinite: bool, max_abs_psi: float):
    if nonfinite:
        raise FloatingPointError("non-finite value detected in the model state")
    if not max_abs_psi <= 1e10:
        raise FloatingPointError(f"solution appears to be blowing up (max |psi| = {max_abs_psi})")


def check_termination_(simulation: Simulation):
    """``_check_termination_conditions!`` (src/simulation.jl:25-44)."""
    bad, m = simulation.model._h.check()
    _raise_termination(bad, m)
    return simulation


_log = logging.getLogger("qgybjplus_amd")


def _report_progress(simulation: Simulation, detailed: bool):
    """The progress report of run! (src/simulation.jl:629-641, 83-97): with
    progress = true the line with the wave and flow speed maxima, else an
    informational log record."""
    c = simulation.clock
    if detailed:
        m = simulation.model.progress_maxima()
        print(f"iteration={c.iteration} | time={c.time} s | max_wave_speed={m['wave_speed']:.3e} m/s | "
              f"max_flow_speed={m['flow_speed']:.3e} m/s", flush=True)
    else:
        _log.info("Simulation progress iteration=%d time=%s", c.iteration, c.time)


def run_(simulation: Simulation, *, stop_iteration=None, stop_time=None, progress: Optional[bool] = None,
         diagnostics_interval: Optional[int] = None, verbose: Optional[bool] = None):
    """``run!(simulation)`` (src/simulation.jl:570-662) without output writers.
    Every diagnostics_interval iterations a verbose run reports progress:
    progress = True prints the reference's line with the speed maxima, else an
    INFO record goes to the ``qgybjplus_amd`` logger."""
    if simulation.state is SimulationState.Finalized:
        raise RuntimeError("cannot run a finalized simulation")
    if simulation.state is SimulationState.Running:
        raise RuntimeError("simulation is already running")
    if simulation.state is SimulationState.Failed:
        raise RuntimeError("simulation is in a failed state")
    if simulation.state is SimulationState.Stopped:
        raise RuntimeError("simulation has stopped; construct a new simulation to continue")
    if stop_iteration is not None:
        simulation.stop_iteration = stop_iteration
    if stop_time is not None:
        simulation.stop_time = stop_time
    if simulation.stop_iteration is None and simulation.stop_time is None:
        raise ValueError("set stop_iteration or stop_time before run!")
    if verbose is not None:
        simulation.verbose = bool(verbose)
    if progress is not None:
        simulation.verbose = bool(progress)
    detailed = progress is True
    if diagnostics_interval is not None:
        if not int(diagnostics_interval) > 0:
            raise ValueError("diagnostics_interval must be positive")
        simulation.diagnostics_interval = int(diagnostics_interval)
    reporting = simulation.verbose and (detailed or _log.isEnabledFor(logging.INFO))
    report = simulation.diagnostics_interval if reporting else 0
    simulation.state = SimulationState.Running
    model = simulation.model
    ts = simulation.timestepper
    try:
        if simulation.diagnostics is not False and simulation.diagnostics is not None:
            if simulation.diagnostics_manager is None:
                simulation.diagnostics_manager = EnergyDiagnostics(simulation.diagnostics, simulation.clock.time)
        mgr = simulation.diagnostics_manager
        if mgr is not None and mgr.due(simulation.clock, initial=True):
            mgr.record(model, simulation.clock)
        if simulation.check_every == 1:
            # every step checked, as run! does: stretches of steps between
            # host events run on the device (qg_run), checks included
            while True:
                n = simulation._steps_to_next_event(mgr, report)
                if n == 0:
                    break
                if model.finalized:
                    raise RuntimeError("cannot use operators after model finalization")
                done, failure, m = model._h.run(ts.dt, n)
                ts.workspace = "device"
                for _ in range(done):
                    simulation.clock.iteration += 1
                    simulation.clock.time += ts.dt
                if failure:
                    _raise_termination(failure == 1, m)
                if mgr is not None and mgr.due(simulation.clock):
                    mgr.record(model, simulation.clock)
                if report and simulation.clock.iteration % report == 0:
                    _report_progress(simulation, detailed)
        else:
            while True:
                it = simulation.clock.iteration
                if simulation.stop_iteration is not None and it >= simulation.stop_iteration:
                    break
                if not simulation._before_stop_time():
                    break
                step_(model, ts)
                if simulation.check_every and (it + 1) % simulation.check_every == 0:
                    check_termination_(simulation)
                simulation.clock.iteration += 1
                simulation.clock.time += ts.dt
                if mgr is not None and mgr.due(simulation.clock):
                    mgr.record(model, simulation.clock)
                if report and simulation.clock.iteration % report == 0:
                    _report_progress(simulation, detailed)
        simulation.state = SimulationState.Stopped
    except Exception:
        simulation.state = SimulationState.Failed
        raise
    return simulation


def finalize_simulation_(simulation: Simulation):
    if simulation.state is not SimulationState.Finalized:
        simulation.model.finalize()
        simulation.state = SimulationState.Finalized
    return simulation
