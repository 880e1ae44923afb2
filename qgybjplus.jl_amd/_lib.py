"""ctypes binding of ``libqgybj.so`` (the C ABI in ``include/qgybj.h``).

Loading fails loudly: there is no CPU fallback for any operation of this
package.  Set ``QGYBJ_LIB`` to load a library from another path.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QGYBJ_LIB", os.path.join(_HERE, "libqgybj.so"))

QG_OK = 0
QG_ERR_ARGUMENT = 1
QG_ERR_DEVICE = 2
QG_ERR_SINGULAR = 3
QG_ERR_UNSUPPORTED = 4
QG_ERR_GTSV = 5

FIELD_IDS = {"q": 0, "B": 1, "psi": 2, "A": 3, "C": 4, "u": 5, "v": 6, "w": 7}
COMPLEX_FIELDS = ("q", "B", "psi", "A", "C")
REAL_FIELDS = ("u", "v", "w")

YBJ_PLUS, PASSIVE, YBJ = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
SPACE_SPECTRAL, SPACE_PHYSICAL = 0, 1

#: every symbol include/qgybj.h declares
EXPORTED = (
    "qg_create", "qg_destroy", "qg_last_error", "qg_set_field", "qg_get_field",
    "qg_step", "qg_run", "qg_run_noalloc", "qg_check", "qg_progress_maxima", "qg_invert_q_to_psi", "qg_invert_B_to_A",
    "qg_compute_velocities", "qg_compute_q_from_psi", "qg_compute_qw", "qg_fft",
    "qg_synchronize", "qg_take_mean_q_warning", "qg_profile_begin", "qg_profile_end",
    "qg_set_field_member", "qg_get_field_member", "qg_check_member",
    "qg_compute_q_from_psi_member", "qg_invert_B_to_A_member", "qg_use_graphs", "qg_set_pruning",
    "qg_set_member_stratification", "qg_memory_required", "qg_device_memory",
    "qg_set_active_members", "qg_move_member", "qg_autotune", "qg_get_tuning", "qg_set_tuning",
    "qg_set_random_psi", "qg_set_surface_wave", "qg_energies",
)


class QgConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("f", C.c_double), ("dz", C.c_double),
        ("kx", C.POINTER(C.c_double)), ("ky", C.POINTER(C.c_double)),
        ("N2", C.POINTER(C.c_double)), ("a_ell", C.POINTER(C.c_double)),
        ("n_flow_terms", C.c_int), ("flow_coefficients", C.POINTER(C.c_double)),
        ("flow_orders", C.POINTER(C.c_int)),
        ("n_wave_terms", C.c_int), ("wave_coefficients", C.POINTER(C.c_double)),
        ("wave_orders", C.POINTER(C.c_int)),
        ("nu_z", C.c_double),
        ("fixed_flow", C.c_int), ("wave_feedback", C.c_int), ("formulation", C.c_int),
        ("inviscid", C.c_int), ("linear", C.c_int), ("no_dispersion", C.c_int),
        ("precision", C.c_int),
        ("members", C.c_int),
    ]


# tunable passes, in the order of qg_get_tuning / qg_set_tuning
TUNE_PASS_NAMES = ("ya_qw", "xa_qw", "ya2_qeff", "yb_adv", "x_c2r", "xb_q", "xb_B", "yb2_q", "yb2_B",
                   "yf_omega", "xf", "y_r2c", "y_herm", "thomas_psi", "thomas_A", "thomas_AC", "omega",
                   "thomas_pair")
TUNE_PASSES = len(TUNE_PASS_NAMES)  # QG_TUNE_PASSES
TUNE_FORCE = 1                      # QG_TUNE_FORCE
TUNE_IN_STEP = 2                    # QG_TUNE_IN_STEP


class QgError(RuntimeError):
    """A failing status returned through the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load the HIP library once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"HIP library not found at {LIB_PATH}; build it with `make -C qgybjplus.jl_amd` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "qg_create": (C.c_int, [C.POINTER(QgConfig), C.POINTER(P)]),
        "qg_destroy": (C.c_int, [P]),
        "qg_last_error": (C.c_char_p, [P]),
        "qg_set_field": (C.c_int, [P, C.c_int, dp, C.c_int]),
        "qg_get_field": (C.c_int, [P, C.c_int, dp]),
        "qg_step": (C.c_int, [P, C.c_double, C.c_int]),
        "qg_check": (C.c_int, [P, C.POINTER(C.c_int), dp]),
        "qg_run": (C.c_int, [P, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
        "qg_run_noalloc": (C.c_int, [P, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
        "qg_progress_maxima": (C.c_int, [P, dp, dp]),
        "qg_set_field_member": (C.c_int, [P, C.c_int, C.c_int, dp, C.c_int]),
        "qg_get_field_member": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "qg_check_member": (C.c_int, [P, C.c_int, C.POINTER(C.c_int), dp]),
        "qg_compute_q_from_psi_member": (C.c_int, [P, C.c_int, C.c_int]),
        "qg_invert_B_to_A_member": (C.c_int, [P, C.c_int]),
        "qg_use_graphs": (C.c_int, [P, C.c_int]),
        "qg_set_pruning": (C.c_int, [P, C.c_int]),
        "qg_set_active_members": (C.c_int, [P, C.c_int]),
        "qg_move_member": (C.c_int, [P, C.c_int, C.c_int]),
        "qg_set_member_stratification": (C.c_int, [P, C.c_int, dp, dp]),
        "qg_memory_required": (C.c_int, [C.POINTER(QgConfig), C.POINTER(C.c_size_t)]),
        "qg_device_memory": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "qg_invert_q_to_psi": (C.c_int, [P]),
        "qg_invert_B_to_A": (C.c_int, [P]),
        "qg_compute_velocities": (C.c_int, [P, C.c_int]),
        "qg_compute_q_from_psi": (C.c_int, [P, C.c_int]),
        "qg_compute_qw": (C.c_int, [P, dp]),
        "qg_fft": (C.c_int, [P, dp, dp, C.c_int]),
        "qg_synchronize": (C.c_int, [P]),
        "qg_take_mean_q_warning": (C.c_int, [P, dp]),
        "qg_profile_begin": (C.c_int, [P, C.c_int]),
        "qg_profile_end": (C.c_int, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "qg_autotune": (C.c_int, [P, C.c_double, C.c_int, C.c_int]),
        "qg_set_random_psi": (C.c_int, [P, C.c_int, C.c_double, C.c_double, C.c_longlong]),
        "qg_set_surface_wave": (C.c_int, [P, C.c_int, dp]),
        "qg_energies": (C.c_int, [P, C.c_int, dp]),
        "qg_get_tuning": (C.c_int, [P, C.POINTER(C.c_int), C.c_int]),
        "qg_set_tuning": (C.c_int, [P, C.POINTER(C.c_int), C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def device_memory():
    """(free, total) bytes of the current device."""
    lib = load()
    f, t = C.c_size_t(0), C.c_size_t(0)
    if lib.qg_device_memory(C.byref(f), C.byref(t)) != QG_OK:
        raise QgError(QG_ERR_DEVICE, "hipMemGetInfo failed")
    return int(f.value), int(t.value)


def memory_required(*, nx, ny, nz, precision, members=1) -> int:
    """Device bytes a model (or ensemble) of this shape allocates."""
    lib = load()
    cfg = QgConfig()
    cfg.nx, cfg.ny, cfg.nz, cfg.precision, cfg.members = int(nx), int(ny), int(nz), int(precision), int(members)
    b = C.c_size_t(0)
    if lib.qg_memory_required(C.byref(cfg), C.byref(b)) != QG_OK:
        raise ValueError("invalid shape")
    return int(b.value)


class Handle:
    """Owns one ``qg_handle`` (one model resident on the current GPU)."""

    def __init__(self, *, nx, ny, nz, f, dz, kx, ky, N2, a_ell, flow_terms, wave_terms,
                 nu_z, fixed_flow, wave_feedback, formulation, inviscid, linear,
                 no_dispersion, precision, members: int = 1):
        lib = load()
        self._lib = lib
        self._keep = []

        def arr(x, dt=np.float64):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt))
            self._keep.append(a)
            return a

        kx_, ky_, N2_, a_ = arr(kx), arr(ky), arr(N2), arr(a_ell)
        fc = arr([c for c, _ in flow_terms] or [0.0])
        fo = arr([o for _, o in flow_terms] or [0], np.int32)
        wc = arr([c for c, _ in wave_terms] or [0.0])
        wo = arr([o for _, o in wave_terms] or [0], np.int32)
        cfg = QgConfig(
            nx, ny, nz, float(f), float(dz), _dptr(kx_), _dptr(ky_), _dptr(N2_), _dptr(a_),
            len(flow_terms), _dptr(fc), _iptr(fo), len(wave_terms), _dptr(wc), _iptr(wo),
            float(nu_z), int(bool(fixed_flow)), int(bool(wave_feedback)), int(formulation),
            int(bool(inviscid)), int(bool(linear)), int(bool(no_dispersion)), int(precision), int(members))
        h = C.c_void_p()
        rc = lib.qg_create(C.byref(cfg), C.byref(h))
        if rc != QG_OK:
            msg = lib.qg_last_error(None).decode()
            raise _exception(rc, msg)
        self.h = h
        self.members = int(members)
        self.member_shape = (nz, nx, ny)
        self.shape = (nz * self.members, nx, ny)

    def _check(self, rc: int):
        if rc != QG_OK:
            raise _exception(rc, self._lib.qg_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value is not None:
            self._lib.qg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- fields ---------------------------------------------------------------
    def set_field(self, name: str, values, space: int = SPACE_SPECTRAL):
        fid = FIELD_IDS[name]
        if name in COMPLEX_FIELDS:
            a = np.asfortranarray(np.asarray(values, dtype=np.complex128))
        else:
            a = np.asfortranarray(np.asarray(values, dtype=np.float64))
        if a.shape != self.shape:
            raise ValueError(f"field has shape {a.shape}; expected {self.shape} in (z, x, y) layout")
        self._check(self._lib.qg_set_field(self.h, fid, a.ctypes.data_as(C.POINTER(C.c_double)), space))

    def get_field(self, name: str) -> np.ndarray:
        fid = FIELD_IDS[name]
        dt = np.complex128 if name in COMPLEX_FIELDS else np.float64
        out = np.empty(self.shape, dtype=dt, order="F")
        self._check(self._lib.qg_get_field(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_double))))
        return np.ascontiguousarray(out)

    # -- one ensemble member -----------------------------------------------------
    def set_field_member(self, member: int, name: str, values, space: int = SPACE_SPECTRAL):
        fid = FIELD_IDS[name]
        dt = np.complex128 if name in COMPLEX_FIELDS else np.float64
        a = np.asfortranarray(np.asarray(values, dtype=dt))
        if a.shape != self.member_shape:
            raise ValueError(f"field has shape {a.shape}; expected {self.member_shape} in (z, x, y) layout")
        self._check(self._lib.qg_set_field_member(self.h, int(member), fid,
                                                  a.ctypes.data_as(C.POINTER(C.c_double)), space))

    def get_field_member(self, member: int, name: str) -> np.ndarray:
        fid = FIELD_IDS[name]
        dt = np.complex128 if name in COMPLEX_FIELDS else np.float64
        out = np.empty(self.member_shape, dtype=dt, order="F")
        self._check(self._lib.qg_get_field_member(self.h, int(member), fid,
                                                  out.ctypes.data_as(C.POINTER(C.c_double))))
        return np.ascontiguousarray(out)

    def check_member(self, member: int):
        bad = C.c_int(0)
        m = C.c_double(0.0)
        self._check(self._lib.qg_check_member(self.h, int(member), C.byref(bad), C.byref(m)))
        return bool(bad.value), float(m.value)

    def compute_q_from_psi_member(self, member: int, barotropic: bool = False):
        self._check(self._lib.qg_compute_q_from_psi_member(self.h, int(member), 1 if barotropic else 0))

    def invert_B_to_A_member(self, member: int):
        self._check(self._lib.qg_invert_B_to_A_member(self.h, int(member)))

    def set_member_stratification(self, member: int, N2, a_ell):
        n2 = np.ascontiguousarray(np.asarray(N2, dtype=np.float64))
        a = np.ascontiguousarray(np.asarray(a_ell, dtype=np.float64))
        if n2.shape != (self.member_shape[0],) or a.shape != n2.shape:
            raise ValueError("N2 and a_ell need one entry per level")
        self._check(self._lib.qg_set_member_stratification(self.h, int(member), _dptr(n2), _dptr(a)))

    # -- operators --------------------------------------------------------------
    def set_active_members(self, active: int):
        self._check(self._lib.qg_set_active_members(self.h, int(active)))

    def move_member(self, src: int, dst: int):
        self._check(self._lib.qg_move_member(self.h, int(src), int(dst)))

    def use_graphs(self, enable: bool = True):
        self._check(self._lib.qg_use_graphs(self.h, int(bool(enable))))

    def set_pruning(self, enable: bool = True):
        """qg_set_pruning: steps pruned to the dealiased disk (default) or not."""
        self._check(self._lib.qg_set_pruning(self.h, int(bool(enable))))

    def step(self, dt: float, nsteps: int = 1):
        self._check(self._lib.qg_step(self.h, float(dt), int(nsteps)))

    def run(self, dt: float, nsteps: int):
        """qg_run: up to `nsteps` steps, each checked on the device; returns
        (steps taken, failure code 0/1/2, max |ifft(psi)|)."""
        done = C.c_int(0)
        fail = C.c_int(0)
        m = C.c_double(0.0)
        self._check(self._lib.qg_run(self.h, float(dt), int(nsteps), C.byref(done), C.byref(fail), C.byref(m)))
        return int(done.value), int(fail.value), float(m.value)

    def run_noalloc(self, dt: float, nsteps: int):
        """qg_run_noalloc: qg_run with its snapshot in the library's static
        device scratch (no device allocation; eager steps); same returns."""
        done = C.c_int(0)
        fail = C.c_int(0)
        m = C.c_double(0.0)
        self._check(self._lib.qg_run_noalloc(self.h, float(dt), int(nsteps), C.byref(done), C.byref(fail),
                                             C.byref(m)))
        return int(done.value), int(fail.value), float(m.value)

    def progress_maxima(self):
        """qg_progress_maxima: (max wave speed |LA|, max flow speed hypot(u, v))."""
        w = C.c_double(0.0)
        f = C.c_double(0.0)
        self._check(self._lib.qg_progress_maxima(self.h, C.byref(w), C.byref(f)))
        return float(w.value), float(f.value)

    def check(self):
        bad = C.c_int(0)
        m = C.c_double(0.0)
        self._check(self._lib.qg_check(self.h, C.byref(bad), C.byref(m)))
        return bool(bad.value), float(m.value)

    def invert_q_to_psi(self):
        self._check(self._lib.qg_invert_q_to_psi(self.h))

    def invert_B_to_A(self):
        self._check(self._lib.qg_invert_B_to_A(self.h))

    def compute_velocities(self, compute_w: bool = True):
        self._check(self._lib.qg_compute_velocities(self.h, int(bool(compute_w))))

    def compute_q_from_psi(self, barotropic: bool = False):
        self._check(self._lib.qg_compute_q_from_psi(self.h, 1 if barotropic else 0))

    def compute_qw(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.complex128, order="F")
        self._check(self._lib.qg_compute_qw(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return np.ascontiguousarray(out)

    def fft(self, a, direction: int) -> np.ndarray:
        src = np.asfortranarray(np.asarray(a, dtype=np.complex128))
        if src.shape != self.shape:
            raise ValueError("array shape does not match the model grid")
        out = np.empty(self.shape, dtype=np.complex128, order="F")
        self._check(self._lib.qg_fft(self.h, src.ctypes.data_as(C.POINTER(C.c_double)),
                                     out.ctypes.data_as(C.POINTER(C.c_double)), int(direction)))
        return np.ascontiguousarray(out)

    def synchronize(self):
        self._check(self._lib.qg_synchronize(self.h))

    def profile_begin(self, kernels: bool = True, steps: bool = True):
        self._check(self._lib.qg_profile_begin(self.h, (1 if kernels else 0) | (2 if steps else 0)))

    def profile_end(self) -> dict:
        import json
        need = C.c_size_t(0)
        buf = C.create_string_buffer(1 << 16)
        rc = self._lib.qg_profile_end(self.h, buf, len(buf), C.byref(need))
        if rc != QG_OK and need.value > len(buf):
            raise RuntimeError("profile summary larger than the buffer")
        self._check(rc)
        return json.loads(buf.value.decode())

    # -- device-side initial conditions -----------------------------------------
    def set_random_psi(self, amplitude: float, slope: float, seed: int, member: int = -1):
        self._check(self._lib.qg_set_random_psi(self.h, int(member), float(amplitude), float(slope), int(seed)))

    def set_surface_wave(self, profile, member: int = -1):
        p = np.ascontiguousarray(profile, dtype=np.float64)
        self._check(self._lib.qg_set_surface_wave(self.h, int(member), _dptr(p)))

    def energies(self, member: int = -1) -> np.ndarray:
        """(members or 1, 5): wave KE, PE, CE, mean-flow KE, PE."""
        n = self.members if member == -1 else 1
        out = np.zeros(5 * n)
        self._check(self._lib.qg_energies(self.h, int(member), _dptr(out)))
        return out.reshape(n, 5)

    # -- autotuner ---------------------------------------------------------------
    def autotune(self, dt: float, steps: int = 3, force: bool = False, in_step: bool = False) -> dict:
        """Time the tile variants / launch shapes of every tunable pass on this
        model's state and keep the fastest (cached per grid class); in_step:
        also refine the passes that take most of the step by whole-step times."""
        flags = (TUNE_FORCE if force else 0) | (TUNE_IN_STEP if in_step else 0)
        self._check(self._lib.qg_autotune(self.h, float(dt), int(steps), flags))
        return self.tuning()

    def tuning(self) -> dict:
        t = np.zeros(TUNE_PASSES, dtype=np.int32)
        self._check(self._lib.qg_get_tuning(self.h, _iptr(t), TUNE_PASSES))
        return dict(zip(TUNE_PASS_NAMES, (int(v) for v in t)))

    def set_tuning(self, choices: dict):
        t = np.array([int(choices.get(n, -1)) for n in TUNE_PASS_NAMES], dtype=np.int32)
        self._check(self._lib.qg_set_tuning(self.h, _iptr(t), TUNE_PASSES))

    def take_mean_q_warning(self) -> float:
        m = C.c_double(0.0)
        self._check(self._lib.qg_take_mean_q_warning(self.h, C.byref(m)))
        return float(m.value)


def _exception(code: int, msg: str) -> Exception:
    if code == QG_ERR_ARGUMENT:
        return ValueError(msg)
    if code == QG_ERR_UNSUPPORTED:
        return NotImplementedError(msg)
    return QgError(code, msg)
