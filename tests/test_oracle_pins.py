"""Closed-form pins of the signs in the oracle's step assembly, and a check
that each of them bites.

The RHS assembly (src/timestep.jl:286-295), the feedback substitution q_eff =
q - q_w (src/timestep.jl:122-140), u = -psi_y, v = psi_x
(src/operators.jl:213-220) and the Neumann ends of the vertical diffusion
(src/nonlinear.jl:1010-1044) are each fixed here by a closed form on single
Fourier modes, worked out by hand below, not by restating the oracle's
formula.  ``test_planted_mistake_fails_a_pin`` then applies one-token
mistakes to a copy of the oracle's source (on the CPU, never the shipped
code) and requires every mutant to fail at least one pin, here or in
tests/test_oracle.py.
"""

import inspect
import math
import os
import sys
import types

import numpy as np
import pytest

from oracle import qgybj_oracle as O_REAL

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                          "qgybj_oracle.py")


def _close(got, exp, tol=1e-12):
    scale = max(float(np.max(np.abs(exp))), 1e-300)
    err = float(np.max(np.abs(got - exp))) / scale
    assert err <= tol, err


def _box(O, n, nz, lz=1.0):
    return O.make_grid((n, n, nz), extent=(2 * math.pi, 2 * math.pi, lz))


def pin_rhs_B(O):
    """FixedFlow, psi = sin 2x (z-uniform), B = e^{3iy} (z-uniform), constant a.
    u = -psi_y = 0, v = psi_x = 2 cos 2x, zeta = -kh^2 psi = -4 sin 2x; the
    z-uniform B has A = -4 B / kh^2 = -4 B / 9 (Neumann ends: the vertical
    term vanishes).  So
      -div(u B)          = -d/dy(2 cos 2x e^{3iy})    = -6i cos 2x e^{3iy}
      i (f/2) kh^2 A     = i (f/2) 9 (-4/9) e^{3iy}   = -2i f e^{3iy}
      -(i/2) zeta B      = 2i sin 2x e^{3iy}
    three distinct Fourier modes, so each term's sign is visible; rhs_q = 0
    with a fixed flow."""
    n, nz, f = 16, 4, 0.5
    g = _box(O, n, nz)
    m = O.OracleModel(g, O.Options(f=f, fixed_flow=True, feedback=False,
                                   closure=O.Closure((0.0,), (4,), (0.0,), (4,))), np.full(nz, 2.0))
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    S = O.Fields.zeros(g)
    S.psi = O.fft_forward(np.broadcast_to(np.sin(2 * X), (nz, n, n)).astype(complex))
    S.B = O.fft_forward(np.broadcast_to(np.exp(3j * Y), (nz, n, n)))
    rq, rb = m._rhs(S)
    assert np.max(np.abs(rq)) == 0
    exp = (-6j * np.cos(2 * X) - 2j * f + 2j * np.sin(2 * X)) * np.exp(3j * Y)
    _close(O.fft_backward(rb), np.broadcast_to(exp, (nz, n, n)))


def pin_rhs_q_advection(O):
    """Evolving flow without feedback, q = -4 sin 2x - 9 sin 3y (z-uniform):
    psi = L^-1 q = sin 2x + sin 3y (z-uniform: psi-hat = -q-hat / kh^2),
    u = -3 cos 3y, v = 2 cos 2x, and u.grad q = -3 cos 3y (-8 cos 2x) +
    2 cos 2x (-27 cos 3y) = -30 cos 2x cos 3y, so rhs_q = +30 cos 2x cos 3y."""
    n, nz = 16, 4
    g = _box(O, n, nz)
    m = O.OracleModel(g, O.Options(f=1.0, feedback=False, closure=O.Closure((0.0,), (4,), (0.0,), (4,))),
                      np.full(nz, 4.0))
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    S = O.Fields.zeros(g)
    S.q = O.fft_forward(np.broadcast_to(-4 * np.sin(2 * X) - 9 * np.sin(3 * Y), (nz, n, n)).astype(complex))
    rq, rb = m._rhs(S)
    _close(O.fft_backward(S.psi), np.broadcast_to(np.sin(2 * X) + np.sin(3 * Y), (nz, n, n)))
    _close(O.fft_backward(rq), np.broadcast_to(30 * np.cos(2 * X) * np.cos(3 * Y), (nz, n, n)))
    assert np.max(np.abs(rb)) == 0


def pin_rhs_q_vertical_diffusion(O):
    """q = cos(m pi (k + 1/2) / nz) cos 2x: its own flow (one horizontal mode)
    does not advect it, and the vertical profile is an eigenvector of the
    Neumann second difference (its reflection across either end is itself)
    with eigenvalue -4 sin^2(m pi / (2 nz)) / dz^2, so rhs_q = nu_z times that
    times q, end rows included."""
    n, nz, nu = 16, 8, 0.37
    g = _box(O, n, nz, lz=3.0)
    m = O.OracleModel(g, O.Options(f=1.0, feedback=False, nu_z=nu, closure=O.Closure((0.0,), (4,), (0.0,), (4,))),
                      np.full(nz, 4.0))
    X, _ = np.meshgrid(g.x, g.y, indexing="ij")
    dz = 3.0 / nz
    for mode in (1, 3, 6):
        prof = np.cos(mode * math.pi * (np.arange(nz) + 0.5) / nz)
        phys = prof[:, None, None] * np.cos(2 * X)[None]
        S = O.Fields.zeros(g)
        S.q = O.fft_forward(phys.astype(complex))
        rq, _ = m._rhs(S)
        lam = -4 * math.sin(mode * math.pi / (2 * nz)) ** 2 / dz ** 2
        _close(O.fft_backward(rq), nu * lam * phys)


def pin_q_etd_step_vertical_diffusion(O):
    """Linear dynamics, no horizontal damping: q = the Neumann eigenmode above
    obeys dq/dt = mu q with mu = nu_z lambda.  With E = 1, h phi_1 = h and
    h phi_2 = h/2, one ETD-RK2 step (stage a = q + h mu q, then
    a + (h/2)(mu a - mu q)) multiplies q by 1 + h mu + (h mu)^2 / 2 exactly."""
    n, nz, nu, h = 16, 8, 0.37, 0.05
    g = _box(O, n, nz, lz=3.0)
    m = O.OracleModel(g, O.Options(f=1.0, feedback=False, linear=True, nu_z=nu,
                                   closure=O.Closure((0.0,), (4,), (0.0,), (4,))), np.full(nz, 4.0))
    X, _ = np.meshgrid(g.x, g.y, indexing="ij")
    mode, dz = 3, 3.0 / nz
    phys = np.cos(mode * math.pi * (np.arange(nz) + 0.5) / nz)[:, None, None] * np.cos(2 * X)[None]
    m.fields.q = O.fft_forward(phys.astype(complex))
    mu = nu * (-4 * math.sin(mode * math.pi / (2 * nz)) ** 2 / dz ** 2)
    m.step(h)
    _close(O.fft_backward(m.fields.q), (1 + h * mu + (h * mu) ** 2 / 2) * phys)


def pin_feedback_substitution(O):
    """test/test_wave_feedback.jl:91-96 run through _diagnose_flow: q = 0,
    B = sin x + i sin y (z-uniform) on [-pi, pi)^2, f = 2e-4.  q_w =
    (-cos x cos y + (cos 2x + cos 2y)/2) / f (test_wave_feedback.jl:34-66),
    q_eff = q - q_w = -q_w, and psi = L^-1 q_eff = q_w-hat / kh^2 per mode:
    psi = (-cos x cos y / 2 + (cos 2x + cos 2y) / 8) / f.  q is restored to
    exactly 0 afterwards."""
    n, nz, f = 16, 4, 2e-4
    g = O.make_grid((n, n, nz), x=(-math.pi, math.pi), y=(-math.pi, math.pi), z=(-1.0, 0.0))
    m = O.OracleModel(g, O.Options(f=f, feedback=True), np.full(nz, 1.0))
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    S = O.Fields.zeros(g)
    S.B = O.fft_forward(np.broadcast_to(np.sin(X) + 1j * np.sin(Y), (nz, n, n)))
    m._diagnose_flow(S, compute_w=False)
    assert np.max(np.abs(S.q)) == 0
    exp = (-0.5 * np.cos(X) * np.cos(Y) + (np.cos(2 * X) + np.cos(2 * Y)) / 8) / f
    _close(O.fft_backward(S.psi), np.broadcast_to(exp, (nz, n, n)), tol=1e-11)


def pin_velocities(O):
    """psi = sin 2x + sin 3y: u = -psi_y = -3 cos 3y, v = psi_x = 2 cos 2x."""
    n, nz = 16, 2
    g = _box(O, n, nz)
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    psi = O.fft_forward(np.broadcast_to(np.sin(2 * X) + np.sin(3 * Y), (nz, n, n)).astype(complex))
    u, v, w = O.compute_velocities(psi, g, 1.0, np.ones(nz), O.dealias_mask(n, n), False)
    _close(u, np.broadcast_to(-3 * np.cos(3 * Y), (nz, n, n)))
    _close(v, np.broadcast_to(2 * np.cos(2 * X), (nz, n, n)))
    assert np.max(np.abs(w)) == 0


PINS = (pin_rhs_B, pin_rhs_q_advection, pin_rhs_q_vertical_diffusion, pin_q_etd_step_vertical_diffusion,
        pin_feedback_substitution, pin_velocities)


@pytest.mark.parametrize("pin", PINS, ids=lambda p: p.__name__)
def test_pin(pin):
    pin(O_REAL)


# One-token mistakes planted in a copy of the oracle: (unique source text,
# replacement, what it breaks).  Each must fail a pin.
MUTANTS = (
    ("- 0.5j * rB", "+ 0.5j * rB", "refraction term of rhs_B"),
    ("rb = -nB +", "rb = +nB +", "advection term of rhs_B"),
    ("+ 1j * alpha * kh2 * S.A", "- 1j * alpha * kh2 * S.A", "dispersion term of rhs_B"),
    ("else -nq + dq", "else +nq + dq", "advection term of rhs_q"),
    ("else -nq + dq", "else -nq - dq", "diffusion term of rhs_q"),
    ("else -nq + dq", "else -nq", "dropped diffusion term"),
    ("q_base - qw", "q_base + qw", "feedback substitution q - q_w"),
    ("uk = -(1j * ky) * psi", "uk = (1j * ky) * psi", "u = -psi_y"),
    ("vk = (1j * kx) * psi", "vk = -(1j * kx) * psi", "v = psi_x"),
    ("out[0] = nu_z * (q[1] - q[0]) * inv", "out[0] = nu_z * (q[1] - 2 * q[0]) * inv",
     "top Neumann row of the diffusion"),
    ("out[nz - 1] = nu_z * (q[nz - 2] - q[nz - 1]) * inv", "out[nz - 1] = nu_z * (q[nz - 2] - 2 * q[nz - 1]) * inv",
     "bottom Neumann row of the diffusion"),
    ("Eq * Sn.q + h1q * rq0\n", "Eq * Sn.q - h1q * rq0\n", "ETD stage of q"),
    ("EB * Sn.B + h1B * rb0, 0.0", "EB * Sn.B - h1B * rb0, 0.0", "ETD stage of B"),
    ("h2q * (rq1 - rq0)", "h2q * (rq1 + rq0)", "ETD final combine of q"),
    ("h2B * (rb1 - rb0)", "h2B * (rb1 + rb0)", "ETD final combine of B"),
    ("rr = 2.0 * (bxr.real * xyr.real - byr.real * xxr.real)",
     "rr = 2.0 * (bxr.real * xyr.real + byr.real * xxr.real)", "omega RHS Jacobian sign"),
    ("psiz[: nz - 1] = (psi[1:] - psi[: nz - 1]) / dz", "psiz[: nz - 1] = (psi[1:] + psi[: nz - 1]) / dz",
     "omega RHS forward difference"),
    ("psiavg[: nz - 1] = 0.5 * (psi[1:] + psi[: nz - 1])", "psiavg[: nz - 1] = 0.5 * (psi[1:] - psi[: nz - 1])",
     "omega RHS two-level average"),
    ("J = np.conj(Bxr) * Byr - np.conj(Byr) * Bxr", "J = np.conj(Bxr) * Byr + np.conj(Byr) * Bxr",
     "q_w Jacobian sign"),
    ("inv_f * (qw - 0.25 * kh2 * temp)", "inv_f * (qw + 0.25 * kh2 * temp)", "q_w |B|^2 term sign"),
    ("zeta = np.where(L[None], -kh2 * psi, 0.0 + 0.0j)", "zeta = np.where(L[None], kh2 * psi, 0.0 + 0.0j)",
     "vorticity sign in refraction"),
)


def _load_mutant(src: str) -> types.ModuleType:
    name = "oracle_mutant"
    mod = types.ModuleType(name)
    mod.__file__ = ORACLE_SRC
    sys.modules[name] = mod  # dataclasses resolve annotations through sys.modules
    try:
        exec(compile(src, "<oracle mutant>", "exec"), mod.__dict__)
    finally:
        del sys.modules[name]
    return mod


def _all_pins():
    """The pins above and every test of tests/test_oracle.py, each run with
    the oracle module it is given in place of the real one."""
    from tests import test_oracle as T
    runs = [(p.__name__, (lambda mod, p=p: p(mod))) for p in PINS]
    for name, fn in inspect.getmembers(T, inspect.isfunction):
        if not name.startswith("test_"):
            continue
        marks = [mk for mk in getattr(fn, "pytestmark", []) if mk.name == "parametrize"]
        arglists = [()]
        if marks:
            arglists = [(v,) for v in marks[0].args[1]]

        def run(mod, fn=fn, arglists=arglists):
            saved = T.O
            T.O = mod
            try:
                for args in arglists:
                    fn(*args)
            finally:
                T.O = saved
        runs.append((name, run))
    return runs


@pytest.mark.parametrize("old,new,what", MUTANTS, ids=[m[2] for m in MUTANTS])
def test_planted_mistake_fails_a_pin(old, new, what):
    with open(ORACLE_SRC) as fh:
        src = fh.read()
    assert src.count(old) == 1, f"mutation site not unique or gone: {old!r}"
    mod = _load_mutant(src.replace(old, new))
    caught = []
    for name, run in _all_pins():
        try:
            with np.errstate(all="ignore"):
                run(mod)
        except Exception:  # an assertion or an error the mistake provokes
            caught.append(name)
    print(f"{what}: caught by {caught}")
    assert caught, f"no pin fails when the oracle has this mistake: {what}"
