"""Error paths of the column solves on the GPU, with the reference's
exception semantics:

* a Thomas pivot below sqrt(eps) raises "Singular matrix in Thomas solver"
  (src/elliptic.jl:1177-1193) from q -> psi, B -> A and the step;
* an exactly zero dgtsv pivot of the omega equation raises "... gtsv failed
  for vertical velocity" (src/operators.jl:447-464) from
  compute_velocities!(compute_w = true) and from the step's finalize;
* a horizontal length without a compiled transform plan is refused when the
  model is created (FFTW would take it; the library names what it takes).

Both conditions come from the inputs (the coefficients the host hands over);
every kernel runs to its normal end and the failure is read back on the
host.  The model stays usable after the error.
"""

import math

import numpy as np
import pytest

from oracle import qgybj_oracle as O

pytestmark = pytest.mark.gpu

QG_ERR_SINGULAR = 3
QG_ERR_GTSV = 5


def test_singular_thomas_pivot_raises(qg):
    """a = f^2/N^2 = 1e-12 and kh^2 dz^2 ~ 4e-17 on the longest horizontal
    wave of a 1e9 m box: the first pivot -(a + kh^2 dz^2) of that column is
    far below sqrt(eps) ~ 1.5e-8."""
    print("case singular thomas pivot", flush=True)
    grid = qg.RectilinearGrid(size=(8, 8, 4), extent=(1e9, 1e9, 4.0))
    m = qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), stratification=qg.ConstantStratification(1e4),
                      feedback="none")
    q = np.zeros((4, 8, 8), complex)
    q[:, 1, 0] = 1.0
    m.fields.q = q
    for call in (m.invert_q_to_psi, m.invert_B_to_A,
                 lambda: qg.step_(m, qg.ExponentialRungeKutta2(dt=1.0))):
        with pytest.raises(qg.QgError) as ei:
            call()
        assert ei.value.code == QG_ERR_SINGULAR
        assert str(ei.value).startswith("Singular matrix in Thomas solver"), str(ei.value)
    # still usable: the fields read back and the handle closes cleanly
    assert m.fields.q[1, 1, 0] == 1.0
    m.finalize()


def test_gtsv_zero_pivot_raises(qg):
    """nz = 3 leaves one interior row, d = -2 f^2/(N^2 dz^2) - kh^2.  With
    f = dz = 1 and N^2 = -1 on the middle level (the omega equation takes N^2
    as given; only a_ell clamps it to sqrt(eps)), the column kh^2 = 2 of a
    2 pi box has d = 2 - 2 = 0 exactly: dgtsv's INFO > 0."""
    print("case gtsv zero pivot", flush=True)
    from qgybjplus_amd import _lib
    n, nz = 8, 3
    k = O.wavenumbers(n, 2 * math.pi)
    assert k[1] == 1.0
    N2 = np.array([1.0, -1.0, 1.0])
    a_ell = 1.0 / np.maximum(N2, math.sqrt(np.finfo(float).eps))
    h = _lib.Handle(nx=n, ny=n, nz=nz, f=1.0, dz=1.0, kx=k, ky=k, N2=N2, a_ell=a_ell, flow_terms=[],
                    wave_terms=[], nu_z=0.0, fixed_flow=False, wave_feedback=False, formulation=0,
                    inviscid=True, linear=False, no_dispersion=False, precision=0)
    psi = np.zeros((nz, n, n), complex)
    psi[1, 1, 1] = 1.0
    h.set_field("psi", psi)
    with pytest.raises(_lib.QgError) as ei:
        h.compute_velocities(True)
    assert ei.value.code == QG_ERR_GTSV
    assert "gtsv failed for vertical velocity" in str(ei.value), str(ei.value)
    # u, v without w still work on the same handle: v = i kx psi on the mode
    h.compute_velocities(False)
    v = h.get_field("v")
    assert np.all(np.isfinite(v)) and np.max(np.abs(v)) > 0
    # the step's finalize needs w: the same error
    with pytest.raises(_lib.QgError) as ei:
        h.step(0.1, 1)
    assert ei.value.code == QG_ERR_GTSV
    h.close()


@pytest.mark.parametrize("size", [(14, 16, 4), (16, 21, 4), (1280, 8, 2), (8, 1536, 2), (11, 11, 2)])
def test_unsupported_transform_length_raises(qg, size):
    """Lengths with a prime factor above 5 (14, 21, 11) or above 1024 (1280,
    1536) have no compiled plan: creating the model fails with
    QG_ERR_UNSUPPORTED (NotImplementedError in Python), the message names
    the supported lengths, and a supported grid works."""
    print(f"case unsupported length {size}", flush=True)
    grid = qg.RectilinearGrid(size=size, extent=(500e3, 500e3, 4000.0))
    with pytest.raises(NotImplementedError) as ei:
        qg.QGYBJModel(grid=grid, coriolis=qg.FPlane(1e-4), stratification=qg.ConstantStratification(1e-4))
    msg = str(ei.value)
    assert "1000" in msg and "1024" in msg and "15" in msg, msg
    ok = qg.RectilinearGrid(size=(15, 16, size[2]), extent=(500e3, 500e3, 4000.0))
    m = qg.QGYBJModel(grid=ok, coriolis=qg.FPlane(1e-4), stratification=qg.ConstantStratification(1e-4))
    qg.step_(m, qg.ExponentialRungeKutta2(dt=60.0), 2)
    assert np.all(np.isfinite(m.fields.q))
    m.finalize()
