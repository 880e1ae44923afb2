"""MI355X-native QG-YBJ+ ETD-RK2 step with the QGYBJplus.jl model API.

The numerical path is ``libqgybj.so`` (HIP, gfx950), reached through the C
ABI in ``include/qgybj.h``; this package is the host-side mirror of the
reference's model/simulation interface on top of it.
"""

from ._lib import QgError, load as load_library, EXPORTED, LIB_PATH  # noqa: F401
from .api import (  # noqa: F401
    RectilinearGrid, dealias_mask, FPlane, ConstantStratification, AnalyticalProfile,
    FixedFlow, EvolvingFlow, NoFeedback, WaveMeanFeedback, NoWaveFeedback, YBJPlus, YBJ,
    PassiveWave, Dissipative, Inviscid, NonlinearDynamics, LinearDynamics, Dispersive,
    NoDispersion, FlowHyperdiffusivity, WaveHyperdiffusivity, HorizontalHyperdiffusivity,
    VerticalDiffusivity, SurfaceWave, RandomStreamfunction, FieldArray, GPU, QGYBJModel, QGYBJEnsemble, finalize_model_, set_,
    ExponentialRungeKutta2, step_, Simulation, SimulationState, Clock, run_, IterationInterval, TimeInterval,
    EnergyDiagnostics, ENERGY_NAMES,
    check_termination_, finalize_simulation_,
)
from .state_io import write_state, restore_, read_prognostic  # noqa: F401,E402
from . import synthetic, serving, state_io  # noqa: F401,E402
