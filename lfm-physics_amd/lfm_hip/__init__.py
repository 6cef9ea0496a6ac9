"""lfm_hip -- MI355X (gfx950) implementation of lfm-physics' leapfrog time step.

The GOV-01/GOV-02 step runs in hand-written HIP kernels (liblfm_hip.so, C ABI
in include/lfm_hip.h) behind the reference's own interfaces:
``HipBackend`` fills the backend slot ``lfm.core.Evolver`` calls, and
``Evolver`` / ``Simulation`` / ``SimulationConfig`` mirror lfm's API.
"""
from .config import (BoundaryType, ChiMode, FieldLevel, PhysicsScale, SimulationConfig,
                     full_physics, gravity_em, gravity_only, spinor_field)
# the step's constants, under the names lfm exports them (lfm/__init__.py:98-125)
from .config import (ALPHA_S, BETA_0, CHI0, D_ST, DT_DEFAULT, E_AMPLITUDE_BY_GRID, EPSILON_CC,
                     EPSILON_W, KAPPA, KAPPA_C, KAPPA_STRING, KAPPA_TUBE, LAMBDA_H, N_COLORS,
                     RANK_SU3, SA_D, SA_GAMMA)
from .backend import HipBackend
from .evolver import Evolver, get_backend
from .simulation import Simulation
from .batch import LatticeBatch, sweep, sweep_2d
from .experiment import Barrier, ContinuousSource, DetectorScreen, Slit, gpu_snapshot_loop
from . import fields

__all__ = ["BoundaryType", "ChiMode", "FieldLevel", "PhysicsScale", "SimulationConfig",
           "full_physics", "gravity_em", "gravity_only", "spinor_field", "HipBackend",
           "Evolver", "get_backend", "Simulation", "LatticeBatch", "sweep", "sweep_2d",
           "gpu_snapshot_loop", "Barrier", "Slit", "ContinuousSource", "DetectorScreen", "fields",
           "ALPHA_S", "BETA_0", "CHI0", "D_ST", "DT_DEFAULT", "E_AMPLITUDE_BY_GRID", "EPSILON_CC",
           "EPSILON_W", "KAPPA", "KAPPA_C", "KAPPA_STRING", "KAPPA_TUBE", "LAMBDA_H", "N_COLORS",
           "RANK_SU3", "SA_D", "SA_GAMMA"]
__version__ = "0.1.0"
