"""lfm_hip -- MI355X (gfx950) implementation of lfm-physics' leapfrog time step.

The GOV-01/GOV-02 step runs in hand-written HIP kernels (liblfm_hip.so, C ABI
in include/lfm_hip.h) behind the reference's own interfaces:
``HipBackend`` fills the backend slot ``lfm.core.Evolver`` calls, and
``Evolver`` / ``Simulation`` / ``SimulationConfig`` mirror lfm's API.
"""
from .config import (BoundaryType, ChiMode, FieldLevel, PhysicsScale, SimulationConfig,
                     full_physics, gravity_em, gravity_only, spinor_field)
from .backend import HipBackend
from .evolver import Evolver, get_backend
from .simulation import Simulation
from .batch import LatticeBatch
from . import fields

__all__ = ["BoundaryType", "ChiMode", "FieldLevel", "PhysicsScale", "SimulationConfig",
           "full_physics", "gravity_em", "gravity_only", "spinor_field", "HipBackend",
           "Evolver", "get_backend", "Simulation", "LatticeBatch", "fields"]
__version__ = "0.1.0"
