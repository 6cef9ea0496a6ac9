"""Simulation parameters: constants, ``SimulationConfig`` and presets.

Same field names, defaults and validation rules as the reference
(lfm/constants.py:27-154, lfm/config.py:36-347, lfm/config_presets.py:29-109)
so a reference config can be rebuilt here field for field.  Fields the
reference's Evolver and backends ignore (``chi_mode``, ``c``, the phase-1 and
cosmology metadata) are accepted and ignored here too.
"""
from __future__ import annotations

import enum
import math
import warnings
from dataclasses import dataclass, field

# -- constants, all derived from chi0 = 3^D - 2^D with D = 3 ---------------
D = 3
D_ST = D + 1
CHI0 = float(3 ** D - 2 ** D)                    # 19
N_COLORS = 3
KAPPA = 1.0 / (4 ** D - 1)                       # 1/63
LAMBDA_H = D_ST / (2 * D_ST ** 2 - 1)            # 4/31
EPSILON_W = 2.0 / (CHI0 + 1)                     # 0.1
ALPHA_S = 2.0 / (CHI0 - 2)                       # 2/17
KAPPA_C = KAPPA / N_COLORS                       # 1/189
EPSILON_CC = ALPHA_S
BETA_0 = int(CHI0) - 12                          # 7
SA_GAMMA = EPSILON_W                             # 0.1
SA_L = BETA_0
SA_D = SA_GAMMA * SA_L ** 2                      # 4.9
Z2_COORD = 2 * D_ST ** 2                         # 32
RANK_SU3 = N_COLORS - 1
KAPPA_TUBE = (Z2_COORD - RANK_SU3) * KAPPA       # 30/63
KAPPA_STRING = KAPPA_C
DT_DEFAULT = 0.02
C_DEFAULT = 1.0
E0_SQ_DEFAULT = 0.0
BOUNDARY_FRACTION_DEFAULT = 0.3
STENCIL_FACE_WEIGHT = 1.0 / 3.0
STENCIL_EDGE_WEIGHT = 1.0 / 6.0
STENCIL_CENTER_WEIGHT = -4.0
CFL_19PT_MASSIVE = 1.0 / math.sqrt(16.0 / 3.0 + CHI0 ** 2)
WELL_THRESHOLD = 17.0
VOID_THRESHOLD = 18.0
E_AMPLITUDE_BY_GRID = {32: 12.0, 64: 6.0, 128: 3.6, 256: 1.8, 512: 0.9}
# age of the universe in Planck times (metadata only; lfm/constants.py:254)
OBSERVABLE_RADIUS_PLANCK = (13.787 * 1e9 * 365.25 * 86400.0) / 5.391e-44


class FieldLevel(enum.IntEnum):
    REAL = 0
    COMPLEX = 1
    COLOR = 2


class ChiMode(enum.Enum):
    WAVE = "wave"
    MEMORY = "memory"
    STATIC = "static"


class BoundaryType(enum.Enum):
    FROZEN = "frozen"
    PERIODIC = "periodic"
    ABSORBING = "absorbing"


class PhysicsScale(enum.Enum):
    PLANCK = "planck"
    NUCLEAR = "nuclear"
    ATOMIC = "atomic"
    MOLECULAR = "molecular"
    CLASSICAL = "classical"
    SOLAR = "solar"
    GALACTIC = "galactic"
    COSMIC = "cosmic"


# (field level, chi mode, lambda_self, kappa_c) per scale (lfm/config.py:120-129)
_SCALE_DEFAULTS = {
    PhysicsScale.PLANCK: (FieldLevel.COLOR, ChiMode.WAVE, LAMBDA_H, KAPPA_C),
    PhysicsScale.NUCLEAR: (FieldLevel.COLOR, ChiMode.WAVE, LAMBDA_H, KAPPA_C),
    PhysicsScale.ATOMIC: (FieldLevel.COMPLEX, ChiMode.WAVE, 0.0, 0.0),
    PhysicsScale.MOLECULAR: (FieldLevel.COMPLEX, ChiMode.WAVE, 0.0, 0.0),
    PhysicsScale.CLASSICAL: (FieldLevel.REAL, ChiMode.WAVE, 0.0, 0.0),
    PhysicsScale.SOLAR: (FieldLevel.REAL, ChiMode.WAVE, 0.0, 0.0),
    PhysicsScale.GALACTIC: (FieldLevel.REAL, ChiMode.MEMORY, 0.0, 0.0),
    PhysicsScale.COSMIC: (FieldLevel.REAL, ChiMode.MEMORY, 0.0, 0.0),
}


@dataclass
class SimulationConfig:
    """All physics and numerics parameters of a run (lfm/config.py:145-347)."""

    grid_size: int = 128
    chi0: float = CHI0
    kappa: float = KAPPA
    lambda_self: float = 0.0
    epsilon_w: float = EPSILON_W
    kappa_c: float = 0.0
    epsilon_cc: float = 0.0
    kappa_string: float = 0.0
    kappa_tube: float = 0.0
    sa_gamma: float = SA_GAMMA
    sa_d: float = SA_D
    e0_sq: float = E0_SQ_DEFAULT
    dt: float = DT_DEFAULT
    c: float = C_DEFAULT
    field_level: FieldLevel = FieldLevel.REAL
    n_colors: int = N_COLORS
    boundary_type: BoundaryType = BoundaryType.FROZEN
    boundary_fraction: float = BOUNDARY_FRACTION_DEFAULT
    e_amplitude: float = 0.0
    blob_sigma_factor: float = 12.0
    phase1_steps: int = 10_000
    phase1_omega: float = 38.0
    phase1_amplitude: float = 0.3
    random_seed: int = 42
    report_interval: int = 5000
    box_planck_radius: float = OBSERVABLE_RADIUS_PLANCK
    physical_scale: PhysicsScale | None = None
    chi_mode: ChiMode = ChiMode.WAVE
    dx: float = field(init=False, default=1.0)
    sigma: float = field(init=False, default=0.0)

    def __post_init__(self) -> None:
        if self.physical_scale is not None:
            (self.field_level, self.chi_mode, self.lambda_self,
             self.kappa_c) = _SCALE_DEFAULTS[self.physical_scale]
        self._validate()
        self.dx = 1.0
        self.sigma = self.grid_size / self.blob_sigma_factor
        if self.e_amplitude == 0.0:
            self.e_amplitude = E_AMPLITUDE_BY_GRID.get(self.grid_size, 3.6)

    def _validate(self) -> None:
        if self.grid_size < 8:
            raise ValueError(f"grid_size must be >= 8, got {self.grid_size}")
        if self.dt <= 0:
            raise ValueError(f"dt must be positive, got {self.dt}")
        cfl = 1.0 / math.sqrt(16.0 / 3.0 + self.chi0 ** 2)
        if self.dt > cfl:
            raise ValueError(f"dt={self.dt} exceeds CFL limit {cfl:.4f} for the 19-point "
                             f"stencil with chi0={self.chi0}")
        if self.chi0 <= 0:
            raise ValueError(f"chi0 must be > 0, got {self.chi0}")
        if self.kappa < 0:
            raise ValueError(f"kappa must be >= 0, got {self.kappa}")
        if self.lambda_self < 0:
            raise ValueError(f"lambda_self must be >= 0, got {self.lambda_self}")
        if not 0 < self.boundary_fraction < 1:
            raise ValueError(f"boundary_fraction must be in (0, 1), got {self.boundary_fraction}")
        if self.field_level == FieldLevel.COLOR and self.n_colors < 1:
            raise ValueError(f"n_colors must be >= 1, got {self.n_colors}")
        if self.physical_scale is not None:
            expected = _SCALE_DEFAULTS[self.physical_scale][0]
            if self.field_level != expected:
                warnings.warn(f"physical_scale={self.physical_scale.value!r} expects "
                              f"FieldLevel.{expected.name}", stacklevel=3)

    @property
    def sa_enabled(self) -> bool:
        return self.kappa_tube > 0.0

    @classmethod
    def for_scale(cls, scale: PhysicsScale, grid_size: int = 128, **kw) -> "SimulationConfig":
        return cls(grid_size=grid_size, physical_scale=scale, **kw)


def gravity_only(grid_size: int = 64, **overrides) -> SimulationConfig:
    """REAL level, gravity only (lfm/config_presets.py:29-40)."""
    p = dict(grid_size=grid_size, field_level=FieldLevel.REAL, boundary_type=BoundaryType.FROZEN)
    p.update(overrides)
    return SimulationConfig(**p)


def gravity_em(grid_size: int = 64, **overrides) -> SimulationConfig:
    """COMPLEX level with the Mexican hat on (lfm/config_presets.py:43-56)."""
    p = dict(grid_size=grid_size, field_level=FieldLevel.COMPLEX,
             boundary_type=BoundaryType.FROZEN, lambda_self=LAMBDA_H)
    p.update(overrides)
    return SimulationConfig(**p)


def full_physics(grid_size: int = 64, **overrides) -> SimulationConfig:
    """COLOR level, every coupling on (lfm/config_presets.py:59-79)."""
    p = dict(grid_size=grid_size, field_level=FieldLevel.COLOR,
             boundary_type=BoundaryType.FROZEN, lambda_self=LAMBDA_H, kappa_c=KAPPA_C,
             epsilon_cc=EPSILON_CC, kappa_string=KAPPA_STRING, kappa_tube=KAPPA_TUBE,
             sa_gamma=SA_GAMMA, sa_d=SA_D)
    p.update(overrides)
    return SimulationConfig(**p)


def spinor_field(grid_size: int = 64, **overrides) -> SimulationConfig:
    """Two-component COLOR preset (lfm/config_presets.py:82-109).  The step
    kernels are built for three colours, as the reference's are, so an
    Evolver refuses this config with a clear error."""
    p = dict(grid_size=grid_size, field_level=FieldLevel.COLOR, n_colors=2,
             boundary_type=BoundaryType.FROZEN, lambda_self=LAMBDA_H)
    p.update(overrides)
    return SimulationConfig(**p)
