"""``Simulation``: the facade over :class:`Evolver` (lfm/simulation.py:47-1565).

Kept: field properties, ``run`` (metrics every ``report_interval``, the
``evolve_chi=False`` freeze), ``run_with_snapshots`` (snapshots and per-step
hooks), the zero-copy ``_native_*`` device views the reference's experiment
hooks write through, ``metrics``, ``total_energy``, S_a access and npz
checkpoints with the reference's keys.  Metrics reductions run in FP64 on
the GPU; only the cluster count (scipy.ndimage.label) pulls |Psi|^2 to the
host, as the reference does.  Initial-condition helpers and experiment
drivers of the reference are outside this package.
"""
from __future__ import annotations

import json
import warnings
from pathlib import Path

import numpy as np
import torch

from .backend import as_device_array
from .config import BoundaryType, FieldLevel, SimulationConfig
from .evolver import Evolver


def psi_sq_f64(pr, pi=None) -> np.ndarray:
    """|Psi|^2 summed over colours in float64 from the float32 fields, as
    compute_metrics forms it for the cluster count (lfm/analysis/metrics.py:
    83-91)."""
    sq = lambda a: (np.sum(a.astype(np.float64) ** 2, axis=0) if a.ndim == 4
                    else a.astype(np.float64) ** 2)
    out = sq(pr)
    if pi is not None:
        out += sq(pi)
    return out


def config_from_checkpoint(data) -> SimulationConfig:
    """The SimulationConfig a checkpoint's ``config_json`` describes, read as
    lfm/simulation.py:1537-1543 reads it: enums from their values, and the
    fields __post_init__ derives (dx, sigma) dropped."""
    from . import config as _c
    cfg = json.loads(str(data["config_json"]))
    cfg["field_level"] = FieldLevel(cfg["field_level"])
    cfg["boundary_type"] = BoundaryType(cfg["boundary_type"])
    for k in ("dx", "sigma"):
        cfg.pop(k, None)
    for k, enum in (("physical_scale", _c.PhysicsScale), ("chi_mode", _c.ChiMode)):
        if cfg.get(k) is not None:
            cfg[k] = enum(cfg[k])
    return SimulationConfig(**cfg)


class Simulation:
    def __init__(self, config: SimulationConfig | None = None, backend="auto",
                 mask_policy: str = "reference", chi_storage: str = "delta") -> None:
        self.config = config if config is not None else SimulationConfig()
        self._evolver = Evolver(self.config, backend=backend, mask_policy=mask_policy,
                                chi_storage=chi_storage)
        self._history: list[dict[str, float]] = []
        # device snapshot of Psi at run() start / last callback (the reference's
        # _psi_r_prev / _psi_i_prev, lfm/simulation.py:1168-1182)
        self._snap_r: torch.Tensor | None = None
        self._snap_i: torch.Tensor | None = None
        self._interior = None
        self._solitons_placed = False
        self._equilibrated = False
        self._velocity_boosts: list = []
        self._poisson = None

    # -- fields ---------------------------------------------------------------
    @property
    def evolver(self) -> Evolver:
        return self._evolver

    @property
    def step(self) -> int:
        return self._evolver.step

    @property
    def history(self) -> list[dict[str, float]]:
        return self._history

    @property
    def chi(self) -> np.ndarray:
        return self._evolver.get_chi()

    @chi.setter
    def chi(self, value) -> None:
        self._evolver.set_chi(value)

    @property
    def psi_real(self) -> np.ndarray:
        return self._evolver.get_psi_real()

    @psi_real.setter
    def psi_real(self, value) -> None:
        self._evolver.set_psi_real(value)

    @property
    def psi_imag(self):
        return self._evolver.get_psi_imag()

    @psi_imag.setter
    def psi_imag(self, value) -> None:
        self._evolver.set_psi_imag(value)

    @property
    def psi_real_prev(self):
        return None if self._snap_r is None else self._snap_r.cpu().numpy().reshape(
            self._evolver.get_psi_real().shape)

    @property
    def psi_imag_prev(self):
        """Psi imaginary part at the last run() start or callback (the
        reference's _psi_i_prev), None at REAL."""
        if self._snap_i is None:
            return None
        return self._snap_i.cpu().numpy().reshape(self._evolver.get_psi_imag().shape)

    @property
    def sa_fields(self):
        return self._evolver.get_sa_fields()

    @sa_fields.setter
    def sa_fields(self, value) -> None:
        self._evolver.set_sa_fields(value)

    @property
    def energy_density(self) -> np.ndarray:
        """|Psi|^2 summed over components, (N, N, N) float32."""
        return self._evolver.get_energy_density()

    # -- zero-copy device views (lfm/simulation.py:209-252) -----------------
    # Live views of the evolver's active buffers (DeviceArray tensors on the
    # GPU, with the reference's .fill / .copy spellings); writes through them
    # are seen by the next step, as the reference's experiment hooks
    # (barriers, sources, detectors) expect.
    def _psi_shape(self):
        N = self.config.grid_size
        return (self._evolver._n_psi, N, N, N) if self._evolver.level == 2 else (N, N, N)

    def _native_psi_real(self):
        """Current Psi real buffer as a device view, (N, N, N) or (3, N, N, N)."""
        return as_device_array(self._evolver.native_psi_real().view(self._psi_shape()))

    def _native_psi_imag(self):
        """Current Psi imaginary buffer as a device view (None at REAL)."""
        pi = self._evolver.native_psi_imag()
        return None if pi is None else as_device_array(pi.view(self._psi_shape()))

    def _native_chi(self):
        """Current chi buffer as an (N, N, N) device view in absolute units.
        The first call switches the evolver to absolute chi storage, the
        reference's representation, so writes of absolute values are right."""
        ev = self._evolver
        ev.use_absolute_chi()
        N = self.config.grid_size
        return as_device_array(ev.native_chi().view(N, N, N))

    def _native_chi_pair(self):
        """The four reference chi buffers as (N, N, N) device views, absolute
        units: (current, other set's current, current prev, other set's prev).
        Writing one state into all four zeroes the chi velocity there, as
        barrier enforcement does (lfm/experiment/barrier.py:295-330)."""
        ev = self._evolver
        ev.use_absolute_chi()
        N = self.config.grid_size
        return tuple(as_device_array(t.view(N, N, N)) for t in ev.native_chi_refs())

    # Experiment components (lfm/simulation.py:984-1121): the barrier, the
    # detector screen and the continuous source, each with the reference's
    # callback form and a fold() that moves its per-step work into the steps.
    def place_barrier(self, axis: int = 2, position: int | None = None,
                      height: float | None = None, thickness: int = 2, slits=None,
                      slit_axis: int | None = None, absorb: bool = True,
                      transit_steps: int = 10):
        """A :class:`~lfm_hip.experiment.Barrier` attached to this simulation
        (applied once now)."""
        from .experiment import Barrier
        return Barrier(self, axis=axis, position=position, height=height, thickness=thickness,
                       slits=slits, slit_axis=slit_axis, absorb=absorb,
                       transit_steps=transit_steps)

    def add_detector(self, axis: int = 2, position: int | None = None,
                     field: str = "energy_density"):
        """A :class:`~lfm_hip.experiment.DetectorScreen` on this simulation."""
        from .experiment import DetectorScreen
        return DetectorScreen(self, axis=axis, position=position, field=field)

    def add_source(self, *, axis: int = 2, position: float = 0.25, omega: float = 2.0,
                   amplitude: float = 2.0, envelope_sigma: float = 0.25, phase: float = 0.0,
                   boost: float = 10.0):
        """A :class:`~lfm_hip.experiment.ContinuousSource` on this simulation."""
        from .experiment import ContinuousSource
        return ContinuousSource(self, axis=axis, position=position, omega=omega,
                                amplitude=amplitude, envelope_sigma=envelope_sigma, phase=phase,
                                boost=boost)

    def _to_device(self, arr):
        """A host array as a flat float32 DeviceArray (the backend's format;
        lfm/simulation.py:253-255)."""
        return self._evolver.backend.from_numpy(np.asarray(arr, np.float32).ravel())

    def get_chi(self):
        return self._evolver.get_chi()

    def get_psi_real(self):
        return self._evolver.get_psi_real()

    def get_psi_imag(self):
        return self._evolver.get_psi_imag()

    def get_energy_density(self):
        return self._evolver.get_energy_density()

    def set_psi_real(self, v):
        self._evolver.set_psi_real(v)

    def set_psi_imag(self, v):
        self._evolver.set_psi_imag(v)

    def set_psi_real_prev(self, v):
        self._evolver.set_psi_real_prev(v)

    def set_psi_imag_prev(self, v):
        self._evolver.set_psi_imag_prev(v)

    def set_chi(self, v):
        self._evolver.set_chi(v)

    def set_psi_real_current(self, v):
        """Only the active current Psi_r buffer (prev untouched), for driven
        sources in step callbacks (lfm/simulation.py:190-197)."""
        self._evolver.set_psi_real_current(v)

    def set_psi_imag_current(self, v):
        self._evolver.set_psi_imag_current(v)

    # -- initial conditions -------------------------------------------------
    def place_soliton(self, position, amplitude=None, sigma=None, phase=0.0, velocity=None):
        """Add a Gaussian soliton to Psi (lfm/simulation.py:495-625).  At the
        COLOR level the single-component soliton is added to every colour
        (NumPy broadcasting in the reference).  A velocity adds the carrier
        phase k = chi0 v / c and a time-shifted prev field."""
        from .fields import gaussian_soliton
        cfg = self.config
        N = cfg.grid_size
        amp = amplitude if amplitude is not None else cfg.e_amplitude
        sig = sigma if sigma is not None else cfg.sigma
        ev = self._evolver
        if velocity is not None:
            import math
            vx, vy, vz = velocity
            speed = math.sqrt(vx * vx + vy * vy + vz * vz)
            if cfg.chi0 * speed / cfg.c > 0.8 * math.pi:
                raise ValueError(f"Velocity |v|={speed:.4f}c gives carrier k={cfg.chi0 * speed / cfg.c:.2f} "
                                 f"rad/cell, exceeding 80% of Nyquist; maximum safe speed is "
                                 f"{0.8 * math.pi * cfg.c / cfg.chi0:.4f}c.")
            kx, ky, kz = (cfg.chi0 * v / cfg.c for v in velocity)
            x = np.arange(N, dtype=np.float32)
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            px, py, pz = position

            def wave(cx, cy, cz, extra):
                r2 = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2
                env = (amp * np.exp(-r2 / (2.0 * sig ** 2))).astype(np.float32)
                ph = (phase + kx * (X - cx) + ky * (Y - cy) + kz * (Z - cz) + extra).astype(np.float32)
                return env, (env * np.cos(ph)).astype(np.float32), (env * np.sin(ph)).astype(np.float32)

            env, pr, pi = wave(px, py, pz, 0.0)
            omega = float(np.sqrt(kx * kx + ky * ky + kz * kz + cfg.chi0 ** 2))
            _, pr_prev, pi_prev = wave(px - vx * cfg.dt, py - vy * cfg.dt, pz - vz * cfg.dt,
                                      omega * cfg.dt)
        else:
            pr, pi = gaussian_soliton(N, position, amp, sig, phase)
        cur_r = ev.get_psi_real()
        ev.set_psi_real(cur_r + pr)
        if velocity is not None:
            ev.set_psi_real_prev(cur_r + pr_prev)
        if cfg.field_level != FieldLevel.REAL:
            cur_i = ev.get_psi_imag()
            ev.set_psi_imag(cur_i + pi)
            if velocity is not None:
                ev.set_psi_imag_prev(cur_i + pi_prev)
        self._solitons_placed = True
        self._equilibrated = False
        if velocity is not None:
            self._velocity_boosts.append((tuple(velocity), env ** 2))

    def place_solitons(self, positions, amplitude=None, sigma=None, phases=None, colors=None):
        """Add several Gaussian solitons (lfm/simulation.py:724-778): at the
        COLOR level soliton s goes to colour colors[s] (round robin by
        default); otherwise all are summed into the single component."""
        from .fields import gaussian_soliton, place_solitons
        cfg = self.config
        N = cfg.grid_size
        amp = amplitude if amplitude is not None else cfg.e_amplitude
        sig = sigma if sigma is not None else cfg.sigma
        if cfg.field_level == FieldLevel.COLOR:
            pr, pi = place_solitons(N, positions, amp, sig, phases=phases, colors=colors,
                                    n_colors=cfg.n_colors)
        else:
            pr = np.zeros((N, N, N), np.float32)
            pi = np.zeros((N, N, N), np.float32)
            for pos, ph in zip(positions, [0.0] * len(positions) if phases is None else phases):
                r, i = gaussian_soliton(N, pos, amp, sig, ph)
                pr += r
                pi += i
        ev = self._evolver
        ev.set_psi_real(ev.get_psi_real() + pr)
        if cfg.field_level != FieldLevel.REAL:
            ev.set_psi_imag(ev.get_psi_imag() + pi)
        self._solitons_placed = True
        self._equilibrated = False

    def equilibrate(self) -> None:
        """Poisson-equilibrate chi from |Psi|^2 (GOV-04) on the GPU
        (lfm/simulation.py:877-963, lfm/fields/equilibrium.py:26-100)."""
        from . import _abi
        from .backend import ScvPlan, stream_handle
        if not self._solitons_placed:
            warnings.warn("equilibrate() called before any solitons were placed -- chi will stay "
                          "flat at chi0.  Call place_soliton() first.", stacklevel=2)
        cfg = self.config
        ev = self._evolver
        if self._poisson is None:
            self._poisson = ScvPlan(cfg.grid_size, 1)
        chi_eq = torch.empty(cfg.grid_size ** 3, dtype=torch.float32, device=ev.device)
        bf = cfg.boundary_fraction if cfg.boundary_type == BoundaryType.FROZEN else -1.0
        pi = ev.native_psi_imag()
        rc = ev._lib.lfm_poisson_chi_f32(self._poisson.handle, ev.native_psi_real().data_ptr(),
                                         pi.data_ptr() if pi is not None else None, ev._n_psi,
                                         float(cfg.kappa), float(cfg.e0_sq), float(cfg.chi0),
                                         float(bf), chi_eq.data_ptr(), stream_handle())
        _abi.check(rc, "lfm_poisson_chi_f32")
        if self._velocity_boosts:
            ce = chi_eq.cpu().numpy().reshape((cfg.grid_size,) * 3)
            ev.set_chi_current(ce)
            g = [np.gradient(ce, axis=a).astype(np.float32) for a in range(3)]
            vdg = np.zeros_like(ce)
            for (vx, vy, vz), rho in self._velocity_boosts:
                vdg += rho * (vx * g[0] + vy * g[1] + vz * g[2])
            total = sum(rho for _, rho in self._velocity_boosts)
            sel = total > 1e-10 * total.max()
            vdg[sel] /= total[sel]
            vdg[~sel] = 0.0
            ev.set_chi_prev(ce + cfg.dt * vdg)
            self._velocity_boosts.clear()
        else:
            ev.set_chi(chi_eq)
        self._equilibrated = True

    def _auto_equilibrate(self) -> None:
        if self._solitons_placed and not self._equilibrated:
            warnings.warn("Solitons were placed but equilibrate() was never called -- "
                          "auto-equilibrating chi now.", stacklevel=3)
            self.equilibrate()

    # -- evolution ------------------------------------------------------------
    def _snapshot(self) -> None:
        ev = self._evolver
        self._snap_r = ev.native_psi_real().clone()
        pi = ev.native_psi_imag()
        self._snap_i = pi.clone() if pi is not None else None

    def run(self, steps: int, callback=None, record_metrics: bool = True,
            evolve_chi: bool = True, device_hook=None) -> None:
        """Advance `steps` steps; a metrics snapshot goes to :attr:`history`
        every report_interval steps (lfm/simulation.py:1123-1188).
        device_hook(evolver, step_tensor) runs after every step as device work
        captured into the replayed graphs (see Evolver.evolve); write
        barriers and sources through the evolver's native views there."""
        self._auto_equilibrate()
        self._snapshot()

        def _cb(ev, step):
            if record_metrics:
                m = self.metrics()
                m["step"] = float(step)
                self._history.append(m)
            if callback is not None:
                callback(self, step)
            self._snapshot()

        self._evolver.evolve(steps, callback=_cb, freeze_chi=not evolve_chi,
                             device_hook=device_hook)

    def run_driven(self, steps: int, chi_forcing, record_metrics: bool = False) -> None:
        """Overwrite chi with chi_forcing(t) (a scalar or an (N, N, N) array or
        device tensor) before every step, t = step * dt (lfm/simulation.py:
        1190-1251); metrics every report_interval steps when record_metrics."""
        ev = self._evolver
        N = self.config.grid_size
        dt = self.config.dt
        base = self.step
        self._equilibrated = True
        self._snapshot()
        for s in range(int(steps)):
            chi = chi_forcing((base + s) * dt)
            if not isinstance(chi, torch.Tensor):
                chi = np.asarray(chi, dtype=np.float32)
                if chi.ndim == 0:
                    chi = np.full((N, N, N), float(chi), dtype=np.float32)
            elif chi.ndim == 0:
                chi = torch.full((N, N, N), float(chi), dtype=torch.float32, device=ev.device)
            ev.set_chi(chi)
            ev.evolve(1)
            self._snapshot()
            if record_metrics and (base + s + 1) % self.config.report_interval == 0:
                m = self.metrics()
                m["step"] = float(base + s + 1)
                self._history.append(m)

    def run_with_snapshots(self, steps: int, snapshot_every: int = 100, fields=None,
                           callback=None, step_callback=None, record_metrics: bool = True,
                           evolve_chi: bool = True, device_hook=None) -> list[dict]:
        """Run and keep host copies of `fields` (any of "chi", "psi_real",
        "psi_imag", "energy_density"; default ["chi"]) every `snapshot_every`
        steps and at the end (lfm/simulation.py:1253-1395).  `callback` fires
        after each batch; `step_callback(sim, step)` fires after every step
        (report_interval is 1 meanwhile, and metrics are then recorded once
        per batch instead of every report_interval).  device_hook: as in
        :meth:`run`, a per-step device-side hook kept inside graph replay."""
        self._auto_equilibrate()
        fields = ["chi"] if fields is None else fields
        snapshots: list[dict] = []
        n_full, remainder = divmod(int(steps), int(snapshot_every))

        def take():
            snap = {"step": self.step}
            if "chi" in fields:
                snap["chi"] = self.chi.copy()
            if "psi_real" in fields:
                snap["psi_real"] = self.psi_real.copy()
            if "psi_imag" in fields and self.psi_imag is not None:
                snap["psi_imag"] = self.psi_imag.copy()
            if "energy_density" in fields:
                snap["energy_density"] = self.energy_density.copy()
            return snap

        batches = [snapshot_every] * n_full + ([remainder] if remainder > 0 else [])
        if step_callback is not None:
            orig = self.config.report_interval
            self.config.report_interval = 1
            try:
                for n in batches:
                    self.run(n, callback=step_callback, record_metrics=False,
                             evolve_chi=evolve_chi, device_hook=device_hook)
                    if callback is not None:
                        callback(self, self.step)
                    snapshots.append(take())
                    if record_metrics:
                        m = self.metrics()
                        m["step"] = float(self.step)
                        self._history.append(m)
            finally:
                self.config.report_interval = orig
        else:
            for n in batches:
                self.run(n, callback=None, record_metrics=record_metrics, evolve_chi=evolve_chi,
                         device_hook=device_hook)
                if callback is not None:
                    callback(self, self.step)
                snapshots.append(take())
        return snapshots

    # -- analysis -------------------------------------------------------------
    def get_interior_mask(self) -> np.ndarray:
        if self._interior is None:
            N = self.config.grid_size
            c = N / 2.0
            x = np.arange(N, dtype=np.float64)
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            d = np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2)
            self._interior = d <= c * (1.0 - self.config.boundary_fraction)
        return self._interior

    def _energy(self):
        ev = self._evolver
        if self._snap_r is None:
            return ev.energy_sums(prev_r=ev.native_psi_real(), prev_i=ev.native_psi_imag())
        return ev.energy_sums(prev_r=self._snap_r, prev_i=self._snap_i)

    def metrics(self, clusters: bool = True) -> dict[str, float]:
        """The reference's compute_metrics keys (lfm/analysis/metrics.py:98-111)."""
        e = self._energy()
        s = self._evolver.chi_stats()
        out = {"energy_kinetic": e["energy_kinetic"], "energy_gradient": e["energy_gradient"],
               "energy_potential": e["energy_potential"], "energy_total": e["energy_total"],
               "chi_min": s["chi_min"], "chi_max": s["chi_max"], "chi_mean": s["chi_mean"],
               "chi_std": s["chi_std"], "well_fraction": s["well_fraction"],
               "void_fraction": s["void_fraction"], "n_clusters": float("nan"),
               "psi_sq_total": e["psi_sq_total"]}
        if clusters:
            out["n_clusters"] = float(self._count_clusters())
        return out

    def _count_clusters(self, percentile: float = 90.0) -> int:
        # lfm/analysis/structure.py:111-145 (host: scipy.ndimage.label)
        from scipy import ndimage
        psq = psi_sq_f64(self._evolver.get_psi_real(), self._evolver.get_psi_imag())
        interior = self.get_interior_mask()
        vals = psq[interior]
        if vals.size == 0:
            return 0
        binary = (psq > np.percentile(vals, percentile)) & interior
        return int(ndimage.label(binary)[1])

    def total_energy(self) -> float:
        return self._energy()["energy_total"]

    # -- checkpoints ----------------------------------------------------------
    def save_checkpoint(self, path) -> None:
        """npz with the reference's keys (lfm/simulation.py:1474-1514) plus the
        true leapfrog prev buffers (psi_real_prev_true, chi_prev_true)."""
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        ev = self._evolver
        cs, cps = ev.get_chi_storage()
        data = {"step": np.int64(self.step), "chi": ev.get_chi(), "psi_real": ev.get_psi_real(),
                "psi_real_prev_true": ev.get_psi_real_prev(), "chi_prev_true": ev.get_chi_prev(),
                "chi_storage": np.array(ev.chi_storage), "chi_stored": cs, "chi_prev_stored": cps}
        if ev.get_psi_imag() is not None:
            data["psi_imag"] = ev.get_psi_imag()
            data["psi_imag_prev_true"] = ev.get_psi_imag_prev()
        if self._snap_r is not None:
            data["psi_real_prev"] = self.psi_real_prev
        if self._snap_i is not None:
            data["psi_imag_prev"] = self._snap_i.cpu().numpy().reshape(ev.get_psi_imag().shape)
        sa = ev.get_sa_fields()
        if sa is not None:
            data["sa_fields"] = sa
        cfg = {k: (v.value if hasattr(v, "value") else v) for k, v in vars(self.config).items()
               if not k.startswith("_")}
        data["config_json"] = np.array(json.dumps(cfg))
        data["history_json"] = np.array(json.dumps(self._history))
        np.savez_compressed(str(path), **data)

    @classmethod
    def load_checkpoint(cls, path, backend="auto", exact: bool = False,
                        **kwargs) -> "Simulation":
        """Restore a checkpoint (lfm/simulation.py:1516-1565).  With
        exact=True and the true prev buffers present the run resumes with
        its time derivatives; otherwise current = prev, as the reference."""
        data = np.load(str(path), allow_pickle=False)
        sim = cls(config_from_checkpoint(data), backend=backend, **kwargs)
        ev = sim._evolver
        ev.set_psi_real(data["psi_real"])
        if "psi_imag" in data:
            ev.set_psi_imag(data["psi_imag"])
        ev.set_chi(data["chi"])
        if exact and "psi_real_prev_true" in data:
            ev.set_psi_real_prev(data["psi_real_prev_true"])
            if "psi_imag_prev_true" in data:
                ev.set_psi_imag_prev(data["psi_imag_prev_true"])
            ev.set_chi_prev(data["chi_prev_true"])
            if "chi_stored" in data and str(data["chi_storage"]) == ev.chi_storage:
                ev.set_chi_storage(data["chi_stored"], data["chi_prev_stored"])
        ev.step = int(data["step"])
        if "sa_fields" in data and ev.sa_enabled:
            ev.set_sa_fields(data["sa_fields"])
        if "psi_real_prev" in data:
            sim._snap_r = ev._dev(data["psi_real_prev"])
        if "psi_imag_prev" in data:
            sim._snap_i = ev._dev(data["psi_imag_prev"])
        sim._history = json.loads(str(data["history_json"]))
        return sim
