"""``gpu_snapshot_loop``: the experiment drivers' run loop
(lfm/experiment/common.py:138-280) over :class:`lfm_hip.Simulation`, and
``ContinuousSource``, the continuous plane-wave source
(lfm/experiment/source.py:80-228) with its per-step injection folded into the
step kernel.

Same arguments, return values and progress lines as the reference.  Fields
stay on the device between the steps; snapshots, movie frames and metrics
are host copies taken at their intervals, the metrics reduced with NumPy
in float32 as the reference reduces them.  A per-step ``step_callback``
(barrier enforcement, a source) runs after every single step, writing
through ``sim._native_*`` views.  Without one, the steps between two
recording points run as one ``Evolver.evolve`` call, replayed from HIP
graphs: the same launches, so the same fields as stepping one at a time.
"""
from __future__ import annotations

import math
import time

import numpy as np


def _movie_frame(sim, names, step: int) -> dict:
    frame: dict = {"step": step}
    for name in names:
        if name == "psi_real":
            frame["psi_real"] = sim.psi_real.copy()
        elif name == "chi":
            frame["chi"] = sim.chi.copy()
        elif name == "energy_density":
            frame["energy_density"] = sim.energy_density.copy()
    return frame


def _next_stop(step: int, total: int, intervals) -> int:
    """The first step after `step` that is a multiple of one of `intervals`
    (or `total`)."""
    stop = total
    for n in intervals:
        if n:
            stop = min(stop, (step // n + 1) * n)
    return stop


def gpu_snapshot_loop(sim, *, total_steps: int, snap_every: int = 200, fields=None,
                      movie_every: int | None = None, movie_fields=None, step_callback=None,
                      verbose: bool = True, label: str = "experiment", evolve_chi: bool = True,
                      metrics_every: int | None = None, device_hook=None):
    """Run `total_steps` leapfrog steps of `sim` and return
    ``(snapshots, metrics, movie_snapshots)``:

    * snapshots every `snap_every` steps of `fields` (default
      ``["energy_density", "chi"]``; also "psi_real", "psi_imag");
    * metrics every `metrics_every` steps: step, chi_min, chi_max,
      energy_total (sum of |Psi|^2) and psi_max (max |Psi_real|);
    * movie frames of `movie_fields` (default ``["psi_real"]``) at step 0
      and every `movie_every` steps.

    Steps are counted from the start of this call.  `step_callback(sim,
    step)` runs after every step; ``evolve_chi=False`` holds chi frozen
    (``Evolver.evolve(freeze_chi=True)``).  Solitons placed without
    ``equilibrate()`` are equilibrated first, as ``Simulation.run`` does.

    `device_hook(evolver, step)` (not in the reference): per-step work that
    queues only device operations, captured with the steps into the replayed
    graphs (see ``Evolver.evolve``); `step` there is the Evolver's own step
    counter as a device tensor."""
    if fields is None:
        fields = ["energy_density", "chi"]
    mfields = movie_fields if movie_fields is not None else ["psi_real"]
    sim._auto_equilibrate()
    evolver = sim.evolver
    snaps: list[dict] = []
    metrics: list[dict] = []
    movie: list[dict] = []
    t0 = time.perf_counter()
    if movie_every:
        movie.append(_movie_frame(sim, mfields, 0))
    progress = max(1, total_steps // 10)
    step = 0
    while step < total_steps:
        if step_callback is None:
            stop = _next_stop(step, total_steps,
                              (snap_every, metrics_every, movie_every, progress if verbose else 0))
        else:
            stop = step + 1
        evolver.evolve(stop - step, freeze_chi=not evolve_chi, device_hook=device_hook)
        step = stop
        if step_callback is not None:
            step_callback(sim, step)

        if step % snap_every == 0:
            snap: dict = {"step": step}
            if "chi" in fields:
                snap["chi"] = sim.chi.copy()
            if "psi_real" in fields:
                snap["psi_real"] = sim.psi_real.copy()
            if "psi_imag" in fields and sim.psi_imag is not None:
                snap["psi_imag"] = sim.psi_imag.copy()
            if "energy_density" in fields:
                snap["energy_density"] = sim.energy_density.copy()
            snaps.append(snap)

        if metrics_every and step % metrics_every == 0:
            chi = sim.chi
            metrics.append({
                "step": step,
                "chi_min": float(np.min(chi)),
                "chi_max": float(np.max(chi)),
                "energy_total": float(np.sum(sim.energy_density)),
                "psi_max": float(np.max(np.abs(sim.psi_real))),
            })

        if movie_every and step % movie_every == 0:
            movie.append(_movie_frame(sim, mfields, step))

        if verbose and step % progress == 0:
            elapsed = time.perf_counter() - t0
            rate = step / elapsed
            remaining = (total_steps - step) / rate
            print(f"  [{label}] step {step}/{total_steps}  "
                  f"({rate:.0f} steps/s, ~{remaining:.0f}s left)", flush=True)

    return snaps, metrics, movie


def _resolve(value, N: int) -> int:
    """Fractional (float < 1: fraction of N, rounded) or absolute cell index,
    clamped to [0, N-1] (lfm/experiment/source.py:57-68)."""
    if isinstance(value, float) and value < 1.0:
        return max(0, min(N - 1, round(value * N)))
    return max(0, min(N - 1, int(value)))


def _resolve_sigma(value, N: int) -> float:
    """Fractional (float < 1: fraction of N) or absolute width in cells
    (lfm/experiment/source.py:71-77)."""
    if isinstance(value, float) and value < 1.0:
        return value * N
    return float(value)


class ContinuousSource:
    """Monochromatic plane-wave source at one lattice plane
    (lfm/experiment/source.py:80-228): after every step, the current Psi on
    the plane `position` across `axis` gets
    ``dt^2 * amplitude * sin(omega * t) * boost * envelope`` (t = step * dt),
    and the imaginary part that amount times sin(phase) when |phase| > 1e-12;
    the envelope is the transverse Gaussian of width `envelope_sigma` about
    N/2.  Positions and widths below 1.0 (as floats) are fractions of N.

    Two ways to run it, with the same results bit for bit:

    * ``step_callback(sim, step)``: the reference's form, a host callback for
      ``Simulation.run_with_snapshots(step_callback=...)`` or
      ``gpu_snapshot_loop``, writing through ``sim._native_psi_real/_imag``
      after each single step;
    * ``fold()``: the injection done by the step kernel as it writes Psi
      (``Evolver.set_plane_source``), so the steps keep running as replayed
      graphs with no per-step work of their own; ``unfold()`` stops it.

    REAL and COMPLEX levels (the reference writes one N^3 component).  The
    dispersion helpers of the reference class (``dispersion``,
    ``transit_steps``, ``recommended_steps``) are not part of this class."""

    def __init__(self, sim, *, axis: int = 2, position: float = 0.25, omega: float = 2.0,
                 amplitude: float = 2.0, envelope_sigma: float = 0.25, phase: float = 0.0,
                 boost: float = 10.0) -> None:
        N = sim.config.grid_size
        self._sim = sim
        self._axis = int(axis)
        if self._axis not in (0, 1, 2):
            raise ValueError(f"axis must be 0, 1 or 2, got {axis}")
        self._position_z = _resolve(position, N)
        self._omega = omega
        self._amplitude = amplitude
        self._phase = phase
        self._boost = boost
        self._dt = sim.config.dt
        self._N = N
        # transverse Gaussian, float32 as the reference builds it
        sigma_cells = _resolve_sigma(envelope_sigma, N)
        centre = N / 2.0
        c0 = (np.arange(N, dtype=np.float32) - centre).reshape(-1, 1)
        c1 = (np.arange(N, dtype=np.float32) - centre).reshape(1, -1)
        self._envelope = np.exp(-(c0 ** 2 + c1 ** 2) / (2.0 * sigma_cells ** 2)).astype(np.float32)
        self._envelope_device = sim._to_device(self._envelope).reshape(N, N)

    @property
    def position(self) -> int:
        """Source plane cell index."""
        return self._position_z

    @property
    def envelope(self) -> np.ndarray:
        """The (N, N) float32 transverse envelope."""
        return self._envelope

    def _injection(self, step: int) -> float:
        t = step * self._dt
        return self._dt ** 2 * self._amplitude * math.sin(self._omega * t) * self._boost

    def amounts(self, step: int) -> tuple[float, float]:
        """(real, imaginary) amounts multiplied into the envelope after the
        step that brings the count to `step`, as float32 values; (0, 0) where
        the reference's callback skips (|injection| < 1e-30)."""
        inj = self._injection(step)
        if abs(inj) < 1e-30:
            return 0.0, 0.0
        imag = 0.0
        if abs(self._phase) > 1e-12 and self._sim.evolver.level >= 1:
            imag = float(np.float32(inj * math.sin(self._phase)))
        return float(np.float32(inj)), imag

    def step_callback(self, sim, step: int) -> None:
        """Inject through the native views (the reference's callback)."""
        injection = self._injection(step)
        if abs(injection) < 1e-30:
            return
        z = self._position_z
        env = self._envelope_device

        def add(psi, amount):
            if self._axis == 0:
                psi[z, :, :] += amount * env
            elif self._axis == 1:
                psi[:, z, :] += amount * env
            else:
                psi[:, :, z] += amount * env
        add(sim._native_psi_real(), injection)
        if abs(self._phase) > 1e-12:
            pi = sim._native_psi_imag()
            if pi is not None:
                add(pi, injection * math.sin(self._phase))

    def fold(self) -> None:
        """Run the injection inside the step kernel from now on."""
        self._sim.evolver.set_plane_source(self._axis, self._position_z, self._envelope,
                                           self.amounts)

    def unfold(self) -> None:
        """Stop the in-kernel injection."""
        self._sim.evolver.set_plane_source(0, 0, None, None)
