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
from dataclasses import dataclass

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


class DetectorScreen:
    """A detector screen recording one lattice plane
    (lfm/experiment/detector.py:53-287): each ``record()`` stores the plane
    `position` across `axis` of `field` ("energy_density" = |Psi|^2,
    "psi_real", "psi_imag" or "chi"), and ``pattern`` is the sum of the
    recorded frames (the interference pattern), with the reference's
    ``snapshots``, ``mean_pattern``, ``click_pattern``, ``line_profile`` and
    ``reset``.

    Two ways to record, with the same frames bit for bit:

    * ``record()`` / ``step_callback(sim, step)``: the reference's form,
      called by the caller (per step, or per snapshot batch); the plane is
      cut on the device and only its N x N values are copied;
    * ``fold(every=1)``: the step kernel writes the plane into a device frame
      buffer after every `every`-th step from now on
      (``Evolver.set_plane_detector``; "energy_density" and "psi_real", REAL
      and COMPLEX), so the steps keep running as replayed graphs; frames are
      copied to the host only when read.  ``unfold()`` stops it, keeping the
      frames recorded so far.

    Like the reference's callback form, a folded detector records after a
    plane source and a barrier's keep factors of the same step."""

    def __init__(self, sim, axis: int = 2, position: int | None = None,
                 field: str = "energy_density") -> None:
        N = sim.config.grid_size
        self._sim = sim
        self._axis = int(axis)
        self._pos = position if position is not None else int(N * 0.80)
        self._field_name = field
        self._N = N
        self._frames: list[np.ndarray] = []
        self._folded = False

    # -- recording -----------------------------------------------------------
    def _plane(self, t):
        idx = [slice(None)] * 3
        idx[self._axis] = self._pos
        return t[tuple(idx)]

    def record(self) -> None:
        """Store the detector plane of the current field."""
        if self._folded:
            raise RuntimeError("this detector is folded into the steps (unfold() first)")
        sim, N = self._sim, self._N
        name = self._field_name
        if name == "energy_density" and sim.evolver.level < 2:
            # |Psi|^2 on the plane, squares and sum rounded as the host array
            # arithmetic rounds them
            pr = self._plane(sim._native_psi_real())
            frame = pr * pr
            pi = sim._native_psi_imag()
            if pi is not None:
                pi = self._plane(pi)
                frame = frame + pi * pi
            self._frames.append(frame.cpu().numpy().astype(np.float32))
            return
        if name in ("psi_real", "psi_imag") and sim.evolver.level < 2:
            t = sim._native_psi_real() if name == "psi_real" else sim._native_psi_imag()
            self._frames.append(np.zeros((N, N), np.float32) if t is None
                                else self._plane(t).cpu().numpy().astype(np.float32))
            return
        field = self._get_field()
        self._frames.append(np.asarray(self._plane(field), dtype=np.float32).copy())

    def step_callback(self, sim, step: int) -> None:
        """Simulation step callback; the same as record()."""
        self.record()

    def reset(self) -> None:
        """Clear all recorded frames (a folded detector records on from the
        next step)."""
        self._frames.clear()
        if self._folded:
            ev = self._sim.evolver
            det = ev._det
            ev.set_plane_detector(det["axis"], det["plane"], det["every"], self._field_name)

    def fold(self, every: int = 1) -> None:
        """Record inside the steps from now on, after every `every`-th step."""
        if self._field_name not in ("energy_density", "psi_real"):
            raise ValueError("only energy_density and psi_real detectors fold into the steps")
        self._sim.evolver.set_plane_detector(self._axis, self._pos, every, self._field_name)
        self._folded = True

    def unfold(self) -> None:
        """Stop recording inside the steps, keeping the frames recorded."""
        if self._folded:
            self._frames = self._all_frames()
            self._sim.evolver.set_plane_detector(None, 0)
            self._folded = False

    def _all_frames(self) -> list[np.ndarray]:
        if not self._folded:
            return self._frames
        dev = self._sim.evolver.detector_frames().cpu().numpy()
        return self._frames + [f for f in dev]

    # -- derived quantities ---------------------------------------------------
    @property
    def pattern(self) -> np.ndarray:
        """Sum of the recorded frames, (N, N) float32 (zeros when none)."""
        frames = self._all_frames()
        if not frames:
            return np.zeros((self._N, self._N), dtype=np.float32)
        return np.sum(frames, axis=0).astype(np.float32)

    @property
    def snapshots(self) -> np.ndarray:
        """All recorded frames, (K, N, N)."""
        frames = self._all_frames()
        if not frames:
            return np.zeros((0, self._N, self._N), dtype=np.float32)
        return np.stack(frames, axis=0)

    def click_pattern(self, n_particles: int = 1000, seed: int | None = None) -> np.ndarray:
        """n_particles detections drawn from the pattern as a probability
        distribution, as an (N, N) histogram."""
        pat = self.pattern
        total = pat.sum()
        if total == 0:
            return np.zeros(pat.shape, dtype=np.int32)
        prob = pat.ravel() / total
        rng = np.random.default_rng(seed)
        indices = rng.choice(len(prob), size=n_particles, p=prob)
        hist, _ = np.histogram(indices, bins=np.arange(len(prob) + 1))
        return hist.reshape(pat.shape).astype(np.int32)

    def mean_pattern(self) -> np.ndarray:
        """Time-averaged intensity, (N, N) float32."""
        frames = self._all_frames()
        if not frames:
            return np.zeros((self._N, self._N), dtype=np.float32)
        return (np.sum(frames, axis=0) / len(frames)).astype(np.float32)

    def line_profile(self, perp_axis: int = 0) -> np.ndarray:
        """The pattern summed along `perp_axis`."""
        return self.pattern.sum(axis=perp_axis).astype(np.float32)

    @property
    def position(self) -> int:
        return self._pos

    @property
    def axis(self) -> int:
        return self._axis

    @property
    def n_frames(self) -> int:
        return len(self._all_frames())

    def _get_field(self):
        sim = self._sim
        name = self._field_name
        if name == "energy_density":
            return sim.energy_density
        if name == "psi_real":
            return sim.psi_real
        if name == "psi_imag":
            f = sim.psi_imag
            return f if f is not None else np.zeros((self._N,) * 3, dtype=np.float32)
        if name == "chi":
            return sim.chi
        raise ValueError(f"Unknown field '{name}'. Choose from 'energy_density', 'psi_real', "
                         "'psi_imag', 'chi'.")

    def __repr__(self) -> str:
        return (f"DetectorScreen(axis={self._axis}, position={self._pos}, "
                f"field='{self._field_name}', n_frames={self.n_frames})")


@dataclass
class Slit:
    """One slit opening of a :class:`Barrier` (lfm/experiment/barrier.py:71-95):
    `center` and `width` in cells along the slit axis; a which-path detector
    damps Psi there by `detector_strength` per slit transit."""

    center: int
    width: int = 4
    detector: bool = False
    detector_strength: float = 1.0


class Barrier:
    """A chi barrier with slit openings (lfm/experiment/barrier.py:98-427):
    chi = `height` (default chi0 + 50) in a slab of `thickness` cells across
    `axis` at `position`, chi0 in the slits; with `absorb` Psi is zeroed in
    the solid cells, and a detector slit multiplies Psi by 1 - gamma per step,
    gamma = 1 - (1 - strength)^(1 / transit_steps).  The barrier is applied
    once on construction.

    Two ways to keep it in place, with the same fields bit for bit:

    * ``step_callback(sim, step)``: the reference's form, after every step:
      all four chi buffers rewritten to the barrier's chi, Psi multiplied by
      the absorber and detector factors through the native views;
    * ``fold()``: chi held at the barrier's chi (every evolve freezes chi,
      ``Evolver.hold_chi``) and the factors applied by the step kernel as it
      writes Psi (``Evolver.set_psi_keep``); ``unfold()`` stops it.  Slits
      whose cells overlap, where two factors would meet at one site, are
      refused by fold().

    A folded barrier applies its factors after a folded source's addition and
    before a folded detector's record, the order of the reference's callbacks
    in its examples."""

    def __init__(self, sim, axis: int = 2, position: int | None = None,
                 height: float | None = None, thickness: int = 2, slits=None,
                 slit_axis: int | None = None, absorb: bool = True,
                 transit_steps: int = 10) -> None:
        from .config import CHI0
        N = sim.config.grid_size
        self._sim = sim
        self._N = N
        self._axis = axis
        self._height = height if height is not None else CHI0 + 50.0
        self._pos = position if position is not None else N // 2
        self._thick = thickness
        self._absorb = absorb
        if slits is None:
            half_sep = max(4, N // 8)
            slit_w = max(2, N // 32)
            slits = [Slit(center=N // 2 - half_sep, width=slit_w),
                     Slit(center=N // 2 + half_sep, width=slit_w)]
        self._slits = list(slits)
        self._slit_axis = slit_axis if slit_axis is not None else (axis + 1) % 3
        self._barrier_mask, self._slit_masks = self._build_masks()
        start = max(0, self._pos - self._thick // 2)
        end = min(N, start + self._thick)
        slab = [slice(None)] * 3
        slab[self._axis] = slice(start, end)
        self._slab = tuple(slab)
        self._slit_slices = []
        for slit in self._slits:
            s0 = max(0, slit.center - slit.width // 2)
            s1 = min(N, s0 + slit.width)
            sl = [slice(None)] * 3
            sl[self._axis] = slice(start, end)
            sl[self._slit_axis] = slice(s0, s1)
            self._slit_slices.append(tuple(sl))
        self._chi0 = float(sim.config.chi0)
        # per-step Psi factors: 0 in the solid cells (absorb), 1 - gamma in
        # detector slits, as float32 device arrays
        td = sim._to_device
        self._psi_mask = td((~self._barrier_mask).astype(np.float32)).reshape(N, N, N)
        self._det_scales = []
        for i, slit in enumerate(self._slits):
            if slit.detector and slit.detector_strength > 0.0:
                transit = max(1, transit_steps)
                strength = float(min(1.0, slit.detector_strength))
                gamma = 1.0 if strength >= 1.0 else 1.0 - (1.0 - strength) ** (1.0 / transit)
                scale = np.ones((N, N, N), dtype=np.float32)
                scale[self._slit_masks[i]] = float(1.0 - gamma)
                self._det_scales.append(scale)
            else:
                self._det_scales.append(None)
        self._det_scales_dev = [td(sc).reshape(N, N, N) if sc is not None else None
                                for sc in self._det_scales]
        self._folded = False
        self.apply()

    def _build_masks(self):
        N, axis = self._N, self._axis
        start = max(0, self._pos - self._thick // 2)
        end = min(N, start + self._thick)
        barrier_mask = np.zeros((N, N, N), dtype=bool)
        sl = [slice(None)] * 3
        sl[axis] = slice(start, end)
        barrier_mask[tuple(sl)] = True
        slit_masks = []
        for slit in self._slits:
            s0 = max(0, slit.center - slit.width // 2)
            s1 = min(N, s0 + slit.width)
            sm = np.zeros((N, N, N), dtype=bool)
            ssl = [slice(None)] * 3
            ssl[axis] = slice(start, end)
            ssl[self._slit_axis] = slice(s0, s1)
            sm[tuple(ssl)] = True
            slit_masks.append(sm)
            barrier_mask[tuple(ssl)] = False
        return barrier_mask, slit_masks

    def chi_template(self) -> np.ndarray:
        """The barrier's chi field, (N, N, N) float32: chi0, the slab at
        `height`, the slits back at chi0."""
        chi = np.full((self._N,) * 3, self._chi0, dtype=np.float32)
        chi[self._slab] = self._height
        for sl in self._slit_slices:
            chi[sl] = self._chi0
        return chi

    def keep_factors(self) -> np.ndarray:
        """The per-site Psi factor of one enforcement, (N, N, N) float32: the
        absorber (0 in the solid cells) times the detector scales."""
        k = (~self._barrier_mask).astype(np.float32) if self._absorb else \
            np.ones((self._N,) * 3, np.float32)
        for sc in self._det_scales:
            if sc is not None:
                k = k * sc
        return k

    def apply(self) -> None:
        """Re-impose the barrier: chi in all four buffers, Psi multiplied by
        the absorber and detector factors (lfm/experiment/barrier.py:321-367)."""
        chi0, h = self._chi0, float(self._height)
        for buf in self._sim._native_chi_pair():
            buf.fill(chi0)
            buf[self._slab] = h
            for sl in self._slit_slices:
                buf[sl] = chi0
        pr, pi = self._sim._native_psi_real(), self._sim._native_psi_imag()
        if self._absorb:
            pr *= self._psi_mask
            if pi is not None:
                pi *= self._psi_mask
        for sc in self._det_scales_dev:
            if sc is not None:
                pr *= sc
        if pi is not None:
            for sc in self._det_scales_dev:
                if sc is not None:
                    pi *= sc

    def attenuate_slits(self) -> None:
        """No-op (the detector damping is part of apply())."""

    def measure_slits(self) -> dict:
        """Mean energy density in each slit opening."""
        ed = self._sim.energy_density
        return {f"slit_{i}": float(ed[sm].mean()) for i, sm in enumerate(self._slit_masks)}

    def step_callback(self, sim, step: int) -> None:
        """Simulation step callback: apply()."""
        if self._folded:
            raise RuntimeError("this barrier is folded into the steps (unfold() first)")
        self.apply()

    def fold(self) -> None:
        """Keep the barrier inside the steps from now on: chi held at the
        barrier's chi, Psi factors applied by the step kernel."""
        n_factors = self._barrier_mask.astype(np.int8) if self._absorb else \
            np.zeros(self._barrier_mask.shape, np.int8)
        for i, sc in enumerate(self._det_scales):
            if sc is not None:
                n_factors += self._slit_masks[i].astype(np.int8)
        if int(n_factors.max(initial=0)) > 1:
            raise ValueError("slits overlap: two Psi factors meet at one site, so a single "
                             "per-site factor would not round as the callback's multiplies")
        ev = self._sim.evolver
        ev.set_chi(self.chi_template())
        ev.set_psi_keep(self.keep_factors())
        ev.hold_chi = True
        self._folded = True

    def unfold(self) -> None:
        """Stop keeping the barrier inside the steps."""
        if self._folded:
            ev = self._sim.evolver
            ev.set_psi_keep(None)
            ev.hold_chi = False
            self._folded = False

    @property
    def mask(self) -> np.ndarray:
        return self._barrier_mask

    @property
    def slit_masks(self) -> list:
        return self._slit_masks

    @property
    def slits(self) -> list:
        return self._slits

    @property
    def position(self) -> int:
        return self._pos

    @property
    def axis(self) -> int:
        return self._axis

    @property
    def height(self) -> float:
        return self._height

    def __repr__(self) -> str:
        return (f"Barrier(axis={self._axis}, position={self._pos}, height={self._height:.1f}, "
                f"thickness={self._thick}, n_slits={len(self._slits)})")
