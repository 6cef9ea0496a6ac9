"""``gpu_snapshot_loop``: the experiment drivers' run loop
(lfm/experiment/common.py:138-280) over :class:`lfm_hip.Simulation`.

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
