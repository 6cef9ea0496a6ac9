"""Soliton eigenmode relaxation on the GPU (lfm/particles/solver.py:828-999).

``relax_eigenmode`` alternates a Poisson solve for chi (the plan-based
GOV-04 transform, then clipped to [0.1, chi0] and held at chi0 on the outer
shell) with imaginary-time relaxation of E in the frozen chi,

    E <- E + dt_E (lap E - chi^2 E),   E <- E * amplitude / max|E|,

until a cycle changes E by less than the tolerance.  The relaxation step and
its renormalisation are HIP kernels (``lfm_relax_step_f32`` /
``lfm_relax_scale_f32``); the lattice never leaves the GPU and the host reads
the max |dE| once per convergence check, where the reference's loop reads
max|E| back every step.

Seeds are the reference's l = 0 Gaussian or a caller-supplied array (the
particle catalogue and the Y_lm seeds of ``lfm.particles`` are not part of
this package).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from .backend import ScvPlan, graph_capture, stream_handle
from .evolver import get_backend

CHI0 = 19.0
KAPPA = 1.0 / 63.0


@dataclass
class SolitonSolution:
    """Result of the relaxation, with the reference's fields
    (lfm/particles/solver.py:65-105).  As there, ``cycles`` holds the total
    number of imaginary-time steps."""
    psi_r: np.ndarray
    psi_i: np.ndarray | None
    chi: np.ndarray
    chi_min: float
    energy: float
    eigenvalue: float
    converged: bool
    cycles: int
    energy_history: list = field(default_factory=list)
    particle: object = None
    N: int = 0


def shell_mask(N: int, boundary_fraction: float = 0.3) -> np.ndarray:
    """True in the frozen outer shell: r >= (N/2)(1 - bf) about (N-1)/2, in
    float32 coordinates (lfm/particles/solver.py:517-524)."""
    c = (N - 1) / 2.0
    x = np.arange(N, dtype=np.float32)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2)
    return r >= (N * 0.5) * (1.0 - boundary_fraction)


def gaussian_seed(N: int, position, amplitude: float, sigma: float) -> np.ndarray:
    """l = 0 seed amplitude * exp(-r^2 / sigma^2) (float64, stored float32;
    lfm/particles/solver.py:917-922)."""
    idx = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    cx, cy, cz = position
    r2 = ((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (sigma ** 2)
    return (amplitude * np.exp(-r2)).astype(np.float32)


class _Relaxer:
    """Device state of one relaxation: E ping-pong buffers, chi^2, the max
    words and the Poisson plan."""

    def __init__(self, E: np.ndarray, N: int, chi0: float, kappa: float, boundary_fraction: float,
                 backend):
        self.N, self.chi0, self.kappa = N, float(chi0), float(kappa)
        self.dev = backend.device
        self.lib = _abi.load()
        self.n3 = N ** 3
        f32 = torch.float32
        self.E = torch.from_numpy(np.ascontiguousarray(E, np.float32).ravel()).to(self.dev)
        self.E_alt = torch.empty_like(self.E)
        self.chi = torch.empty(self.n3, dtype=f32, device=self.dev)
        self.chi2 = torch.empty_like(self.chi)
        self.shell = torch.from_numpy(shell_mask(N, boundary_fraction).ravel()).to(self.dev)
        # slots 0-2: max|E_new| of consecutive steps; slot 3: max|dE|
        self.sw = _abi.LFM_RELAX_SLOT_WORDS
        self.slots = torch.zeros(4 * self.sw, dtype=torch.int32, device=self.dev)
        self.plan = ScvPlan(N, 1)
        self._graphs = {}

    def slot(self, i: int) -> int:
        return self.slots.data_ptr() + 4 * self.sw * i

    def slot_max(self, i: int) -> float:
        """The max a slot holds (host read; a synchronisation)."""
        words = self.slots[i * self.sw:(i + 1) * self.sw].cpu().numpy().astype(np.uint32)
        return float(words.view(np.float32).max())

    def poisson(self) -> None:
        """chi := clip(chi0 + Poisson(kappa E^2), 0.1, chi0), chi0 on the
        shell (lfm/particles/solver.py:760-772); chi2 := chi * chi."""
        rc = self.lib.lfm_poisson_chi_f32(self.plan.handle, self.E.data_ptr(), None, 1,
                                          self.kappa, 0.0, self.chi0, -1.0, self.chi.data_ptr(),
                                          stream_handle())
        _abi.check(rc, "lfm_poisson_chi_f32")
        self.chi.clamp_(0.1, self.chi0)
        self.chi.masked_fill_(self.shell, self.chi0)
        torch.mul(self.chi, self.chi, out=self.chi2)

    # Steps between convergence checks run as fused launches (the previous
    # step's rescale folded into the update), so E holds the last output
    # before its rescale; six steps return the three max slots and the two
    # E buffers to where they started, so blocks of six replay from graphs.
    GRAPH_STEPS = 6
    GRAPH_CACHE = 12      # six phase / buffer-order keys per (amplitude, dt)

    def _step(self, k: int, scale_k, amplitude: float, dt: float) -> None:
        """One update into slot k; scale_k: slot of the pending rescale of E
        (None when E holds rescaled values)."""
        _abi.check(self.lib.lfm_relax_step_f32(
            self.E.data_ptr(), self.chi2.data_ptr(), self.E_alt.data_ptr(), self.slot(k),
            None if scale_k is None else self.slot(scale_k), self.slot((k + 1) % 3),
            float(amplitude), self.N, float(dt), stream_handle()), "lfm_relax_step_f32")
        self.E, self.E_alt = self.E_alt, self.E

    def _rescale(self, k: int, amplitude: float, E_prev=None) -> None:
        """E *= amplitude / max(slot k); with E_prev, max|E - E_prev| into slot 3."""
        _abi.check(self.lib.lfm_relax_scale_f32(
            self.E.data_ptr(), None if E_prev is None else E_prev.data_ptr(), self.slot(k), None,
            None if E_prev is None else self.slot(3), float(amplitude), self.n3, stream_handle()),
            "lfm_relax_scale_f32")

    def _block_graph(self, k0: int, amplitude: float, dt: float):
        """Graph of GRAPH_STEPS fused steps from slot phase k0 (E pending the
        rescale held in slot (k0 - 1) % 3), keyed by the phase, the buffer
        order and the scalars."""
        key = (k0, self.E.data_ptr(), float(amplitude), float(dt))
        g = self._graphs.get(key)
        if g is None:
            while len(self._graphs) >= self.GRAPH_CACHE:
                self._graphs.pop(next(iter(self._graphs)))   # oldest first
            s = torch.cuda.Stream(device=self.dev)
            s.wait_stream(torch.cuda.current_stream(self.dev))
            g = torch.cuda.CUDAGraph()
            E0, E1 = self.E, self.E_alt
            with graph_capture(g, s):
                for u in range(self.GRAPH_STEPS):
                    k = (k0 + u) % 3
                    self._step(k, (k - 1) % 3, amplitude, dt)
            torch.cuda.current_stream(self.dev).wait_stream(s)
            # capture recorded the launches without running them; six swaps
            # leave the buffers in their starting order
            assert self.E is E0 and self.E_alt is E1
            self._graphs[key] = g
        return g

    def relax(self, amplitude: float, dt: float, steps: int, check_interval: int,
              tolerance: float, graphs: bool = True) -> tuple[bool, int]:
        """Up to `steps` relaxation steps; (converged, steps taken).  E holds
        rescaled values on return."""
        self.slots.zero_()
        pending = None         # slot whose max E still has to be rescaled by
        step = 0
        while step < steps:
            step += 1
            k = step % 3
            if step % check_interval == 0:
                # convergence check: max|E_n - E_{n-1}| over rescaled fields
                if pending is not None:
                    self._rescale(pending, amplitude)
                    pending = None
                self._step(k, None, amplitude, dt)
                self.slots[3 * self.sw:].zero_()
                self._rescale(k, amplitude, E_prev=self.E_alt)
                if self.slot_max(3) < tolerance:
                    return True, step
                continue
            # fused steps up to the next check; whole blocks of six from graphs
            nxt = min(steps, (step // check_interval + 1) * check_interval - 1)
            if graphs and pending is not None:
                while nxt - step + 1 >= self.GRAPH_STEPS:
                    self._block_graph(k, amplitude, dt).replay()
                    step += self.GRAPH_STEPS
                    k = step % 3
                    pending = (k - 1) % 3
                if step > nxt:
                    step -= 1
                    continue
            self._step(k, pending, amplitude, dt)
            pending = k
        if pending is not None:
            self._rescale(pending, amplitude)
        return False, steps


def relax_eigenmode(particle=None, N: int = 64, position=None, amplitude: float | None = None,
                    sigma: float | None = None, l: int = 0, m: int = 0,  # noqa: E741
                    chi0: float = CHI0, kappa: float = KAPPA, floor_lambda: float = 10.0,
                    relax_dt: float = 0.0, max_cycles: int = 20, steps_per_cycle: int = 1000,
                    check_interval: int = 200, tolerance: float = 1e-4,
                    boundary_fraction: float = 0.3, verbose: bool = False, seed=None,
                    backend="auto") -> SolitonSolution:
    """Poisson / imaginary-time relaxation to a stable soliton eigenmode, with
    the reference's arguments and results (lfm/particles/solver.py:828-999).
    ``seed`` (an (N, N, N) array) replaces the Gaussian seed; a particle or
    l > 0 needs the reference's particle catalogue and raises here.
    ``floor_lambda`` is accepted and unused, as in the reference."""
    if particle is not None or l != 0:
        raise NotImplementedError("particle catalogue and Y_lm seeds are not part of lfm_hip; "
                                  "pass seed=<(N, N, N) array> instead")
    if position is None:
        position = (N // 2, N // 2, N // 2)
    if amplitude is None:
        amplitude = 14.0
    if sigma is None:
        sigma = max(3.0, N / 16.0)
    E0 = gaussian_seed(N, position, amplitude, sigma) if seed is None else \
        np.asarray(seed, np.float32).reshape(N, N, N)
    dt_E = 0.8 / (chi0 * chi0 + 12.0)
    if relax_dt > 0:
        dt_E = min(relax_dt, dt_E)
    dt32 = float(np.float32(dt_E))

    R = _Relaxer(E0, N, chi0, kappa, boundary_fraction, get_backend(backend))
    converged, total, cycle = False, 0, 0
    for cycle in range(max_cycles):
        R.poisson()
        E_start = R.E.clone()
        if verbose:
            print(f"  cycle {cycle + 1}/{max_cycles} Poisson (GPU) | "
                  f"chi_min={float(R.chi.min()):.4f} | E_sum={float((R.E.double() ** 2).sum()):.1f}")
        e_conv, n = R.relax(amplitude, dt32, steps_per_cycle, check_interval, tolerance)
        total += n
        d_cycle = float((R.E - E_start).abs().max())
        if verbose:
            print(f"  cycle {cycle + 1} done | dE_cycle={d_cycle:.3e} | e_converged={e_conv}")
        if d_cycle < tolerance and cycle > 0:
            converged = True
            break
    R.poisson()
    E = R.E.cpu().numpy().reshape(N, N, N)
    chi = R.chi.cpu().numpy().reshape(N, N, N)
    # omega^2 = <E|(-lap + chi^2)|E> / <E|E> on the host in float64
    from .fields import laplacian_19pt_host
    E64, chi64 = E.astype(np.float64), chi.astype(np.float64)
    HE = -laplacian_19pt_host(E64) + chi64 * chi64 * E64
    e_sq = float(np.sum(E64 * E64))
    omega = math.sqrt(max(float(np.sum(E64 * HE)) / max(e_sq, 1e-30), 0.0))
    if verbose:
        print(f"  {'CONVERGED' if converged else 'DID NOT CONVERGE'} (GPU) after {total} E-steps "
              f"({cycle + 1} cycles) | chi_min={float(chi.min()):.3f} | omega={omega:.4f}")
    return SolitonSolution(psi_r=E, psi_i=None, chi=chi, chi_min=float(chi.min()), energy=e_sq,
                           eigenvalue=omega, converged=converged, cycles=total, particle=None, N=N)
