"""Initial-condition builders (host NumPy, run once before the hot path).

Same formulas and float32 outputs as lfm/fields/soliton.py:23-117 and
lfm/fields/random.py:19-54, so seeded inputs match the reference's.
"""
from __future__ import annotations

import numpy as np


def _grid(N: int):
    x = np.arange(N, dtype=np.float32)
    return np.meshgrid(x, x, x, indexing="ij")


def gaussian_soliton(N, position, amplitude, sigma, phase=0.0):
    """amplitude * exp(-r^2 / 2 sigma^2) * e^{i phase} as (real, imag) float32."""
    X, Y, Z = _grid(N)
    px, py, pz = position
    env = amplitude * np.exp(-((X - px) ** 2 + (Y - py) ** 2 + (Z - pz) ** 2) / (2.0 * sigma ** 2))
    return (env * np.cos(phase)).astype(np.float32), (env * np.sin(phase)).astype(np.float32)


def place_solitons(N, positions, amplitude, sigma, phases=None, colors=None, n_colors=3):
    """Superpose Gaussian solitons onto an (n_colors, N, N, N) field; soliton
    s goes to colour colors[s] (round robin when not given)."""
    n = len(positions)
    phases = np.zeros(n) if phases is None else np.asarray(phases, dtype=np.float64)
    colors = [s % n_colors for s in range(n)] if colors is None else colors
    pr = np.zeros((n_colors, N, N, N), dtype=np.float32)
    pi = np.zeros((n_colors, N, N, N), dtype=np.float32)
    X, Y, Z = _grid(N)
    for s, (px, py, pz) in enumerate(positions):
        env = amplitude * np.exp(-((X - px) ** 2 + (Y - py) ** 2 + (Z - pz) ** 2)
                                 / (2.0 * sigma ** 2))
        pr[colors[s]] += env * np.cos(phases[s])
        pi[colors[s]] += env * np.sin(phases[s])
    return pr, pi


def seed_noise(N, amplitude=1e-6, n_colors=1, rng=None):
    """Gaussian noise of rms `amplitude` for (real, imag); real drawn first."""
    if not isinstance(rng, np.random.Generator):
        rng = np.random.default_rng(rng)
    shape = (N, N, N) if n_colors == 1 else (n_colors, N, N, N)
    r = (rng.standard_normal(shape) * amplitude).astype(np.float32)
    i = (rng.standard_normal(shape) * amplitude).astype(np.float32)
    return r, i


def boundary_mask_host(N: int, boundary_fraction: float) -> np.ndarray:
    """cos^2 absorbing mask, flat float32 of N^3, built plane by plane.

    The element-wise float32 NumPy operations are those of
    lfm/core/backends/numpy_backend.py:49-69 (and cupy_backend.py:106-125,
    which also builds the mask on the host), so the values are bit-identical
    to the reference's; building one i-plane at a time bounds host memory at
    large N.
    """
    center = N / 2.0
    r_max = N / 2.0
    r_freeze = (1.0 - boundary_fraction) * r_max
    coords = np.arange(N, dtype=np.float32) - center + 0.5
    Y, Z = np.meshgrid(coords, coords, indexing="ij")
    yz = Y ** 2
    out = np.empty((N, N, N), dtype=np.float32)
    for i in range(N):
        X = np.full((N, N), coords[i], dtype=np.float32)
        R = np.sqrt(X ** 2 + yz + Z ** 2)
        t = np.clip((R - r_freeze) / (r_max - r_freeze), 0.0, 1.0)
        out[i] = (np.sin(0.5 * np.pi * t) ** 2).astype(np.float32)
    return out.ravel()


def laplacian_19pt_host(f: np.ndarray) -> np.ndarray:
    """Periodic 19-point Laplacian of an (N, N, N) host array: faces 1/3,
    edges 1/6, centre -4 (lfm/core/stencils.py:26-71), for host-side
    analysis such as the relaxation's eigenvalue estimate."""
    sh = lambda a, s, ax: np.roll(a, s, axis=ax)
    faces = sum(sh(f, s, ax) for ax in range(3) for s in (1, -1))
    edges = sum(sh(sh(f, s1, a1), s2, a2) for a1, a2 in ((0, 1), (0, 2), (1, 2))
                for s1 in (1, -1) for s2 in (1, -1))
    return faces / 3.0 + edges / 6.0 - 4.0 * f
