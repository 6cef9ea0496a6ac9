"""GPU checks at the benchmark sizes (512^3), through properties that do not
need the CPU oracle to run at that size.

* Translation equivariance (periodic, mask-free): rolling the initial
  fields by (5, 17, 33) and stepping gives the rolled result bit for bit.
  Every site's arithmetic is the same wherever the tile boundaries fall, so
  any indexing or halo mistake at a tile, chunk or lattice edge shows up.
* The tiled kernel against the untiled (direct) kernel, with the mask and,
  for COLOR, the S_a transform on: relative L2 <= 1e-6 after 2 steps (the two
  kernels sum the Laplacian in different orders).
* The two-step kernel against single steps at 512^3 COMPLEX: bit-identical.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
N = 512


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(level, seed):
    """Seeded random Psi (unit normal) and chi = 19 + 0.1 normal on the GPU."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    shape = (3, N, N, N) if level == 2 else (N, N, N)
    pr = torch.randn(shape, generator=g, device="cuda")
    pi = torch.randn(shape, generator=g, device="cuda") if level else None
    chi = 19.0 + 0.1 * torch.randn((N, N, N), generator=g, device="cuda")
    return pr, pi, chi


def _state(ev):
    out = [ev.native_psi_real().clone(), ev.chi_tensor()]
    if ev.native_psi_imag() is not None:
        out.append(ev.native_psi_imag().clone())
    return out


def _rel_l2(a, b):
    import torch
    a, b = a.double(), b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("level", [0, 1])
def test_translation_equivariance_512(gpu, level):
    import torch
    import lfm_hip as L
    log(f"translation equivariance {N}^3 level {level}")
    mk = L.gravity_only if level == 0 else L.gravity_em
    cfg = mk(N, boundary_type=L.BoundaryType.PERIODIC)
    shift = (5, 17, 33)
    pr, pi, chi = _fields(level, 7)
    res = []
    for rolled in (False, True):
        roll = (lambda t: torch.roll(t, shift, dims=(-3, -2, -1))) if rolled else (lambda t: t)
        ev = L.Evolver(cfg, backend="hip", mask_policy="boundary_type")
        ev.set_psi_real(roll(pr))
        if pi is not None:
            ev.set_psi_imag(roll(pi))
        ev.set_chi(roll(chi))
        ev.evolve(4)
        torch.cuda.synchronize()
        res.append(_state(ev))
        del ev
    shape = (N, N, N)
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi", "psi_i")):
        ra = torch.roll(a.view(shape), shift, dims=(0, 1, 2))
        same = bool(torch.equal(ra, b.view(shape)))
        log(f"  {name}: rolled-then-stepped equals stepped-then-rolled: {same}")
        assert same, name
    del res, pr, pi, chi
    torch.cuda.empty_cache()


@pytest.mark.parametrize("level", [1, 2])
def test_tiled_vs_direct_512(gpu, level):
    import torch
    import lfm_hip as L
    log(f"tiled vs direct {N}^3 level {level}")
    cfg = L.gravity_em(N) if level == 1 else L.full_physics(N)
    pr, pi, chi = _fields(level, 11)
    pr *= 0.5
    pi *= 0.5
    res = []
    for direct in (False, True):
        ev = L.Evolver(cfg, backend="hip")
        ev.set_psi_real(pr)
        ev.set_psi_imag(pi)
        ev.set_chi(chi)
        ev._step(direct=direct)
        ev._step(direct=direct)
        torch.cuda.synchronize()
        res.append(_state(ev))
        del ev
        torch.cuda.empty_cache()
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi", "psi_i")):
        e = _rel_l2(a, b)
        log(f"  {name}: rel L2 {e:.2e}")
        assert bool(torch.isfinite(a).all()) and e <= 1e-6, (name, e)
    del res, pr, pi, chi
    torch.cuda.empty_cache()


def test_two_step_equals_single_512(gpu):
    import torch
    import lfm_hip as L
    log(f"two-step vs single {N}^3 COMPLEX")
    cfg = L.gravity_em(N)
    pr, pi, chi = _fields(1, 13)
    res = []
    for two in (True, False):
        ev = L.Evolver(cfg, backend="hip", two_step=two)
        ev.set_psi_real(pr)
        ev.set_psi_imag(pi)
        ev.set_chi(chi)
        ev.evolve(5)
        torch.cuda.synchronize()
        res.append(_state(ev))
        del ev
        torch.cuda.empty_cache()
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi", "psi_i")):
        same = bool(torch.equal(a, b))
        log(f"  {name}: bit-identical {same}")
        assert same, name
    del res, pr, pi, chi
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [1023, 1024])
def test_largest_sizes_real(gpu, n):
    """The step kernels at the top of their size ranges (REAL gravity_only,
    periodic, seeded random fields, 2 steps).  N = 1023 is the largest size the
    tiled kernel's 32-bit byte offsets cover (ragged 64-wide tiles at the
    edge): tiled against the untiled kernel within 1e-6 relative L2.  At
    N = 1024 only the untiled kernel runs (64-bit indices): rolling the
    initial fields by (5, 17, 33) and stepping gives the rolled result bit
    for bit."""
    import torch
    import lfm_hip as L
    log(f"largest sizes: REAL {n}^3")
    cfg = L.gravity_only(n, boundary_type=L.BoundaryType.PERIODIC)
    g = torch.Generator(device="cuda").manual_seed(n)
    pr = torch.randn((n, n, n), generator=g, device="cuda")
    chi = 19.0 + 0.1 * torch.randn((n, n, n), generator=g, device="cuda")
    shift = (5, 17, 33)
    runs = [(False, False), (True, False)] if n <= 1023 else [(False, False), (False, True)]
    res = []
    for direct, rolled in runs:
        roll = (lambda t: torch.roll(t, shift, dims=(0, 1, 2))) if rolled else (lambda t: t)
        ev = L.Evolver(cfg, backend="hip", mask_policy="boundary_type", graphs=False)
        assert not ev.has_mask
        ev.set_psi_real(roll(pr))
        ev.set_chi(roll(chi))
        ev._step(direct=direct)
        ev._step(direct=direct)
        torch.cuda.synchronize()
        res.append(_state(ev))
        del ev
        torch.cuda.empty_cache()
    shape = (n, n, n)
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi")):
        assert bool(torch.isfinite(a).all()), name
        if n <= 1023:
            e = _rel_l2(b, a)
            log(f"  {name}: direct vs tiled rel L2 {e:.2e}")
            assert e <= 1e-6, (name, e)
        else:
            same = bool(torch.equal(torch.roll(a.view(shape), shift, dims=(0, 1, 2)), b.view(shape)))
            log(f"  {name}: rolled-then-stepped equals stepped-then-rolled: {same}")
            assert same, name
    del res, pr, chi
    torch.cuda.empty_cache()
