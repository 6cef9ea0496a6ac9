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
* Periodic tiling (periodic, mask-free): a 512^3 lattice holding 2 x 2 x 2
  copies of a 256^3 state stays the tiling of the 256^3 lattice's evolution:
  bit for bit at COMPLEX; at COLOR, where the 512^3 S_a transform sees the
  tiling as a spectrum on even wavenumbers only (the same filter values, other
  roundings), within relative L2 1e-7 (measured 2e-10 for Psi, 4e-9 for chi
  after 6 steps).
* The top of the size ranges: REAL at 1023^3 and 1024^3 (tiled against
  untiled, and translation equivariance of the tiled kernel at 1024), 1025^3
  on the untiled kernel; the two-step kernel and the COLOR step with S_a at
  1024^3.
"""
import gc
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
def test_periodic_tiling_512(gpu, level):
    import torch
    import lfm_hip as L
    n2 = N // 2
    log(f"periodic tiling {n2}^3 -> {N}^3, level {level}")
    g = torch.Generator(device="cuda").manual_seed(11)
    shape = (3, n2, n2, n2) if level == 2 else (n2, n2, n2)
    pr = 0.5 * torch.randn(shape, generator=g, device="cuda")
    pi = 0.5 * torch.randn(shape, generator=g, device="cuda")
    chi = 19.0 + 0.1 * torch.randn((n2, n2, n2), generator=g, device="cuda")
    reps = (1, 2, 2, 2) if level == 2 else (2, 2, 2)
    mk = L.gravity_em if level == 1 else L.full_physics
    res = []
    for n, tile in ((n2, False), (N, True)):
        cfg = mk(n, boundary_type=L.BoundaryType.PERIODIC)
        ev = L.Evolver(cfg, backend="hip", mask_policy="boundary_type")
        ev.set_psi_real(pr.repeat(*reps) if tile else pr)
        ev.set_psi_imag(pi.repeat(*reps) if tile else pi)
        ev.set_chi(chi.repeat(2, 2, 2) if tile else chi)
        ev.evolve(6)
        torch.cuda.synchronize()
        # chi in storage units (chi - chi0): its float32 rounding does not
        # blur the comparison of the deviations
        assert ev.chi_storage == "delta"
        res.append([ev.native_psi_real().clone(), ev.native_chi().clone(),
                    ev.native_psi_imag().clone()])
        del ev
        torch.cuda.empty_cache()
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi", "psi_i")):
        sh = (3, n2, n2, n2) if (level == 2 and name != "chi") else (n2, n2, n2)
        r = (1, 2, 2, 2) if len(sh) == 4 else (2, 2, 2)
        tiled = a.view(sh).repeat(*r).reshape(-1)
        if level == 1:
            same = bool(torch.equal(tiled, b))
            log(f"  {name}: {N}^3 equals the tiled {n2}^3 result: {same}")
            assert same, name
        else:
            e = _rel_l2(b, tiled)
            log(f"  {name}: rel-L2 {e:.2e}")
            assert e <= 1e-7, name
    del res
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


@pytest.mark.parametrize("n", [1023, 1024, 1025])
def test_largest_sizes_real(gpu, n):
    """The step kernels at the top of their size ranges (REAL gravity_only,
    periodic, seeded random fields, 2 steps).  The tiled kernel's 32-bit byte
    offsets cover N <= 1024 (the largest, 4 N^3 - 4, is 2^32 - 4 at 1024):
    at N = 1023 (ragged 64-wide tiles at the edge) and 1024 it agrees with the
    untiled kernel within 1e-6 relative L2, and at 1024 rolling the initial
    fields by (5, 17, 33) and stepping gives the rolled result bit for bit.
    N = 1025 runs the untiled kernel (64-bit indices): rolled-then-stepped
    equals stepped-then-rolled bit for bit."""
    import torch
    import lfm_hip as L
    log(f"largest sizes: REAL {n}^3")
    cfg = L.gravity_only(n, boundary_type=L.BoundaryType.PERIODIC)
    g = torch.Generator(device="cuda").manual_seed(n)
    pr = torch.randn((n, n, n), generator=g, device="cuda")
    chi = 19.0 + 0.1 * torch.randn((n, n, n), generator=g, device="cuda")
    shift = (5, 17, 33)
    tiled = n <= 1024
    # (direct kernel, rolled initial fields)
    runs = [(False, False), (True, False)] if tiled else [(True, False)]
    if n >= 1024:
        runs.append((not tiled, True))
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
        gc.collect()   # the Evolver's transform plan frees its device memory
        torch.cuda.empty_cache()
    shape = (n, n, n)
    for q, name in enumerate(("psi_r", "chi")):
        a = res[0][q]
        assert bool(torch.isfinite(a).all()), name
        if tiled:
            e = _rel_l2(res[1][q], a)
            log(f"  {name}: direct vs tiled rel L2 {e:.2e}")
            assert e <= 1e-6, (name, e)
        if n >= 1024:
            same = bool(torch.equal(torch.roll(a.view(shape), shift, dims=(0, 1, 2)),
                                    res[-1][q].view(shape)))
            log(f"  {name}: rolled-then-stepped equals stepped-then-rolled "
                f"({'tiled' if tiled else 'untiled'}): {same}")
            assert same, name
    del res, pr, chi
    torch.cuda.empty_cache()


def test_two_step_equals_single_1024(gpu):
    """The two-step kernel at N = 1024 (COMPLEX gravity_em, mask on, 5 steps:
    two pairs and one single step) against single steps: bit-identical."""
    import torch
    import lfm_hip as L
    n = 1024
    log(f"two-step vs single {n}^3 COMPLEX")
    cfg = L.gravity_em(n)
    g = torch.Generator(device="cuda").manual_seed(17)
    pr = torch.randn((n, n, n), generator=g, device="cuda")
    pi = torch.randn((n, n, n), generator=g, device="cuda")
    chi = 19.0 + 0.1 * torch.randn((n, n, n), generator=g, device="cuda")
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
        gc.collect()   # the Evolver's transform plan frees its device memory
        torch.cuda.empty_cache()
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi", "psi_i")):
        same = bool(torch.equal(a, b))
        log(f"  {name}: bit-identical {same}")
        assert same, name
    del res, pr, pi, chi
    torch.cuda.empty_cache()


def test_color_tiled_vs_direct_1024(gpu):
    """COLOR full_physics at 1024^3 (mask on, S_a through the five transform
    passes, whose column passes run 16 values per thread only at this size):
    the tiled step against the untiled one, relative L2 <= 1e-6 after 2
    steps."""
    import torch
    import lfm_hip as L
    n = 1024
    log(f"tiled vs direct {n}^3 COLOR")
    cfg = L.full_physics(n)
    g = torch.Generator(device="cuda").manual_seed(19)
    pr = 0.5 * torch.randn((3, n, n, n), generator=g, device="cuda")
    pi = 0.5 * torch.randn((3, n, n, n), generator=g, device="cuda")
    chi = 19.0 + 0.1 * torch.randn((n, n, n), generator=g, device="cuda")
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
        gc.collect()   # the Evolver's transform plan frees its device memory
        torch.cuda.empty_cache()
    for a, b, name in zip(res[0], res[1], ("psi_r", "chi", "psi_i")):
        e = _rel_l2(a, b)
        log(f"  {name}: rel L2 {e:.2e}")
        assert bool(torch.isfinite(a).all()) and e <= 1e-6, (name, e)
    del res, pr, pi, chi
    torch.cuda.empty_cache()
