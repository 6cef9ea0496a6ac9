"""The per-site Psi factor folded into the step (``Evolver.set_psi_keep``)
against the driver form it replaces, bit for bit.

The driver form is the reference barrier's per-step enforcement
(lfm/experiment/barrier.py:313-353): after every step all four chi buffers are
rewritten to the barrier's chi template and the current Psi is multiplied in
place by the absorber / which-path-detector factors.  Run here as a device hook
through the native views, it is compared with set_chi(template) +
set_psi_keep(factors) + evolve(freeze_chi=True), which does the same work inside
the step kernel: chi frozen at the template (the rewritten buffers make chi's
update irrelevant), Psi' multiplied as it is written.
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _barrier(N):
    """A slab of raised chi (25 against 19) across axis 0 with two slits along
    axis 1; Psi absorbed (factor 0) in the solid part, a which-path detector
    (factor 0.75) in the first slit.  chi values are exact in float32 as
    chi0 + (chi - chi0), so both chi storages hold the same template."""
    x0, x1 = N // 2 - 2, N // 2 + 2
    chi = np.full((N, N, N), 19.0, np.float32)
    keep = np.ones((N, N, N), np.float32)
    chi[x0:x1] = 25.0
    keep[x0:x1] = 0.0
    for s0, w, k in ((N // 2 - 7, 3, 0.75), (N // 2 + 4, 3, 1.0)):
        chi[x0:x1, s0:s0 + w, :] = 19.0
        keep[x0:x1, s0:s0 + w, :] = k
    return chi, keep


def _fields(level, N, keep):
    """Seeded Gaussian packets moving towards the barrier, already multiplied
    by the factors (the barrier applied once before the first step)."""
    rng = np.random.default_rng(5)
    x = np.arange(N, dtype=np.float32)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    nc = 3 if level == 2 else 1
    pr = np.zeros((nc, N, N, N), np.float32)
    pi = np.zeros((nc, N, N, N), np.float32)
    for a in range(nc):
        c = N / 4 + rng.uniform(-2, 2, 3)
        env = 2.0 * np.exp(-((X - c[0]) ** 2 + (Y - N / 2) ** 2 + (Z - N / 2) ** 2) / 18.0)
        ph = 0.6 * X + rng.uniform(0, 2 * np.pi)
        pr[a] = env * np.cos(ph) * keep
        pi[a] = env * np.sin(ph) * keep
    if level == 2:
        return pr, pi
    return pr[0], (pi[0] if level == 1 else None)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_keep_factor_equals_driver_form(gpu, level):
    import torch
    import lfm_hip as L
    N, steps = 48, 40
    log(f"keep factor vs driver form, level {level}, {N}^3, {steps} steps")
    cfg = (L.gravity_only, L.gravity_em, L.full_physics)[level](N)
    chi_t, keep = _barrier(N)
    pr, pi = _fields(level, N, keep)
    res = {}
    for form in ("driver", "fused"):
        ev = L.Evolver(cfg, backend="hip")
        ev.set_psi_real(pr)
        if pi is not None:
            ev.set_psi_imag(pi)
        ev.set_chi(chi_t)
        if form == "driver":
            ev.use_absolute_chi()
            chi_d = torch.from_numpy(chi_t.reshape(-1)).to(ev.device)
            k_d = torch.from_numpy(np.tile(keep.reshape(-1), 3 if level == 2 else 1)).to(ev.device)

            def hook(e, step):
                for buf in e.native_chi_refs():
                    buf.copy_(chi_d)
                e.native_psi_real().mul_(k_d)
                if e.native_psi_imag() is not None:
                    e.native_psi_imag().mul_(k_d)
            ev.evolve(2, device_hook=hook)     # warm-up: capture the graphs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ev.evolve(steps - 2, device_hook=hook)
        else:
            ev.set_psi_keep(keep)
            ev.evolve(2, freeze_chi=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ev.evolve(steps - 2, freeze_chi=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / (steps - 2) * 1e3
        res[form] = (ev.native_psi_real().clone(),
                     ev.native_psi_imag().clone() if pi is not None else None,
                     torch.from_numpy(ev.get_chi()))
        log(f"  {form}: {dt:.3f} ms per step")
        del ev
    a, b = res["driver"], res["fused"]
    assert bool(torch.isfinite(a[0]).all())
    assert float(a[0].abs().max()) > 1e-3          # the packets are still there
    assert torch.equal(a[0], b[0]), "Psi real"
    if pi is not None:
        assert torch.equal(a[1], b[1]), "Psi imag"
    assert torch.equal(a[2], b[2]), "chi"


def test_keep_factor_cleared_and_rejected_where_unsupported(gpu):
    """set_psi_keep(None) restores the plain step (same bits as an Evolver
    that never had one); a factor array of the wrong size is refused."""
    import torch
    import lfm_hip as L
    N = 32
    cfg = L.gravity_em(N)
    chi_t, keep = _barrier(N)
    pr, pi = _fields(1, N, np.ones_like(keep))
    out = []
    for use in (False, True):
        ev = L.Evolver(cfg, backend="hip")
        ev.set_psi_real(pr)
        ev.set_psi_imag(pi)
        if use:
            ev.set_psi_keep(keep)
            ev.evolve(3)
            ev.set_psi_keep(None)
            ev.set_psi_real(pr)
            ev.set_psi_imag(pi)
            ev.set_chi(np.full((N, N, N), cfg.chi0, np.float32))
            ev.step = 0
        ev.evolve(10)
        out.append(ev.native_psi_real().clone())
    assert torch.equal(out[0], out[1])
    with pytest.raises(ValueError):
        L.Evolver(cfg, backend="hip").set_psi_keep(np.ones(N ** 3 - 1, np.float32))
