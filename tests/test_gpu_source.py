"""The plane source folded into the step (``ContinuousSource.fold``,
``Evolver.set_plane_source``, LFM_F_SOURCE) against the form it replaces.

The reference's continuous source (lfm/experiment/source.py:196-228) is a
per-step callback that adds ``amount * envelope`` to the current Psi on one
plane through the live device buffers.  Folded into the step kernel, the same
additions happen as the step writes Psi, with each step's amounts read from a
device table at the step counter, so the steps keep running as replayed
graphs.  Both forms must give the same bits:

* the C entry points against the plain step plus the addition done with
  torch float32 operations (every axis, both components, the table bounds);
* ``Simulation.run_with_snapshots`` with the callback against the folded
  source, REAL and COMPLEX, every axis, frozen and periodic boundaries, over
  several calls that cross the amount table's growth;
* with a barrier's per-step enforcement (chi template in all four buffers,
  Psi multiplied by keep factors), the callback pair against the folded
  source with ``set_psi_keep`` and frozen chi.
"""
import ctypes
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ics(level, N, seed):
    import lfm_hip as L
    rng = np.random.default_rng(seed)
    pr = np.zeros((N,) * 3, np.float32)
    pi = np.zeros((N,) * 3, np.float32)
    for _ in range(2):
        pos = rng.uniform(0.35 * N, 0.65 * N, 3)
        r, i = L.fields.gaussian_soliton(N, tuple(pos), rng.uniform(2, 5), rng.uniform(2, 4),
                                         rng.uniform(0, 2 * np.pi))
        pr += r
        pi += i
    return pr, (pi if level else None)


@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_source_flag_equals_plain_step_plus_addition(gpu, level, axis):
    """One step with LFM_F_SOURCE equals the plain step followed by
    psi[plane] += amount * envelope in torch float32 (product and sum each
    rounded once); amounts come from the table at counter + step_off, and
    indexes outside the table add nothing.  Tiled and untiled kernels."""
    import torch
    from lfm_hip import _abi
    lib = _abi.load()
    N, plane = 40, 13
    log(f"source flag vs plain step + torch addition, level {level}, axis {axis}")
    g = torch.Generator(device="cuda").manual_seed(31 + 3 * level + axis)
    f = lambda s: s * torch.randn((N, N, N), generator=g, device="cuda")
    pr, prp, chi, chip = f(0.3), f(0.3), f(0.01), f(0.01)
    pi, pip = (f(0.3), f(0.3)) if level else (None, None)
    env = torch.rand((N, N), generator=g, device="cuda") + 0.5
    table = torch.randn(2 * 64, generator=g, device="cuda") * 1e-2
    table[2 * 9] = 0.0                       # step 9: no real-part addition
    counter = torch.tensor(7, dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: t.data_ptr() if t is not None else None

    def step(flags, off=0, length=64, direct=False):
        p = _abi.LfmParams()
        p.N, p.dt2, p.kappa, p.lambda_self, p.chi0, p.epsilon_w = N, 4e-4, 1 / 63, 4 / 31, 19.0, 0.1
        p.flags = flags | _abi.LFM_F_CHI_DELTA | (_abi.LFM_F_DIRECT if direct else 0)
        p.src_env, p.src_amp, p.step_counter = env.data_ptr(), table.data_ptr(), counter.data_ptr()
        p.src_len, p.src_axis, p.src_plane, p.step_off = length, axis, plane, off
        o = [torch.empty_like(pr), torch.empty_like(pr) if level else None, torch.empty_like(chi)]
        if level == 0:
            rc = lib.lfm_step_real_f32(ptr(pr), ptr(prp), ptr(chi), ptr(chip), None, ptr(o[0]),
                                       None, ptr(o[2]), None, ctypes.byref(p), st)
        else:
            rc = lib.lfm_step_complex_f32(ptr(pr), ptr(prp), ptr(pi), ptr(pip), ptr(chi),
                                          ptr(chip), None, ptr(o[0]), None, ptr(o[1]), None,
                                          ptr(o[2]), None, ctypes.byref(p), st)
        _abi.check(rc, "step")
        return o

    def on_plane(t):
        return t[plane] if axis == 0 else t[:, plane] if axis == 1 else t[:, :, plane]

    for direct in (False, True):
        plain = step(0, direct=direct)
        for off in (1, 2, 60, -8, 57):       # indexes 8, 9, 67 (past the end), -1, 64
            got = step(_abi.LFM_F_SOURCE, off=off, direct=direct)
            s = 7 + off
            want = [t.clone() if t is not None else None for t in plain]
            if 0 <= s < 64:
                for c in range(2 if level else 1):
                    a = table[2 * s + c]
                    if float(a) != 0.0:
                        view = on_plane(want[c])
                        view.copy_(view + a * env)
            for c, (x, y) in enumerate(zip(got, want)):
                if x is not None:
                    assert torch.equal(x, y), (direct, off, c)
            # the real part changes only at index 8 (9 has a zero real amount)
            assert (not torch.equal(got[0], plain[0])) == (s == 8), (direct, off)
        torch.cuda.synchronize()
    log("  equal for offsets inside and outside the table, tiled and untiled")


def _source_run(level, axis, boundary, fold, steps, phase, seed):
    import lfm_hip as L
    N = 32
    mk = (L.gravity_only, L.gravity_em)[level]
    cfg = mk(N, boundary_type=L.BoundaryType.PERIODIC) if boundary == "periodic" else mk(N)
    pr, pi = _ics(level, N, seed)
    sim = L.Simulation(cfg, backend="hip")
    sim.set_psi_real(pr)
    if pi is not None:
        sim.set_psi_imag(pi)
    src = L.ContinuousSource(sim, axis=axis, position=0.25, omega=21.0, amplitude=2.0,
                             envelope_sigma=0.25, phase=phase, boost=10.0)
    if fold:
        src.fold()
    snaps = []
    for n, every in steps:
        fields = ["psi_real", "psi_imag", "chi"]
        if fold:
            snaps += sim.run_with_snapshots(n, snapshot_every=every, fields=fields,
                                            record_metrics=False)
        else:
            snaps += sim.run_with_snapshots(n, snapshot_every=every, fields=fields,
                                            step_callback=src.step_callback, record_metrics=False)
    return sim, snaps


@pytest.mark.parametrize("level,axis,boundary,phase", [
    (0, 2, "frozen", 0.0), (0, 0, "periodic", 0.0),
    (1, 0, "frozen", 0.8), (1, 1, "periodic", 0.8), (1, 2, "frozen", 0.0)])
def test_folded_source_equals_callback(gpu, level, axis, boundary, phase):
    """Simulation.run_with_snapshots with the reference's callback against the
    folded source: every snapshot equal bit for bit over three calls (the last
    one past the amount table's first 1024 steps, so the table grows and the
    graphs are captured again)."""
    import torch
    steps = [(120, 50), (700, 350), (330, 330)]
    log(f"folded source vs callback, level {level}, axis {axis}, {boundary}, phase {phase}")
    t0 = time.perf_counter()
    a_sim, a = _source_run(level, axis, boundary, False, steps, phase, seed=3 + axis)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    b_sim, b = _source_run(level, axis, boundary, True, steps, phase, seed=3 + axis)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"  callback {(t1 - t0) / 1150 * 1e6:.0f} us per step, folded "
        f"{(t2 - t1) / 1150 * 1e6:.0f} us per step (incl. snapshots and captures)")
    assert [s["step"] for s in a] == [s["step"] for s in b] == [50, 100, 120, 470, 820, 1150]
    for x, y in zip(a, b):
        for k in ("psi_real", "psi_imag", "chi"):
            if k in x:
                np.testing.assert_array_equal(x[k], y[k], err_msg=f"{k} at step {x['step']}")
    # the source did inject: the source plane moved away from the run without it
    assert float(np.abs(a[-1]["psi_real"]).max()) > 1e-3
    assert b_sim.evolver._src["table"].numel() // 2 > 1024


def test_folded_source_with_barrier_equals_callbacks(gpu):
    """A source and a barrier (chi template held in all four buffers, Psi
    multiplied by keep factors after every step), as the reference's two
    callbacks in that order, against the folded source with set_psi_keep and
    frozen chi: the same bits (COMPLEX 48^3, 200 steps)."""
    import torch
    import lfm_hip as L
    N, T = 48, 200
    log(f"folded source + keep factors vs source and barrier callbacks, COMPLEX {N}^3")
    cfg = L.gravity_em(N)
    x0, x1 = N // 2 - 2, N // 2 + 2
    chi_t = np.full((N, N, N), 19.0, np.float32)
    keep = np.ones((N, N, N), np.float32)
    chi_t[:, :, x0:x1] = 25.0
    keep[:, :, x0:x1] = 0.0
    for s0, k in ((N // 2 - 7, 0.75), (N // 2 + 4, 1.0)):
        chi_t[:, s0:s0 + 3, x0:x1] = 19.0
        keep[:, s0:s0 + 3, x0:x1] = k
    pr, pi = _ics(1, N, seed=9)
    res = {}
    for form in ("callbacks", "folded"):
        sim = L.Simulation(cfg, backend="hip")
        sim.set_psi_real(pr * keep)
        sim.set_psi_imag(pi * keep)
        sim.set_chi(chi_t)
        ev = sim.evolver
        src = L.ContinuousSource(sim, axis=2, position=0.25, omega=21.0, amplitude=2.0,
                                 envelope_sigma=0.2, phase=0.5, boost=10.0)
        if form == "callbacks":
            ev.use_absolute_chi()
            chi_d = torch.from_numpy(chi_t.reshape(-1)).to(ev.device)
            k_d = torch.from_numpy(keep.reshape(-1)).to(ev.device)

            def both(s, step):
                src.step_callback(s, step)
                for buf in ev.native_chi_refs():
                    buf.copy_(chi_d)
                ev.native_psi_real().mul_(k_d)
                ev.native_psi_imag().mul_(k_d)
            snaps = sim.run_with_snapshots(T, snapshot_every=T, fields=["psi_real", "psi_imag",
                                                                        "chi"],
                                           step_callback=both, record_metrics=False)
        else:
            ev.set_psi_keep(keep)
            src.fold()
            snaps = sim.run_with_snapshots(T, snapshot_every=T, fields=["psi_real", "psi_imag",
                                                                        "chi"],
                                           record_metrics=False, evolve_chi=False)
        res[form] = snaps[-1]
    a, b = res["callbacks"], res["folded"]
    assert float(np.abs(a["psi_real"]).max()) > 1e-3
    for k in ("psi_real", "psi_imag", "chi"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_folded_source_cost(gpu):
    """Per-step time of a 64^3 COMPLEX run: plain, with the folded source,
    with the same source as a graph-captured device hook, and with the
    reference's host callback (logged; the folded source must stay within
    a small margin of the plain step)."""
    import torch
    import lfm_hip as L
    N, T = 64, 600
    log(f"plane source cost, COMPLEX {N}^3, {T} steps")
    cfg = L.gravity_em(N)
    pr, pi = _ics(1, N, seed=2)
    out = {}
    for form in ("plain", "folded", "device hook", "callback"):
        sim = L.Simulation(cfg, backend="hip")
        sim.set_psi_real(pr)
        sim.set_psi_imag(pi)
        ev = sim.evolver
        src = L.ContinuousSource(sim, axis=2, position=0.25, omega=21.0, phase=0.5)
        kw = {}
        if form == "folded":
            src.fold()
        elif form == "device hook":
            env = torch.from_numpy(src.envelope).cuda()
            dt = cfg.dt

            def hook(e, step):
                a = (dt * dt * 2.0 * 10.0) * torch.sin(step.double() * (21.0 * dt))
                e.native_psi_real().view(N, N, N)[:, :, src.position] += a.float() * env
                e.native_psi_imag().view(N, N, N)[:, :, src.position] += \
                    (a * np.sin(0.5)).float() * env
            kw = {"device_hook": hook}
        elif form == "callback":
            kw = {"step_callback": src.step_callback}
        runs = (lambda n: sim.run_with_snapshots(n, snapshot_every=n, record_metrics=False, **kw))
        runs(60)                       # warm-up and graph captures
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        runs(T)
        torch.cuda.synchronize()
        out[form] = (time.perf_counter() - t0) / T * 1e6
        log(f"  {form:12s} {out[form]:8.1f} us per step")
    assert out["folded"] < out["device hook"]
