"""The detector plane folded into the step (``DetectorScreen.fold``,
``Evolver.set_plane_detector``, LFM_F_DETECTOR) against the form it replaces.

The reference's detector screen (lfm/experiment/detector.py:97-115) stores the
plane of |Psi|^2 (or another field) each time ``record()`` is called, usually
as a per-step callback.  Folded into the step kernel, the kernel writes the
plane of the new Psi into a device frame buffer on the recorded steps.  Both
must give the same frames bit for bit:

* the C entry points against the plain step with the plane cut and squared
  in torch float32 (every axis, both fields, interval, first step and the
  buffer bounds; untouched frames keep their contents);
* ``Simulation.run_with_snapshots`` with a source, a barrier and the screen
  as the reference's callbacks against the three folded into the steps, at
  REAL and COMPLEX, recording every step and every 5th step;
* the same with the experiment classes from the Simulation factories
  (``place_barrier`` with a which-path detector slit, ``add_source``,
  ``add_detector``) and their ``fold()``.
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


@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_detector_flag_equals_plain_step_plane(gpu, level, axis):
    """One step with LFM_F_DETECTOR writes frame (s - first) / every of the
    plane of Pr'^2 + Pi'^2 (or Pr'), equal to the plain step's plane squared
    in torch float32; steps off the interval, before `first` or past the
    buffer write nothing, and the step's own outputs are unchanged.  Tiled
    and untiled kernels."""
    import torch
    from lfm_hip import _abi
    lib = _abi.load()
    N, plane = 40, 29
    log(f"detector flag vs plain step plane, level {level}, axis {axis}")
    g = torch.Generator(device="cuda").manual_seed(41 + 3 * level + axis)
    f = lambda s: s * torch.randn((N, N, N), generator=g, device="cuda")
    pr, prp, chi, chip = f(0.3), f(0.3), f(0.01), f(0.01)
    pi, pip = (f(0.3), f(0.3)) if level else (None, None)
    counter = torch.tensor(20, dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: t.data_ptr() if t is not None else None

    def step(flags, frames=None, off=0, first=0, every=1, field=0, direct=False):
        p = _abi.LfmParams()
        p.N, p.dt2, p.kappa, p.lambda_self, p.chi0, p.epsilon_w = N, 4e-4, 1 / 63, 4 / 31, 19.0, 0.1
        p.flags = flags | _abi.LFM_F_CHI_DELTA | (_abi.LFM_F_DIRECT if direct else 0)
        p.step_counter, p.step_off = counter.data_ptr(), off
        if frames is not None:
            p.det_frames, p.det_len = frames.data_ptr(), frames.shape[0]
        p.det_axis, p.det_plane, p.det_every, p.det_field, p.det_first = axis, plane, every, field, first
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
        for field in (0, 1):
            want_plane = on_plane(plain[0]) * on_plane(plain[0])
            if level and field == 0:
                want_plane = want_plane + on_plane(plain[1]) * on_plane(plain[1])
            if field == 1:
                want_plane = on_plane(plain[0]).clone()
            # (step_off, first, every) -> frame written or None; counter 20
            for off, first, every, frame in ((1, 21, 1, 0), (3, 15, 2, 4), (2, 15, 2, None),
                                             (1, 22, 1, None), (1, 5, 4, 4), (1, 1, 1, None)):
                frames = torch.full((5, N, N), 7.0, device="cuda")
                got = step(_abi.LFM_F_DETECTOR, frames, off, first, every, field, direct)
                for x, y in zip(got, plain):
                    if x is not None:
                        assert torch.equal(x, y), "step outputs changed"
                for q in range(5):
                    if q == frame:
                        assert torch.equal(frames[q], want_plane), (direct, field, off, first, q)
                    else:
                        assert bool((frames[q] == 7.0).all()), (direct, field, off, first, q)
    torch.cuda.synchronize()
    log("  frames equal the plane of the plain step; other frames untouched")


def _experiment(level, fold, steps, every, seed):
    """A source, a barrier (chi template in all four buffers, Psi multiplied
    by keep factors) and a screen, as the reference's callbacks in that order
    or folded into the steps."""
    import torch
    import lfm_hip as L
    N = 40
    cfg = (L.gravity_only, L.gravity_em)[level](N)
    x0, x1 = N // 2 - 2, N // 2 + 2
    chi_t = np.full((N, N, N), 19.0, np.float32)
    keep = np.ones((N, N, N), np.float32)
    chi_t[:, :, x0:x1] = 25.0
    keep[:, :, x0:x1] = 0.0
    for s0 in (N // 2 - 6, N // 2 + 3):
        chi_t[:, s0:s0 + 3, x0:x1] = 19.0
        keep[:, s0:s0 + 3, x0:x1] = 1.0
    rng = np.random.default_rng(seed)
    pr = (0.05 * rng.standard_normal((N, N, N))).astype(np.float32) * keep
    pi = (0.05 * rng.standard_normal((N, N, N))).astype(np.float32) * keep
    sim = L.Simulation(cfg, backend="hip")
    sim.set_psi_real(pr)
    if level:
        sim.set_psi_imag(pi)
    sim.set_chi(chi_t)
    ev = sim.evolver
    src = L.ContinuousSource(sim, axis=2, position=0.2, omega=21.0, amplitude=2.0,
                             envelope_sigma=0.2, phase=0.5)
    screen = L.DetectorScreen(sim, axis=2, position=int(N * 0.8))
    if fold:
        ev.set_psi_keep(keep)
        src.fold()
        screen.fold(every)
        for n in steps:
            sim.run_with_snapshots(n, snapshot_every=n, record_metrics=False, evolve_chi=False)
    else:
        ev.use_absolute_chi()
        chi_d = torch.from_numpy(chi_t.reshape(-1)).to(ev.device)
        k_d = torch.from_numpy(keep.reshape(-1)).to(ev.device)

        def callbacks(s, step):
            src.step_callback(s, step)
            for buf in ev.native_chi_refs():
                buf.copy_(chi_d)
            ev.native_psi_real().mul_(k_d)
            if level:
                ev.native_psi_imag().mul_(k_d)
            if step % every == 0:
                screen.record()
        for n in steps:
            sim.run_with_snapshots(n, snapshot_every=n, step_callback=callbacks,
                                   record_metrics=False)
    return sim, screen


@pytest.mark.parametrize("level,every", [(0, 1), (1, 1), (1, 5)])
def test_folded_detector_equals_record(gpu, level, every):
    """Source, barrier and screen folded into the steps against the
    reference's three callbacks: every frame, the pattern and the fields equal
    bit for bit over two run calls (200 + 97 steps)."""
    import torch
    steps = (200, 97)
    log(f"folded source + barrier + screen vs callbacks, level {level}, every {every}")
    t0 = time.perf_counter()
    a_sim, a = _experiment(level, False, steps, every, seed=7)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    b_sim, b = _experiment(level, True, steps, every, seed=7)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"  callbacks {(t1 - t0) / sum(steps) * 1e6:.0f} us per step, folded "
        f"{(t2 - t1) / sum(steps) * 1e6:.0f} us per step (incl. graph captures)")
    assert a.n_frames == b.n_frames == sum(steps) // every
    np.testing.assert_array_equal(a.snapshots, b.snapshots)
    np.testing.assert_array_equal(a.pattern, b.pattern)
    assert float(a.pattern.max()) > 0.0
    np.testing.assert_array_equal(a_sim.psi_real, b_sim.psi_real)
    np.testing.assert_array_equal(a_sim.chi, b_sim.chi)
    # unfold keeps the frames; reset starts again
    b.unfold()
    assert b.n_frames == a.n_frames
    np.testing.assert_array_equal(a.mean_pattern(), b.mean_pattern())
    np.testing.assert_array_equal(a.line_profile(1), b.line_profile(1))
    b.fold(every)
    b.reset()
    assert b.n_frames == 0
    b_sim.run(2 * every)
    assert b.n_frames == 2


@pytest.mark.parametrize("level", [0, 1])
def test_folded_experiment_classes_equal_callbacks(gpu, level):
    """The reference's double-slit set-up through the Simulation factories:
    place_barrier (two slits, the second with a which-path detector),
    add_source and add_detector, run with the three callbacks in the
    reference's order (source, barrier, screen) against the three folded into
    the steps: frames, pattern and fields equal bit for bit (48^3, 240 steps
    in two run calls)."""
    import torch
    import lfm_hip as L
    N = 48
    log(f"Barrier + ContinuousSource + DetectorScreen folded vs callbacks, level {level}")
    out = {}
    for fold in (False, True):
        cfg = (L.gravity_only, L.gravity_em)[level](N)
        sim = L.Simulation(cfg, backend="hip")
        rng = np.random.default_rng(3)
        sim.set_psi_real((0.02 * rng.standard_normal((N, N, N))).astype(np.float32))
        if level:
            sim.set_psi_imag((0.02 * rng.standard_normal((N, N, N))).astype(np.float32))
        barrier = sim.place_barrier(axis=2, slits=[L.Slit(center=18, width=3),
                                                   L.Slit(center=30, width=3, detector=True,
                                                          detector_strength=0.9)])
        source = sim.add_source(axis=2, position=0.2, omega=21.0, phase=0.7)
        screen = sim.add_detector(axis=2, position=40)
        t0 = time.perf_counter()
        if fold:
            source.fold()
            barrier.fold()
            screen.fold()
            for n in (160, 80):
                sim.run_with_snapshots(n, snapshot_every=n, record_metrics=False)
        else:
            def callbacks(s, step):
                source.step_callback(s, step)
                barrier.step_callback(s, step)
                screen.record()
            for n in (160, 80):
                sim.run_with_snapshots(n, snapshot_every=n, step_callback=callbacks,
                                       record_metrics=False)
        torch.cuda.synchronize()
        log(f"  {'folded   ' if fold else 'callbacks'} {(time.perf_counter() - t0) / 240 * 1e6:.0f} "
            f"us per step")
        out[fold] = (screen.snapshots, screen.pattern, sim.psi_real, sim.chi,
                     barrier.measure_slits())
    a, b = out[False], out[True]
    assert a[0].shape == (240, N, N)
    for x, y, name in zip(a[:4], b[:4], ("frames", "pattern", "psi_real", "chi")):
        np.testing.assert_array_equal(x, y, err_msg=name)
    assert a[4] == b[4]
    assert float(a[1].max()) > 0.0
