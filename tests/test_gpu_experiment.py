"""GPU tests of ``gpu_snapshot_loop`` (lfm/experiment/common.py:138-280).

* Without a per-step callback the loop runs the steps between recording
  points as one evolve call (HIP-graph replay); its snapshots, metrics and
  movie frames must equal, bit for bit, a restatement of the reference's
  loop (one ``evolve(1)`` per step, host copies at the intervals).
* With a per-step source callback writing through ``sim._native_psi_real()``
  (as lfm/experiment/source.py:199-230 does), the last snapshot is checked
  against the FP64 oracle running the same injection between its steps:
  relative L2 <= 1e-5 of |Psi_0|.
* ``evolve_chi=False`` keeps chi at its start value bit for bit.
"""
import math
import sys

import numpy as np
import pytest

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-5


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ics(N, seed):
    import lfm_hip as L
    rng = np.random.default_rng(seed)
    pr = np.zeros((N,) * 3, np.float32)
    pi = np.zeros((N,) * 3, np.float32)
    for _ in range(2):
        pos = rng.uniform(0.3 * N, 0.7 * N, 3)
        r, i = L.fields.gaussian_soliton(N, tuple(pos), rng.uniform(2, 6), rng.uniform(2, 4),
                                         rng.uniform(0, 2 * np.pi))
        pr += r
        pi += i
    return pr, pi


def _reference_loop(sim, total_steps, snap_every, fields, movie_every, movie_fields,
                    evolve_chi, metrics_every):
    """The reference's loop body, one step per evolve call."""
    snaps, metrics, movie = [], [], []

    def frame(step):
        f = {"step": step}
        for name in movie_fields:
            f[name] = getattr(sim, name).copy()
        return f

    if movie_every:
        movie.append(frame(0))
    for step in range(1, total_steps + 1):
        sim.evolver.evolve(1, freeze_chi=not evolve_chi)
        if step % snap_every == 0:
            snap = {"step": step}
            for name in ("chi", "psi_real", "psi_imag", "energy_density"):
                if name in fields and getattr(sim, name) is not None:
                    snap[name] = getattr(sim, name).copy()
            snaps.append(snap)
        if metrics_every and step % metrics_every == 0:
            metrics.append({"step": step, "chi_min": float(np.min(sim.chi)),
                            "chi_max": float(np.max(sim.chi)),
                            "energy_total": float(np.sum(sim.energy_density)),
                            "psi_max": float(np.max(np.abs(sim.psi_real)))})
        if movie_every and step % movie_every == 0:
            movie.append(frame(step))
    return snaps, metrics, movie


def _assert_records_equal(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.keys() == y.keys()
        for k in x:
            if isinstance(x[k], np.ndarray):
                np.testing.assert_array_equal(x[k], y[k], err_msg=k)
            else:
                assert x[k] == y[k], k


@pytest.mark.parametrize("level,evolve_chi", [(1, True), (2, True), (0, False)])
def test_snapshot_loop_equals_stepwise_loop(gpu, level, evolve_chi):
    import lfm_hip as L
    N, T = 32, 47
    log(f"gpu_snapshot_loop vs one evolve per step, level {level}, evolve_chi={evolve_chi}")
    cfg = (L.gravity_only(N), L.gravity_em(N), L.full_physics(N))[level]
    pr, pi = _ics(N, seed=11 + level)
    kw = dict(snap_every=10, fields=["chi", "psi_real", "psi_imag", "energy_density"],
              movie_every=7, movie_fields=["psi_real", "chi", "energy_density"],
              evolve_chi=evolve_chi, metrics_every=4)
    sims = []
    for _ in range(2):
        sim = L.Simulation(cfg, backend="hip")
        sim.set_psi_real(pr if level != 2 else np.stack([pr, pr[::-1], pr[:, ::-1]]))
        if level:
            sim.set_psi_imag(pi if level != 2 else np.stack([pi, pi[::-1], pi[:, ::-1]]))
        sims.append(sim)
    chi_start = sims[0].chi.copy()
    got = L.gpu_snapshot_loop(sims[0], total_steps=T, verbose=True, label="test", **kw)
    want = _reference_loop(sims[1], T, kw["snap_every"], kw["fields"], kw["movie_every"],
                           kw["movie_fields"], evolve_chi, kw["metrics_every"])
    log(f"  snapshots {len(got[0])}, metrics {len(got[1])}, movie frames {len(got[2])}")
    assert [s["step"] for s in got[0]] == [10, 20, 30, 40]
    assert [m["step"] for m in got[1]] == list(range(4, T + 1, 4))
    assert [f["step"] for f in got[2]] == list(range(0, T + 1, 7))
    for a, b in zip(got, want):
        _assert_records_equal(a, b)
    assert sims[0].step == T
    if not evolve_chi:
        np.testing.assert_array_equal(sims[0].chi, chi_start)


def test_snapshot_loop_source_callback_matches_oracle(gpu):
    import torch
    import lfm_hip as L
    N, T = 32, 40
    log(f"gpu_snapshot_loop with a per-step source callback vs FP64 oracle, COMPLEX {N}^3")
    cfg = L.gravity_em(N)
    pr, pi = _ics(N, seed=5)
    dt, z0 = cfg.dt, 8
    zz = np.arange(N) - N / 2
    env = np.exp(-(zz[:, None] ** 2) / 18.0 - (zz[None, :] ** 2) / 18.0)
    env_d = torch.tensor(env, dtype=torch.float32, device="cuda")

    def amp(step):
        return dt * dt * 0.7 * math.sin(step * 1.3 * dt)

    def source(sim, step):
        # the reference's CW source: psi[z, :, :] += injection * env on the
        # live device buffers (lfm/experiment/source.py:199-230)
        a = amp(step)
        psi = sim._native_psi_real()
        psi[z0, :, :] += a * env_d
        sim._native_psi_imag()[z0, :, :] += 0.5 * a * env_d

    sim = L.Simulation(cfg, backend="hip")
    sim.set_psi_real(pr)
    sim.set_psi_imag(pi)
    snaps, metrics, _ = L.gpu_snapshot_loop(sim, total_steps=T, snap_every=T,
                                            fields=["psi_real", "psi_imag", "chi"],
                                            step_callback=source, verbose=False,
                                            metrics_every=10)
    st = O.OracleState(1, N, pr, pi, np.full(N ** 3, cfg.chi0),
                       O.boundary_mask(N, cfg.boundary_fraction),
                       dict(dt2=dt ** 2, kappa=cfg.kappa, lambda_self=cfg.lambda_self,
                            chi0=cfg.chi0, e0_sq=cfg.e0_sq, epsilon_w=cfg.epsilon_w))
    for step in range(1, T + 1):
        st.step(1)
        a = amp(step)
        st.r[st.cur].reshape(N, N, N)[z0] += np.float32(a) * env.astype(np.float32)
        st.i[st.cur].reshape(N, N, N)[z0] += np.float32(0.5 * a) * env.astype(np.float32)
    norm0 = float(np.linalg.norm(np.concatenate([pr.ravel(), pi.ravel()]).astype(np.float64)))
    got = np.concatenate([snaps[-1]["psi_real"].ravel(), snaps[-1]["psi_imag"].ravel()])
    e = float(np.linalg.norm(got - np.concatenate([st.psi_r, st.psi_i])) / norm0)
    ec = O.rel_l2(snaps[-1]["chi"], st.chi)
    log(f"  psi error / |psi_0| {e:.2e}, chi rel-L2 {ec:.2e}")
    assert e <= TOL64 and ec <= TOL64
    # the metrics are the reference's float32 reductions of the host fields
    assert [m["step"] for m in metrics] == [10, 20, 30, 40]
    last = metrics[-1]
    e2 = snaps[-1]["psi_real"] ** 2 + snaps[-1]["psi_imag"] ** 2
    assert last["energy_total"] == float(np.sum(e2))
    assert last["psi_max"] == float(np.max(np.abs(snaps[-1]["psi_real"])))
    assert last["chi_min"] == float(np.min(snaps[-1]["chi"]))
    assert last["chi_max"] == float(np.max(snaps[-1]["chi"]))


def test_snapshot_loop_device_hook_equals_stepwise(gpu):
    """A device hook (a sine-driven plane source on the live Psi buffer) in
    the block-replayed loop against one hooked evolve(1) per step: the same
    bits."""
    import torch
    import lfm_hip as L
    N, T = 32, 30
    log(f"gpu_snapshot_loop with a device hook vs one hooked evolve per step, COMPLEX {N}^3")
    cfg = L.gravity_em(N)
    pr, pi = _ics(N, seed=21)
    zz = np.arange(N) - N / 2
    env = torch.tensor(np.exp(-(zz[:, None] ** 2 + zz[None, :] ** 2) / 18.0),
                       dtype=torch.float32, device="cuda")
    dt = cfg.dt

    def hook(ev, step):
        amp = (dt * dt * 0.7) * torch.sin(step.float() * (1.3 * dt))
        ev.native_psi_real().view(N, N, N)[8] += amp * env

    sims = []
    for _ in range(2):
        sim = L.Simulation(cfg, backend="hip")
        sim.set_psi_real(pr)
        sim.set_psi_imag(pi)
        sims.append(sim)
    snaps, _, _ = L.gpu_snapshot_loop(sims[0], total_steps=T, snap_every=10,
                                      fields=["psi_real", "psi_imag", "chi"], verbose=False,
                                      device_hook=hook)
    want = []
    for step in range(1, T + 1):
        sims[1].evolver.evolve(1, device_hook=hook)
        if step % 10 == 0:
            want.append({"step": step, "chi": sims[1].chi.copy(),
                         "psi_real": sims[1].psi_real.copy(),
                         "psi_imag": sims[1].psi_imag.copy()})
    _assert_records_equal(snaps, want)

