"""GPU tests of the batched form (LatticeBatch / batched sweep).

* Every lattice of a batch must equal the same lattice run alone through
  Evolver bit for bit: both use the same per-site arithmetic (the batched
  kernel only changes which workgroup computes which tile).
* HIP-graph replays must equal eager launches bit for bit over several
  replays with eager work in between.
* The batch is also checked against the FP64 oracle (relative L2 <= 1e-5).
"""
import sys

import numpy as np
import pytest

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ics(level, N, rng):
    import lfm_hip as L
    c = N / 2
    amp, sig, ph = rng.uniform(2, 6), rng.uniform(2, 3), rng.uniform(0, 6.28)
    if level == 2:
        pos = [tuple(c + rng.uniform(-2, 2, 3)) for _ in range(3)]
        return L.fields.place_solitons(N, pos, amp, sig, phases=rng.uniform(0, 6.28, 3),
                                       colors=[0, 1, 2])
    r, i = L.fields.gaussian_soliton(N, (c, c - 1, c + 0.5), amp, sig, ph)
    nr, ni = L.fields.seed_noise(N, 1e-3, rng=rng)
    return r + nr, (i + ni if level == 1 else None)


def _configs(level, sizes, rng):
    import lfm_hip as L
    mk = {0: L.gravity_only, 1: L.gravity_em,
          2: lambda n, **k: L.full_physics(n, kappa_tube=0.0, **k)}[level]
    return [mk(N, kappa=(1 / 63) * rng.uniform(0.5, 2)) for N in sizes]


@pytest.mark.parametrize("level,group_bytes,cycles",
                         [(0, None, 1), (1, None, 1), (2, None, 1), (0, 1 << 21, 2),
                          (1, 1 << 21, 1), (2, 1 << 20, 2)])
def test_batch_equals_single_lattice_runs(gpu, level, group_bytes, cycles):
    import torch
    import lfm_hip as L
    sizes = [24, 40, 16, 72, 33] if level < 2 else [16, 24, 40]
    log(f"batch vs single level {level} sizes {sizes} group_bytes {group_bytes} cycles {cycles}")
    rng = np.random.default_rng(11 + level)
    cfgs = _configs(level, sizes, rng)
    ics = [_ics(level, N, rng) for N in sizes]
    batch = L.LatticeBatch(cfgs, backend="hip", group_bytes=group_bytes, cycles=cycles)
    if group_bytes is not None:
        assert len(batch.groups) >= 3, batch.groups
    singles = []
    for b, (c, (r, i)) in enumerate(zip(cfgs, ics)):
        batch.set_psi_real(b, r)
        ev = L.Evolver(c, backend="hip")
        ev.set_psi_real(r)
        if i is not None:
            batch.set_psi_imag(b, i)
            ev.set_psi_imag(i)
        singles.append(ev)
    steps = 40
    batch.evolve(steps)
    for ev in singles:
        ev.evolve(steps)
    torch.cuda.synchronize()
    for b, ev in enumerate(singles):
        np.testing.assert_array_equal(batch.get_psi_real(b), ev.get_psi_real(), err_msg=f"psi_r {b}")
        np.testing.assert_array_equal(batch.get_chi(b), ev.get_chi(), err_msg=f"chi {b}")
        if level:
            np.testing.assert_array_equal(batch.get_psi_imag(b), ev.get_psi_imag())
    # one lattice against the FP64 oracle
    b = 1
    c = cfgs[b]
    p = dict(dt2=c.dt ** 2, kappa=c.kappa, lambda_self=c.lambda_self, chi0=c.chi0,
             e0_sq=c.e0_sq, epsilon_w=c.epsilon_w, kappa_c=c.kappa_c, epsilon_cc=c.epsilon_cc,
             kappa_string=c.kappa_string, kappa_tube=0.0)
    st = O.OracleState(level, c.grid_size, ics[b][0], ics[b][1], np.full(c.grid_size ** 3, 19.0),
                       O.boundary_mask(c.grid_size, c.boundary_fraction), p)
    st.step(steps)
    e = O.rel_l2(batch.get_chi(b), st.chi)
    e2 = O.rel_l2(batch.get_psi_real(b), st.psi_r)
    log(f"  lattice {b} vs FP64 oracle: chi {e:.2e} psi_r {e2:.2e}")
    assert e <= 1e-5 and e2 <= 1e-5


@pytest.mark.parametrize("group_bytes,cycles", [(None, 1), (1 << 20, 2)])
def test_graph_replay_equals_eager(gpu, group_bytes, cycles):
    import torch
    import lfm_hip as L
    log(f"graph replay vs eager, group_bytes {group_bytes} cycles {cycles}")
    rng = np.random.default_rng(3)
    sizes = [32, 48, 20]
    cfgs = _configs(1, sizes, rng)
    ics = [_ics(1, N, rng) for N in sizes]
    out = []
    for graph in (True, False):
        batch = L.LatticeBatch(cfgs, backend="hip", group_bytes=group_bytes, cycles=cycles)
        for b, (r, i) in enumerate(ics):
            batch.set_psi_real(b, r)
            batch.set_psi_imag(b, i)
        for rep in range(4):
            batch.evolve(10 if cycles == 1 else 14, graph=graph)   # eager steps + replays
            torch.cuda.synchronize()
            _ = batch.get_chi(0).sum()       # eager host work between replays
        out.append([batch.get_psi_real(b) for b in range(3)] + [batch.get_chi(b) for b in range(3)])
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("color,report,steps", [(False, 5000, 30), (True, 5000, 30),
                                                (False, 20, 30), (False, 15, 30)])
def test_batched_sweep_rows_match_simulation(gpu, color, report, steps):
    """Every row of the batched sweep equals what the reference's sweep loop
    builds for that value (lfm/sweep.py:53-78): a Simulation with the soliton
    at the centre, equilibrate(), run(steps, record_metrics=False), then
    metrics().  Its kinetic term is taken against the Psi snapshot of run
    start (report_interval 5000 > steps), of step 20 (report_interval 20), or
    of step 30 itself (report_interval 15: the last snapshot is the final
    state, so the kinetic energy is 0), as Simulation.run keeps it.  The
    kinetic="leapfrog" rows equal the true-prev energy sums instead."""
    import lfm_hip as L
    from lfm_hip.batch import sweep
    log(f"batched sweep vs per-value Simulation ({'COLOR full_physics' if color else 'COMPLEX'}),"
        f" report_interval {report}, {steps} steps")
    N = 16 if color else 24
    mk = (lambda **k: L.full_physics(N, report_interval=report, **k)) if color else \
        (lambda **k: L.SimulationConfig(grid_size=N, field_level=L.FieldLevel.COMPLEX,
                                        report_interval=report, **k))
    base = mk()
    vals = [1 / 63, 1.5 / 63, 2 / 63]
    rows = sweep(base, "kappa", vals, steps, soliton={"amplitude": 4.0}, backend="hip")
    rows_lf = sweep(base, "kappa", vals, steps, soliton={"amplitude": 4.0}, backend="hip",
                    kinetic="leapfrog", metric_names=["energy_total", "chi_min"])
    assert [r["kappa"] for r in rows] == vals
    for v, row, row_lf in zip(vals, rows, rows_lf):
        cfg = mk(kappa=v)
        sim = L.Simulation(cfg, backend="hip")
        sim.place_soliton((N // 2, N // 2, N // 2), amplitude=4.0)
        sim.equilibrate()
        sim.run(steps, record_metrics=False)
        m = sim.metrics()
        assert set(row) == {"kappa"} | set(m)
        for k, x in m.items():
            assert row[k] == x, (k, row[k], x)
        if report == 15:
            assert row["energy_kinetic"] == 0.0
        e = sim.evolver.energy_sums()
        assert row_lf["energy_total"] == e["energy_total"]
        assert row_lf["chi_min"] == m["chi_min"]
        log(f"  kappa {v:.5f}: E_total {row['energy_total']:.6e} (snapshot) "
            f"{row_lf['energy_total']:.6e} (leapfrog)")


@pytest.mark.parametrize("sizes,group_bytes", [([16, 32, 24], None), ([32, 16, 64, 32], 1 << 22)])
def test_batch_color_with_sa_equals_single_runs(gpu, sizes, group_bytes):
    """COLOR full_physics lattices with the S_a transform (kappa_tube > 0):
    per-lattice transforms between batched launches give every lattice
    exactly its single-lattice Evolver run (24^3 uses a rocFFT plan and stays
    eager; the power-of-two batch replays graphs, in groups for the second
    case)."""
    import torch
    import lfm_hip as L
    log(f"batch COLOR + S_a vs single, sizes {sizes} group_bytes {group_bytes}")
    rng = np.random.default_rng(23)
    cfgs = [L.full_physics(N, kappa=(1 / 63) * rng.uniform(0.5, 2)) for N in sizes]
    assert all(c.kappa_tube > 0 for c in cfgs)
    ics = [_ics(2, N, rng) for N in sizes]
    batch = L.LatticeBatch(cfgs, backend="hip", group_bytes=group_bytes)
    singles = []
    for b, (c, (r, i)) in enumerate(zip(cfgs, ics)):
        batch.set_psi_real(b, r)
        batch.set_psi_imag(b, i)
        ev = L.Evolver(c, backend="hip")
        ev.set_psi_real(r)
        ev.set_psi_imag(i)
        singles.append(ev)
    for n in (10, 14):
        batch.evolve(n)
        for ev in singles:
            ev.evolve(n)
        torch.cuda.synchronize()
        for b, ev in enumerate(singles):
            np.testing.assert_array_equal(batch.get_psi_real(b), ev.get_psi_real(), err_msg=f"psi_r {b}")
            np.testing.assert_array_equal(batch.get_psi_imag(b), ev.get_psi_imag(), err_msg=f"psi_i {b}")
            np.testing.assert_array_equal(batch.get_chi(b), ev.get_chi(), err_msg=f"chi {b}")


def test_batched_sweep_2d_rows(gpu):
    """sweep_2d (lfm/sweep.py:85-157): one row per (value1, value2) pair in
    the reference's order, each equal to its own batched lattice run alone."""
    import lfm_hip as L
    from lfm_hip.batch import sweep, sweep_2d
    log("batched sweep_2d")
    base = L.SimulationConfig(grid_size=16, field_level=L.FieldLevel.COMPLEX)
    k_vals, l_vals = [1 / 63, 2 / 63], [0.0, 4 / 31]
    rows = sweep_2d(base, "kappa", k_vals, "lambda_self", l_vals, 20, soliton={"amplitude": 3.0},
                    backend="hip", metric_names=["chi_min", "energy_total"])
    assert [(r["kappa"], r["lambda_self"]) for r in rows] == [(k, l) for k in k_vals for l in l_vals]
    assert all(set(r) == {"kappa", "lambda_self", "chi_min", "energy_total"} for r in rows)
    for r in rows:
        from copy import deepcopy
        c = deepcopy(base)
        c.lambda_self = r["lambda_self"]
        one = sweep(c, "kappa", [r["kappa"]], 20, soliton={"amplitude": 3.0}, backend="hip",
                    metric_names=["chi_min", "energy_total"])[0]
        assert one["chi_min"] == r["chi_min"] and one["energy_total"] == r["energy_total"]


@pytest.mark.parametrize("level,group_bytes,cycles,tile,periodic",
                         [(0, None, 1, 0, False), (1, None, 1, 0, False), (1, None, 2, 4, True),
                          (1, 1 << 21, 1, 0, False), (0, 1 << 21, 2, 4, True)])
def test_two_step_batch_equals_single_step_batch(gpu, level, group_bytes, cycles, tile, periodic):
    """Pairs of steps per launch (lfm_batch_step2_f32, four-slot rotation)
    against the one-step batch: ragged sizes, a mid-run setter, odd step
    counts, graph replay and eager launches, lattice groups; every lattice bit
    for bit (Psi, chi, and the leapfrog prev through metrics' kinetic term)."""
    import torch
    import lfm_hip as L
    sizes = [24, 40, 16, 72, 33, 20]
    log(f"two-step batch vs one-step batch, level {level}, groups {group_bytes}, cycles {cycles}, "
        f"tile {tile}, periodic {periodic}")
    rng = np.random.default_rng(31 + level)
    cfgs = _configs(level, sizes, rng)
    if periodic:
        for c in cfgs[::2]:
            c.boundary_type = L.BoundaryType.PERIODIC
    ics = [_ics(level, N, rng) for N in sizes]
    batches = {}
    for two in (True, False):
        bt = L.LatticeBatch(cfgs, backend="hip", group_bytes=group_bytes, cycles=cycles,
                            two_step=two, tile=tile if two else 0,
                            mask_policy="boundary_type" if periodic else "reference")
        assert bt.two_step == two
        for b, (r, i) in enumerate(ics):
            bt.set_psi_real(b, r)
            if i is not None:
                bt.set_psi_imag(b, i)
        batches[two] = bt
    for n, graph in ((7, True), (10, False), (33, True), (1, True), (24, True)):
        for bt in batches.values():
            bt.evolve(n, graph=graph)
        if n == 10:
            for bt in batches.values():
                bt.set_psi_real(2, bt.get_psi_real(2) * 0.99)
        torch.cuda.synchronize()
        a, b2 = batches[True], batches[False]
        for b in range(len(sizes)):
            np.testing.assert_array_equal(a.get_psi_real(b), b2.get_psi_real(b), err_msg=f"psi_r {b}")
            np.testing.assert_array_equal(a.get_chi(b), b2.get_chi(b), err_msg=f"chi {b}")
            if level:
                np.testing.assert_array_equal(a.get_psi_imag(b), b2.get_psi_imag(b))
        ka = [a.metrics(b, clusters=False)["energy_kinetic"] for b in range(len(sizes))]
        kb = [b2.metrics(b, clusters=False)["energy_kinetic"] for b in range(len(sizes))]
        assert ka == kb
    if group_bytes is not None:
        assert len(batches[True].groups) >= 3
    used = [k for k, v in batches[True]._graphs.items() if v is not None]
    log(f"  two-step graphs captured: {len(used)}")
    assert used


@pytest.mark.parametrize("group_bytes", [None, 1 << 22])
def test_batch_autotune_kernel_choice_keeps_results(gpu, group_bytes):
    """LatticeBatch.autotune times the single-step tile shapes and the
    two-step kernel (negative tiles), sets two_step to the faster, restores
    the state, and the run that follows equals an untuned one-step batch bit
    for bit, whichever kernel it picked; with choose_two_step=False a
    two-step batch tunes only its own kernel."""
    import torch
    import lfm_hip as L
    sizes = [24, 40, 16, 72, 33, 20]
    log(f"batch autotune kernel choice, groups {group_bytes}")
    rng = np.random.default_rng(57)
    cfgs = _configs(1, sizes, rng)
    ics = [_ics(1, N, rng) for N in sizes]
    batches = []
    for tune in (True, False):
        bt = L.LatticeBatch(cfgs, backend="hip", group_bytes=group_bytes)
        for b, (r, i) in enumerate(ics):
            bt.set_psi_real(b, r)
            bt.set_psi_imag(b, i)
        bt.evolve(5)
        if tune:
            res = bt.autotune(steps=4)
            best = min(res, key=res.get)
            assert any(k[0] < 0 for k in res) and any(k[0] > 0 for k in res)
            assert bt.two_step == (best[0] < 0) and (bt.tile, bt.chunk) == (abs(best[0]), best[1])
            assert len(bt._pr) == (4 if bt.two_step else 3)
            log(f"  best {best}: {res[best]:.4f} ms per step (two_step={bt.two_step})")
        batches.append(bt)
    for n in (9, 24):
        for bt in batches:
            bt.evolve(n)
        torch.cuda.synchronize()
        for b in range(len(sizes)):
            np.testing.assert_array_equal(batches[0].get_psi_real(b), batches[1].get_psi_real(b))
            np.testing.assert_array_equal(batches[0].get_psi_imag(b), batches[1].get_psi_imag(b))
            np.testing.assert_array_equal(batches[0].get_chi(b), batches[1].get_chi(b))
    two = L.LatticeBatch(cfgs, backend="hip", two_step=True)
    res = two.autotune(steps=4, choose_two_step=False)
    assert all(k[0] < 0 for k in res) and two.two_step
