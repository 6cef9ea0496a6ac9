"""The reference's Evolver, backend and S_a contract tests, restated on the HIP path.

Each case follows one of the reference's own tests (cited per test) with the
same configuration and the same assertion, through ``lfm_hip.Evolver`` /
``lfm_hip.Simulation`` on the GPU.  Where the HIP path holds a value exactly
(the empty universe: zero fields stay zero and chi is stored as chi - chi0,
whose Laplacian is exactly zero), the check is tightened to bit equality.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
N = 16


def _cfg(level="REAL", **kw):
    import lfm_hip as L
    return L.SimulationConfig(grid_size=N, e_amplitude=12.0,
                              field_level=getattr(L.FieldLevel, level), **kw)


def _ev(level="REAL", **kw):
    import lfm_hip as L
    return L.Evolver(_cfg(level, **kw), backend="hip")


# --- tests/test_evolver.py:10-55 (initial state) ----------------------------

def test_initial_state(gpu):
    ev = _ev()
    assert ev.step == 0
    np.testing.assert_allclose(ev.get_chi(), ev.config.chi0)
    np.testing.assert_allclose(ev.get_psi_real(), 0.0)
    assert ev.get_psi_imag() is None
    evc = _ev("COMPLEX")
    assert evc.get_psi_imag() is not None
    np.testing.assert_allclose(evc.get_psi_imag(), 0.0)
    evk = _ev("COLOR")
    assert evk.get_psi_real().shape == (3, N, N, N)
    assert evk.get_psi_imag().shape == (3, N, N, N)


# --- tests/test_evolver.py:64-67, 115-119, 142-153 (empty universe) ---------

@pytest.mark.parametrize("level", ["REAL", "COMPLEX", "COLOR"])
def test_empty_universe_stays_vacuum(gpu, level):
    """The reference asserts |Psi| <= 1e-10 and |chi - chi0| <= 1e-5 after
    20 steps (5 at COLOR); here every value is exact, also with the COLOR S_a
    transform on (full couplings, kappa_tube > 0)."""
    import lfm_hip as L
    cfg = L.full_physics(N) if level == "COLOR" else _cfg(level)
    ev = L.Evolver(cfg, backend="hip")
    steps = 5 if level == "COLOR" else 20
    ev.evolve(steps)
    assert ev.step == steps
    assert np.all(ev.get_psi_real() == 0.0)
    if level != "REAL":
        assert np.all(ev.get_psi_imag() == 0.0)
    assert np.all(ev.get_chi() == np.float32(cfg.chi0))


# --- tests/test_evolver.py:69-79 (step counter, callback cadence) -----------

def test_step_counter_and_callback(gpu):
    ev = _ev()
    ev.evolve(10)
    assert ev.step == 10
    ev.evolve(5)
    assert ev.step == 15
    ev2 = _ev()
    calls = []
    ev2.config.report_interval = 5
    ev2.evolve(10, callback=lambda e, s: calls.append(s))
    assert calls == [5, 10]
    # the cadence follows the global step count across calls
    ev2.evolve(7, callback=lambda e, s: calls.append(s))
    assert calls == [5, 10, 15] and ev2.step == 17


# --- tests/test_evolver.py:81-104, 121-131 (chi well, energy density) -------

def test_energy_creates_chi_well(gpu):
    ev = _ev()
    c = N // 2
    E = np.zeros((N, N, N), dtype=np.float32)
    E[c, c, c] = 5.0
    ev.set_psi_real(E)
    ev.evolve(50)
    assert ev.get_chi()[c, c, c] < ev.config.chi0


def test_energy_density(gpu):
    c = N // 2
    ev = _ev()
    E = np.zeros((N, N, N), dtype=np.float32)
    E[c, c, c] = 3.0
    ev.set_psi_real(E)
    ed = ev.get_energy_density()
    assert ed.shape == (N, N, N)
    np.testing.assert_allclose(ed[c, c, c], 9.0)
    evc = _ev("COMPLEX")
    Pr = np.zeros((N, N, N), dtype=np.float32)
    Pi = np.zeros((N, N, N), dtype=np.float32)
    Pr[c, c, c], Pi[c, c, c] = 3.0, 4.0
    evc.set_psi_real(Pr)
    evc.set_psi_imag(Pi)
    np.testing.assert_allclose(evc.get_energy_density()[c, c, c], 25.0)


# --- tests/test_evolver.py:134-139, 156-165 (setters) ------------------------

def test_setters(gpu):
    ev = _ev()
    with pytest.raises(ValueError, match="imaginary"):
        ev.set_psi_imag(np.zeros((N, N, N)))
    ev.set_chi(np.full((N, N, N), 18.5, dtype=np.float32))
    np.testing.assert_allclose(ev.get_chi(), 18.5)


# --- tests/test_sa_fields.py:54-99, 205-224 (S_a buffers) -------------------

def _sa_cfg(kappa_tube):
    import lfm_hip as L
    from lfm_hip.config import KAPPA, KAPPA_C
    return L.SimulationConfig(grid_size=N, field_level=L.FieldLevel.COLOR,
                              boundary_type=L.BoundaryType.FROZEN, dt=0.02,
                              report_interval=9999, kappa_c=KAPPA_C,
                              kappa_tube=kappa_tube * KAPPA)


def test_sa_fields_allocation(gpu):
    import lfm_hip as L
    assert L.Simulation(_sa_cfg(0.0), backend="hip").sa_fields is None
    sa = L.Simulation(_sa_cfg(10.0), backend="hip").sa_fields
    assert sa is not None and sa.shape == (3, N, N, N)
    assert np.all(sa == 0.0)


def test_sa_fields_grow_and_survive_checkpoint(gpu, tmp_path):
    import lfm_hip as L
    sim = L.Simulation(_sa_cfg(10.0), backend="hip")
    sim.place_soliton((N // 2, N // 2, N // 2), amplitude=8.0)
    sim.run(steps=200, record_metrics=False)
    sa = sim.sa_fields
    assert sa.max() > 1e-6, "S_a never grew from zero"
    path = tmp_path / "ckpt.npz"
    sim.save_checkpoint(str(path))
    sa_after = L.Simulation.load_checkpoint(str(path), backend="hip").sa_fields
    assert sa_after is not None
    np.testing.assert_allclose(sa_after, sa, atol=1e-6)


# --- tests/test_backends.py:27-290 (the backend object itself) --------------

def test_backend_mask_and_conversions(gpu):
    import lfm_hip as L
    b = L.HipBackend()
    m = b.create_boundary_mask(N, 0.15)
    h = b.to_numpy(m)
    assert h.shape == (N ** 3,) and h.dtype == np.float32
    c = N // 2
    assert h[c * N * N + c * N + c] == 0.0, "centre must not be absorbing"
    assert h[0] == 1.0, "corner must be absorbing"
    arr = np.array([1.0, 2.0, 3.0], dtype=np.float32)
    np.testing.assert_array_equal(b.to_numpy(b.from_numpy(arr)), arr)
    assert b.to_numpy(b.from_numpy(np.array([1.0, 2.0]))).dtype == np.float32


def _ab_buffers(b, ncomp, energy=0.0):
    """The reference tests' hand-made A/B buffers: Psi = 0 (or an energy blob
    at the centre of component 0), chi = 19, as flat float32 device arrays."""
    total = N ** 3
    psi = np.zeros(ncomp * total, dtype=np.float32)
    c = N // 2
    psi[c * N * N + c * N + c] = energy
    chi = np.full(total, 19.0, dtype=np.float32)
    return {k: b.from_numpy(v) for k, v in (
        ("pr", psi), ("prp", psi), ("pi", psi * 0), ("pip", psi * 0), ("chi", chi), ("chip", chi))}


@pytest.mark.parametrize("level", ["REAL", "COMPLEX", "COLOR"])
def test_backend_empty_universe_stable(gpu, level):
    """10 A/B-toggled steps of the backend's own step function from Psi = 0,
    chi = 19 with the 0.15 mask: Psi stays 0 (atol 1e-10), chi stays 19
    (atol 1e-5), the reference's bars on its absolute-chi buffers."""
    import lfm_hip as L
    b = L.HipBackend()
    mask = b.create_boundary_mask(N, 0.15)
    nc = 3 if level == "COLOR" else 1
    A, B = _ab_buffers(b, nc), _ab_buffers(b, nc)
    dt2 = 0.02 ** 2
    for _ in range(10):
        if level == "REAL":
            b.step_real(A["pr"], A["prp"], A["chi"], A["chip"], mask, B["pr"], B["prp"],
                        B["chi"], B["chip"], N, dt2, 1 / 63, 0.0, 19.0, 0.0)
        else:
            fn = b.step_complex if level == "COMPLEX" else b.step_color
            fn(A["pr"], A["prp"], A["pi"], A["pip"], A["chi"], A["chip"], mask, B["pr"], B["prp"],
               B["pi"], B["pip"], B["chi"], B["chip"], N, dt2, 1 / 63, 0.0, 19.0, 0.0, 0.1)
        A, B = B, A
    np.testing.assert_allclose(b.to_numpy(A["pr"]), 0.0, atol=1e-10)
    if level != "REAL":
        np.testing.assert_allclose(b.to_numpy(A["pi"]), 0.0, atol=1e-10)
    np.testing.assert_allclose(b.to_numpy(A["chi"]), 19.0, atol=1e-5)


def test_backend_energy_reduces_chi(gpu):
    """A blob of E = 5 at the centre (E_prev too), 50 A/B-toggled REAL steps:
    chi at the centre drops below chi0."""
    import lfm_hip as L
    b = L.HipBackend()
    mask = b.create_boundary_mask(N, 0.15)
    A, B = _ab_buffers(b, 1, energy=5.0), _ab_buffers(b, 1)
    for _ in range(50):
        b.step_real(A["pr"], A["prp"], A["chi"], A["chip"], mask, B["pr"], B["prp"], B["chi"],
                    B["chip"], N, 0.02 ** 2, 1 / 63, 0.0, 19.0, 0.0)
        A, B = B, A
    c = N // 2
    assert b.to_numpy(A["chi"])[c * N * N + c * N + c] < 19.0


# --- tests/test_gpu_parity.py:428-443 (allocate) -----------------------------

def test_backend_allocate(gpu):
    """allocate(N, n_psi_arrays, chi0): the eight A/B buffers, Psi zeros of
    n_psi_arrays * N^3 and chi filled with chi0 (numpy_backend.py:28-47),
    each its own storage."""
    import lfm_hip as L
    b = L.HipBackend()
    bufs = b.allocate(N, 2, 19.0)
    assert sorted(bufs) == sorted(["psi_A", "psi_prev_A", "chi_A", "chi_prev_A",
                                   "psi_B", "psi_prev_B", "chi_B", "chi_prev_B"])
    for k, v in bufs.items():
        h = b.to_numpy(v)
        assert h.dtype == np.float32
        if k.startswith("psi"):
            assert h.shape == (2 * N ** 3,) and np.all(h == 0.0), k
        else:
            assert h.shape == (N ** 3,) and np.all(h == np.float32(19.0)), k
    assert len({v.data_ptr() for v in bufs.values()}) == 8


# --- tests/test_integration.py:29-106 (sweep API) ---------------------------

def test_sweep_api(gpu):
    import lfm_hip as L
    cfg = L.SimulationConfig(grid_size=8)
    rows = L.sweep(cfg, param="kappa", values=[0.01, 0.02], steps=5, equilibrate=False,
                   backend="hip")
    assert isinstance(rows, list) and len(rows) == 2
    assert rows[0]["kappa"] == 0.01 and rows[1]["kappa"] == 0.02
    assert "chi_min" in rows[0] and "energy_total" in rows[0]
    only = L.sweep(cfg, param="kappa", values=[0.016], steps=5, metric_names=["chi_min"],
                   equilibrate=False, backend="hip")
    assert "chi_min" in only[0] and "total_energy" not in only[0]
    cfg16 = L.SimulationConfig(grid_size=16)
    for eq in (False, True):
        r = L.sweep(cfg16, param="kappa", values=[1 / 63], steps=5, soliton={"amplitude": 3.0},
                    equilibrate=eq, backend="hip")
        assert len(r) == 1
        if not eq:
            assert r[0]["chi_min"] < L.CHI0


# --- tests/test_simulation.py:298-389 (checkpoint round trips) ---------------

def _small(**kw):
    import lfm_hip as L
    d = dict(grid_size=N, field_level=L.FieldLevel.REAL, boundary_type=L.BoundaryType.FROZEN,
             report_interval=100, dt=0.02)
    d.update(kw)
    return L.SimulationConfig(**d)


def test_checkpoint_round_trips(gpu, tmp_path):
    import lfm_hip as L
    c = (N // 2, N // 2, N // 2)
    sim = L.Simulation(_small(), backend="hip")
    sim.place_soliton(c, amplitude=3.0, sigma=2.0)
    sim.equilibrate()
    sim.run(10, record_metrics=False)
    path = tmp_path / "ckpt.npz"
    sim.save_checkpoint(path)
    assert path.exists() and path.stat().st_size > 0
    loaded = L.Simulation.load_checkpoint(path, backend="hip")
    assert loaded.step == sim.step
    np.testing.assert_allclose(loaded.get_chi(), sim.get_chi(), atol=1e-6)
    np.testing.assert_allclose(loaded.get_psi_real(), sim.get_psi_real(), atol=1e-6)

    cfg = _small(chi0=18.5, kappa=0.02)
    s2 = L.Simulation(cfg, backend="hip")
    s2.run(10, record_metrics=False)
    s2.save_checkpoint(tmp_path / "c2.npz")
    l2 = L.Simulation.load_checkpoint(tmp_path / "c2.npz", backend="hip")
    assert l2.config.grid_size == cfg.grid_size
    assert l2.config.chi0 == pytest.approx(cfg.chi0)
    assert l2.config.kappa == pytest.approx(cfg.kappa)

    s3 = L.Simulation(_small(report_interval=5), backend="hip")
    s3.place_soliton(c, amplitude=3.0, sigma=2.0)
    s3.run(10)
    s3.save_checkpoint(tmp_path / "c3.npz")
    assert len(L.Simulation.load_checkpoint(tmp_path / "c3.npz", backend="hip").history) == \
        len(s3.history)

    s4 = L.Simulation(_small(field_level=L.FieldLevel.COMPLEX), backend="hip")
    s4.place_soliton(c, amplitude=3.0, sigma=2.0, phase=0.5)
    s4.equilibrate()
    s4.run(10, record_metrics=False)
    s4.save_checkpoint(tmp_path / "c4.npz")
    l4 = L.Simulation.load_checkpoint(tmp_path / "c4.npz", backend="hip")
    assert l4.config.field_level == L.FieldLevel.COMPLEX
    np.testing.assert_allclose(l4.get_psi_imag(), s4.get_psi_imag(), atol=1e-6)


# --- tests/test_integration.py:210-229, tests/test_edge_cases.py:55-82 ------

def test_no_nan_runs_and_metric_keys(gpu):
    """COLOR with kappa_c (amplitude 6, 20 steps), an empty REAL run, COMPLEX
    and COLOR with a soliton (10 steps): no NaN; metrics() carries chi_min,
    chi_max and energy_total."""
    import lfm_hip as L
    sim = L.Simulation(L.SimulationConfig(grid_size=16, field_level=L.FieldLevel.COLOR,
                                          kappa_c=1 / 189), backend="hip")
    sim.place_soliton((8, 8, 8), amplitude=6.0)
    sim.run(steps=20)
    assert not np.any(np.isnan(sim.chi)) and not np.any(np.isnan(sim.psi_real))
    sim = L.Simulation(L.SimulationConfig(grid_size=16, e_amplitude=0.0), backend="hip")
    sim.run(steps=10)
    assert not np.any(np.isnan(sim.chi))
    for lvl in (L.FieldLevel.COMPLEX, L.FieldLevel.COLOR):
        sim = L.Simulation(L.SimulationConfig(grid_size=16, field_level=lvl), backend="hip")
        sim.place_soliton((8, 8, 8), amplitude=2.0)
        sim.run(steps=10)
        assert not np.any(np.isnan(sim.chi))
    sim = L.Simulation(L.SimulationConfig(grid_size=16), backend="hip")
    sim.place_soliton((8, 8, 8), amplitude=2.0)
    sim.run(steps=5)
    m = sim.metrics()
    assert "chi_min" in m and "chi_max" in m and "energy_total" in m
