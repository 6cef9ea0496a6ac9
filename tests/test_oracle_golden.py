"""CPU pins of the oracle against vectors the reference itself wrote, and
closed forms for the branches the older pins left open.

* ``golden/step_real_complex.npz``: the reference's ``step_real`` /
  ``step_complex`` (numpy_backend.py:80-205) with the cos^2 mask and the
  damped prev, three A/B steps, float32 and float64.  The float32 oracle
  must reproduce the float32 vectors bit for bit (same NumPy operations in
  the same order); the float64 oracle the float64 vectors to 1e-14.
* masks of N = 16 bitwise, N = 64 / 128 by SHA-256, for the oracle and for
  the product's host mask builder (lfm_hip.fields.boundary_mask_host).
* ``golden/checkpoint_*_N16.npz``: the reference's ``metrics()``
  (lfm/simulation.py:1407-1440, analysis/metrics.py:29-111) at save time,
  recomputed from the checkpoint's arrays with the oracle's FP64 energy and
  chi statistics (this pins ``energy_components`` with the imaginary and
  colour branches), and the checkpoint's config read into
  lfm_hip.SimulationConfig.
* closed forms: a complex plane wave's T, G, V (np.gradient's one-sided
  edges included), and the freeze-chi A/B alternation of
  ``OracleState.step(freeze_chi=True)`` as a scalar mode recurrence.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from oracle import lfm_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")


def _gold(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _run(level, g, tag, dt, steps=3):
    """A/B steps of the oracle from the golden inputs, returning the state
    after each step as {step: {key: array}}."""
    N = int(tag[1:])
    p = {k[6:]: float(g[k]) for k in g.files if k.startswith("param_")}
    c = lambda k: g[f"{tag}_{k}"].astype(dt)
    A = dict(r=c("pr"), rp=c("prp"), c=c("chi"), cp=c("chip"))
    if level == 1:
        A.update(i=c("pi"), ip=c("pip"))
    B = {k: np.zeros_like(v) for k, v in A.items()}
    m = c("mask")
    dt2 = float(g["dt2"])
    out = {}
    for s in range(1, steps + 1):
        if level == 0:
            O.step_real(A["r"], A["rp"], A["c"], A["cp"], m, B["r"], B["rp"], B["c"], B["cp"], N,
                        dt2, p["kappa"], p["lambda_self"], p["chi0"], p["e0_sq"])
        else:
            O.step_complex(A["r"], A["rp"], A["i"], A["ip"], A["c"], A["cp"], m, B["r"], B["rp"],
                           B["i"], B["ip"], B["c"], B["cp"], N, dt2, p["kappa"], p["lambda_self"],
                           p["chi0"], p["e0_sq"], p["epsilon_w"])
        A, B = B, A
        out[s] = {k: v.copy() for k, v in A.items()}
    return out


@pytest.mark.parametrize("level", [0, 1], ids=["real", "complex"])
def test_masked_steps_match_reference_float32_bitwise(level):
    g = _gold("step_real_complex.npz")
    name = ("real", "complex")[level]
    got = _run(level, g, "N8", np.float32)
    n = 0
    for s in (1, 3):
        for k, v in got[s].items():
            np.testing.assert_array_equal(v, g[f"N8_{name}_f32_s{s}_{k}"], err_msg=f"{name} s{s} {k}")
            n += 1
    assert n == (8 if level == 0 else 12)


@pytest.mark.parametrize("tag", ["N8", "N16"])
@pytest.mark.parametrize("level", [0, 1], ids=["real", "complex"])
def test_masked_steps_match_reference_float64(level, tag):
    g = _gold("step_real_complex.npz")
    name = ("real", "complex")[level]
    got = _run(level, g, tag, np.float64)
    keys = [k for k in g.files if k.startswith(f"{tag}_{name}_f64_")]
    assert keys
    for key in keys:
        s, k = int(key.split("_s")[1][0]), key.rsplit("_", 1)[1]
        ref = g[key]
        err = float(np.abs(got[s][k] - ref).max() / np.abs(ref).max())
        assert err <= 1e-14, (key, err)


def test_masks_match_reference():
    from lfm_hip.fields import boundary_mask_host
    g = _gold("step_real_complex.npz")
    for key, N, bf in (("mask_N16", 16, 0.3), ("mask_N16_bf05", 16, 0.5)):
        np.testing.assert_array_equal(O.boundary_mask(N, bf), g[key])
        np.testing.assert_array_equal(np.asarray(boundary_mask_host(N, bf)).ravel(), g[key])
    for N in (64, 128):
        want = str(g[f"mask_N{N}_sha256"])
        for m in (O.boundary_mask(N, 0.3), np.asarray(boundary_mask_host(N, 0.3)).ravel()):
            m = np.ascontiguousarray(m, dtype=np.float32)
            assert hashlib.sha256(m.tobytes()).hexdigest() == want, N


def _oracle_metrics(psi_r, psi_r_prev, chi, dt, psi_i, psi_i_prev, bf):
    """compute_metrics keys from the oracle's FP64 reductions (clusters left out)."""
    T, Gr, V = O.energy_components(psi_r, psi_r_prev, chi, dt, 1.0, psi_i, psi_i_prev)
    N = chi.shape[0]
    inter = O.interior_mask(N, bf)
    st = O.chi_statistics(chi.astype(np.float64), inter)
    wf, vf = O.well_void_fraction(chi, 17.0, 18.0, inter)
    psq = np.sum(np.asarray(psi_r, np.float64) ** 2, axis=0) if psi_r.ndim == 4 else \
        np.asarray(psi_r, np.float64) ** 2
    psq = psq + (np.sum(np.asarray(psi_i, np.float64) ** 2, axis=0) if psi_i.ndim == 4
                 else np.asarray(psi_i, np.float64) ** 2)
    e = dict(energy_kinetic=float(T.sum()), energy_gradient=float(Gr.sum()),
             energy_potential=float(V.sum()))
    e["energy_total"] = e["energy_kinetic"] + e["energy_gradient"] + e["energy_potential"]
    e.update(chi_min=st["min"], chi_max=st["max"], chi_mean=st["mean"], chi_std=st["std"],
             well_fraction=wf, void_fraction=vf, psi_sq_total=float(psq.sum()))
    return e


@pytest.mark.parametrize("name", ["complex", "color"])
def test_reference_checkpoint_metrics_from_oracle(name):
    d = _gold(f"checkpoint_{name}_N16.npz")
    ex = _gold("checkpoint_expect.npz")
    cfg = json.loads(str(d["config_json"]))
    ref = json.loads(str(ex[f"{name}_metrics_json"]))
    got = _oracle_metrics(d["psi_real"], d["psi_real_prev"], d["chi"], cfg["dt"], d["psi_imag"],
                          d["psi_imag_prev"], cfg["boundary_fraction"])
    assert set(got) | {"n_clusters"} == set(ref)
    check_metrics(got, ref)
    assert ref["energy_kinetic"] > 0      # the snapshot is two steps older than the fields


def check_metrics(got, ref):
    """Energies and |Psi|^2 are FP64 sums on both sides (1e-12); chi min/max
    and the well/void fractions are exact; the reference takes chi's mean
    and std with NumPy's float32 accumulation of a float32 array
    (structure.py:43-51), so those agree to float32 rounding (2e-7 of chi0)."""
    for k, v in got.items():
        if k in ("chi_mean", "chi_std"):
            assert abs(v - ref[k]) <= 2e-7 * 19.0, (k, v, ref[k])
        elif k in ("chi_min", "chi_max", "well_fraction", "void_fraction", "n_clusters"):
            assert v == ref[k], (k, v, ref[k])
        else:
            assert abs(v - ref[k]) <= 1e-12 * max(abs(ref[k]), 1e-300), (k, v, ref[k])


@pytest.mark.parametrize("name", ["complex", "color"])
def test_reference_checkpoint_config_reads(name):
    from lfm_hip.simulation import config_from_checkpoint
    d = _gold(f"checkpoint_{name}_N16.npz")
    raw = json.loads(str(d["config_json"]))
    cfg = config_from_checkpoint(d)
    for k, v in raw.items():
        got = getattr(cfg, k)
        got = got.value if hasattr(got, "value") else got
        assert got == v or (isinstance(v, list) and list(got) == v), (k, got, v)
    assert int(d["step"]) == 12
    hist = json.loads(str(d["history_json"]))
    assert [h["step"] for h in hist] == [5.0, 10.0]


# ------------------------------------------------------------- closed forms
def test_complex_plane_wave_energy_closed_form():
    """psi = A exp(i(k.x - w t)): |psi|^2 = A^2 (V = chi^2 A^2 / 2 per site),
    |grad psi|^2 per axis = A^2 sin^2 k on interior points (np.gradient's
    central difference) and A^2 4 sin^2(k/2) on the two edge planes
    (one-sided difference), |psi - psi_prev|^2 = 4 A^2 sin^2(w dt / 2)."""
    N, A, dt, w, chi = 12, 0.7, 0.02, 19.3, 18.5
    n = (1, 2, 3)
    k = [2 * np.pi * q / N for q in n]
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ph = k[0] * X + k[1] * Y + k[2] * Z
    pr, pi = A * np.cos(ph), A * np.sin(ph)
    prp, pip = A * np.cos(ph - w * dt), A * np.sin(ph - w * dt)
    chif = np.full((N, N, N), chi)
    T, Gd, V = O.energy_components(pr, prp, chif, dt, 1.0, pi, pip)
    n3 = N ** 3
    assert np.isclose(V.sum(), 0.5 * chi ** 2 * A ** 2 * n3, rtol=1e-13)
    assert np.isclose(T.sum(), 0.5 * 4 * A ** 2 * math.sin(w * dt / 2) ** 2 / dt ** 2 * n3, rtol=1e-12)
    g = sum((N - 2) * N * N * math.sin(kd) ** 2 + 2 * N * N * 4 * math.sin(kd / 2) ** 2 for kd in k)
    assert np.isclose(Gd.sum(), 0.5 * A ** 2 * g, rtol=1e-12)
    # the colour axis is summed: three copies give three times each term
    T3, G3, V3 = O.energy_components(np.stack([pr] * 3), np.stack([prp] * 3), chif, dt, 1.0,
                                     np.stack([pi] * 3), np.stack([pip] * 3))
    for a, b in ((T3, T), (G3, Gd), (V3, V)):
        assert np.isclose(a.sum(), 3 * b.sum(), rtol=1e-13)


def test_freeze_chi_alternates_ab_sets():
    """With chi frozen, GOV-01 with spatially uniform chi is a scalar
    recurrence per plane-wave mode.  After one free step the A and B chi sets
    differ; freeze_chi restores both each step, so successive steps read
    chi_A, chi_B, chi_A, ... (lfm/core/evolver.py:160-176): the mode obeys
    a_{n+1} = (2 - dt^2 (chi_s(n)^2 - lambda(k))) a_n - a_{n-1} with the set
    index alternating, and chi never changes (kappa != 0 here)."""
    N, dt = 8, 0.02
    n = (1, 0, 2)
    k = [2 * np.pi * q / N for q in n]
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    mode = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    c = [math.cos(q) for q in k]
    lam = (2 / 3) * sum(c) + (2 / 3) * (c[0] * c[1] + c[0] * c[2] + c[1] * c[2]) - 4.0
    p = dict(dt2=dt * dt, kappa=1 / 63, lambda_self=0.0, chi0=19.0, e0_sq=0.0)
    st = O.OracleState(0, N, 1e-3 * mode, None, np.full(N ** 3, 19.0), np.zeros(N ** 3), p)
    chi_a, chi_b = 19.0, 18.2
    st.c[0][:] = chi_a     # set A (current) and set B chi made uniform but different
    st.c[1][:] = chi_b
    st.cp[0][:] = chi_a
    st.cp[1][:] = chi_b
    a_prev = a = 1e-3
    seq = []
    for s in range(10):
        chi = chi_a if s % 2 == 0 else chi_b
        a, a_prev = (2 - dt * dt * (chi * chi - lam)) * a - a_prev, a
        seq.append(a)
    got = []
    for s in range(10):
        st.step(1, freeze_chi=True)
        got.append(float(np.sum(st.psi_r * mode.ravel()) / np.sum(mode ** 2)))
        assert np.all(st.c[0] == chi_a) and np.all(st.c[1] == chi_b)
    np.testing.assert_allclose(got, seq, rtol=1e-11)
