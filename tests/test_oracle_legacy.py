"""The oracle's backend-path steps against the reference's legacy FP64
integrator (lfm/core/integrator.py:103-300), restated in
oracle/legacy_integrator.py: a second source in the reference for the same
physics.

Where the two paths coincide -- no mask (PERIODIC), epsilon_w = 0, and at
COLOR no CCV / SCV -- 50 float64 steps of each must agree to rounding
(relative max error <= 1e-12).  A planted mistake in either restatement (a
dropped Mexican hat, a wrong colour-variance sign, a missing cross-colour
term) moves them apart by many orders more; the last test shows that for
the colour-variance term.
"""
import numpy as np
import pytest

from oracle import lfm_oracle as O
from oracle.legacy_integrator import LegacyState, step_leapfrog

N, STEPS, DT = 12, 50, 0.02
CHI0 = 19.0


def _fields(level, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    c = N / 2
    nc = 3 if level == 2 else 1
    out = []
    for a in range(nc):
        env = 3.0 * np.exp(-((X - c - a) ** 2 + (Y - c + 0.5 * a) ** 2 + (Z - c) ** 2) / 8.0)
        ph = rng.uniform(0, 2 * np.pi) + 0.4 * X
        out.append((env * np.cos(ph) + 1e-2 * rng.standard_normal(X.shape),
                    env * np.sin(ph) + 1e-2 * rng.standard_normal(X.shape)))
    chi = CHI0 - 0.3 * np.exp(-((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2) / 10.0)
    return out, chi


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("level", [0, 1, 2], ids=["real", "complex", "color"])
def test_legacy_integrator_equals_oracle_without_mask(level):
    comps, chi = _fields(level, 5 + level)
    p = dict(dt2=DT * DT, kappa=1 / 63, lambda_self=4 / 31, chi0=CHI0, e0_sq=0.02,
             epsilon_w=0.0, kappa_c=1 / 189 if level == 2 else 0.0,
             epsilon_cc=2 / 17 if level == 2 else 0.0, kappa_string=0.0, kappa_tube=0.0)
    if level == 0:
        psi = comps[0][0]
        st_o = O.OracleState(0, N, psi, None, chi, np.zeros(N ** 3), p)
        legacy = LegacyState(psi, psi, chi, chi)
    elif level == 1:
        pr, pi = comps[0]
        st_o = O.OracleState(1, N, pr, pi, chi, np.zeros(N ** 3), p)
        legacy = LegacyState(np.stack([pr, pi]), np.stack([pr, pi]), chi, chi)
    else:
        pr = np.stack([c[0] for c in comps])
        pi = np.stack([c[1] for c in comps])
        st_o = O.OracleState(2, N, pr, pi, chi, np.zeros(N ** 3), p)
        psi = np.stack([np.stack([pr[a], pi[a]]) for a in range(3)])
        legacy = LegacyState(psi, psi, chi, chi)
    st_o.step(STEPS)
    for _ in range(STEPS):
        step_leapfrog(legacy, level, DT, 1.0, p["kappa"], CHI0, p["e0_sq"], p["lambda_self"],
                      p["kappa_c"], p["epsilon_cc"])
    sh = (N, N, N)
    if level == 0:
        pairs = [(legacy.psi, st_o.psi_r.reshape(sh)), (legacy.psi_prev, st_o.psi_r_prev.reshape(sh))]
    elif level == 1:
        pairs = [(legacy.psi[0], st_o.psi_r.reshape(sh)), (legacy.psi[1], st_o.psi_i.reshape(sh))]
    else:
        pairs = [(legacy.psi[:, 0], st_o.psi_r.reshape((3,) + sh)),
                 (legacy.psi[:, 1], st_o.psi_i.reshape((3,) + sh))]
    pairs += [(legacy.chi, st_o.chi.reshape(sh)), (legacy.chi_prev, st_o.chi_prev.reshape(sh))]
    errs = [_rel(a, b) for a, b in pairs]
    assert max(errs) <= 1e-12, errs
    # the run moved: chi is no longer the initial field
    assert _rel(legacy.chi, chi) > 1e-6


def test_legacy_cross_check_detects_a_planted_error():
    """The same COLOR comparison with a planted mistake on one side (the
    legacy step run with twice the colour-variance coupling) separates the
    runs far beyond the 1e-12 bar, so the check can see an error of that
    term."""
    comps, chi = _fields(2, 7)
    pr = np.stack([c[0] for c in comps])
    pi = np.stack([c[1] for c in comps])
    p = dict(dt2=DT * DT, kappa=1 / 63, lambda_self=0.0, chi0=CHI0, e0_sq=0.0, epsilon_w=0.0,
             kappa_c=1 / 189, epsilon_cc=0.0, kappa_string=0.0, kappa_tube=0.0)
    st_o = O.OracleState(2, N, pr, pi, chi, np.zeros(N ** 3), p)
    psi = np.stack([np.stack([pr[a], pi[a]]) for a in range(3)])
    legacy = LegacyState(psi, psi, chi, chi)
    st_o.step(STEPS)
    for _ in range(STEPS):
        step_leapfrog(legacy, 2, DT, 1.0, p["kappa"], CHI0, kappa_c=2 * p["kappa_c"])
    assert _rel(legacy.chi, st_o.chi.reshape(N, N, N)) > 1e-9
