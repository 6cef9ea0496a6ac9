"""Pins for the CPU oracle: closed forms, special cases and golden vectors.

Each test derives its expected value independently of the oracle's
vectorised code (plane-wave symbols, scalar recurrences, colour special
cases, the reference's own recorded outputs), so a dropped term, a wrong sign
or a transposed index in oracle/lfm_oracle.py fails here.
"""
import math
import os

import numpy as np
import pytest

from oracle import lfm_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CHI0 = 19.0
DT = 0.02


def lam_symbol(kx, ky, kz):
    """Plane-wave eigenvalue of the 19-point stencil (faces 1/3, edges 1/6)."""
    c = (math.cos(kx), math.cos(ky), math.cos(kz))
    return (2 / 3) * sum(c) + (2 / 3) * (c[0] * c[1] + c[0] * c[2] + c[1] * c[2]) - 4.0


def plane(N, n, phase=0.0, fn=np.cos):
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    k = [2 * np.pi * q / N for q in n]
    return fn(k[0] * X + k[1] * Y + k[2] * Z + phase), k


# ---------------------------------------------------------------- Laplacian
@pytest.mark.parametrize("n", [(1, 0, 0), (0, 2, 1), (3, 1, 2), (4, 4, 4), (1, 7, 5)])
def test_laplacian_plane_wave_symbol(n):
    N = 8
    f, k = plane(N, n, 0.3)
    lap = O.laplacian_19pt(f)
    np.testing.assert_allclose(lap, lam_symbol(*k) * f, atol=1e-12)


def test_laplacian_spectrum_extremes():
    N = 8
    lams = [lam_symbol(*(2 * np.pi * q / N for q in (a, b, c)))
            for a in range(N) for b in range(N) for c in range(N)]
    assert abs(max(lams)) < 1e-15          # lambda(0) = 0 is the top of the spectrum
    assert abs(min(lams) + 16 / 3) < 1e-12  # 16/3 of the CFL bound (config.py:318)


def test_laplacian_quadratic_interior():
    N = 12
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    for f in (X ** 2, Y ** 2, Z ** 2):
        lap = O.laplacian_19pt(f)
        np.testing.assert_allclose(lap[2:-2, 2:-2, 2:-2], 2.0, atol=1e-10)
    np.testing.assert_allclose(O.laplacian_19pt(np.full((N,) * 3, 3.7)), 0.0, atol=1e-13)


def test_laplacian_cubic_symmetry():
    rng = np.random.default_rng(0)
    f = rng.standard_normal((8, 8, 8))
    lap = O.laplacian_19pt(f)
    for perm in ((1, 0, 2), (2, 1, 0), (0, 2, 1)):
        np.testing.assert_allclose(O.laplacian_19pt(f.transpose(perm)), lap.transpose(perm),
                                   atol=1e-12)
    np.testing.assert_allclose(O.laplacian_19pt(f[::-1]), lap[::-1], atol=1e-12)


# --------------------------------------------------------------------- mask
def test_mask_against_reference_golden():
    g = np.load(os.path.join(HERE, "golden", "step_color_N8.npz"))
    np.testing.assert_array_equal(O.boundary_mask(8, 0.3), g["mask"])


def test_mask_known_answers():
    # tests/test_backends.py:35-52: shape, dtype, centre 0, corner 1
    N = 16
    m = O.boundary_mask(N, 0.15)
    assert m.shape == (N ** 3,) and m.dtype == np.float32
    c = N // 2
    assert m[c * N * N + c * N + c] == 0.0
    assert m[0] == 1.0
    # 0 inside r_freeze, 1 outside r_max, monotone in R in between
    co = np.arange(N) - N / 2 + 0.5
    R = np.sqrt(co[:, None, None] ** 2 + co[None, :, None] ** 2 + co[None, None, :] ** 2).ravel()
    rf, rm = (1 - 0.15) * N / 2, N / 2
    assert np.all(m[R <= rf - 1e-4] == 0.0)
    assert np.all(m[R >= rm + 1e-4] == 1.0)
    order = np.argsort(R)
    assert np.all(np.diff(m[order].astype(np.float64)) >= -1e-6)


# ------------------------------------------------------------ REAL level
def _real_state(N, E, Ep, chi, chip, mask, **p):
    base = dict(dt2=DT ** 2, kappa=0.0, lambda_self=0.0, chi0=CHI0, e0_sq=0.0)
    base.update(p)
    return O.OracleState(0, N, E, None, chi, mask, base, psi_r_prev=Ep, chi_prev=chip)


def test_gov01_exact_dispersion():
    N, steps = 8, 40
    f, k = plane(N, (1, 2, 3), 0.2)
    a0, am1 = 1.3, 1.1
    st = _real_state(N, a0 * f, am1 * f, np.full(N ** 3, CHI0), None, np.zeros(N ** 3))
    coef = 2 - DT ** 2 * (CHI0 ** 2 - lam_symbol(*k))
    a = [am1, a0]
    for _ in range(steps):
        a.append(coef * a[-1] - a[-2])
    st.step(steps)
    np.testing.assert_allclose(st.psi_r, a[-1] * f.ravel(), rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(st.chi, CHI0)


def _scalar_real(e, ep, u, up, steps, kappa, lam, e0):
    for _ in range(steps):
        en = 2 * e - ep - DT ** 2 * u * u * e
        un = 2 * u - up + DT ** 2 * (-(kappa / CHI0) * u * (e * e - e0) - 4 * lam * u * (u * u - CHI0 ** 2))
        un = max(un, -CHI0)
        e, ep, u, up = en, e, un, u
    return e, u


@pytest.mark.parametrize("kappa,lam,e0", [(0.0, 4 / 31, 0.0), (1 / 63, 0.0, 0.0), (1 / 63, 4 / 31, 0.7)])
def test_gov02_uniform_recurrence(kappa, lam, e0):
    N, steps = 8, 25
    one = np.ones(N ** 3)
    st = _real_state(N, 2.5 * one, 2.4 * one, 18.7 * one, 18.75 * one, np.zeros(N ** 3),
                     kappa=kappa, lambda_self=lam, e0_sq=e0)
    st.step(steps)
    e, u = _scalar_real(2.5, 2.4, 18.7, 18.75, steps, kappa, lam, e0)
    np.testing.assert_allclose(st.psi_r, e, rtol=1e-12)
    np.testing.assert_allclose(st.chi, u, rtol=1e-12)


def test_bh_clamp():
    N = 8
    one = np.ones(N ** 3)
    st = _real_state(N, 0 * one, 0 * one, -18.9 * one, -18.0 * one, np.zeros(N ** 3))
    st.step(1)
    np.testing.assert_array_equal(st.chi, -CHI0)


def test_mask_blend_and_prev_outputs():
    N = 8
    rng = np.random.default_rng(3)
    m = rng.uniform(0, 1, N ** 3)
    one = np.ones(N ** 3)
    E, Ep, u, up = 1.5, 1.4, 18.9, 18.95
    outs = [np.zeros(N ** 3) for _ in range(4)]
    O.step_real(E * one, Ep * one, u * one, up * one, m, *outs, N=N, dt2=DT ** 2,
                kappa=1 / 63, lambda_self=0.0, chi0=CHI0, e0_sq=0.0)
    en = 2 * E - Ep - DT ** 2 * u * u * E
    un = 2 * u - up - DT ** 2 * (1 / 63 / CHI0) * u * E * E
    np.testing.assert_allclose(outs[0], (1 - m) * en, rtol=1e-13)
    np.testing.assert_allclose(outs[1], (1 - m) * E, rtol=1e-13)
    np.testing.assert_allclose(outs[2], m * CHI0 + (1 - m) * un, rtol=1e-13)
    np.testing.assert_array_equal(outs[3], u)


# --------------------------------------------------------- COMPLEX level
def test_complex_plane_wave_current():
    """Psi = A e^{i k.x}: |Psi|^2 = A^2, j_total = A^2 sum_d sin k_d (uniform)."""
    N, A, eps = 8, 1.7, 0.1
    c, k = plane(N, (1, 2, 3), 0.0, np.cos)
    s, _ = plane(N, (1, 2, 3), 0.0, np.sin)
    p = dict(dt2=DT ** 2, kappa=1 / 63, lambda_self=0.0, chi0=CHI0, e0_sq=0.2, epsilon_w=eps)
    st = O.OracleState(1, N, A * c, A * s, np.full(N ** 3, CHI0), np.zeros(N ** 3), p)
    st.step(1)
    j = A * A * sum(math.sin(q) for q in k)
    chi1 = CHI0 + DT ** 2 * (-(1 / 63 / CHI0) * CHI0 * (A * A + eps * j - 0.2))
    np.testing.assert_allclose(st.chi, chi1, rtol=1e-13)
    coef = 2 - DT ** 2 * (CHI0 ** 2 - lam_symbol(*k))
    np.testing.assert_allclose(st.psi_r, (coef - 1) * A * c.ravel(), atol=1e-12)
    np.testing.assert_allclose(st.psi_i, (coef - 1) * A * s.ravel(), atol=1e-12)


# ----------------------------------------------------------- COLOR level
FULL = dict(kappa=1 / 63, lambda_self=4 / 31, chi0=CHI0, e0_sq=0.0, epsilon_w=0.1,
            kappa_c=1 / 189, epsilon_cc=2 / 17, kappa_string=1 / 189, kappa_tube=30 / 63,
            sa_gamma=0.1, sa_d=4.9)


def _color_state(N, pr, pi, p, prp=None, pip=None):
    q = dict(p, dt2=DT ** 2)
    return O.OracleState(2, N, np.concatenate([x.ravel() for x in pr]),
                         np.concatenate([x.ravel() for x in pi]), np.full(N ** 3, CHI0),
                         np.zeros(N ** 3), q,
                         psi_r_prev=None if prp is None else np.concatenate([x.ravel() for x in prp]),
                         psi_i_prev=None if pip is None else np.concatenate([x.ravel() for x in pip]))


def test_color_singlet_has_no_colour_terms():
    """Equal colours: f_c = 0, CCV = 0, SCV = 0, eps_cc term = 0."""
    N = 8
    c, k = plane(N, (1, 0, 2), 0.0, np.cos)
    s, _ = plane(N, (1, 0, 2), 0.0, np.sin)
    A = 1.2
    st = _color_state(N, [A * c] * 3, [A * s] * 3, FULL)
    st.step(1)
    e = A * A
    j = A * A * sum(math.sin(q) for q in k)      # per colour, already halved
    chi1 = CHI0 + DT ** 2 * (-(FULL["kappa"] / CHI0) * CHI0 * (3 * e + 0.1 * 3 * j))
    np.testing.assert_allclose(st.chi, chi1, rtol=1e-13)
    np.testing.assert_allclose(st.sa_fields, e, rtol=1e-12)   # Helmholtz DC gain 1


def test_color_single_colour_terms():
    """One colour carries a plane wave: f_c = 2/3, CCV = 2/3 sum_d j_d^2,
    SCV = 2/3 S^2 with S = A^2."""
    N = 8
    c, k = plane(N, (1, 0, 0), 0.0, np.cos)
    s, _ = plane(N, (1, 0, 0), 0.0, np.sin)
    A = 1.4
    z = np.zeros_like(c)
    st = _color_state(N, [A * c, z, z], [A * s, z, z], FULL)
    st.step(1)
    e = A * A
    jx = 2 * A * A * math.sin(k[0])
    src = (FULL["kappa"] / CHI0) * CHI0 * (e + 0.1 * 0.5 * jx)
    cv = (FULL["kappa_c"] / CHI0) * CHI0 * (2 / 3) * e
    ccv = (2 / 3) * jx * jx
    scv = (2 / 3) * e * e
    chi1 = CHI0 + DT ** 2 * (-src - cv - FULL["kappa_string"] * ccv - FULL["kappa_tube"] * scv)
    np.testing.assert_allclose(st.chi, chi1, rtol=1e-13)


def test_color_cross_colour_coupling_uniform():
    N = 8
    one = np.ones((N, N, N))
    vals_r, vals_i = [0.9, -0.3, 0.5], [0.2, 0.4, -0.6]
    p = dict(FULL, kappa_string=0.0, kappa_tube=0.0, kappa_c=0.0)
    st = _color_state(N, [v * one for v in vals_r], [v * one for v in vals_i], p)
    st.step(1)
    for vals, got in ((vals_r, st.psi_r), (vals_i, st.psi_i)):
        bar = sum(vals) / 3
        for a, v in enumerate(vals):
            exp = v - DT ** 2 * CHI0 ** 2 * v - DT ** 2 * (2 / 17) * CHI0 ** 2 * (v - bar)
            np.testing.assert_allclose(got[a * N ** 3:(a + 1) * N ** 3], exp, rtol=1e-13)


def test_helmholtz_single_mode():
    N, g, D = 16, 0.1, 4.9
    c, k = plane(N, (2, 1, 3), 0.4)
    Aa, B = 2.0, 0.7
    S = O.helmholtz_sa((Aa + B * c).ravel(), N, g, D)
    k2 = sum(q * q for q in k)
    np.testing.assert_allclose(S, (Aa + B * g / (g + D * k2) * c).ravel(), atol=1e-12)


@pytest.mark.parametrize("name", ["full", "kc"])
def test_step_color_golden(name):
    """The reference's own step_color output on seeded N=8 inputs."""
    g = np.load(os.path.join(HERE, "golden", "step_color_N8.npz"))
    N = int(g["N"])
    keys = ("kappa", "lambda_self", "chi0", "e0_sq", "epsilon_w", "kappa_c", "epsilon_cc",
            "kappa_string", "kappa_tube", "sa_gamma", "sa_d")
    p = {k: float(g[f"{name}_param_{k}"]) for k in keys}
    p["dt2"] = 0.02 ** 2
    for dtype, tol in ((np.float32, 0.0), (np.float64, 2e-6)):
        st = O.OracleState(2, N, g["pr"], g["pi"], g["chi"], g["mask"], p, psi_r_prev=g["prp"],
                           psi_i_prev=g["pip"], chi_prev=g["chip"], dtype=dtype)
        for s in (1, 3):
            st.step(1 if s == 1 else 2)
            got = dict(r=st.psi_r, i=st.psi_i, c=st.chi, rp=st.psi_r_prev, cp=st.chi_prev)
            if st.sa_fields is not None:
                got["sa"] = st.sa_fields
            for key, v in got.items():
                ref = g[f"{name}_s{s}_{key}"]
                assert O.rel_l2(v, ref) <= tol, (name, dtype, s, key, O.rel_l2(v, ref))


# ------------------------------------------------------------- analysis
def test_energy_linear_field_closed_form():
    N, chi = 6, 19.0
    x = np.arange(N, dtype=np.float64)
    X = np.meshgrid(x, x, x, indexing="ij")[0]
    E, delta = X.copy(), 0.01
    T, G, V = O.energy_components(E, E - delta, np.full_like(E, chi), DT)
    assert math.isclose(T.sum(), N ** 3 * 0.5 * (delta / DT) ** 2, rel_tol=1e-12)
    assert math.isclose(G.sum(), N ** 3 * 0.5, rel_tol=1e-12)   # d/dx x = 1, edges too
    assert math.isclose(V.sum(), 0.5 * chi * chi * N * N * (x * x).sum(), rel_tol=1e-12)


def test_interior_mask_and_stats_brute_force():
    N, bf = 8, 0.3
    m = O.interior_mask(N, bf)
    rng = np.random.default_rng(1)
    chi = rng.uniform(16, 20, (N, N, N))
    vals = []
    for i in range(N):
        for j in range(N):
            for k in range(N):
                if math.sqrt((i - 4) ** 2 + (j - 4) ** 2 + (k - 4) ** 2) <= 4 * 0.7:
                    vals.append(chi[i, j, k])
    assert m.sum() == len(vals)
    s = O.chi_statistics(chi, m)
    assert math.isclose(s["mean"], sum(vals) / len(vals), rel_tol=1e-13)
    assert s["min"] == min(vals) and s["max"] == max(vals)
    w, v = O.well_void_fraction(chi, 17.0, 18.0, m)
    assert w == sum(x < 17 for x in vals) / len(vals)
    assert v == sum(x > 18 for x in vals) / len(vals)


def test_poisson_single_mode():
    """psi_sq = A + B cos(k.x): delta chi = -kappa B / |k|^2 cos(k.x), DC dropped."""
    N, kappa = 16, 1 / 63
    c, k = plane(N, (1, 3, 2), 0.7)
    A, B = 3.0, 1.5
    chi = O.poisson_chi(A + B * c, N, 19.0, kappa)
    k2 = sum(q * q for q in k)
    np.testing.assert_allclose(chi, 19.0 - kappa * B / k2 * c, atol=1e-12)
    frozen = O.poisson_chi(A + B * c, N, 19.0, kappa, frozen_bf=0.3)
    outside = ~O.interior_mask(N, 0.3)
    assert np.all(frozen[outside] == 19.0)
    np.testing.assert_allclose(frozen[~outside], chi[~outside], atol=0)


# ---------------------------------------------------- eigenmode relaxation
def test_relax_plane_wave_is_a_fixed_shape():
    """kappa = 0 keeps chi = chi0 everywhere, so a plane wave of amplitude A
    is multiplied by 1 + dt (lambda(k) - chi0^2) per step and rescaled back:
    it is a fixed point of the rescaled step, and its eigenvalue estimate is
    sqrt(chi0^2 - lambda(k)) exactly."""
    N, chi0, A = 16, 19.0, 5.0
    f, k = plane(N, (2, 1, 0), 0.4)
    E0 = A * f / np.abs(f).max()
    dt = 0.8 / (chi0 * chi0 + 12.0)
    E, chi, conv, total, cyc = O.relax_eigenmode(E0, N, chi0, 0.0, A, dt, max_cycles=2,
                                                 steps_per_cycle=40, check_interval=20,
                                                 tolerance=1e-300, boundary_fraction=0.3)
    assert np.all(chi == chi0) and total == 80 and not conv
    np.testing.assert_allclose(E, E0, atol=1e-11)
    lam = lam_symbol(*k)
    assert math.isclose(O.relax_eigenvalue(E, chi), math.sqrt(chi0 * chi0 - lam), rel_tol=1e-12)


def test_relax_mode_decay_closed_form():
    """E = a + b cos(k.x) in uniform chi: the step multiplies the k = 0 part
    by 1 - dt chi^2 and the k part by 1 + dt (lambda(k) - chi^2); the rescale
    multiplies both alike, so b/a shrinks by their ratio every step."""
    N, chi0 = 16, 19.0
    f, k = plane(N, (1, 2, 3))
    a, b = 4.0, 1.0
    dt = 0.8 / (chi0 * chi0 + 12.0)
    steps = 25
    E, _, _, total, _ = O.relax_eigenmode(a + b * f, N, chi0, 0.0, a + b, dt, max_cycles=1,
                                          steps_per_cycle=steps, check_interval=1000,
                                          tolerance=1e-300, boundary_fraction=0.3)
    assert total == steps
    lam = lam_symbol(*k)
    r = (1 + dt * (lam - chi0 * chi0)) / (1 - dt * chi0 * chi0)
    a_n = E.mean()
    b_n = 2.0 * np.sum(E * f) / f.size     # projection on cos(k.x) (<cos^2> = 1/2)
    assert math.isclose(b_n / a_n, (b / a) * r ** steps, rel_tol=1e-10)
    # the rescale holds max|E| at the amplitude
    assert math.isclose(float(np.abs(E).max()), a + b, rel_tol=1e-12)


def test_relax_shell_mask_brute_force():
    N, bf = 9, 0.3
    m = O.relax_shell_mask(N, bf)
    c = (N - 1) / 2.0
    cnt = sum(1 for i in range(N) for j in range(N) for k in range(N)
              if math.sqrt((i - c) ** 2 + (j - c) ** 2 + (k - c) ** 2) >= (N * 0.5) * (1 - bf))
    assert m.sum() == cnt and m[0, 0, 0] and not m[4, 4, 4]


def test_relax_chi_clip_and_shell():
    """A deep E^2 source drives chi below 0.1 at the centre: the Poisson chi is
    clipped to [0.1, chi0] and held at chi0 on the shell."""
    N, chi0 = 16, 19.0
    E = np.zeros((N, N, N))
    E[6:10, 6:10, 6:10] = 60.0
    shell = O.relax_shell_mask(N, 0.3)
    chi = O.relax_chi(E, N, chi0, 1 / 63, shell)
    raw = O.poisson_chi(E ** 2, N, chi0, 1 / 63)
    assert raw.min() < 0.1 and chi.min() == 0.1 and chi.max() <= chi0
    assert np.all(chi[shell] == chi0)
    inner = ~shell
    np.testing.assert_allclose(chi[inner], np.clip(raw[inner], 0.1, chi0), atol=0)
