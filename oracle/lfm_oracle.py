"""CPU oracle for the GOV-01/GOV-02 leapfrog step of lfm-physics.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import this
module: it is the checker for the HIP kernels (``tests/``), the body of
``__graft_entry__.smoke()``'s comparison and the timed CPU leg of
``bench.py`` (``cpu_baseline``).  It is never the thing measured or shipped.

It restates, in NumPy, the reference CPU backend op for op.  Every function
is dtype-generic: fed float64 arrays it runs entirely in float64 (the parity
oracle); fed float32 arrays it mirrors the reference's float32 NumPy path
(the CPU baseline).  Unlike the reference ``step_color`` it never drops to
float32 scratch, so the float64 run is float64 end to end.

Pinning (what checks each function independently of this restatement;
tests/test_oracle_pins.py, test_oracle_golden.py, test_oracle_legacy.py):

* ``laplacian_19pt``   -- plane-wave symbol closed form, x^2 -> 2, constants
                          -> 0, cubic symmetry.
* ``boundary_mask``    -- centre 0 / corner 1 / taper values at known radii;
                          bit-equal to the reference's create_boundary_mask at
                          N = 8, 16 (two boundary fractions) and by SHA-256 at
                          N = 64, 128 (tests/golden/step_real_complex.npz).
* ``step_real``        -- exact discrete dispersion recurrence for GOV-01,
                          scalar GOV-02 recurrences (gravity, Mexican hat,
                          clamp), mask blend; the reference's own masked,
                          damped-prev outputs (three steps, N = 8 float32 bit
                          for bit, N = 8 and 16 float64 to 1e-14); the legacy
                          FP64 integrator (lfm/core/integrator.py:128-160,
                          restated in legacy_integrator.py) with no mask.
* ``step_complex``     -- as REAL plus the plane-wave current closed form
                          j = A^2 * sum_d sin(k_d); the reference's masked,
                          damped-prev outputs as for REAL; the legacy
                          integrator with epsilon_w = 0 and no mask.
* ``step_color``       -- colour special cases (singlet f_c = 0, one colour
                          f_c = 2/3, CCV/SCV/eps_cc vanish for identical
                          colours, one-colour CCV = 2/3 sum j^2); golden
                          vectors from the reference ``step_color`` (all
                          terms on, including SCV, N = 8); the legacy
                          integrator (f_c and cross-colour terms) with no
                          mask, no currents, CCV or SCV.
* ``helmholtz_sa``     -- single-mode closed form S = A + B*g/(g+D k^2) cos;
                          DC gain 1; inside the step_color golden vectors.
* ``OracleState``      -- the A/B toggle through all the step vectors above;
                          ``step(freeze_chi=True)`` by the alternating-chi
                          mode recurrence (uniform A and B chi sets).
* ``energy_components``-- real plane-wave and complex plane-wave closed forms
                          (T, G with np.gradient's one-sided edges, V; the
                          colour axis summed); the reference's metrics() of
                          two reference-written checkpoints (COMPLEX and COLOR,
                          1e-12); np.gradient is the NumPy routine the
                          reference calls.
* ``chi_statistics`` / ``interior_mask`` -- counted by brute force; the
                          reference checkpoints' metrics (the reference takes
                          mean and std in float32, so those agree to its
                          rounding).
* ``poisson_chi``      -- single-mode closed form, DC dropped.
* ``relax_*``          -- a plane wave in uniform chi is a fixed shape of the
                          rescaled imaginary-time step and its eigenvalue is
                          sqrt(chi^2 - lambda(k)); a mode's amplitude against
                          the k = 0 mode decays by the closed-form factor
                          (1 + dt (lambda(k) - chi^2)) / (1 - dt chi^2) per
                          step; shell mask counted by brute force.
"""

from __future__ import annotations

import numpy as np

# Stencil weights: lfm/constants.py:135-142 (face 1/3, edge 1/6, centre -4).
W_FACE = 1.0 / 3.0
W_EDGE = 1.0 / 6.0
W_CENTER = -4.0


def laplacian_19pt(f: np.ndarray) -> np.ndarray:
    """19-point periodic Laplacian of a (N,N,N) field.

    Follows lfm/core/stencils.py:26-71: six face neighbours weighted 1/3,
    twelve edge neighbours weighted 1/6, centre -4, all by np.roll.
    """
    r = np.roll
    faces = (r(f, 1, 0) + r(f, -1, 0) + r(f, 1, 1) + r(f, -1, 1)
             + r(f, 1, 2) + r(f, -1, 2))
    edges = (r(r(f, 1, 0), 1, 1) + r(r(f, 1, 0), -1, 1)
             + r(r(f, -1, 0), 1, 1) + r(r(f, -1, 0), -1, 1)
             + r(r(f, 1, 0), 1, 2) + r(r(f, 1, 0), -1, 2)
             + r(r(f, -1, 0), 1, 2) + r(r(f, -1, 0), -1, 2)
             + r(r(f, 1, 1), 1, 2) + r(r(f, 1, 1), -1, 2)
             + r(r(f, -1, 1), 1, 2) + r(r(f, -1, 1), -1, 2))
    return W_FACE * faces + W_EDGE * edges + W_CENTER * f


def _lap_flat(flat: np.ndarray, N: int) -> np.ndarray:
    # lfm/core/backends/numpy_backend.py:71-78
    return laplacian_19pt(flat.reshape(N, N, N)).ravel()


def boundary_mask(N: int, boundary_fraction: float) -> np.ndarray:
    """cos^2 absorbing mask, float32, flat N^3.

    Follows lfm/core/backends/numpy_backend.py:49-69 (same float32 NumPy ops,
    so the values are bit-identical to the reference's host-built mask).
    """
    center = N / 2.0
    r_max = N / 2.0
    r_freeze = (1.0 - boundary_fraction) * r_max
    coords = np.arange(N, dtype=np.float32) - center + 0.5
    X, Y, Z = np.meshgrid(coords, coords, coords, indexing="ij")
    R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    t = np.clip((R - r_freeze) / (r_max - r_freeze), 0.0, 1.0)
    return (np.sin(0.5 * np.pi * t) ** 2).astype(np.float32).ravel()


def step_real(E, E_prev, chi, chi_prev, mask, E_out, E_prev_out, chi_out,
              chi_prev_out, N, dt2, kappa, lambda_self, chi0, e0_sq):
    """REAL-level step (GOV-01 for E, GOV-02 for chi).

    Follows lfm/core/backends/numpy_backend.py:80-130; GPU operation order
    lfm/core/backends/kernel_source.py:377-417.
    """
    lap_E = _lap_flat(E, N)
    lap_chi = _lap_flat(chi, N)
    chi_sq = chi * chi
    E_new = 2.0 * E - E_prev + dt2 * (lap_E - chi_sq * E)
    chi_accel = lap_chi - (kappa / chi0) * chi * (E * E - e0_sq)
    if lambda_self > 0:
        chi_accel = chi_accel - 4.0 * lambda_self * chi * (chi_sq - chi0 * chi0)
    chi_new = 2.0 * chi - chi_prev + dt2 * chi_accel
    chi_new = np.maximum(chi_new, -chi0)               # BH excision clamp
    absorb = 1.0 - mask
    np.copyto(E_out, E_new * absorb)
    np.copyto(E_prev_out, E * absorb)
    np.copyto(chi_out, mask * chi0 + absorb * chi_new)
    np.copyto(chi_prev_out, chi)


def _currents(Pr: np.ndarray, Pi: np.ndarray, N: int):
    """Per-direction currents j_d = Pr*(Pi[+d]-Pi[-d]) - Pi*(Pr[+d]-Pr[-d]).

    Unscaled periodic central differences, numpy_backend.py:171-182.
    """
    Pr3 = Pr.reshape(N, N, N)
    Pi3 = Pi.reshape(N, N, N)
    out = []
    for ax in range(3):
        dPr = np.roll(Pr3, -1, ax) - np.roll(Pr3, 1, ax)
        dPi = np.roll(Pi3, -1, ax) - np.roll(Pi3, 1, ax)
        out.append((Pr3 * dPi - Pi3 * dPr).ravel())
    return out


def step_complex(Pr, Pr_prev, Pi, Pi_prev, chi, chi_prev, mask,
                 Pr_out, Pr_prev_out, Pi_out, Pi_prev_out, chi_out, chi_prev_out,
                 N, dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w):
    """COMPLEX-level step.  Follows numpy_backend.py:132-205."""
    chi_sq = chi * chi
    lap_Pr = _lap_flat(Pr, N)
    lap_Pi = _lap_flat(Pi, N)
    lap_chi = _lap_flat(chi, N)
    Pr_new = 2.0 * Pr - Pr_prev + dt2 * (lap_Pr - chi_sq * Pr)
    Pi_new = 2.0 * Pi - Pi_prev + dt2 * (lap_Pi - chi_sq * Pi)
    psi_sq = Pr * Pr + Pi * Pi
    jx, jy, jz = _currents(Pr, Pi, N)
    j_total = 0.5 * (jx + jy + jz)
    chi_accel = lap_chi - (kappa / chi0) * chi * (psi_sq + epsilon_w * j_total - e0_sq)
    if lambda_self > 0:
        chi_accel = chi_accel - 4.0 * lambda_self * chi * (chi_sq - chi0 * chi0)
    chi_new = np.maximum(2.0 * chi - chi_prev + dt2 * chi_accel, -chi0)
    absorb = 1.0 - mask
    np.copyto(Pr_out, Pr_new * absorb)
    np.copyto(Pi_out, Pi_new * absorb)
    np.copyto(Pr_prev_out, Pr * absorb)
    np.copyto(Pi_prev_out, Pi * absorb)
    np.copyto(chi_out, mask * chi0 + absorb * chi_new)
    np.copyto(chi_prev_out, chi)


def helmholtz_sa(e: np.ndarray, N: int, gamma: float, D: float) -> np.ndarray:
    """S = max(0, irfftn(gamma/(gamma + D k^2) * rfftn(e))) for one colour.

    k_d = 2*pi*fftfreq(N) (continuous k^2, not the stencil symbol).
    Follows numpy_backend.py:342-353 and cupy_backend.py:262-275; computed
    in the dtype of ``e`` promoted to at least float64 for the transform.
    """
    kx = np.fft.fftfreq(N) * (2.0 * np.pi)
    kz = np.fft.rfftfreq(N) * (2.0 * np.pi)
    k_sq = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    h = np.float64(gamma) / (np.float64(gamma) + np.float64(D) * k_sq)
    s = np.fft.irfftn(h * np.fft.rfftn(e.reshape(N, N, N)), s=(N, N, N), axes=(0, 1, 2))
    return np.clip(s, 0.0, None).astype(e.dtype).ravel()


def step_color(Pr_in, Pr_prev_in, Pi_in, Pi_prev_in, chi, chi_prev, mask,
               Pr_out, Pr_prev_out, Pi_out, Pi_prev_out, chi_out, chi_prev_out,
               N, dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w,
               kappa_c=0.0, epsilon_cc=0.0, kappa_string=0.0, kappa_tube=0.0,
               sa_fields_in=None, sa_fields_out=None, sa_gamma=0.1, sa_d=4.9,
               dt=0.02, n_colors=3):
    """3-colour step (GOV-01 with cross-colour coupling; GOV-02 with f_c,
    CCV and SCV).  Follows numpy_backend.py:207-382 with the term guards of
    kernel_source.py:103-233 (f_c guard on total_sq > 1e-30).

    S_a is computed from the *input* Psi whenever kappa_tube > 0 and
    ``sa_fields_out`` is given (numpy_backend.py:341); ``sa_fields_in`` is
    never read, as in the reference.
    """
    total = N ** 3
    dt_ = chi.dtype
    chi_sq = chi * chi
    C = n_colors
    Pr = [Pr_in[a * total:(a + 1) * total] for a in range(C)]
    Pi = [Pi_in[a * total:(a + 1) * total] for a in range(C)]
    if epsilon_cc > 0:
        Pr_avg = sum(Pr) / C
        Pi_avg = sum(Pi) / C
    e = []
    psi_sq_total = np.zeros(total, dtype=dt_)
    j_total = np.zeros(total, dtype=dt_)
    jdir = []
    for a in range(C):
        s = slice(a * total, (a + 1) * total)
        lap_Pr = _lap_flat(Pr[a], N)
        lap_Pi = _lap_flat(Pi[a], N)
        Pr_new = 2.0 * Pr[a] - Pr_prev_in[s] + dt2 * (lap_Pr - chi_sq * Pr[a])
        Pi_new = 2.0 * Pi[a] - Pi_prev_in[s] + dt2 * (lap_Pi - chi_sq * Pi[a])
        if epsilon_cc > 0:
            Pr_new = Pr_new - dt2 * epsilon_cc * chi_sq * (Pr[a] - Pr_avg)
            Pi_new = Pi_new - dt2 * epsilon_cc * chi_sq * (Pi[a] - Pi_avg)
        np.copyto(Pr_out[s], Pr_new)
        np.copyto(Pi_out[s], Pi_new)
        np.copyto(Pr_prev_out[s], Pr[a])
        np.copyto(Pi_prev_out[s], Pi[a])
        ea = Pr[a] * Pr[a] + Pi[a] * Pi[a]
        e.append(ea)
        psi_sq_total += ea
        jx, jy, jz = _currents(Pr[a], Pi[a], N)
        j_total += 0.5 * (jx + jy + jz)
        jdir.append((jx, jy, jz))

    cv_term = np.zeros(total, dtype=dt_)
    if kappa_c > 0:
        total_sq = psi_sq_total * psi_sq_total
        safe = total_sq > 1e-30
        sum_sq = sum(ea * ea for ea in e)
        ratio = np.divide(sum_sq, total_sq, out=np.zeros_like(sum_sq), where=safe)
        f_c = np.where(safe, ratio - 1.0 / C, 0.0).astype(dt_)
        cv_term = (kappa_c / chi0) * chi * f_c * psi_sq_total

    ccv = np.zeros(total, dtype=dt_)
    if kappa_string > 0:
        for d in range(3):
            jd = [jdir[a][d] for a in range(C)]
            ccv += sum(x * x for x in jd) - (1.0 / C) * sum(jd) ** 2

    scv = np.zeros(total, dtype=dt_)
    if kappa_tube > 0 and sa_fields_out is not None:
        sa_sum = np.zeros(total, dtype=dt_)
        sa_sq = np.zeros(total, dtype=dt_)
        for a in range(C):
            sa = helmholtz_sa(e[a], N, sa_gamma, sa_d)
            np.copyto(sa_fields_out[a * total:(a + 1) * total], sa)
            sa_sum += sa
            sa_sq += sa * sa
        scv = sa_sq - (1.0 / C) * sa_sum ** 2

    lap_chi = _lap_flat(chi, N)
    chi_accel = (lap_chi - (kappa / chi0) * chi * (psi_sq_total + epsilon_w * j_total - e0_sq)
                 - cv_term - kappa_string * ccv - kappa_tube * scv)
    if lambda_self > 0:
        chi_accel = chi_accel - 4.0 * lambda_self * chi * (chi_sq - chi0 * chi0)
    chi_new = np.maximum(2.0 * chi - chi_prev + dt2 * chi_accel, -chi0)

    absorb3 = 1.0 - np.tile(mask, C)
    Pr_out *= absorb3
    Pi_out *= absorb3
    Pr_prev_out *= absorb3
    Pi_prev_out *= absorb3
    np.copyto(chi_out, mask * chi0 + (1.0 - mask) * chi_new)
    np.copyto(chi_prev_out, chi)


# ---------------------------------------------------------------------------
# Double-buffered driver (the Evolver's A/B toggle, lfm/core/evolver.py:182-281)
# ---------------------------------------------------------------------------
class OracleState:
    """A/B double-buffered state, exactly as ``Evolver`` holds it.

    level: 0 REAL, 1 COMPLEX, 2 COLOR.  ``params`` carries the scalar
    arguments the Evolver passes to ``step_*`` (evolver.py:201-279).
    """

    def __init__(self, level, N, psi_r, psi_i, chi, mask, params,
                 psi_r_prev=None, psi_i_prev=None, chi_prev=None,
                 dtype=np.float64, n_colors=3):
        self.level, self.N, self.p, self.C = level, N, dict(params), n_colors
        cv = lambda a: np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel())
        pr = cv(psi_r)
        self.r = [pr.copy(), pr.copy()]
        self.rp = [cv(psi_r_prev) if psi_r_prev is not None else pr.copy() for _ in range(2)]
        if level > 0:
            pi = cv(psi_i)
            self.i = [pi.copy(), pi.copy()]
            self.ip = [cv(psi_i_prev) if psi_i_prev is not None else pi.copy() for _ in range(2)]
        c = cv(chi)
        self.c = [c.copy(), c.copy()]
        self.cp = [cv(chi_prev) if chi_prev is not None else c.copy() for _ in range(2)]
        self.mask = cv(mask)
        self.sa = None
        if level == 2 and self.p.get("kappa_tube", 0.0) > 0:
            self.sa = [np.zeros(n_colors * N ** 3, dtype=dtype) for _ in range(2)]
        self.cur = 0
        self.steps = 0

    def step(self, n=1, freeze_chi=False):
        """n steps; freeze_chi restores all four chi buffers after every step
        from a snapshot taken at the call (lfm/core/evolver.py:160-176)."""
        p = self.p
        snap = [x.copy() for x in (self.c[0], self.c[1], self.cp[0], self.cp[1])] if freeze_chi else None
        for _ in range(n):
            a, b = self.cur, 1 - self.cur
            common = dict(N=self.N, dt2=p["dt2"], kappa=p["kappa"],
                          lambda_self=p["lambda_self"], chi0=p["chi0"], e0_sq=p["e0_sq"])
            if self.level == 0:
                step_real(self.r[a], self.rp[a], self.c[a], self.cp[a], self.mask,
                          self.r[b], self.rp[b], self.c[b], self.cp[b], **common)
            elif self.level == 1:
                step_complex(self.r[a], self.rp[a], self.i[a], self.ip[a], self.c[a],
                             self.cp[a], self.mask, self.r[b], self.rp[b], self.i[b],
                             self.ip[b], self.c[b], self.cp[b], epsilon_w=p["epsilon_w"],
                             **common)
            else:
                step_color(self.r[a], self.rp[a], self.i[a], self.ip[a], self.c[a],
                           self.cp[a], self.mask, self.r[b], self.rp[b], self.i[b],
                           self.ip[b], self.c[b], self.cp[b], epsilon_w=p["epsilon_w"],
                           kappa_c=p.get("kappa_c", 0.0), epsilon_cc=p.get("epsilon_cc", 0.0),
                           kappa_string=p.get("kappa_string", 0.0),
                           kappa_tube=p.get("kappa_tube", 0.0),
                           sa_fields_in=self.sa[a] if self.sa else None,
                           sa_fields_out=self.sa[b] if self.sa else None,
                           sa_gamma=p.get("sa_gamma", 0.1), sa_d=p.get("sa_d", 4.9),
                           n_colors=self.C, **common)
            if snap is not None:
                for dst, src in zip((self.c[0], self.c[1], self.cp[0], self.cp[1]), snap):
                    np.copyto(dst, src)
            self.cur = b
            self.steps += 1
        return self

    def set_psi_real(self, a):
        """All four Psi_r buffers := a (evolver.py:355-375)."""
        for buf in self.r + self.rp:
            np.copyto(buf, np.asarray(a, dtype=buf.dtype).ravel())

    def set_chi_current(self, a):
        for buf in self.c:
            np.copyto(buf, np.asarray(a, dtype=buf.dtype).ravel())

    def set_chi_prev(self, a):
        for buf in self.cp:
            np.copyto(buf, np.asarray(a, dtype=buf.dtype).ravel())

    def set_psi_real_prev(self, a):
        for buf in self.rp:
            np.copyto(buf, np.asarray(a, dtype=buf.dtype).ravel())

    # current-buffer views
    @property
    def psi_r(self):
        return self.r[self.cur]

    @property
    def psi_r_prev(self):
        return self.rp[self.cur]

    @property
    def psi_i(self):
        return self.i[self.cur] if self.level > 0 else None

    @property
    def psi_i_prev(self):
        return self.ip[self.cur] if self.level > 0 else None

    @property
    def chi(self):
        return self.c[self.cur]

    @property
    def chi_prev(self):
        return self.cp[self.cur]

    @property
    def sa_fields(self):
        return self.sa[self.cur] if self.sa else None


def poisson_chi(psi_sq, N, chi0, kappa, e0_sq=0.0, frozen_bf=None):
    """GOV-04 equilibrated chi: lap(delta) = kappa (|Psi|^2 - E0^2) solved with
    -1/k^2 (DC dropped), chi = chi0 + delta, chi = chi0 outside the interior
    ball for a FROZEN boundary.  Follows lfm/fields/equilibrium.py:26-100.
    """
    src = kappa * (np.asarray(psi_sq, dtype=np.float64).reshape(N, N, N) - e0_sq)
    kx = np.fft.fftfreq(N) * 2.0 * np.pi
    kz = np.fft.rfftfreq(N) * 2.0 * np.pi
    K2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    K2[0, 0, 0] = 1.0
    phi_hat = -np.fft.rfftn(src) / K2
    phi_hat[0, 0, 0] = 0.0
    chi = chi0 + np.fft.irfftn(phi_hat, s=(N, N, N), axes=(0, 1, 2))
    if frozen_bf is not None:
        chi[~interior_mask(N, frozen_bf)] = chi0
    return chi


# ---------------------------------------------------------------------------
# Eigenmode relaxation (lfm/particles/solver.py:454-534,620-826,828-999)
# ---------------------------------------------------------------------------
def relax_shell_mask(N, boundary_fraction=0.3):
    """True in the frozen outer shell: r >= (N/2)(1 - bf) about the centre
    (N-1)/2, with the reference's float32 coordinates
    (particles/solver.py:517-524)."""
    c = (N - 1) / 2.0
    x = np.arange(N, dtype=np.float32)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2)
    return r >= (N * 0.5) * (1.0 - boundary_fraction)


def relax_seed(N, position, amplitude, sigma):
    """l = 0 seed amplitude * exp(-r^2 / sigma^2), computed in float64 and
    stored as float32 (particles/solver.py:917-922)."""
    idx = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    cx, cy, cz = position
    r2 = ((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (sigma ** 2)
    return (amplitude * np.exp(-r2)).astype(np.float32)


def relax_chi(E, N, chi0, kappa, shell):
    """Poisson chi of one relaxation cycle: chi0 + (-kappa E^2 / k^2, DC
    dropped), clipped to [0.1, chi0], chi0 on the shell
    (particles/solver.py:760-772)."""
    chi = poisson_chi(np.asarray(E, np.float64) ** 2, N, chi0, kappa)
    chi = np.clip(np.where(chi < 0.0, 0.1, chi), 0.1, chi0)
    chi[shell] = chi0
    return chi


def relax_eigenmode(E, N, chi0, kappa, amplitude, dt_E, max_cycles, steps_per_cycle,
                    check_interval, tolerance, boundary_fraction, dtype=np.float64):
    """Poisson / imaginary-time relaxation cycles of particles/solver.py:742-826:
    per cycle chi from E, then up to steps_per_cycle steps of
    E <- E + dt_E (lap E - chi^2 E) rescaled to max|E| = amplitude, stopping
    early when max|dE| < tolerance at a check; converged when a whole cycle
    changes E by less than tolerance (after the first cycle).  Returns
    (E, chi, converged, total_steps, cycle) with a final Poisson chi."""
    shell = relax_shell_mask(N, boundary_fraction)
    E = np.asarray(E, dtype=dtype).reshape(N, N, N).copy()
    converged, total, cycle = False, 0, 0
    for cycle in range(max_cycles):
        chi = relax_chi(E, N, chi0, kappa, shell).astype(dtype)
        chi2 = chi * chi
        E_prev = E.copy()
        e_conv = False
        for step in range(1, steps_per_cycle + 1):
            E_new = E + dt_E * (laplacian_19pt(E) - chi2 * E)
            m = float(np.max(np.abs(E_new)))
            if m > 1e-8:
                E_new *= amplitude / m
            if step % check_interval == 0:
                if float(np.max(np.abs(E_new - E))) < tolerance:
                    e_conv = True
                    E = E_new
                    total += step
                    break
            E = E_new
        if not e_conv:
            total += steps_per_cycle
        if float(np.max(np.abs(E - E_prev))) < tolerance and cycle > 0:
            converged = True
            break
    chi = relax_chi(E, N, chi0, kappa, shell)
    return E, chi, converged, total, cycle


def relax_eigenvalue(E, chi):
    """omega = sqrt(<E|(-lap + chi^2)|E> / <E|E>) in float64
    (particles/solver.py:973-977)."""
    E = np.asarray(E, np.float64)
    chi = np.asarray(chi, np.float64).reshape(E.shape)
    HE = -laplacian_19pt(E) + chi * chi * E
    return float(np.sqrt(max(float(np.sum(E * HE) / max(float(np.sum(E * E)), 1e-30)), 0.0)))


# ---------------------------------------------------------------------------
# FP64 analysis reductions (lfm/analysis/energy.py, structure.py)
# ---------------------------------------------------------------------------
def energy_components(psi_r, psi_r_prev, chi, dt, c=1.0, psi_i=None, psi_i_prev=None):
    """Kinetic, gradient and potential energy densities in float64.

    Follows lfm/analysis/energy.py:26-113: T = 1/2 (dPsi/dt)^2 by backward
    difference, G = 1/2 c^2 |grad Psi|^2 with np.gradient (second-order
    central interior, one-sided first-order edges, not periodic),
    V = 1/2 chi^2 |Psi|^2.  Colour axis (ndim 4) summed.
    """
    f64 = lambda a: np.asarray(a, dtype=np.float64)

    def csum(x):
        return x.sum(axis=0) if x.ndim == 4 else x

    T = 0.5 * csum(((f64(psi_r) - f64(psi_r_prev)) / dt) ** 2)
    if psi_i is not None and psi_i_prev is not None:
        T = T + 0.5 * csum(((f64(psi_i) - f64(psi_i_prev)) / dt) ** 2)

    def grad_sq(f):
        f = f64(f)
        fs = f if f.ndim == 4 else f[None]
        tot = np.zeros(fs.shape[1:])
        for a in range(fs.shape[0]):
            for ax in range(3):
                tot += np.gradient(fs[a], axis=ax) ** 2
        return tot

    G = 0.5 * c * c * grad_sq(psi_r)
    if psi_i is not None:
        G = G + 0.5 * c * c * grad_sq(psi_i)
    psq = csum(f64(psi_r) ** 2)
    if psi_i is not None:
        psq = psq + csum(f64(psi_i) ** 2)
    V = 0.5 * f64(chi) ** 2 * psq
    return T, G, V


def total_energy(psi_r, psi_r_prev, chi, dt, c=1.0, psi_i=None, psi_i_prev=None):
    """Sum of T + G + V (lfm/analysis/energy.py:116-135)."""
    T, G, V = energy_components(psi_r, psi_r_prev, chi, dt, c, psi_i, psi_i_prev)
    return float(np.sum(T + G + V))


def interior_mask(N, boundary_fraction=0.3):
    """dist((x,y,z), N/2) <= (1-bf) N/2 (lfm/analysis/structure.py:148-171)."""
    center = N / 2.0
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    d = np.sqrt((X - center) ** 2 + (Y - center) ** 2 + (Z - center) ** 2)
    return d <= center * (1.0 - boundary_fraction)


def chi_statistics(chi, interior=None):
    """min/max/mean/std of chi over the region (structure.py:26-54)."""
    region = chi[interior] if interior is not None else np.asarray(chi).ravel()
    return {"min": float(np.min(region)), "max": float(np.max(region)),
            "mean": float(np.mean(region)), "std": float(np.std(region))}


def well_void_fraction(chi, well=17.0, void=18.0, interior=None):
    """Fractions chi < well and chi > void (structure.py:57-108)."""
    region = chi[interior] if interior is not None else np.asarray(chi).ravel()
    if region.size == 0:
        return 0.0, 0.0
    return float(np.mean(region < well)), float(np.mean(region > void))


def rel_l2(a, b):
    """Relative L2 distance ||a-b|| / ||b|| in float64."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
