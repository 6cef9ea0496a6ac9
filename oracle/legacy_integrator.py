"""Restatement of the reference's legacy FP64 leapfrog integrator.

TEST INFRASTRUCTURE ONLY (like lfm_oracle.py): a second, independent source
for the oracle's step formulas, used by tests/test_oracle_legacy.py.

It follows lfm/core/integrator.py:103-300 (``step_leapfrog`` and its
``_step_real`` / ``_step_complex`` / ``_step_color``), which differs from the
backend path in ways that make it a cross-check only where they coincide:

* state layout: Psi is (N,N,N) at REAL, (2,N,N,N) = [re, im] at COMPLEX and
  (n_colors,2,N,N,N) at COLOR; chi is (N,N,N); float64 throughout;
* boundary: a hard boolean mask (Psi' = 0, chi' = chi0 inside) when FROZEN,
  no cos^2 taper, no damping of prev and no clamp at -chi0 (:150-156);
* COMPLEX has no epsilon_w * j term; COLOR has no currents, no CCV and no
  SCV (:215-293).

With no mask, epsilon_w = 0 (and at COLOR kappa_string = kappa_tube = 0) and
chi staying above -chi0, a step here equals the oracle's backend-path step
with m = 0, term for term; c multiplies the Laplacians (c = 1 in every
preset).
"""
from __future__ import annotations

import numpy as np

from .lfm_oracle import laplacian_19pt


class LegacyState:
    """psi, psi_prev, chi, chi_prev in the legacy layout (integrator.py:29-56)."""

    def __init__(self, psi, psi_prev, chi, chi_prev, frozen=None):
        f = lambda a: np.array(a, dtype=np.float64, copy=True)
        self.psi, self.psi_prev = f(psi), f(psi_prev)
        self.chi, self.chi_prev = f(chi), f(chi_prev)
        self.frozen = None if frozen is None else np.asarray(frozen, dtype=bool)
        self.step = 0


def _gov02(chi, chi_prev, rho, dt2, c2, kappa, chi0, e0_sq, lambda_self, extra=None):
    """chi' = 2 chi - chi_prev + dt^2 (c^2 lap chi - (kappa/chi0) chi (rho - E0^2)
    - extra - 4 lambda chi (chi^2 - chi0^2)) (integrator.py:143-152,195-201,283-288)."""
    acc = c2 * laplacian_19pt(chi) - (kappa / chi0) * chi * (rho - e0_sq)
    if extra is not None:
        acc = acc - extra
    if lambda_self > 0:
        acc = acc - 4.0 * lambda_self * chi * (chi * chi - chi0 * chi0)
    return 2.0 * chi - chi_prev + dt2 * acc


def _gov01(f, f_prev, chi_sq, dt2, c2):
    """f' = 2 f - f_prev + dt^2 (c^2 lap f - chi^2 f) (integrator.py:137-138)."""
    return 2.0 * f - f_prev + dt2 * (c2 * laplacian_19pt(f) - chi_sq * f)


def step_leapfrog(st: LegacyState, level: int, dt: float, c: float, kappa: float, chi0: float,
                  e0_sq: float = 0.0, lambda_self: float = 0.0, kappa_c: float = 0.0,
                  epsilon_cc: float = 0.0) -> None:
    """One step of integrator.step_leapfrog (level 0 REAL, 1 COMPLEX, 2 COLOR)."""
    dt2, c2 = dt * dt, c * c
    psi, chi = st.psi, st.chi
    chi_sq = chi * chi
    if level == 0:
        new = _gov01(psi, st.psi_prev, chi_sq, dt2, c2)
        rho, extra = psi * psi, None
    elif level == 1:
        new = np.stack([_gov01(psi[q], st.psi_prev[q], chi_sq, dt2, c2) for q in range(2)])
        rho, extra = psi[0] * psi[0] + psi[1] * psi[1], None
    else:
        nc = psi.shape[0]
        new = np.empty_like(psi)
        mean = psi.mean(axis=0) if epsilon_cc > 0 else None     # colour average of [re, im]
        e = np.empty((nc,) + chi.shape)
        for a in range(nc):
            for q in range(2):
                v = _gov01(psi[a, q], st.psi_prev[a, q], chi_sq, dt2, c2)
                if epsilon_cc > 0:
                    v = v - dt2 * epsilon_cc * chi_sq * (psi[a, q] - mean[q])
                new[a, q] = v
            e[a] = psi[a, 0] ** 2 + psi[a, 1] ** 2
        rho = e.sum(axis=0)
        extra = None
        if kappa_c > 0:
            tot2 = rho * rho
            ok = tot2 > 1e-30
            fc = np.where(ok, (e * e).sum(axis=0) / np.where(ok, tot2, 1.0) - 1.0 / nc, 0.0)
            extra = (kappa_c / chi0) * chi * fc * rho
    chi_new = _gov02(chi, st.chi_prev, rho, dt2, c2, kappa, chi0, e0_sq, lambda_self, extra)
    if st.frozen is not None:
        new[..., st.frozen] = 0.0
        chi_new[st.frozen] = chi0
    st.psi_prev, st.psi = psi, new
    st.chi_prev, st.chi = chi, chi_new
    st.step += 1
