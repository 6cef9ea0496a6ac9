"""FP64 autograd reference for gradients of the COMPLEX leapfrog.

TEST INFRASTRUCTURE ONLY (see lfm_oracle.py's header).  The forward map is
the reference's COMPLEX step (lfm/core/backends/numpy_backend.py:132-205)
with the Evolver's buffer semantics (lfm/core/evolver.py:182-281):

    cur_{n+1}  = (1-m) * (2 cur_n - prev_n + dt2 (L cur_n - chi_n^2 cur_n))
    prev_{n+1} = (1-m) * cur_n                       (prev_0 = Psi_0, undamped)
    chi_{n+1}  = m chi0 + (1-m) max(2 chi_n - chi_prev_n + dt2 acc_n, -chi0)
    chi_prev_{n+1} = chi_n
    acc = L chi - (kappa/chi0) chi (|Psi|^2 + eps_w j - E0^2) - 4 lam chi (chi^2 - chi0^2)

written with torch float64 tensor ops on the CPU so torch.autograd gives the
exact FP64 gradient of any loss of the final state with respect to Psi_0 and
kappa.  It is itself pinned against central finite differences of the NumPy
oracle (tests/test_oracle_adjoint.py).
"""
from __future__ import annotations

import numpy as np
import torch


def lap19(f: torch.Tensor) -> torch.Tensor:
    """19-point periodic Laplacian (faces 1/3, edges 1/6, centre -4)."""
    r = torch.roll
    faces = sum(r(f, s, d) for d in range(3) for s in (1, -1))
    edges = 0
    for d1, d2 in ((0, 1), (0, 2), (1, 2)):
        for s1 in (1, -1):
            for s2 in (1, -1):
                edges = edges + r(r(f, s1, d1), s2, d2)
    return faces / 3.0 + edges / 6.0 - 4.0 * f


def current(pr: torch.Tensor, pi: torch.Tensor) -> torch.Tensor:
    """j = 1/2 sum_d [Pr (Pi[+d] - Pi[-d]) - Pi (Pr[+d] - Pr[-d])]."""
    j = 0
    for d in range(3):
        dpi = torch.roll(pi, -1, d) - torch.roll(pi, 1, d)
        dpr = torch.roll(pr, -1, d) - torch.roll(pr, 1, d)
        j = j + pr * dpi - pi * dpr
    return 0.5 * j


def forward_complex(pr0, pi0, chi_init, kappa, mask, steps, dt=0.02, lam=0.0, chi0=19.0,
                    e0_sq=0.0, eps_w=0.1):
    """Run `steps` COMPLEX steps from Psi_0 = (pr0, pi0) (prev tied to Psi_0)
    and chi = chi_prev = chi_init; return (pr_T, pi_T, chi_T)."""
    dt2 = dt * dt
    a = 1.0 - mask
    cr, ci, pr_, pi_ = pr0, pi0, pr0, pi0
    c, cp = chi_init, chi_init
    for _ in range(steps):
        csq = c * c
        nr = a * (2 * cr - pr_ + dt2 * (lap19(cr) - csq * cr))
        ni = a * (2 * ci - pi_ + dt2 * (lap19(ci) - csq * ci))
        S = cr * cr + ci * ci + eps_w * current(cr, ci) - e0_sq
        acc = lap19(c) - (kappa / chi0) * c * S - 4.0 * lam * c * (csq - chi0 * chi0)
        h = 2 * c - cp + dt2 * acc
        h = torch.where(h > -chi0, h, torch.full_like(h, -chi0))
        cn = mask * chi0 + a * h
        pr_, pi_ = a * cr, a * ci
        cr, ci = nr, ni
        cp, c = c, cn
    return cr, ci, c


def grads(pr0, pi0, chi_init, kappa, mask, steps, chi_target, **kw):
    """FP64 autograd gradient of sum((chi_T - chi_target)^2) w.r.t. Psi_0, kappa."""
    t = lambda x: torch.as_tensor(np.asarray(x, dtype=np.float64))
    pr = t(pr0).requires_grad_(True)
    pi = t(pi0).requires_grad_(True)
    k = torch.tensor(float(kappa), dtype=torch.float64, requires_grad=True)
    _, _, chiT = forward_complex(pr, pi, t(chi_init), k, t(mask), steps, **kw)
    loss = ((chiT - t(chi_target)) ** 2).sum()
    loss.backward()
    return float(loss.detach()), pr.grad.numpy(), pi.grad.numpy(), float(k.grad)


def vjp(pr0, pi0, chi_init, kappa, mask, steps, cotangent, **kw):
    """FP64 vector-Jacobian product of chi_T with a fixed cotangent:
    gradients of sum(cotangent * chi_T) w.r.t. Psi_0 and kappa."""
    t = lambda x: torch.as_tensor(np.asarray(x, dtype=np.float64))
    pr = t(pr0).requires_grad_(True)
    pi = t(pi0).requires_grad_(True)
    k = torch.tensor(float(kappa), dtype=torch.float64, requires_grad=True)
    _, _, chiT = forward_complex(pr, pi, t(chi_init), k, t(mask), steps, **kw)
    (chiT * t(cotangent)).sum().backward()
    return pr.grad.numpy(), pi.grad.numpy(), float(k.grad)
