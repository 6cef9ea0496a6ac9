"""Differentiable form: gradients of a final-chi loss through the COMPLEX leapfrog.

``chi_T = evolve_chi(pr0, pi0, kappa, config, steps)`` is a
``torch.autograd.Function``: the forward runs the hand-written step kernel
(``lfm_step_complex_f32``) and keeps a tape of states; the backward runs the
hand-written adjoint kernel (``lfm_adjoint_complex_step_f32``) from step T
down to 0 and returns dL/dPsi_0 (real and imaginary parts, Psi_prev tied to
Psi_0 as ``set_psi_*`` ties them), dL/dkappa and dL/dchi_0.

Tape.  With stride K the forward keeps a checkpoint every K steps (current
Psi, the prev Psi it steps from, chi and chi_prev); the backward recomputes
each K-step segment from its checkpoint into a K-state buffer and runs the
adjoint through it.  K = ceil(sqrt(T)) keeps O(sqrt(T)) states (long
horizons fit in HBM) for one extra forward pass; K = 1 keeps every state
and skips the recompute.  ``stride="auto"`` keeps every state when they fit
in `memory_budget` bytes, else uses sqrt(T).

The dynamics are the reference's COMPLEX step with the Evolver's buffer
semantics (first step reads the undamped prev); chi is stored as chi - chi0
as in :class:`Evolver`.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _abi
from .backend import params, stream_handle
from .config import BoundaryType, FieldLevel, SimulationConfig
from .evolver import get_backend


class ComplexTape:
    """Forward stepping with a checkpoint tape and the adjoint sweep."""

    def __init__(self, config: SimulationConfig, steps: int, stride="auto",
                 memory_budget: float = 64e9, mask_policy: str = "reference", backend="auto"):
        if config.field_level != FieldLevel.COMPLEX:
            raise ValueError("the adjoint is implemented for FieldLevel.COMPLEX")
        self.cfg = config
        self.N = config.grid_size
        self.T = int(steps)
        self.backend = get_backend(backend)
        self.device = self.backend.device
        self._lib = _abi.load()
        n3 = self.N ** 3
        if stride == "auto":
            stride = 1 if 3 * 4 * n3 * (self.T + 2) <= memory_budget else math.ceil(math.sqrt(self.T))
        self.K = max(1, int(stride))
        self.has_mask = not (mask_policy == "boundary_type"
                             and config.boundary_type == BoundaryType.PERIODIC)
        self.mask = self.backend.create_boundary_mask(self.N, config.boundary_fraction)
        self.absorb = 1.0 - self.mask
        self.chi_off = float(np.float32(config.chi0))
        self.kappa = float(config.kappa)

    # ------------------------------------------------------------ forward
    def _flags(self, damped: bool) -> int:
        f = _abi.LFM_F_CHI_DELTA
        if self.has_mask:
            f |= _abi.LFM_F_MASK | (_abi.LFM_F_DAMP_PREV if damped else 0)
        return f

    def _p(self, flags):
        c = self.cfg
        return params(self.N, c.dt ** 2, self.kappa, c.lambda_self, c.chi0, c.e0_sq, c.epsilon_w,
                      flags=flags)

    def _fwd(self, cr, ci, pr_, pi_, c, cp, damped, out_r, out_i, out_c):
        p = self._p(self._flags(damped))
        rc = self._lib.lfm_step_complex_f32(
            cr.data_ptr(), pr_.data_ptr(), ci.data_ptr(), pi_.data_ptr(), c.data_ptr(),
            cp.data_ptr(), self.mask.data_ptr() if self.has_mask else None, out_r.data_ptr(),
            None, out_i.data_ptr(), None, out_c.data_ptr(), None, ctypes.byref(p),
            stream_handle())
        _abi.check(rc, "lfm_step_complex_f32")

    def forward(self, pr0: torch.Tensor, pi0: torch.Tensor, kappa: float, chi_init: torch.Tensor):
        """Run T steps from Psi_0 (prev tied) and chi_init; return chi_T."""
        self.kappa = float(kappa)
        n3 = self.N ** 3
        f32 = dict(dtype=torch.float32, device=self.device)
        pr0 = pr0.detach().to(**f32).reshape(-1)
        pi0 = pi0.detach().to(**f32).reshape(-1)
        chi_d = chi_init.detach().to(**f32).reshape(-1) - self.chi_off
        if self.K == 1:
            # every state kept: one allocation, each step writes into its slot
            R = torch.empty((self.T + 1, n3), **f32)
            I = torch.empty((self.T + 1, n3), **f32)
            C = torch.empty((self.T + 1, n3), **f32)
            R[0].copy_(pr0)
            I[0].copy_(pi0)
            C[0].copy_(chi_d)
            for n in range(self.T):
                m = n - 1 if n else 0
                self._fwd(R[n], I[n], R[m], I[m], C[n], C[m], n > 0, R[n + 1], I[n + 1], C[n + 1])
            self.tape = (R, I, C)
            return C[self.T] + self.chi_off
        # three-slot rotation; checkpoint (cur, prev, chi, chi_prev) every K steps
        nck = (self.T + self.K - 1) // self.K
        self.ck = torch.empty((nck, 6, n3), **f32)
        self.ck_damped = [False] * nck
        S = torch.empty((3, 3, n3), **f32)     # [slot][pr, pi, chi]
        S[0, 0].copy_(pr0); S[1, 0].copy_(pr0)
        S[0, 1].copy_(pi0); S[1, 1].copy_(pi0)
        S[0, 2].copy_(chi_d); S[1, 2].copy_(chi_d)
        cur, prev, cc, cp = 0, 1, 0, 1
        damped = False
        for n in range(self.T):
            if n % self.K == 0:
                c = self.ck[n // self.K]
                c[0].copy_(S[cur, 0]); c[1].copy_(S[cur, 1])
                c[2].copy_(S[prev, 0]); c[3].copy_(S[prev, 1])
                c[4].copy_(S[cc, 2]); c[5].copy_(S[cp, 2])
                self.ck_damped[n // self.K] = damped
            nxt = 3 - cur - prev
            cn = 3 - cc - cp
            self._fwd(S[cur, 0], S[cur, 1], S[prev, 0], S[prev, 1], S[cc, 2], S[cp, 2], damped,
                      S[nxt, 0], S[nxt, 1], S[cn, 2])
            prev, cur, cp, cc = cur, nxt, cc, cn
            damped = True
        self.seg = torch.empty((self.K + 1, 3, n3), **f32)
        return S[cc, 2] + self.chi_off

    def _segment(self, s: int, e: int):
        """States s..e-1 as (pr, pi, chi, chi_prev): from the tape, or
        recomputed from checkpoint s into the segment buffer."""
        if self.K == 1:
            R, I, C = self.tape
            return [(R[n], I[n], C[n], C[n - 1] if n else C[0]) for n in range(s, e)]
        c = self.ck[s // self.K]
        B = self.seg
        B[0, 0].copy_(c[0]); B[0, 1].copy_(c[1]); B[0, 2].copy_(c[4])
        states = [(B[0, 0], B[0, 1], B[0, 2], c[5])]
        pr_, pi_, damped = c[2], c[3], self.ck_damped[s // self.K]
        for t in range(e - s - 1):
            cr, ci, ch, cp = states[-1]
            self._fwd(cr, ci, pr_, pi_, ch, cp, damped, B[t + 1, 0], B[t + 1, 1], B[t + 1, 2])
            pr_, pi_ = cr, ci
            damped = True
            states.append((B[t + 1, 0], B[t + 1, 1], B[t + 1, 2], ch))
        return states

    # ----------------------------------------------------------- backward
    def backward(self, g_chiT: torch.Tensor, g_prT=None, g_piT=None):
        """Adjoint sweep; returns (dL/dpr0, dL/dpi0, dL/dkappa, dL/dchi_init)."""
        n3 = self.N ** 3
        f32 = dict(dtype=torch.float32, device=self.device)
        A = torch.empty((11, n3), **f32)
        Ur, Ui, X, Unr, Uni, Xn = A[0], A[1], A[2], A[3], A[4], A[5]
        Ut = [(A[6], A[7], A[8]), (A[9], A[10], None)]
        Xt1 = torch.empty(n3, **f32)
        Ut[1] = (A[9], A[10], Xt1)
        X.copy_(g_chiT.detach().to(**f32).reshape(-1))
        if g_prT is not None:
            Ur.copy_(g_prT.detach().to(**f32).reshape(-1))
        else:
            Ur.zero_()
        if g_piT is not None:
            Ui.copy_(g_piT.detach().to(**f32).reshape(-1))
        else:
            Ui.zero_()
        nb = int(self._lib.lfm_adjoint_partials_len(self.N))
        part = torch.empty((self.T, nb), dtype=torch.float64, device=self.device)
        p = self._p(self._flags(False) & ~_abi.LFM_F_DAMP_PREV)
        mask = self.mask.data_ptr() if self.has_mask else None
        st = stream_handle()
        slot, have2 = 0, False
        for s in reversed(range(0, self.T, self.K)):
            e = min(s + self.K, self.T)
            states = self._segment(s, e)
            for n in range(e - 1, s - 1, -1):
                cr, ci, c, cp = states[n - s]
                u1 = Ut[slot]
                u2 = Ut[1 - slot]
                rc = self._lib.lfm_adjoint_complex_step_f32(
                    cr.data_ptr(), ci.data_ptr(), c.data_ptr(), cp.data_ptr(), mask,
                    Ur.data_ptr(), Ui.data_ptr(), X.data_ptr(), u1[0].data_ptr(),
                    u1[1].data_ptr(), u1[2].data_ptr(), u2[0].data_ptr() if have2 else None,
                    u2[1].data_ptr() if have2 else None, u2[2].data_ptr() if have2 else None,
                    Unr.data_ptr(), Uni.data_ptr(), Xn.data_ptr(), part[n].data_ptr(), None,
                    ctypes.byref(p), st)
                _abi.check(rc, "lfm_adjoint_complex_step_f32")
                Ur, Unr = Unr, Ur
                Ui, Uni = Uni, Ui
                X, Xn = Xn, X
                slot, have2 = 1 - slot, True
        # the last call wrote Ut, Xt of step 1 into Ut[1 - slot]
        u1 = Ut[1 - slot]
        g_pr0 = Ur - u1[0]        # dL/dcur_0 + dL/dprev_0, prev_0 = Psi_0
        g_pi0 = Ui - u1[1]
        g_chi = X - u1[2]         # dL/dchi_0 + dL/dchi_prev_0
        g_k = float(part.sum().item())
        return g_pr0, g_pi0, g_k, g_chi


class _ChiRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pr0, pi0, kappa, chi_init, tape):
        chiT = tape.forward(pr0, pi0, float(kappa), chi_init)
        ctx.tape = tape
        ctx.meta = (pr0.shape, pi0.shape, chi_init.shape, kappa.dtype, kappa.device, pr0.dtype)
        return chiT.reshape(chi_init.shape)

    @staticmethod
    def backward(ctx, g):
        s_r, s_i, s_c, kd, kdev, dt = ctx.meta
        gr, gi, gk, gc = ctx.tape.backward(g)
        return (gr.reshape(s_r).to(dt), gi.reshape(s_i).to(dt),
                torch.tensor(gk, dtype=kd, device=kdev), gc.reshape(s_c).to(dt), None)


def evolve_chi(pr0: torch.Tensor, pi0: torch.Tensor, kappa, config: SimulationConfig,
               steps: int, chi_init: torch.Tensor | None = None, stride="auto",
               memory_budget: float = 64e9, mask_policy: str = "reference") -> torch.Tensor:
    """chi after `steps` COMPLEX steps, differentiable w.r.t. pr0, pi0,
    kappa (a 0-d tensor) and chi_init."""
    tape = ComplexTape(config, steps, stride=stride, memory_budget=memory_budget,
                       mask_policy=mask_policy)
    if not isinstance(kappa, torch.Tensor):
        kappa = torch.tensor(float(kappa), dtype=torch.float64, device=tape.device)
    if chi_init is None:
        N = config.grid_size
        chi_init = torch.full((N, N, N), float(config.chi0), dtype=torch.float32, device=tape.device)
    return _ChiRun.apply(pr0, pi0, kappa, chi_init, tape)
