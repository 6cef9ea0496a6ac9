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
        self.chi_off = float(np.float32(config.chi0))
        self.kappa = float(config.kappa)
        self._pcache = {}
        self._mask_ptr = self.mask.data_ptr() if self.has_mask else None
        self._stream = None

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
        """One forward step; every buffer argument is a device address (int)."""
        key = (bool(damped), self.kappa)
        p = self._pcache.get(key)
        if p is None:
            p = self._pcache[key] = ctypes.byref(self._p(self._flags(damped)))
        rc = self._lib.lfm_step_complex_f32(cr, pr_, ci, pi_, c, cp, self._mask_ptr, out_r, None,
                                            out_i, None, out_c, None, p, self._stream)
        if rc:
            _abi.check(rc, "lfm_step_complex_f32")

    def forward(self, pr0: torch.Tensor, pi0: torch.Tensor, kappa: float, chi_init: torch.Tensor):
        """Run T steps from Psi_0 (prev tied) and chi_init; return chi_T."""
        self.kappa = float(kappa)
        self._stream = stream_handle()
        n3 = self.N ** 3
        f32 = dict(dtype=torch.float32, device=self.device)
        pr0 = pr0.detach().to(**f32).reshape(-1)
        pi0 = pi0.detach().to(**f32).reshape(-1)
        chi_d = chi_init.detach().to(**f32).reshape(-1) - self.chi_off
        row = 4 * n3
        if self.K == 1:
            # every state kept: one allocation, each step writes into its slot
            R = torch.empty((self.T + 1, n3), **f32)
            I = torch.empty((self.T + 1, n3), **f32)
            C = torch.empty((self.T + 1, n3), **f32)
            R[0].copy_(pr0)
            I[0].copy_(pi0)
            C[0].copy_(chi_d)
            r0, i0, c0 = R.data_ptr(), I.data_ptr(), C.data_ptr()
            for n in range(self.T):
                m = n - 1 if n else 0
                self._fwd(r0 + n * row, i0 + n * row, r0 + m * row, i0 + m * row, c0 + n * row,
                          c0 + m * row, n > 0, r0 + (n + 1) * row, i0 + (n + 1) * row,
                          c0 + (n + 1) * row)
            self.tape = (R, I, C)
            self.chiT = C[self.T]
            return C[self.T] + self.chi_off
        # three-slot rotation; checkpoint (cur, prev, chi, chi_prev) every K steps
        nck = (self.T + self.K - 1) // self.K
        self.ck = torch.empty((nck, 6, n3), **f32)
        self.ck_damped = [False] * nck
        S = torch.empty((3, 3, n3), **f32)     # [slot][pr, pi, chi]
        S[0, 0].copy_(pr0); S[1, 0].copy_(pr0)
        S[0, 1].copy_(pi0); S[1, 1].copy_(pi0)
        S[0, 2].copy_(chi_d); S[1, 2].copy_(chi_d)
        flat = S.view(9, n3)
        sp = lambda slot, f: S.data_ptr() + (3 * slot + f) * row
        gathers = {}
        cur, prev, cc, cp = 0, 1, 0, 1
        damped = False
        for n in range(self.T):
            if n % self.K == 0:
                key = (cur, prev, cc, cp)
                if key not in gathers:
                    gathers[key] = torch.tensor([3 * cur, 3 * cur + 1, 3 * prev, 3 * prev + 1,
                                                 3 * cc + 2, 3 * cp + 2], device=self.device)
                torch.index_select(flat, 0, gathers[key], out=self.ck[n // self.K])
                self.ck_damped[n // self.K] = damped
            nxt = 3 - cur - prev
            cn = 3 - cc - cp
            self._fwd(sp(cur, 0), sp(cur, 1), sp(prev, 0), sp(prev, 1), sp(cc, 2), sp(cp, 2),
                      damped, sp(nxt, 0), sp(nxt, 1), sp(cn, 2))
            prev, cur, cp, cc = cur, nxt, cc, cn
            damped = True
        self.seg = torch.empty((self.K + 1, 3, n3), **f32)
        self.chiT = S[cc, 2]
        return S[cc, 2] + self.chi_off

    def _segment(self, s: int, e: int):
        """Device addresses (pr, pi, chi) of states s..e-1: from the tape, or
        recomputed from checkpoint s into the segment buffer."""
        row = 4 * self.N ** 3
        if self.K == 1:
            R, I, C = self.tape
            r0, i0, c0 = R.data_ptr(), I.data_ptr(), C.data_ptr()
            return [(r0 + n * row, i0 + n * row, c0 + n * row) for n in range(s, e)]
        c = self.ck[s // self.K]
        B = self.seg
        B[0].copy_(c[[0, 1, 4]])
        b0, k0 = B.data_ptr(), c.data_ptr()
        bp = lambda t, f: b0 + (3 * t + f) * row
        states = [(bp(0, 0), bp(0, 1), bp(0, 2))]
        pr_, pi_, cp = k0 + 2 * row, k0 + 3 * row, k0 + 5 * row
        damped = self.ck_damped[s // self.K]
        for t in range(e - s - 1):
            cr, ci, ch = states[-1]
            self._fwd(cr, ci, pr_, pi_, ch, cp, damped, bp(t + 1, 0), bp(t + 1, 1), bp(t + 1, 2))
            pr_, pi_, cp = cr, ci, ch
            damped = True
            states.append((bp(t + 1, 0), bp(t + 1, 1), bp(t + 1, 2)))
        return states

    # ----------------------------------------------------------- backward
    def backward(self, g_chiT: torch.Tensor, g_prT=None, g_piT=None):
        """Adjoint sweep; returns (dL/dpr0, dL/dpi0, dL/dkappa, dL/dchi_init)."""
        n3 = self.N ** 3
        row = 4 * n3
        f32 = dict(dtype=torch.float32, device=self.device)
        # three (Ut_r, Ut_i, Xt) slots: step n+1 (stencil input), n+2, output
        A = torch.empty((3, 3, n3), **f32)
        G = torch.zeros((3, n3), **f32)          # raw adjoints of the final state
        G[2].copy_(g_chiT.detach().to(**f32).reshape(-1))
        if g_prT is not None:
            G[0].copy_(g_prT.detach().to(**f32).reshape(-1))
        if g_piT is not None:
            G[1].copy_(g_piT.detach().to(**f32).reshape(-1))
        nb = int(self._lib.lfm_adjoint_partials_len(self.N))
        part = torch.empty((self.T, nb), dtype=torch.float64, device=self.device)
        p = ctypes.byref(self._p(self._flags(False) & ~_abi.LFM_F_DAMP_PREV))
        mask = self._mask_ptr
        st = self._stream = stream_handle()
        ap = lambda slot, f: A.data_ptr() + (3 * slot + f) * row
        gp = lambda f: G.data_ptr() + f * row
        rc = self._lib.lfm_adjoint_complex_seed_f32(
            self.chiT.data_ptr(), mask, gp(0), gp(1), gp(2), ap(0, 0), ap(0, 1), ap(0, 2), p, st)
        _abi.check(rc, "lfm_adjoint_complex_seed_f32")
        step = self._lib.lfm_adjoint_complex_step_f32
        part0 = part.data_ptr()
        c1, c2, c3, have2 = 0, 1, 2, False
        for s in reversed(range(0, self.T, self.K)):
            e = min(s + self.K, self.T)
            states = self._segment(s, e)
            for n in range(e - 1, s - 1, -1):
                cr, ci, c = states[n - s]
                rc = step(cr, ci, c, mask, ap(c1, 0), ap(c1, 1), ap(c1, 2),
                          ap(c2, 0) if have2 else None, ap(c2, 1) if have2 else None,
                          ap(c2, 2) if have2 else None, ap(c3, 0), ap(c3, 1), ap(c3, 2),
                          1 if n > 0 else 0, part0 + 8 * nb * n, None, p, st)
                if rc:
                    _abi.check(rc, "lfm_adjoint_complex_step_f32")
                c1, c2, c3, have2 = c3, c1, c2, True
        # the last call wrote the raw U_0, X_0 into A[c1]; A[c2] holds Ut_1, Xt_1
        U0, T1 = A[c1], A[c2]
        g_pr0 = U0[0] - T1[0]     # dL/dcur_0 + dL/dprev_0, prev_0 = Psi_0
        g_pi0 = U0[1] - T1[1]
        g_chi = U0[2] - T1[2]     # dL/dchi_0 + dL/dchi_prev_0
        g_k = float(part.sum().item())
        self.release()
        return g_pr0, g_pi0, g_k, g_chi

    def release(self):
        """Drop the tape (every kept state, checkpoints, segment buffer)."""
        self.tape = self.ck = self.seg = self.chiT = None


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
