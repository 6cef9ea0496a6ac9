"""Batched form: many independent lattices advanced by one launch per step.

The reference's parameter sweep (lfm/sweep.py:14-157) builds one
``Simulation`` per value and runs them one after another, one kernel launch
per step per lattice.  :class:`LatticeBatch` packs every lattice of a sweep
into shared arenas; ``lfm_batch_step_f32`` advances all of them with a single
launch whose workgroups each take one (lattice, column tile, plane chunk)
descriptor, and the per-step launches are captured in a HIP graph and
replayed.  :func:`sweep` / :func:`sweep_2d` keep the reference's call
signature and result rows.

Optionally (``group_bytes`` > 0) the lattices are split into contiguous
groups whose three-slot working set fits the GPU's 256 MB last-level cache,
and a replay advances group 0 by whole rotation cycles (3 steps, or a
multiple), then group 1, and so on, so a group's later steps can read its
fields from the cache.  Lattices are independent, so the order changes no
result bit.  One group (the default) measured fastest: the extra, smaller
launches cost more than the cache saves.

Lattices of one batch share a field level.  COLOR lattices with
kappa_tube > 0 keep their S_a input and |Psi'|^2 in two more arenas; before
each step the S_a transform runs per lattice (one plan per lattice size, the
|Psi'|^2 the previous step wrote as its input), then the one batched launch
reads the S_a input and writes the next |Psi'|^2.  With the hand-written
transforms (power-of-two sizes) the S_a input is the SCV term the transform's
last pass forms (one field per lattice), else the three S_a fields.
"""
from __future__ import annotations

import ctypes
import itertools
from copy import deepcopy

import numpy as np
import torch

from . import _abi
from .backend import ScvPlan, graph_capture, params, stream_handle
from .config import BoundaryType, SimulationConfig
from .evolver import get_backend
from .fields import boundary_mask_host, gaussian_soliton


class LatticeBatch:
    """State of `len(configs)` lattices in packed arenas (three-slot rotation,
    chi stored as chi - chi0 per lattice, as :class:`Evolver` does)."""

    #: default three-slot working-set bound of one lattice group (bytes);
    #: 0 = every lattice in one group, which measured fastest for the 64
    #: COMPLEX lattice sweep (0.39 ms per step against 0.41-0.58 ms for
    #: 256-64 MB groups: each extra launch costs more than the cache saves)
    GROUP_BYTES = 0

    def __init__(self, configs, backend="auto", tile: int = 0, chunk: int = 0,
                 mask_policy: str = "reference", group_bytes: int | None = None,
                 cycles: int = 1, two_step: bool = False):
        if not configs:
            raise ValueError("LatticeBatch needs at least one config")
        levels = {int(c.field_level) for c in configs}
        if len(levels) != 1:
            raise ValueError("all lattices of a batch must share one field level")
        self.level = levels.pop()
        if self.level == 2 and any(c.n_colors != 3 for c in configs):
            raise NotImplementedError("the COLOR step is defined for 3 colours")
        # two_step: REAL / COMPLEX pairs of steps in one launch
        # (lfm_batch_step2_f32, bit-identical to single steps) over a fourth
        # rotation slot, since both outputs must differ from both inputs
        if two_step and self.level == 2:
            raise ValueError("two_step: REAL and COMPLEX batches only (a COLOR step needs its "
                             "S_a transform between steps)")
        self.two_step = bool(two_step)
        self._ns = 4 if self.two_step else 3
        self.backend = get_backend(backend)
        self.device = self.backend.device
        self._lib = _abi.load()
        self.configs = list(configs)
        self.nlat = len(configs)
        self.sizes = [c.grid_size for c in configs]
        self.ncomp = 3 if self.level == 2 else 1
        n3 = [N ** 3 for N in self.sizes]
        self.offsets = np.concatenate([[0], np.cumsum(n3)[:-1]]).astype(np.int64)
        self.total = int(sum(n3))
        f32 = torch.float32
        dev = self.device
        npsi = self.ncomp * self.total
        ns = self._ns
        self._pr = [torch.zeros(npsi, dtype=f32, device=dev) for _ in range(ns)]
        self._pi = ([torch.zeros(npsi, dtype=f32, device=dev) for _ in range(ns)]
                    if self.level > 0 else None)
        self._chi = [torch.zeros(self.total, dtype=f32, device=dev) for _ in range(ns)]
        self._chi0 = [float(np.float32(c.chi0)) for c in configs]
        self.mask = torch.zeros(self.total, dtype=f32, device=dev)
        self.has_mask = []
        for b, c in enumerate(configs):
            m = not (mask_policy == "boundary_type" and c.boundary_type == BoundaryType.PERIODIC)
            self.has_mask.append(m)
            if m:
                self._view(self.mask, b).copy_(torch.from_numpy(
                    boundary_mask_host(c.grid_size, c.boundary_fraction)))
        # S_a (COLOR lattices with kappa_tube > 0): raw S_a and |Psi'|^2 arenas,
        # one transform plan per lattice size
        self.scv = [self.level == 2 and c.kappa_tube > 0 for c in configs]
        self._sa = self._psisq = None
        self._scv_plans = {}
        self.sa_scv = False
        if any(self.scv):
            self._psisq = torch.empty(npsi, dtype=f32, device=dev)
            for b, on in enumerate(self.scv):
                if on and self.sizes[b] not in self._scv_plans:
                    self._scv_plans[self.sizes[b]] = ScvPlan(self.sizes[b], 3)
            # when every transform is the hand-written one, its last pass
            # writes the step's SCV term (one field per lattice; the same
            # bits as the three S_a fields, 16 B per site less traffic),
            # else the three S_a fields travel between transform and step
            self.sa_scv = all(p.method == 2 for p in self._scv_plans.values())
            self._sa = torch.zeros(self.total if self.sa_scv else npsi, dtype=f32, device=dev)
        # per lattice: its |Psi'|^2 arena span matches its current Psi (it has
        # stepped since its Psi was last set), so S_a comes from it, as the
        # single-lattice Evolver does after its first step
        self._psisq_ok = [False] * self.nlat
        # Psi snapshot arenas (allocated by the first snapshot_psi call)
        self._snap_r = self._snap_i = None
        self._snap_ok = [False] * self.nlat
        self._cur, self._prev, self._chi_cur, self._chi_prev = 0, 1, 0, 1
        self._damped = False
        self.step = 0
        self._graphs = {}
        self._plans = []
        self.tile, self.chunk = 0, 0
        if int(cycles) < 1:
            raise ValueError("cycles must be >= 1")
        self.cycles = int(cycles)
        self._set_groups(self.GROUP_BYTES if group_bytes is None else int(group_bytes))
        self._make_plan(tile, chunk)

    # ----------------------------------------------------------- groups
    @property
    def bytes_per_site(self) -> int:
        """The rotation slots (three, four with two_step) of Psi (and Psi_i)
        and chi plus the mask, float32."""
        per_slot = self.ncomp * (2 if self.level > 0 else 1) + 1
        return 4 * (self._ns * per_slot + 1)

    def _set_groups(self, group_bytes: int) -> None:
        """Contiguous lattice ranges [b0, b1) of at most `group_bytes` of
        working set each (a single larger lattice forms its own group);
        group_bytes <= 0 puts every lattice in one group."""
        cap = group_bytes // self.bytes_per_site if group_bytes > 0 else self.total
        groups, b0, acc = [], 0, 0
        for b, N in enumerate(self.sizes):
            n3 = N ** 3
            if b > b0 and acc + n3 > cap:
                groups.append((b0, b))
                b0, acc = b, 0
            acc += n3
        groups.append((b0, self.nlat))
        self.group_bytes = group_bytes
        self.groups = groups
        self._graphs = {}

    # ------------------------------------------------------------- plan
    def _make_plan(self, tile: int, chunk: int) -> None:
        """One batch plan per lattice group; plan g addresses the arenas from
        the group's first site on (pointer offsets at launch).  With two_step
        the column tiles are 64 x 8 (tile variant 2, the default, or 4)."""
        if self.two_step:
            tile = 2 if int(tile) <= 0 else int(tile)
            if tile not in (2, 4):
                raise ValueError(f"two_step batches take tile variant 2 or 4 (64 x 8 column "
                                 f"tiles), got {tile}")
        plans = []
        try:
            for b0, b1 in self.groups:
                arr = (_abi.LfmParams * (b1 - b0))()
                for b in range(b0, b1):
                    c = self.configs[b]
                    arr[b - b0] = params(c.grid_size, c.dt ** 2, c.kappa, c.lambda_self, c.chi0,
                                         c.e0_sq, c.epsilon_w, c.kappa_c, c.epsilon_cc,
                                         c.kappa_string, c.kappa_tube if self.scv[b] else 0.0,
                                         _abi.LFM_F_MASK if self.has_mask[b] else 0)
                h = ctypes.c_void_p()
                _abi.check(self._lib.lfm_batch_plan_create(self.level, b1 - b0, arr, int(tile),
                                                           int(chunk), ctypes.byref(h)),
                           "lfm_batch_plan_create")
                plans.append(h)
        except Exception:
            for h in plans:
                self._lib.lfm_batch_plan_destroy(h)
            raise
        self._free_plan()
        self._plans = plans
        self.tile, self.chunk = tile, chunk
        self._graphs = {}

    def _free_plan(self):
        for h in self._plans:
            if h is not None and h.value:
                self._lib.lfm_batch_plan_destroy(h)
        self._plans = []

    def __del__(self):
        try:
            self._free_plan()
        except Exception:
            pass

    @property
    def workgroups(self) -> int:
        """Workgroups of one step over all groups."""
        return sum(int(self._lib.lfm_batch_workgroups(h)) for h in self._plans)

    # ------------------------------------------------------------ views
    def _view(self, arena: torch.Tensor, b: int, psi: bool = False) -> torch.Tensor:
        n3 = self.sizes[b] ** 3
        k = self.ncomp if psi else 1
        o = int(self.offsets[b]) * k
        return arena[o:o + k * n3]

    def _shape(self, b):
        N = self.sizes[b]
        return (3, N, N, N) if self.level == 2 else (N, N, N)

    def _undamp(self):
        if self._damped:
            for b in range(self.nlat):
                if self.has_mask[b]:
                    a = 1.0 - self._view(self.mask, b)
                    if self.ncomp > 1:
                        a = a.repeat(self.ncomp)
                    self._view(self._pr[self._prev], b, True).mul_(a)
                    if self._pi is not None:
                        self._view(self._pi[self._prev], b, True).mul_(a)
        self._damped = False

    def _dev(self, arr) -> torch.Tensor:
        if isinstance(arr, torch.Tensor):
            return arr.detach().to(self.device, torch.float32).reshape(-1)
        return torch.from_numpy(np.ascontiguousarray(np.asarray(arr, np.float32).ravel())).to(self.device)

    # ---------------------------------------------------- setters/getters
    def set_psi_real(self, b: int, arr) -> None:
        """Current and prev Psi_r of lattice b := arr (evolver.py:355-375)."""
        d = self._dev(arr)
        self._undamp()
        self._view(self._pr[self._cur], b, True).copy_(d)
        self._view(self._pr[self._prev], b, True).copy_(d)
        self._psisq_ok[b] = False

    def set_psi_imag(self, b: int, arr) -> None:
        if self._pi is None:
            raise ValueError("Cannot set imaginary part for REAL field level")
        d = self._dev(arr)
        self._undamp()
        self._view(self._pi[self._cur], b, True).copy_(d)
        self._view(self._pi[self._prev], b, True).copy_(d)
        self._psisq_ok[b] = False

    def set_chi(self, b: int, arr) -> None:
        """Current and prev chi of lattice b := arr."""
        d = self._dev(arr) - self._chi0[b]
        self._view(self._chi[self._chi_cur], b).copy_(d)
        self._view(self._chi[self._chi_prev], b).copy_(d)

    def get_psi_real(self, b: int) -> np.ndarray:
        return self._view(self._pr[self._cur], b, True).cpu().numpy().reshape(self._shape(b))

    def get_psi_imag(self, b: int):
        if self._pi is None:
            return None
        return self._view(self._pi[self._cur], b, True).cpu().numpy().reshape(self._shape(b))

    def get_chi(self, b: int) -> np.ndarray:
        N = self.sizes[b]
        return (self._view(self._chi[self._chi_cur], b) + self._chi0[b]).cpu().numpy().reshape(N, N, N)

    def get_energy_density(self, b: int) -> np.ndarray:
        pr = self.get_psi_real(b).astype(np.float32)
        e = np.sum(pr ** 2, axis=0) if pr.ndim == 4 else pr ** 2
        pi = self.get_psi_imag(b)
        if pi is not None:
            e = e + (np.sum(pi ** 2, axis=0) if pi.ndim == 4 else pi ** 2)
        return e

    # -------------------------------------------------------------- steps
    def _slots(self):
        """(cur, prev, next, chi cur, chi prev, chi next, flags) of the next step."""
        cur, prev = self._cur, self._prev
        cc, cp = self._chi_cur, self._chi_prev
        ns = self._ns
        nxt = next(s for s in range(ns) if s not in (cur, prev))
        cn = next(s for s in range(ns) if s not in (cc, cp)) if cc != cp else (cc + 1) % ns
        flags = _abi.LFM_F_CHI_DELTA | (_abi.LFM_F_DAMP_PREV if self._damped else 0)
        return cur, prev, nxt, cc, cp, cn, flags

    def _slots2(self):
        """(cur, prev, out1, out2, chi cur, chi prev, chi out1, chi out2, flags)
        of the next pair of steps (two_step: the two free slots of four)."""
        cur, prev = self._cur, self._prev
        cc, cp = self._chi_cur, self._chi_prev
        assert self.two_step and cc != cp
        o1, o2 = (s for s in range(4) if s not in (cur, prev))
        c1, c2 = (s for s in range(4) if s not in (cc, cp))
        flags = _abi.LFM_F_CHI_DELTA | (_abi.LFM_F_DAMP_PREV if self._damped else 0)
        return cur, prev, o1, o2, cc, cp, c1, c2, flags

    def _advance2(self, slots2) -> None:
        _, _, o1, o2, _, _, c1, c2, _ = slots2
        self._prev, self._cur = o1, o2
        self._chi_prev, self._chi_cur = c1, c2
        self._damped = True

    def _advance(self, slots) -> None:
        cur, _, nxt, cc, _, cn, _ = slots
        self._prev, self._cur = cur, nxt
        self._chi_prev, self._chi_cur = cc, cn
        self._damped = True

    def _transforms(self, g: int, cur: int) -> None:
        """S_a of every lattice of group g with kappa_tube > 0, from the
        |Psi|^2 the previous step wrote (or from Psi after a set)."""
        b0, b1 = self.groups[g]
        for b in range(b0, b1):
            if not self.scv[b]:
                continue
            c, plan = self.configs[b], self._scv_plans[self.sizes[b]]
            if self.sa_scv:
                e = self._view(self._psisq, b, True).data_ptr() if self._psisq_ok[b] else None
                _abi.check(self._lib.lfm_scv_apply_scv_f32(
                    plan.handle, self._view(self._pr[cur], b, True).data_ptr(),
                    self._view(self._pi[cur], b, True).data_ptr(), e, c.sa_gamma, c.sa_d,
                    self._view(self._sa, b).data_ptr(), stream_handle()), "lfm_scv_apply_scv_f32")
                continue
            sa = self._view(self._sa, b, True)
            if self._psisq_ok[b]:
                plan.apply_psisq(self._view(self._psisq, b, True), c.sa_gamma, c.sa_d, sa)
            else:
                plan.apply(self._view(self._pr[cur], b, True), self._view(self._pi[cur], b, True),
                           c.sa_gamma, c.sa_d, sa)

    def _launch(self, g: int, slots) -> None:
        cur, prev, nxt, cc, cp, cn, flags = slots
        if self._sa is not None:
            self._transforms(g, cur)
        o = int(self.offsets[self.groups[g][0]]) * 4
        op = o * self.ncomp
        pr, pi, chi = self._pr, self._pi, self._chi
        sa = self._sa.data_ptr() + (o if self.sa_scv else op) if self._sa is not None else None
        sq = self._psisq.data_ptr() + op if self._sa is not None else None
        if self.sa_scv:
            flags |= _abi.LFM_F_SA_SCV
        rc = self._lib.lfm_batch_step_f32(
            self._plans[g], pr[cur].data_ptr() + op, pr[prev].data_ptr() + op,
            pi[cur].data_ptr() + op if pi else None, pi[prev].data_ptr() + op if pi else None,
            chi[cc].data_ptr() + o, chi[cp].data_ptr() + o, self.mask.data_ptr() + o,
            pr[nxt].data_ptr() + op, pi[nxt].data_ptr() + op if pi else None,
            chi[cn].data_ptr() + o, sa, sq, flags, stream_handle())
        _abi.check(rc, "lfm_batch_step_f32")
        b0, b1 = self.groups[g]
        self._psisq_ok[b0:b1] = [True] * (b1 - b0)

    def _launch2(self, g: int, slots2) -> None:
        cur, prev, o1, o2, cc, cp, c1, c2, flags = slots2
        o = int(self.offsets[self.groups[g][0]]) * 4
        pr, pi, chi = self._pr, self._pi, self._chi
        ptr = lambda arena, slot: arena[slot].data_ptr() + o if arena else None
        rc = self._lib.lfm_batch_step2_f32(
            self._plans[g], ptr(pr, cur), ptr(pr, prev), ptr(pi, cur), ptr(pi, prev),
            ptr(chi, cc), ptr(chi, cp), self.mask.data_ptr() + o,
            ptr(pr, o1), ptr(pi, o1), ptr(chi, c1), ptr(pr, o2), ptr(pi, o2), ptr(chi, c2),
            flags, stream_handle())
        _abi.check(rc, "lfm_batch_step2_f32")

    def _step2(self) -> None:
        """Two steps of every lattice (two_step: one launch per group)."""
        slots2 = self._slots2()
        for g in range(len(self.groups)):
            self._launch2(g, slots2)
        self._advance2(slots2)

    def _step(self) -> None:
        """One step of every lattice (one launch per group)."""
        slots = self._slots()
        for g in range(len(self.groups)):
            self._launch(g, slots)
        self._advance(slots)

    def _cycle(self) -> None:
        """Three steps of every lattice launched eagerly in the replay's
        order: each group's three steps, group after group."""
        seq = []
        for _ in range(3):
            seq.append(self._slots())
            self._advance(seq[-1])
        for grp in range(len(self.groups)):
            for slots in seq:
                self._launch(grp, slots)

    def evolve(self, steps: int, graph: bool = True) -> None:
        """Advance every lattice `steps` steps.  With graph=True the launches
        of `cycles` full three-slot rotations of every group, group after
        group, are captured once in a HIP graph and replayed
        (steps // (3 cycles) replays, the remainder launched directly)."""
        steps = int(steps)
        if steps <= 0:
            return
        if self.two_step:
            self._evolve2(steps, graph)
            return
        if not graph or any(p.method != 2 for p in self._scv_plans.values()):
            # rocFFT plans (non-power-of-two S_a sizes) stay out of graphs
            for _ in range(steps):
                self._step()
            self.step += steps
            return
        done = 0
        # reach the steady state: prev damped, the chi slots distinct and every
        # lattice's S_a taken from its |Psi'|^2
        while done < steps and (not self._damped or self._chi_cur == self._chi_prev
                                or not all(self._psisq_ok)):
            self._step()
            done += 1
        per = 3 * self.cycles
        reps = (steps - done) // per
        if reps:
            g = self._graphs.get(self._state_key())
            if g is None:
                g = self._capture()
            for _ in range(reps):
                g.replay()
            done += per * reps
        while done < steps:
            self._step()
            done += 1
        self.step += steps

    def _evolve2(self, steps: int, graph: bool) -> None:
        """evolve() with two_step: pairs of steps per launch, replayed from HIP
        graphs of `cycles` x 2 pairs (a pair's outputs take the two free slots,
        so two pairs return every slot to where it was); an odd last step is a
        single launch."""
        done = 0
        if graph:
            # steady state: prev damped and the current slot above the prev
            # slot, from which every two pairs come back to the same slots
            while steps - done >= 2 and (not self._damped or self._cur < self._prev):
                self._step2()
                done += 2
            per = 4 * self.cycles
            reps = (steps - done) // per
            if reps:
                g = self._graphs.get(self._state_key())
                if g is None:
                    g = self._capture2()
                for _ in range(reps):
                    g.replay()
                done += per * reps
        while steps - done >= 2:
            self._step2()
            done += 2
        if done < steps:
            self._step()
        self.step += steps

    def _capture2(self):
        """Capture 2 x `cycles` pair launches of every group from the current
        (steady) state; see _capture."""
        key = self._state_key()
        saved = (self._cur, self._prev, self._chi_cur, self._chi_prev)
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        seq = []
        for _ in range(2 * self.cycles):
            seq.append(self._slots2())
            self._advance2(seq[-1])
        if (self._cur, self._prev, self._chi_cur, self._chi_prev) != saved:
            raise RuntimeError("two-step slot rotation did not return to its start")
        with graph_capture(g, s):
            for grp in range(len(self.groups)):
                for slots2 in seq:
                    self._launch2(grp, slots2)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graphs[key] = g
        return g

    def _state_key(self):
        return (self._cur, self._prev, self._chi_cur, self._chi_prev, self._damped)

    def _capture(self):
        """Capture the launches of one replay from the current (steady) state;
        graphs are kept per starting state, whose slots fix their pointers."""
        key = self._state_key()
        saved = (self._cur, self._prev, self._chi_cur, self._chi_prev)
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        seq = []
        for _ in range(3 * self.cycles):
            seq.append(self._slots())
            self._advance(seq[-1])
        with graph_capture(g, s):
            for grp in range(len(self.groups)):
                for slots in seq:
                    self._launch(grp, slots)
        torch.cuda.current_stream(self.device).wait_stream(s)
        # a whole number of rotations returns every slot to where it was
        self._cur, self._prev, self._chi_cur, self._chi_prev = saved
        self._graphs[key] = g
        return g

    def autotune(self, steps: int = 6, tiles=None, chunks=None,
                 choose_two_step: bool = True) -> dict:
        """Time tile shapes x plane chunks of the batched kernel on the
        current state (restored afterwards) and keep the fastest.  REAL and
        COMPLEX batches with `choose_two_step` also time the two-step kernel
        (tile variants 2 and 4) and set `two_step` to whichever advances a
        step faster (the single-step kernel is bound by the box's HBM rate,
        the two-step kernel by its latency); with it False only the kernel
        `two_step` names is tuned.  Returns {(tile, chunk): ms per step}, a
        negative tile standing for the two-step kernel at that tile shape."""
        fams = []
        if not self.two_step or choose_two_step:
            fams.append((False, tuple(tiles or range(1, _abi.LFM_TILE_VARIANTS + 1)),
                         tuple(chunks or (8, 16, 32))))
        if self.level < 2 and (self.two_step or choose_two_step):
            fams.append((True, tuple(t for t in (tiles or (2, 4)) if t in (2, 4)) or (2, 4),
                         tuple(chunks or (16, 32, 64))))
        wanted = self.two_step
        grown = False
        if any(pair for pair, _, _ in fams) and len(self._pr) < 4:
            # the two-step rotation needs a fourth slot
            self._pr.append(torch.zeros_like(self._pr[0]))
            if self._pi is not None:
                self._pi.append(torch.zeros_like(self._pi[0]))
            self._chi.append(torch.zeros_like(self._chi[0]))
            self._ns = 4
            self._set_groups(self.group_bytes)
            grown = True
        snap = ([t.clone() for t in self._pr], [t.clone() for t in self._pi] if self._pi else None,
                [t.clone() for t in self._chi], self._state_key())
        res = {}
        for pair, ftiles, fchunks in fams:
            self.two_step = pair
            for tile, ch in itertools.product(ftiles, fchunks):
                self._make_plan(tile, ch)
                self._step2() if pair else self._step()
                per = 4 if pair else 3
                reps = max(1, -(-steps // per))
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ev0.record()
                for _ in range(reps):
                    if pair:
                        self._step2()
                        self._step2()
                    else:
                        self._cycle()
                ev1.record()
                torch.cuda.synchronize(self.device)
                res[(-tile if pair else tile, ch)] = ev0.elapsed_time(ev1) / (per * reps)
                for dst, src in zip(self._pr, snap[0]):
                    dst.copy_(src)
                if snap[1]:
                    for dst, src in zip(self._pi, snap[1]):
                        dst.copy_(src)
                for dst, src in zip(self._chi, snap[2]):
                    dst.copy_(src)
                self._cur, self._prev, self._chi_cur, self._chi_prev, self._damped = snap[3]
                self._psisq_ok = [False] * self.nlat    # the |Psi'|^2 arena holds tuning steps
        best = min(res, key=res.get) if res else (self.tile, self.chunk)
        self.two_step = best[0] < 0 if res else wanted
        if grown and not self.two_step:
            # single steps rotate through the first three slots, where the
            # restored state lies: drop the fourth again
            self._pr.pop()
            if self._pi is not None:
                self._pi.pop()
            self._chi.pop()
            self._ns = 3
            self._set_groups(self.group_bytes)
        self._make_plan(abs(best[0]), best[1])
        return res

    # ----------------------------------------------------------- analysis
    def _red(self):
        if not hasattr(self, "_partials"):
            n = int(self._lib.lfm_reduce_partials_len(max(self.sizes)))
            self._partials = torch.empty(n, dtype=torch.float64, device=self.device)
            self._red_out = torch.empty(8, dtype=torch.float64, device=self.device)
        return self._partials, self._red_out

    def snapshot_psi(self, lattices=None) -> None:
        """Copy the current Psi of the given lattices (default: all) into the
        snapshot arenas: the kinetic-energy reference :meth:`metrics` uses
        with ``kinetic="snapshot"``, as Simulation.run keeps Psi at run start
        and at every report_interval (lfm/simulation.py:1168-1182)."""
        if self._snap_r is None:
            self._snap_r = torch.empty_like(self._pr[0])
            self._snap_i = torch.empty_like(self._pi[0]) if self._pi is not None else None
        for b in (range(self.nlat) if lattices is None else lattices):
            self._view(self._snap_r, b, True).copy_(self._view(self._pr[self._cur], b, True))
            if self._pi is not None:
                self._view(self._snap_i, b, True).copy_(self._view(self._pi[self._cur], b, True))
            self._snap_ok[b] = True

    def metrics(self, b: int, clusters: bool = True, kinetic: str = "leapfrog") -> dict:
        """compute_metrics keys for lattice b (FP64 reductions on the GPU).

        kinetic: "leapfrog" takes the kinetic term against the true leapfrog
        prev; "snapshot" against the Psi that :meth:`snapshot_psi` stored,
        which is what the reference's Simulation.metrics() reports
        (lfm/simulation.py:1438-1440, its prev being a run-start or
        report_interval snapshot)."""
        part, out = self._red()
        c = self.configs[b]
        N = self.sizes[b]
        pr = self._view(self._pr[self._cur], b, True)
        pi = self._view(self._pi[self._cur], b, True) if self._pi else None
        if kinetic == "snapshot":
            if not self._snap_ok[b]:
                raise ValueError(f"lattice {b} has no Psi snapshot (call snapshot_psi first)")
            prp = self._view(self._snap_r, b, True)
            pip = self._view(self._snap_i, b, True) if self._pi else None
            damp = 0
        elif kinetic == "leapfrog":
            prp = self._view(self._pr[self._prev], b, True)
            pip = self._view(self._pi[self._prev], b, True) if self._pi else None
            damp = int(self._damped and self.has_mask[b])
        else:
            raise ValueError(f"kinetic must be 'leapfrog' or 'snapshot', got {kinetic!r}")
        rc = self._lib.lfm_energy_f64(
            pr.data_ptr(), prp.data_ptr(), pi.data_ptr() if pi is not None else None,
            pip.data_ptr() if pip is not None else None,
            self._view(self._chi[self._chi_cur], b).data_ptr(), self._chi0[b],
            self._view(self.mask, b).data_ptr(), damp,
            self.ncomp, N, float(c.dt), float(c.c), part.data_ptr(), out.data_ptr(),
            stream_handle())
        _abi.check(rc, "lfm_energy_f64")
        T, G, V, S = out[:4].tolist()
        rc = self._lib.lfm_chi_stats_f64(self._view(self._chi[self._chi_cur], b).data_ptr(),
                                         self._chi0[b], N, float(c.boundary_fraction), 17.0,
                                         18.0, part.data_ptr(), out.data_ptr(), stream_handle())
        _abi.check(rc, "lfm_chi_stats_f64")
        v = out[:7].tolist()
        m = {"energy_kinetic": T, "energy_gradient": G, "energy_potential": V,
             "energy_total": T + G + V, "chi_min": v[0], "chi_max": v[1], "chi_mean": v[2],
             "chi_std": v[3], "well_fraction": v[4], "void_fraction": v[5],
             "n_clusters": float("nan"), "psi_sq_total": S}
        if clusters:
            from scipy import ndimage
            from .simulation import psi_sq_f64
            psq = psi_sq_f64(self.get_psi_real(b), self.get_psi_imag(b))
            x = np.arange(N, dtype=np.float64)
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            h = N / 2.0
            interior = np.sqrt((X - h) ** 2 + (Y - h) ** 2 + (Z - h) ** 2) <= h * (1 - c.boundary_fraction)
            vals = psq[interior]
            if vals.size:
                binary = (psq > np.percentile(vals, 90.0)) & interior
                m["n_clusters"] = float(ndimage.label(binary)[1])
            else:
                m["n_clusters"] = 0.0
        return m

    # ------------------------------------------------------------- setup
    def place_soliton(self, b: int, position=None, amplitude=None, sigma=None, phase=0.0) -> None:
        """Add a static Gaussian soliton to lattice b (Simulation.place_soliton,
        lfm/simulation.py:495-625, without velocity)."""
        c = self.configs[b]
        N = c.grid_size
        if position is None:
            position = (N // 2, N // 2, N // 2)
        pr, pi = gaussian_soliton(N, position, c.e_amplitude if amplitude is None else amplitude,
                                  c.sigma if sigma is None else sigma, phase)
        self.set_psi_real(b, self.get_psi_real(b) + pr)
        if self._pi is not None:
            self.set_psi_imag(b, self.get_psi_imag(b) + pi)

    def equilibrate(self, b: int) -> None:
        """GOV-04 Poisson chi for lattice b on the GPU (lfm/simulation.py:877-963)."""
        c = self.configs[b]
        N = c.grid_size
        if not hasattr(self, "_poisson"):
            self._poisson = {}
        if N not in self._poisson:
            self._poisson[N] = ScvPlan(N, 1)
        chi = torch.empty(N ** 3, dtype=torch.float32, device=self.device)
        bf = c.boundary_fraction if c.boundary_type == BoundaryType.FROZEN else -1.0
        pi = self._view(self._pi[self._cur], b, True) if self._pi else None
        rc = self._lib.lfm_poisson_chi_f32(
            self._poisson[N].handle, self._view(self._pr[self._cur], b, True).data_ptr(),
            pi.data_ptr() if pi is not None else None, self.ncomp, float(c.kappa),
            float(c.e0_sq), float(c.chi0), float(bf), chi.data_ptr(), stream_handle())
        _abi.check(rc, "lfm_poisson_chi_f32")
        self.set_chi(b, chi)


def _batched_sweep(cfgs, rows, steps, metric_names, equilibrate, soliton, backend, tune,
                   kinetic):
    batch = LatticeBatch(cfgs, backend=backend)
    for b, c in enumerate(cfgs):
        if soliton is not None:
            kw = dict(soliton)
            pos = kw.pop("position", None)
            if "velocity" in kw:
                raise NotImplementedError("moving solitons are not batched")
            batch.place_soliton(b, pos, **kw)
        # Simulation.run auto-equilibrates when solitons were placed and
        # equilibrate() was not called (lfm/simulation.py:965-982), so a
        # soliton run is equilibrated whatever `equilibrate` says
        if equilibrate or soliton is not None:
            batch.equilibrate(b)
    if tune:
        batch.autotune()
    steps = int(steps)
    if kinetic == "snapshot":
        # Simulation.run snapshots Psi at its start and again whenever the
        # evolver's step count reaches a multiple of report_interval
        # (lfm/simulation.py:1168-1182, lfm/core/evolver.py:178-179); the
        # last snapshot is the kinetic reference of metrics()
        snap_at = {}
        for b, c in enumerate(cfgs):
            r = int(c.report_interval)
            s = (steps // r) * r if r > 0 else 0
            snap_at.setdefault(s, []).append(b)
        done = 0
        for s in sorted(snap_at):
            batch.evolve(s - done)
            batch.snapshot_psi(snap_at[s])
            done = s
        batch.evolve(steps - done)
    elif kinetic == "leapfrog":
        batch.evolve(steps)
    else:
        raise ValueError(f"kinetic must be 'snapshot' or 'leapfrog', got {kinetic!r}")
    out = []
    for b, row in enumerate(rows):
        m = batch.metrics(b, kinetic=kinetic)
        if metric_names is not None:
            row.update({k: m.get(k) for k in metric_names})
        else:
            row.update(m)
        out.append(row)
    return out


def sweep(config: SimulationConfig, param: str, values, steps: int, metric_names=None, *,
          equilibrate: bool = True, soliton=None, backend="auto", tune: bool = False,
          kinetic: str = "snapshot"):
    """lfm.sweep.sweep with every value run in one batch (lfm/sweep.py:14-82).

    As in the reference, each value is set on a deep copy of `config` with
    setattr (derived fields such as sigma keep the base config's value), an
    optional soliton is placed (at the centre when no position is given),
    chi is equilibrated (always after a soliton, as Simulation.run's
    auto-equilibration does), `steps` steps are run and the final metrics
    are returned as {param: value, **metrics} rows.

    kinetic: "snapshot" (default) gives the reference's rows: the kinetic
    term is taken against Psi at run start, or at the last multiple of
    report_interval when `steps` reaches one, as the reference's
    ``sim.run(...); sim.metrics()`` does.  "leapfrog" takes it against the
    true leapfrog prev (the physical kinetic energy at the final step).
    """
    cfgs, rows = [], []
    for v in values:
        c = deepcopy(config)
        setattr(c, param, v)
        cfgs.append(c)
        rows.append({param: v})
    return _batched_sweep(cfgs, rows, steps, metric_names, equilibrate, soliton, backend, tune,
                          kinetic)


def sweep_2d(config: SimulationConfig, param1: str, values1, param2: str, values2, steps: int,
             metric_names=None, *, equilibrate: bool = True, soliton=None, backend="auto",
             tune: bool = False, kinetic: str = "snapshot"):
    """lfm.sweep.sweep_2d as one batch (lfm/sweep.py:85-157); `kinetic` as
    in :func:`sweep`."""
    cfgs, rows = [], []
    for v1 in values1:
        for v2 in values2:
            c = deepcopy(config)
            setattr(c, param1, v1)
            setattr(c, param2, v2)
            cfgs.append(c)
            rows.append({param1: v1, param2: v2})
    return _batched_sweep(cfgs, rows, steps, metric_names, equilibrate, soliton, backend, tune,
                          kinetic)
