"""``Evolver``: device-resident leapfrog state and the step loop.

Same public surface and observable semantics as lfm/core/evolver.py:42-536
(getters, setters, ``evolve(steps, callback, freeze_chi)``, ``step``), but
the buffers are arranged for the GPU:

* **Three-buffer rotation.**  The reference keeps A/B sets and writes four
  outputs per step (new Psi, damped prev, new chi, chi prev).  Here each field
  has a pool of three slots; a step writes only the new Psi and chi into the
  free slot and the old current slot becomes prev.  The reference's
  ``psi_prev_out = (1-m) Psi`` is reproduced exactly by multiplying prev by
  (1-m) when the kernel reads it (the mask does not change in time, and the
  float32 product is the same single rounding).  The first step after a
  ``set_psi_*`` call reads prev undamped, as the reference does.  This saves
  1 (REAL), 2 (COMPLEX) or 6 (COLOR) float writes plus one chi write per site.
* **chi references.**  The reference's four chi buffers (current-set chi and
  prev, other-set chi and prev) are four references into the three-slot pool,
  so ``freeze_chi`` alternation and ``set_chi_current`` / ``set_chi_prev``
  behave exactly as in the reference (evolver.py:160-176,467-505).
* **S_a.**  S_a carries no state (it is recomputed from the current Psi every
  step, numpy_backend.py:341), so one raw buffer is kept; the step kernel
  clips it at 0 where it is used and the getter clips on read.  The step
  kernel also emits |Psi_a'|^2 of the new fields, which is the next step's
  transform input, so the S_a pipeline never re-reads Psi.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from .backend import HipBackend, ScvPlan, graph_capture, params, stream_handle
from .config import BoundaryType, FieldLevel, SimulationConfig


def get_backend(preference="auto"):
    """Backend registry for this package: 'auto', 'gpu' and 'hip' give a
    :class:`HipBackend`; 'cpu' is refused (no CPU path ships here)."""
    if isinstance(preference, HipBackend):
        return preference
    pref = str(preference).lower()
    if pref in ("auto", "gpu", "hip"):
        return HipBackend()
    if pref == "cpu":
        raise ValueError("lfm_hip has no CPU backend; use the reference lfm package for CPU runs")
    raise ValueError(f"Unknown backend preference '{preference}'. Use 'auto', 'gpu' or 'hip'.")


class Evolver:
    """Double-buffer-equivalent leapfrog evolver on the GPU.

    graphs: evolve() replays whole slot-rotation cycles from HIP graphs once
    the rotation is steady (a graph of GRAPH_CYCLES cycles for long runs);
    the launches are the eager ones, so the results are the same bits.

    mask_policy: "reference" (default) applies the cos^2 mask whatever the
    boundary type, as lfm/core/evolver.py:126 does; "boundary_type" drops it
    for BoundaryType.PERIODIC (truly periodic, mask-free kernel).

    chi_storage: "delta" (default) keeps chi - chi0 in the device buffers.
    float32 resolves chi ~ 19 only to ~2e-6 absolute, and the leapfrog
    accumulates that rounding step after step (the dominant float32 error
    at the REAL level); the deviation is resolved ~1e6 times finer.  Getters
    and setters still speak absolute chi.  "absolute" stores chi as the
    reference does.
    """

    GRAPH_CYCLES = 8
    DIRECT = -1          # `tile` value selecting the one-thread-per-site kernel
    GRAPH_CACHE = 24     # captured graphs kept per Evolver

    def __init__(self, config: SimulationConfig, backend="auto", mask_policy="reference",
                 chi_storage="delta", sa_transform="auto", two_step=False, graphs=True):
        self.config = config
        self.backend = get_backend(backend)
        self._lib = _abi.load()
        self.N = config.grid_size
        self.total = self.N ** 3
        self.step = 0
        lvl = config.field_level
        if lvl == FieldLevel.COLOR and config.n_colors != 3:
            raise NotImplementedError(
                f"n_colors={config.n_colors}: the COLOR step is defined for 3 colours "
                "(lfm/core/backends/numpy_backend.py:240 hard-wires 3 as well)")
        self.level = int(lvl)
        self._n_psi = config.n_colors if lvl == FieldLevel.COLOR else 1
        self._has_imag = lvl != FieldLevel.REAL
        self._dt2 = config.dt ** 2
        dev = self.backend.device
        self.device = dev
        n = self._n_psi * self.total
        f32 = torch.float32
        # two_step: REAL / COMPLEX advance two steps per launch (lfm_step2_*,
        # bit-identical to single steps), which needs a fourth rotation slot:
        # both outputs must differ from both inputs.  Off by default: measured
        # on MI355X it moves fewer HBM bytes (COMPLEX 512^3: 26.3 against 39.1
        # B per site and step) but its ring recomputation leaves it
        # latency-bound at 15 waves per CU, so it wins only on boxes with
        # slower HBM (0.903 against 1.07 ms per step there; 0.86-0.92 against
        # 0.85-0.92 elsewhere) and loses below 128^3, where a chunk's four
        # extra planes and the few workgroups dominate.  autotune() picks
        # between the two from 256^3 up.
        self.two_step = bool(two_step) and lvl != FieldLevel.COLOR
        ns = 4 if self.two_step else 3
        self._pr = [torch.zeros(n, dtype=f32, device=dev) for _ in range(ns)]
        self._pi = [torch.zeros(n, dtype=f32, device=dev) for _ in range(ns)] if self._has_imag else None
        if chi_storage not in ("delta", "absolute"):
            raise ValueError(f"chi_storage must be 'delta' or 'absolute', got {chi_storage!r}")
        self.chi_storage = chi_storage
        self._chi_off = float(np.float32(config.chi0)) if chi_storage == "delta" else 0.0
        self._chi = [torch.full((self.total,), float(config.chi0) - self._chi_off, dtype=f32,
                                device=dev) for _ in range(ns)]
        self._cur, self._prev = 0, 1          # Psi slots
        self._damped = False                   # Psi prev needs (1-m) on read
        self._cc, self._cp, self._oc, self._op = 0, 1, 0, 1   # chi references
        if mask_policy not in ("reference", "boundary_type"):
            raise ValueError(f"mask_policy must be 'reference' or 'boundary_type', got {mask_policy!r}")
        self.mask_policy = mask_policy
        self.has_mask = not (mask_policy == "boundary_type"
                             and config.boundary_type == BoundaryType.PERIODIC)
        # held as a plain tensor: the per-step pointer reads skip the
        # DeviceArray subclass dispatch (the backend's buffers are for callers)
        self.boundary_mask = self.backend.create_boundary_mask(
            self.N, config.boundary_fraction).as_subclass(torch.Tensor)
        self._absorb = 1.0 - self.boundary_mask
        self.sa_enabled = bool(config.sa_enabled and lvl == FieldLevel.COLOR)
        if sa_transform not in ("auto", "fused", "separate", "fields"):
            raise ValueError("sa_transform must be 'auto', 'fused', 'separate' or 'fields', "
                             f"got {sa_transform!r}")
        self._fused = False
        self._sa_scv = False
        if self.sa_enabled:
            self._sa = torch.zeros(3 * self.total, dtype=f32, device=dev)
            self._psisq = torch.empty(3 * self.total, dtype=f32, device=dev)
            self._psisq_valid = False
            self._scv = ScvPlan(self.N, 3)
            # "separate" / "auto" on the hand-written passes: the last pass
            # writes the step's SCV term (one field) instead of the three S_a
            # fields (lfm_scv_apply_scv_f32 + LFM_F_SA_SCV; same bits, 16 B
            # per site less traffic between the two); "fields" keeps the
            # three S_a fields between them, as the drop-in backend does.
            # S_a itself is then recomputed on demand (get_sa_fields).
            self._sa_scv = self._scv.method == 2 and sa_transform in ("auto", "separate")
            self._scvf = torch.empty(self.total, dtype=f32, device=dev) if self._sa_scv else None
            # "fused": the k-axis transform passes folded into the step kernel
            # (power-of-two N up to 512).  Its full-row workgroups hold one
            # workgroup per CU at N = 512, so it is latency-bound there; measured
            # per step (MI355X), fused vs separate: N=64 0.044 vs 0.041 ms,
            # N=128 0.179 vs 0.116, N=256 0.88 vs 0.78, N=512 8.4 vs 6.1 ms.
            # "auto" therefore keeps the separate passes.
            can_fuse = self._scv.method == 2 and self.N in (64, 128, 256, 512)
            if sa_transform == "fused" and not can_fuse:
                raise ValueError(f"sa_transform='fused' needs N in (64, 128, 256, 512), got {self.N}")
            self._fused = can_fuse and sa_transform == "fused"
            self._spec_valid = False    # plan spectrum = k-axis spectra of S_a for current Psi
            self._sa_stale = False      # self._sa not updated by fused steps
        else:
            self._sa = None
        self._params = {}
        self._keep = None      # set_psi_keep: per-site factor folded into every step
        self._src = None       # set_plane_source: plane source folded into every step
        self._det = None       # set_plane_detector: detector plane recorded by the steps
        # True: every evolve holds chi frozen (a barrier folded into the steps,
        # whose per-step chi rewrite makes chi's own update irrelevant)
        self.hold_chi = False
        self._pending = 0      # steps launched since the device step counter was advanced
        self.tile = 0      # tiled-kernel shape (0 = default, see autotune())
        self.chunk = 0     # i planes per workgroup (0 = automatic)
        self._partials = None
        self._red_out = None
        # Optional list: when set, each launch appends a (start, end) event pair
        # recorded on the launch stream around the stencil kernel alone, and
        # kernel_event_steps the number of steps that launch advanced.
        self.kernel_events: list | None = None
        self.kernel_event_steps: list = []
        # evolve() replays whole slot-rotation cycles from HIP graphs once the
        # rotation is in its steady state (see _cycle_graph).
        self.graphs = bool(graphs)
        self._graphs = {}

    # ------------------------------------------------------------------ steps
    def _p(self, flags):
        keep = self._keep.data_ptr() if self._keep is not None else 0
        if keep:
            flags |= _abi.LFM_F_KEEP_PSI
        src, det = self._src, self._det
        if src is not None:
            flags |= _abi.LFM_F_SOURCE
        if det is not None:
            flags |= _abi.LFM_F_DETECTOR
        key = (flags, self.tile, self.chunk, keep, src["table"].data_ptr() if src else 0,
               (det["frames"].data_ptr(), det["first"]) if det else 0)
        p = self._params.get(key)
        if p is None:
            c = self.config
            p = params(self.N, self._dt2, c.kappa, c.lambda_self, c.chi0, c.e0_sq, c.epsilon_w,
                       c.kappa_c, c.epsilon_cc, c.kappa_string,
                       c.kappa_tube if self.sa_enabled else 0.0, flags, self.tile, self.chunk)
            p.keep = keep or None
            if src is not None or det is not None:
                p.step_counter = self._step_dev.data_ptr()
            if src is not None:
                p.src_env = src["env"].data_ptr()
                p.src_amp = src["table"].data_ptr()
                p.src_len = src["table"].numel() // 2
                p.src_axis, p.src_plane = src["axis"], src["plane"]
            if det is not None:
                p.det_frames = det["frames"].data_ptr()
                p.det_len = det["frames"].shape[0]
                p.det_axis, p.det_plane = det["axis"], det["plane"]
                p.det_every, p.det_field, p.det_first = det["every"], det["field"], det["first"]
            self._params[key] = p
        if src is not None or det is not None:
            # this launch's step index is the device counter + step_off
            p.step_off = self._pending + 1
        return p

    def set_plane_source(self, axis: int, plane: int, envelope, amounts) -> None:
        """Add a plane source to Psi inside each step (REAL and COMPLEX), or
        stop it with envelope=None.  After the step that brings the step count
        to s, every site whose coordinate along `axis` is `plane` gets
        Pr += a_r(s) * envelope[t0, t1] and Pi += a_i(s) * envelope[t0, t1]
        (t0 < t1 the other two coordinates; products and sums in float32, a
        zero amount adds nothing), where amounts(s) -> (a_r, a_i), rounded to
        float32 on the host.  It is the `psi[plane] += amount * envelope` a
        continuous source's step callback applies to the current buffers
        (lfm/experiment/source.py:196-228), done by the step kernel as it
        writes Psi, bit for bit the same (before a set_psi_keep factor, as a
        source callback runs before a barrier's).  The amounts are tabulated
        on the device ahead of each evolve call, so steps replayed from graphs
        read their own step's amounts.  Steps then run one at a time."""
        if envelope is None:
            self._src = None
            self._graphs.clear()
            return
        if self.level == 2:
            raise ValueError("a plane source is for the REAL and COMPLEX levels (the reference's "
                             "source writes one N^3 component)")
        N = self.N
        axis, plane = int(axis), int(plane)
        if axis not in (0, 1, 2) or not 0 <= plane < N:
            raise ValueError(f"source axis {axis} / plane {plane} out of range for N={N}")
        env = torch.as_tensor(np.asarray(envelope, dtype=np.float32) if not torch.is_tensor(envelope)
                              else envelope, dtype=torch.float32, device=self.device)
        if env.numel() != N * N:
            raise ValueError(f"envelope must hold N^2 = {N * N} values, got {env.numel()}")
        if not callable(amounts):
            raise ValueError("amounts must be a callable step -> (real amount, imaginary amount)")
        if not hasattr(self, "_step_dev"):
            self._step_dev = torch.zeros((), dtype=torch.int64, device=self.device)
        self._src = {"axis": axis, "plane": plane, "env": env.reshape(-1).contiguous().clone(),
                     "fn": amounts, "table": torch.zeros(2, device=self.device)}
        self._graphs.clear()

    def _source_table(self, last: int) -> None:
        """Make the source's amount table cover step indexes 0..last."""
        src = self._src
        have = src["table"].numel() // 2
        if last < have:
            return
        cap = max(2 * have, last + 1, 1024)
        amt = np.empty((cap, 2), np.float32)
        old = src["table"].cpu().numpy().reshape(-1, 2)
        amt[:have] = old
        fn = src["fn"]
        for s in range(have, cap):
            a_r, a_i = fn(s)
            amt[s] = (a_r, a_i)
        src["table"] = torch.from_numpy(amt.reshape(-1)).to(self.device)
        self._graphs.clear()     # the captured launches point at the old table

    def set_plane_detector(self, axis: int, plane: int, every: int = 1,
                           field: str = "energy_density") -> None:
        """Record the plane `plane` across `axis` inside the steps (REAL and
        COMPLEX), or stop with axis=None.  From now on, after every step that
        brings the step count to start + every, start + 2 every, ..., the
        step kernel writes the plane of the new Psi into the next N x N frame
        of a device buffer: |Psi|^2 = Pr^2 + Pi^2 ("energy_density", as
        lfm/core/evolver.py:346-353 forms it) or Pr ("psi_real"), after a
        plane source and a set_psi_keep factor.  It is a detector screen's
        record() called as a per-step callback (lfm/experiment/detector.py:
        97-115), without the callback's per-step launches and copies;
        detector_frames() returns the frames recorded so far.  Steps then run
        one at a time."""
        if axis is None:
            self._det = None
            self._graphs.clear()
            return
        if self.level == 2:
            raise ValueError("a detector plane is for the REAL and COMPLEX levels")
        kinds = {"energy_density": 0, "psi_real": 1}
        if field not in kinds:
            raise ValueError(f"field must be one of {sorted(kinds)}, got {field!r}")
        axis, plane, every = int(axis), int(plane), int(every)
        if axis not in (0, 1, 2) or not 0 <= plane < self.N or every < 1:
            raise ValueError(f"detector axis {axis} / plane {plane} / every {every} out of range "
                             f"for N={self.N}")
        if not hasattr(self, "_step_dev"):
            self._step_dev = torch.zeros((), dtype=torch.int64, device=self.device)
        N = self.N
        self._det = {"axis": axis, "plane": plane, "every": every, "field": kinds[field],
                     "first": self.step + every,
                     "frames": torch.zeros((0, N, N), device=self.device)}
        self._graphs.clear()

    def _detector_frames(self, last: int) -> None:
        """Make the detector's frame buffer hold every frame up to step index
        `last` (frames already recorded are kept)."""
        det = self._det
        need = (last - det["first"]) // det["every"] + 1
        have = det["frames"].shape[0]
        if need <= have:
            return
        cap = max(2 * have, need, 64)
        frames = torch.zeros((cap,) + tuple(det["frames"].shape[1:]), device=self.device)
        frames[:have].copy_(det["frames"])
        det["frames"] = frames
        self._graphs.clear()     # the captured launches point at the old buffer

    def detector_frames(self) -> torch.Tensor:
        """The frames the detector plane has recorded, (n, N, N) on the
        device, oldest first."""
        det = self._det
        if det is None:
            raise ValueError("no detector plane is set (set_plane_detector)")
        n = max(0, (self.step - det["first"]) // det["every"] + 1)
        return det["frames"][:min(n, det["frames"].shape[0])]

    def _counted(self) -> bool:
        """Whether the steps read the device step counter (a plane source or
        a detector plane)."""
        return self._src is not None or self._det is not None

    def _advance_counter(self) -> None:
        """Add the steps launched since the last call to the device step
        counter (read by a device hook and by the plane source)."""
        if self._pending and hasattr(self, "_step_dev"):
            self._step_dev.add_(self._pending)
        self._pending = 0

    def set_psi_keep(self, keep) -> None:
        """Multiply every new Psi value by keep[site] inside each step (all
        components; N^3 factors, or None to stop).  It is the in-place
        `psi *= factor` an experiment driver applies to the current buffers
        after every step (lfm/experiment/barrier.py:327-353: the barrier's
        absorber and which-path detectors), done by the step kernel as it
        writes Psi, bit for bit the same.  With set_chi(template) and
        evolve(..., freeze_chi=True) it reproduces a barrier's per-step
        enforcement (chi rewritten to the template in all four buffers, Psi
        multiplied) without its per-step passes.  Steps then run one at a
        time (no two-step launches) with the separate S_a transform."""
        if keep is None:
            self._keep = None
        else:
            t = torch.as_tensor(np.asarray(keep, dtype=np.float32) if not torch.is_tensor(keep)
                                else keep, dtype=torch.float32, device=self.device)
            if t.numel() != self.total:
                raise ValueError(f"keep must hold N^3 = {self.total} values, got {t.numel()}")
            self._keep = t.reshape(-1).contiguous().clone()
        self._graphs.clear()

    def autotune(self, reps: int = 6, chunks=(0, 1, 2, 4, 8, 16, 32, 64, 128),
                 choose_two_step: bool = True) -> dict:
        """Time every tile shape and the given plane chunks (0: the library's
        choice) of the stencil kernel on the current state and keep the
        fastest.  At REAL and COMPLEX (N <= 1024, no keep factor, source or
        detector) from 256^3 up with `choose_two_step`, the two-step kernel's
        chunks are timed as well and `two_step` is set to whichever kernel
        advances a step faster: the single-step kernel is bound by the box's HBM rate,
        the two-step kernel by its latency, so the faster one differs from
        box to box.  With `choose_two_step` False, only the kernel `two_step`
        selects is tuned.  The state is saved and restored, so tuning does
        not advance the simulation.  Returns {(tile, chunk): ms per step},
        tile 0 standing for the two-step kernel (or the fused COLOR kernel)."""
        saved = ([t.clone() for t in self._pr], [t.clone() for t in self._pi] if self._pi else None,
                 [t.clone() for t in self._chi], (self._cur, self._prev, self._damped, self._cc,
                                                  self._cp, self._oc, self._op))
        results = {}
        old = (self.tile, self.chunk, self.kernel_events)
        N = self.N
        pair_ok = (self.level < 2 and N <= 1024 and self._keep is None and self._src is None
                   and self._det is None)
        families = []
        if self._fused:
            # the fused COLOR kernel has one tile shape (a full k-row); tune its chunk
            families.append((False, [(0, ch) for ch in dict.fromkeys((0, N, N // 2, N // 4, N // 8))
                                     if ch == 0 or ch >= 4]))
        else:
            if not (pair_ok and self.two_step and not choose_two_step):
                # one-plane chunks only pay on small cache-resident lattices
                # (REAL 64^3: 4.47 us per step at tile 2 / chunk 1 against 4.68)
                families.append((False, [(t, ch) for t in range(1, _abi.LFM_TILE_VARIANTS + 1)
                                         for ch in dict.fromkeys(min(c, N) for c in chunks)
                                         if ch != 1 or N <= 128]))
            # below 256^3 the kernel events of single launches include launch
            # gaps that graph replay hides, and replayed single steps beat
            # pairs there (REAL 64^3: 4.65 against 5.81 us per step), so the
            # two-step kernel is a candidate only from 256^3 up
            if pair_ok and (self.two_step or (choose_two_step and N >= 256)):
                # the two-step kernel has one tile shape; tune its chunk
                families.append((True, [(0, ch) for ch in dict.fromkeys((0, 2, 4, 8, 16, 32, 64))
                                        if ch <= N]))
        grown = False
        if any(pairs for pairs, _ in families) and len(self._pr) < 4:
            # the two-step kernel's rotation needs a fourth slot
            self._pr.append(torch.zeros_like(self._pr[0]))
            if self._pi is not None:
                self._pi.append(torch.zeros_like(self._pi[0]))
            self._chi.append(torch.zeros_like(self._chi[0]))
            grown = True
        kind = {}
        for pairs, combos in families:
            for tile, ch in combos:
                self.tile, self.chunk = tile, ch
                self.kernel_events, self.kernel_event_steps = [], []
                for _ in range(reps + 1):
                    self._step2() if pairs else self._step()
                torch.cuda.synchronize(self.device)
                ts = [a.elapsed_time(b) / n for (a, b), n in
                      zip(self.kernel_events[1:], self.kernel_event_steps[1:])]
                results[(tile, ch)] = sorted(ts)[len(ts) // 2]
                kind[(tile, ch)] = pairs
        best = min(results, key=results.get)
        self.tile, self.chunk = best
        if pair_ok and not self._fused and (self.two_step or (choose_two_step and N >= 256)):
            self.two_step = kind[best]
        self.kernel_events = old[2]
        if grown and not self.two_step:
            # single steps rotate through the first three slots (the restored
            # state below lies in them): drop the fourth again
            self._pr.pop()
            if self._pi is not None:
                self._pi.pop()
            self._chi.pop()
        for dst, src in zip(self._pr, saved[0]):
            dst.copy_(src)
        if saved[1]:
            for dst, src in zip(self._pi, saved[1]):
                dst.copy_(src)
        for dst, src in zip(self._chi, saved[2]):
            dst.copy_(src)
        (self._cur, self._prev, self._damped, self._cc, self._cp, self._oc, self._op) = saved[3]
        if self.sa_enabled:
            self._psisq_valid = False
            self._spec_valid = False
            self._sa_stale = True    # S_a of the restored state's last step: recomputed on read
        return results

    def _free(self, pool, *used):
        return next(q for q in range(len(pool)) if q not in used)

    def _step(self, freeze_chi: bool = False, direct: bool = False) -> None:
        lib = self._lib
        direct = direct or self.tile == self.DIRECT
        cur, prev = self._cur, self._prev
        nxt = self._free(self._pr, cur, prev)
        flags = _abi.LFM_F_CHI_DELTA if self._chi_off else 0
        if self.has_mask:
            flags |= _abi.LFM_F_MASK
            if self._damped:
                flags |= _abi.LFM_F_DAMP_PREV
        if freeze_chi:
            flags |= _abi.LFM_F_FREEZE_CHI
        if direct:
            flags |= _abi.LFM_F_DIRECT
        fused = self._fused and not (freeze_chi or direct) and self._keep is None
        separate_scv = self.level == 2 and self._sa_scv and not fused
        if separate_scv:
            flags |= _abi.LFM_F_SA_SCV
        p = ctypes.byref(self._p(flags))
        st = stream_handle()
        cc, cp = self._cc, self._cp
        cn = self._free(self._chi, cc, cp)
        chi_in, chip_in = self._chi[cc].data_ptr(), self._chi[cp].data_ptr()
        chi_out = self._chi[cn].data_ptr()
        mask = self.boundary_mask.data_ptr() if self.has_mask else None
        pr, pi = self._pr, self._pi
        if self.level < 2:
            ev0 = self._mark()
        if self.level == 0:
            rc = lib.lfm_step_real_f32(pr[cur].data_ptr(), pr[prev].data_ptr(), chi_in, chip_in,
                                       mask, pr[nxt].data_ptr(), None, chi_out, None, p, st)
        elif self.level == 1:
            rc = lib.lfm_step_complex_f32(pr[cur].data_ptr(), pr[prev].data_ptr(),
                                          pi[cur].data_ptr(), pi[prev].data_ptr(), chi_in,
                                          chip_in, mask, pr[nxt].data_ptr(), None,
                                          pi[nxt].data_ptr(), None, chi_out, None, p, st)
        elif fused:
            cfg = self.config
            if not self._spec_valid:
                _abi.check(lib.lfm_scv_prepare_f32(self._scv.handle, pr[cur].data_ptr(),
                                                   pi[cur].data_ptr(), cfg.sa_gamma, cfg.sa_d,
                                                   st), "lfm_scv_prepare_f32")
            ev0 = self._mark()
            rc = lib.lfm_step_color_fused_f32(pr[cur].data_ptr(), pr[prev].data_ptr(),
                                              pi[cur].data_ptr(), pi[prev].data_ptr(), chi_in,
                                              chip_in, mask, pr[nxt].data_ptr(),
                                              pi[nxt].data_ptr(), chi_out, self._scv.handle, p, st)
            self._mark(ev0)
            _abi.check(rc, "lfm_step_color_fused_f32")
            _abi.check(lib.lfm_scv_columns_f32(self._scv.handle, cfg.sa_gamma, cfg.sa_d, st),
                       "lfm_scv_columns_f32")
            ev0 = None
            self._spec_valid = True
            self._sa_stale = True
            self._psisq_valid = False
        else:
            sa = psisq = None
            if separate_scv:
                # the transform writes the step's SCV term; S_a is left stale
                self._spec_valid = False
                e = self._psisq.data_ptr() if self._psisq_valid else None
                _abi.check(lib.lfm_scv_apply_scv_f32(self._scv.handle, pr[cur].data_ptr(),
                                                     pi[cur].data_ptr(), e, self.config.sa_gamma,
                                                     self.config.sa_d, self._scvf.data_ptr(), st),
                           "lfm_scv_apply_scv_f32")
                self._sa_stale = True
                sa, psisq = self._scvf.data_ptr(), self._psisq.data_ptr()
            elif self.sa_enabled:
                self._spec_valid = False
                self._sa_stale = False
                if self._psisq_valid:
                    self._scv.apply_psisq(self._psisq, self.config.sa_gamma, self.config.sa_d,
                                          self._sa)
                else:
                    self._scv.apply(pr[cur], pi[cur], self.config.sa_gamma, self.config.sa_d,
                                    self._sa)
                sa, psisq = self._sa.data_ptr(), self._psisq.data_ptr()
            ev0 = self._mark()
            rc = lib.lfm_step_color_f32(pr[cur].data_ptr(), pr[prev].data_ptr(),
                                        pi[cur].data_ptr(), pi[prev].data_ptr(), chi_in, chip_in,
                                        mask, pr[nxt].data_ptr(), None, pi[nxt].data_ptr(), None,
                                        chi_out, None, sa, None, psisq, p, st)
            if self.sa_enabled:
                self._psisq_valid = True
        if ev0 is not None:
            self._mark(ev0)
        _abi.check(rc, "step")
        self._pending += 1
        self._prev, self._cur = cur, nxt
        self._damped = True
        if freeze_chi:
            # chi untouched; the reference's A/B sets swap roles
            self._cc, self._cp, self._oc, self._op = self._oc, self._op, cc, cp
        else:
            self._cc, self._cp, self._oc, self._op = cn, cc, cc, cp

    def _step2(self) -> None:
        """Two steps in one launch (REAL / COMPLEX, chi evolving)."""
        lib = self._lib
        cur, prev = self._cur, self._prev
        n1 = self._free(self._pr, cur, prev)
        n2 = self._free(self._pr, cur, prev, n1)
        flags = _abi.LFM_F_CHI_DELTA if self._chi_off else 0
        if self.has_mask:
            flags |= _abi.LFM_F_MASK
            if self._damped:
                flags |= _abi.LFM_F_DAMP_PREV
        p = ctypes.byref(self._p(flags))
        st = stream_handle()
        cc, cp = self._cc, self._cp
        c1 = self._free(self._chi, cc, cp)
        c2 = self._free(self._chi, cc, cp, c1)
        ch = self._chi
        mask = self.boundary_mask.data_ptr() if self.has_mask else None
        pr, pi = self._pr, self._pi
        ev0 = self._mark()
        if self.level == 0:
            rc = lib.lfm_step2_real_f32(pr[cur].data_ptr(), pr[prev].data_ptr(), ch[cc].data_ptr(),
                                        ch[cp].data_ptr(), mask, pr[n1].data_ptr(),
                                        ch[c1].data_ptr(), pr[n2].data_ptr(), ch[c2].data_ptr(),
                                        p, st)
        else:
            rc = lib.lfm_step2_complex_f32(pr[cur].data_ptr(), pr[prev].data_ptr(),
                                           pi[cur].data_ptr(), pi[prev].data_ptr(),
                                           ch[cc].data_ptr(), ch[cp].data_ptr(), mask,
                                           pr[n1].data_ptr(), pi[n1].data_ptr(), ch[c1].data_ptr(),
                                           pr[n2].data_ptr(), pi[n2].data_ptr(), ch[c2].data_ptr(),
                                           p, st)
        if ev0 is not None:
            self._mark(ev0, steps=2)
        _abi.check(rc, "step2")
        self._prev, self._cur = n1, n2
        self._damped = True
        self._cc, self._cp, self._oc, self._op = c2, c1, c1, cc

    def _mark(self, start=None, steps=1):
        """Record a timing event on the current stream (kernel_events only)."""
        if self.kernel_events is None:
            return None
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        if start is not None:
            self.kernel_events.append((start, e))
            self.kernel_event_steps.append(steps)
        return e

    # ------------------------------------------------------- graph replay
    def _key(self):
        return (self._cur, self._prev, self._cc, self._cp, self._oc, self._op, self._damped)

    def _rotate(self, key, pair, freeze=False):
        """The slot bookkeeping of one _step (pair=False, chi evolving; with
        freeze, chi frozen: the two chi reference sets swap) or one _step2
        (pair=True), applied to a state key without launching."""
        cur, prev, cc, cp, oc, op, _ = key
        n = len(self._pr)
        free = lambda *u: next(q for q in range(n) if q not in u)
        if freeze:
            return (free(cur, prev), cur, oc, op, cc, cp, True)
        if pair:
            n1 = free(cur, prev)
            c1 = free(cc, cp)
            return (free(cur, prev, n1), n1, free(cc, cp, c1), c1, c1, cc, True)
        cn = free(cc, cp)
        return (free(cur, prev), cur, cn, cc, cc, cp, True)

    def _sa_flags(self):
        return (self._psisq_valid, self._spec_valid, self._sa_stale) if self.sa_enabled else ()

    def _hooked_step(self, hook, freeze_chi: bool = False) -> None:
        """One step, then (with a device hook or a plane source) the device
        step counter and the hook, as evolve runs them."""
        self._step(freeze_chi=freeze_chi)
        if hook is not None or self._counted():
            self._advance_counter()
        if hook is not None:
            hook(self, self._step_dev)

    def _cycle_graph(self, cycles: int = 1, hook=None, freeze: bool = False):
        """(graph, launches, steps) replaying `cycles` whole rotation cycles
        from the current state, or None when the state is not steady or
        graphs are off.  A cycle is the shortest launch sequence that returns
        every slot reference to where it started, so a replay leaves the
        bookkeeping unchanged; the S_a flags must also come back to their
        values (every step leaves them alike), else the capture is dropped.
        Graphs are kept per starting state (the pointers, flags and
        parameters of every launch are fixed by it), cycle count and hook."""
        if not self.graphs or self.kernel_events is not None or not self._damped:
            return None
        if self._cc == self._cp and not freeze:
            return None
        if self.level == 2 and self.sa_enabled and self._scv.method != 2:
            return None              # rocFFT launches stay out of graphs
        pair = (self.two_step and self.N <= 1024 and hook is None and self._keep is None
                and not self._counted() and not freeze)
        key0 = self._key()
        k, m = key0, 0
        while True:
            k, m = self._rotate(k, pair, freeze), m + 1
            if k == key0 or m > 6:
                break
        if k != key0:
            return None
        m *= cycles
        flags0 = self._sa_flags()
        gkey = key0 + flags0 + (self.tile, self.chunk, pair, cycles, hook, freeze)
        if gkey in self._graphs:
            return self._graphs[gkey]
        while len(self._graphs) >= self.GRAPH_CACHE:
            # bounded cache (each hook object keys its own graphs): drop the oldest
            self._graphs.pop(next(iter(self._graphs)))
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        self._pending = 0
        with graph_capture(g, s):
            for _ in range(m):
                if pair:
                    self._step2()
                elif hook is not None:
                    self._hooked_step(hook)
                else:
                    # source / detector launches read counter + 1 .. m;
                    # the counter advances once per replay
                    self._step(freeze_chi=freeze)
            if self._counted():
                self._advance_counter()
        self._pending = 0
        torch.cuda.current_stream(self.device).wait_stream(s)
        # Capture recorded the launches without running them.  The bookkeeping
        # it advanced is back at the start (a whole cycle); when the S_a flags
        # are too, the state still describes the buffers and every replay
        # repeats the same launches.  Otherwise the flags are restored and the
        # state is marked as not replayable.
        if self._key() != key0:
            raise RuntimeError("slot rotation did not return to its start after a captured cycle")
        if self._sa_flags() != flags0:
            (self._psisq_valid, self._spec_valid, self._sa_stale) = flags0
            self._graphs[gkey] = None
            return None
        hit = (g, m, 2 * m if pair else m)
        self._graphs[gkey] = hit
        return hit

    def evolve(self, steps: int, callback=None, freeze_chi: bool = False,
               device_hook=None) -> None:
        """Run `steps` leapfrog steps; callback(evolver, step) every
        config.report_interval steps (lfm/core/evolver.py:137-180).  REAL and
        COMPLEX run pairs of steps per launch when no callback falls between
        them; the result is the same as single steps.  Once the slot rotation
        is steady, whole rotation cycles are replayed from HIP graphs (same
        launches, so the same bits as eager stepping).  With freeze_chi a
        cycle is six steps: Psi rotates through three buffers while the two
        chi reference sets swap every step.

        device_hook(evolver, step): called after every step with `step` a
        0-d int64 device tensor holding the step number.  It must only queue
        device work (torch ops on the native views, no host reads), so it is
        captured into the replayed graphs with the steps: a per-step barrier
        or source then costs no host launches."""
        report = self.config.report_interval
        left = int(steps)
        hook = device_hook
        freeze_chi = freeze_chi or self.hold_chi
        pair_ok = (self.two_step and not freeze_chi and self.N <= 1024 and hook is None
                   and self._keep is None and not self._counted())
        if hook is not None and freeze_chi:
            raise ValueError("device_hook with freeze_chi is not supported")
        if self._src is not None:
            self._source_table(self.step + left)
        if self._det is not None:
            self._detector_frames(self.step + left)
        if hook is not None or self._counted():
            if not hasattr(self, "_step_dev"):
                self._step_dev = torch.zeros((), dtype=torch.int64, device=self.device)
            self._step_dev.fill_(self.step)
        self._pending = 0
        while left > 0:
            to_report = report - self.step % report if (callback is not None and report > 0) else left
            cyc = self._cycle_graph(hook=hook, freeze=freeze_chi)
            if cyc is not None and min(left, to_report) >= cyc[2]:
                # long runs replay a graph of several cycles (fewer host
                # launches per step), then single cycles
                big = self._cycle_graph(self.GRAPH_CYCLES, hook, freeze_chi) \
                    if min(left, to_report) >= self.GRAPH_CYCLES * cyc[2] else None
                for g, _, n in ((big or cyc), cyc):
                    while min(left, to_report) >= n:
                        g.replay()
                        self.step += n
                        left -= n
                        to_report -= n
                if callback is not None and report > 0 and self.step % report == 0:
                    callback(self, self.step)
                continue
            if pair_ok and left >= 2 and to_report >= 2:
                self._step2()
                self.step += 2
                left -= 2
            else:
                self._hooked_step(hook, freeze_chi)
                self.step += 1
                left -= 1
            if callback is not None and report > 0 and self.step % report == 0:
                callback(self, self.step)

    # -------------------------------------------------------- device views
    def native_psi_real(self) -> torch.Tensor:
        """Live device view of the current Psi real parts.  Writes through it
        are allowed, so the cached |Psi|^2 / S_a spectra are dropped."""
        self._touch_psi()
        return self._pr[self._cur]

    def native_psi_imag(self) -> torch.Tensor | None:
        self._touch_psi()
        return self._pi[self._cur] if self._has_imag else None

    def native_chi(self) -> torch.Tensor:
        """Live device view of the current chi buffer, in storage units
        (chi - chi0 when chi_storage == "delta")."""
        return self._chi[self._cc]

    def native_chi_prev(self) -> torch.Tensor:
        return self._chi[self._cp]

    def native_chi_refs(self) -> tuple:
        """The four reference chi buffers (current set's chi, other set's chi,
        current set's prev, other set's prev) as device views, in storage
        units; some may alias one slot."""
        ch = self._chi
        return ch[self._cc], ch[self._oc], ch[self._cp], ch[self._op]

    def use_absolute_chi(self) -> None:
        """Switch the chi buffers to absolute storage (the reference's
        representation), in place, so that code writing absolute chi through
        native views sees what it expects.  No-op when already absolute."""
        if not self._chi_off:
            return
        for t in self._chi:
            t.add_(self._chi_off)
        self._chi_off = 0.0
        self.chi_storage = "absolute"
        self._graphs.clear()

    def chi_tensor(self) -> torch.Tensor:
        """Absolute chi as a new device tensor."""
        return self._chi[self._cc] + self._chi_off if self._chi_off else self._chi[self._cc].clone()

    def _prev_view(self, pool) -> torch.Tensor:
        t = pool[self._prev]
        if self._damped and self.has_mask:
            return t * self._absorb.repeat(self._n_psi) if self._n_psi > 1 else t * self._absorb
        return t

    # ------------------------------------------------------------- getters
    def _shape(self, flat: np.ndarray) -> np.ndarray:
        N = self.N
        if self.level == 2:
            return flat.reshape(self._n_psi, N, N, N)
        return flat.reshape(N, N, N)

    def _abs_chi(self, slot) -> np.ndarray:
        t = self._chi[slot]
        if self._chi_off:
            t = t + self._chi_off
        return t.cpu().numpy().reshape(self.N, self.N, self.N)

    def get_chi(self) -> np.ndarray:
        return self._abs_chi(self._cc)

    def get_chi_prev(self) -> np.ndarray:
        return self._abs_chi(self._cp)

    def get_psi_real(self) -> np.ndarray:
        return self._shape(self._pr[self._cur].cpu().numpy())

    def get_psi_imag(self) -> np.ndarray | None:
        return self._shape(self._pi[self._cur].cpu().numpy()) if self._has_imag else None

    def get_psi_real_prev(self) -> np.ndarray:
        return self._shape(self._prev_view(self._pr).cpu().numpy())

    def get_psi_imag_prev(self) -> np.ndarray | None:
        return self._shape(self._prev_view(self._pi).cpu().numpy()) if self._has_imag else None

    def get_energy_density(self) -> np.ndarray:
        pr = self.get_psi_real()
        e2 = np.sum(pr ** 2, axis=0) if pr.ndim == 4 else pr ** 2
        pi = self.get_psi_imag()
        if pi is not None:
            e2 += np.sum(pi ** 2, axis=0) if pi.ndim == 4 else pi ** 2
        return e2

    def get_sa_fields(self) -> np.ndarray | None:
        """S_a used by the last step: the transform of |Psi|^2 of that step's
        input Psi (numpy_backend.py:341), clipped at 0."""
        if self._sa is None:
            return None
        self._materialize_sa()
        return torch.clamp_min(self._sa, 0.0).cpu().numpy().reshape(3, self.N, self.N, self.N)

    def _materialize_sa(self) -> None:
        """Fused steps keep S_a only in k space and SCV steps only as the SCV
        term: recompute S_a of the last step from that step's input Psi,
        which the rotation keeps (undamped) in the prev slot.  Called before
        anything overwrites or damps that slot."""
        if self.sa_enabled and self._sa_stale:
            self._scv.apply(self._pr[self._prev], self._pi[self._prev], self.config.sa_gamma,
                            self.config.sa_d, self._sa)
            self._sa_stale = False

    # ------------------------------------------------------------- setters
    def _dev_chi(self, arr) -> torch.Tensor:
        d = self._dev(arr)
        return d - self._chi_off if self._chi_off else d

    def _dev(self, arr) -> torch.Tensor:
        if isinstance(arr, torch.Tensor):
            return arr.detach().to(device=self.device, dtype=torch.float32).reshape(-1).contiguous()
        flat = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).ravel())
        return torch.from_numpy(flat).to(self.device)

    def _undamp(self) -> None:
        """Store the damped prev values explicitly (before a partial write)."""
        self._materialize_sa()
        if self._damped and self.has_mask:
            a = self._absorb.repeat(self._n_psi) if self._n_psi > 1 else self._absorb
            self._pr[self._prev].mul_(a)
            if self._has_imag:
                self._pi[self._prev].mul_(a)
        self._damped = False

    def _touch_psi(self) -> None:
        if self.sa_enabled:
            self._psisq_valid = False
            self._spec_valid = False

    def set_psi_real(self, arr) -> None:
        d = self._dev(arr)
        self._undamp()
        self._pr[self._cur].copy_(d)
        self._pr[self._prev].copy_(d)
        self._touch_psi()

    def set_psi_imag(self, arr) -> None:
        if not self._has_imag:
            raise ValueError("Cannot set imaginary part for REAL field level")
        d = self._dev(arr)
        self._undamp()
        self._pi[self._cur].copy_(d)
        self._pi[self._prev].copy_(d)
        self._touch_psi()

    def set_psi_real_prev(self, arr) -> None:
        d = self._dev(arr)
        self._undamp()
        self._pr[self._prev].copy_(d)

    def set_psi_imag_prev(self, arr) -> None:
        if not self._has_imag:
            raise ValueError("Cannot set imaginary part for REAL field level")
        d = self._dev(arr)
        self._undamp()
        self._pi[self._prev].copy_(d)

    def set_psi_real_current(self, arr) -> None:
        self._pr[self._cur].copy_(self._dev(arr))
        self._touch_psi()

    def set_psi_imag_current(self, arr) -> None:
        if not self._has_imag:
            raise ValueError("Cannot set imaginary part for REAL field level")
        self._pi[self._cur].copy_(self._dev(arr))
        self._touch_psi()

    def set_chi(self, arr) -> None:
        """All four reference chi buffers := arr."""
        self._chi[0].copy_(self._dev_chi(arr))
        self._cc = self._cp = self._oc = self._op = 0

    def set_chi_current(self, arr) -> None:
        """Both sets' current chi := arr; prev buffers unchanged."""
        s = self._free(self._chi, self._cp, self._op)
        self._chi[s].copy_(self._dev_chi(arr))
        self._cc = self._oc = s

    def set_chi_prev(self, arr) -> None:
        """Both sets' prev chi := arr; current buffers unchanged."""
        s = self._free(self._chi, self._cc, self._oc)
        self._chi[s].copy_(self._dev_chi(arr))
        self._cp = self._op = s

    def get_chi_storage(self) -> tuple[np.ndarray, np.ndarray]:
        """Raw stored (current, prev) chi buffers (chi - chi0 for deviation
        storage): the exact state, for bit-exact checkpoints."""
        return self._chi[self._cc].cpu().numpy(), self._chi[self._cp].cpu().numpy()

    def set_chi_storage(self, cur, prev) -> None:
        """Restore raw stored chi buffers written by get_chi_storage."""
        self._chi[0].copy_(self._dev(cur))
        self._chi[1].copy_(self._dev(prev))
        self._cc, self._cp, self._oc, self._op = 0, 1, 1, 1

    def set_sa_fields(self, arr) -> None:
        if self._sa is None:
            raise ValueError("SA fields not allocated -- set kappa_tube > 0 in SimulationConfig"
                             " before creating the Evolver.")
        self._sa.copy_(self._dev(arr))
        self._sa_stale = False

    # --------------------------------------------------------- reductions
    def _red_bufs(self):
        if self._partials is None:
            n = int(self._lib.lfm_reduce_partials_len(self.N))
            self._partials = torch.empty(n, dtype=torch.float64, device=self.device)
            self._red_out = torch.empty(8, dtype=torch.float64, device=self.device)
        return self._partials, self._red_out

    def energy_sums(self, prev_r=None, prev_i=None, c=None) -> dict[str, float]:
        """FP64 sums of kinetic, gradient and potential energy and of |Psi|^2
        (lfm/analysis/energy.py:26-135) on the GPU.  By default the kinetic
        term uses the true leapfrog prev buffers (as the reference stores
        them: damped); pass `prev_r`/`prev_i` device tensors to use others."""
        part, out = self._red_bufs()
        damp = 0
        if prev_r is None:
            prev_r = self._pr[self._prev]
            prev_i = self._pi[self._prev] if self._has_imag else None
            damp = int(self._damped and self.has_mask)
        pi = self._pi[self._cur] if self._has_imag else None
        c = self.config.c if c is None else c
        rc = self._lib.lfm_energy_f64(
            self._pr[self._cur].data_ptr(), prev_r.data_ptr(), pi.data_ptr() if pi is not None else None,
            prev_i.data_ptr() if prev_i is not None else None, self._chi[self._cc].data_ptr(),
            self._chi_off, self.boundary_mask.data_ptr(), damp, self._n_psi, self.N, float(self.config.dt),
            float(c), part.data_ptr(), out.data_ptr(), stream_handle())
        _abi.check(rc, "lfm_energy_f64")
        T, G, V, S = out[:4].tolist()
        return {"energy_kinetic": T, "energy_gradient": G, "energy_potential": V,
                "energy_total": T + G + V, "psi_sq_total": S}

    def total_energy(self) -> float:
        return self.energy_sums()["energy_total"]

    def chi_stats(self, well_threshold=17.0, void_threshold=18.0) -> dict[str, float]:
        """chi min/max/mean/std and well/void fractions over the interior ball
        (lfm/analysis/structure.py:26-108,148-171), FP64 on the GPU."""
        part, out = self._red_bufs()
        rc = self._lib.lfm_chi_stats_f64(self._chi[self._cc].data_ptr(), self._chi_off, self.N,
                                         float(self.config.boundary_fraction),
                                         float(well_threshold), float(void_threshold),
                                         part.data_ptr(), out.data_ptr(), stream_handle())
        _abi.check(rc, "lfm_chi_stats_f64")
        v = out[:7].tolist()
        return {"chi_min": v[0], "chi_max": v[1], "chi_mean": v[2], "chi_std": v[3],
                "well_fraction": v[4], "void_fraction": v[5], "interior_count": v[6]}
