#!/usr/bin/env python3
"""Throughput of the GOV-01/GOV-02 leapfrog step on one MI355X.

Default workload: COLOR level, full_physics preset, 512^3, float32, cos^2 mask
on (FROZEN), S_a Helmholtz transform every step (gamma 0.1, D 4.9); three
seeded Gaussian blobs (one colour dominant each), amplitude 4, width 3,
random phases, chi = 19.  A "step" is one full leapfrog step: the five
hand-written S_a transform passes (k-axis real-to-half-spectrum, j forward, i
forward + Helmholtz filter + inverse, j inverse, k-axis back to real), then
the fused stencil kernel.

Prints one JSON line.  `value` is site updates per second of the whole step
(GLUPS, inputs resident in HBM).  `roofline` is for the dominant kernel, the
fused COLOR stencil, timed with HIP events on its own launch stream, on its
algorithmic 100 B per site update (the |Psi'|^2 it also writes for the next
transform is reported beside it, 112 B); `step_roofline` is the whole step on
its 160 B per site.  `batched` and `adjoint` carry the metric's other two
clauses, each at its own configuration: the 64-lattice batched sweep against
64 sequential plain runs (lattices/s, 500 steps) and the COMPLEX 128^3
400-step adjoint against a plain forward run (cost ratios).  `cpu_baseline`
times the NumPy oracle (float32, one core) on bounded samples of the REAL
64^3, COMPLEX 128^3 and COLOR 64^3 workloads, with the host's CPU model,
NumPy version and thread variables.  --quick leaves out the batched and
adjoint sub-results.

Other workloads (--workload): complex512 (gravity_em, periodic, mask-free
kernel), real64 (gravity_only 64^3), batch64 (batched sweep of 64 COMPLEX
lattices against sequential runs), adjoint128 (forward + VJP against the
forward alone), relax64 (eigenmode relaxation), copy (device copy used to
scale the HBM counters).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lfm-physics_amd"))
sys.path.insert(0, ROOT)

METRIC = "GLUPS + %HBM roofline at 512³ COLOR fp32; batched & adjoint speedup vs plain"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per site update (float32): every state value read
# once, every new value written once (three-buffer rotation).
#   COLOR, mask, fused stencil kernel: reads 6 Psi + 6 Psi_prev + chi + chi_prev
#                + mask + 3 S_a = 72 B, writes 6 Psi' + chi' = 28 B      -> 100 B
#                (b_extra: the kernel also writes |Psi_a'|^2, 12 B, the next
#                S_a transform's input, which the whole-step figure books to
#                the transform; b_scv_saving: the Evolver's kernel reads the
#                SCV term, 4 B, which the last transform pass forms from the
#                three S_a values, instead of the 12 B of S_a, so it moves
#                100 - 8 + 12 = 104 B per site itself)
#   COLOR whole step: the kernel's 100 B + |Psi'|^2 written (12) + real-to-half-
#                spectrum read and write (12 + 12) + half-spectrum-to-real read
#                and write (12 + 12)                                     -> 160 B
#   COMPLEX, no mask: reads Pr, Pi, Pr_p, Pi_p, chi, chi_p; writes Pr', Pi', chi' -> 36 B
#   REAL, mask: reads E, E_p, chi, chi_p, mask; writes E', chi'         -> 28 B
WORKLOADS = {
    "color512": dict(level=2, N=512, preset="full_physics", mask=True, b_alg=100, b_extra=12,
                     b_scv_saving=8, b_step=160, streams=(18, 10), streams_scv=(16, 10),
                     desc="COLOR full_physics 512^3 fp32, mask on, S_a FFT every step"),
    # b_alg2 / streams2: per step of the two-step kernel (one launch reads the
    # fields once and writes both steps' Psi and chi)
    "complex512": dict(level=1, N=512, preset="gravity_em", mask=False, b_alg=36, streams=(6, 3),
                       b_alg2=24, streams2=(6, 6),
                       desc="COMPLEX gravity_em 512^3 fp32, periodic (mask-free kernel)"),
    "real64": dict(level=0, N=64, preset="gravity_only", mask=True, b_alg=28, streams=(5, 2),
                   b_alg2=18, streams2=(5, 4),
                   desc="REAL gravity_only 64^3 fp32, mask on"),
}


WARMUP_CALLS = (25, 25, 25, 26, 26, 27)


def stream_ceiling(nin, nout, n, reps=10):
    """HBM rate of a pure streaming kernel (lfm_calib_stream_f32) with the step
    kernel's count of planar float32 read and write streams, n floats each:
    what that access pattern reaches on this device with no stencil, halo or
    arithmetic in the way.  Returns GB/s, the best over plain / non-temporal
    accesses, 4- / 8-byte lanes and 4096 / 65536 workgroups."""
    import ctypes
    import torch
    from lfm_hip import _abi
    lib = _abi.load()
    bufs = [torch.zeros(n, device="cuda") for _ in range(nin + nout)]
    ins = (ctypes.c_void_p * nin)(*[b.data_ptr() for b in bufs[:nin]])
    outs = (ctypes.c_void_p * nout)(*[b.data_ptr() for b in bufs[nin:]])
    st = torch.cuda.current_stream().cuda_stream
    best = 0.0
    variants = [(f, g) for f in (0, 1, 2, 3) for g in (4096, 65536) if n % 2 == 0 or f < 2]
    for flags, grid in variants:
        for _ in range(2):
            _abi.check(lib.lfm_calib_stream_f32(ins, nin, outs, nout, n, flags, grid, st),
                       "lfm_calib_stream_f32")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            lib.lfm_calib_stream_f32(ins, nin, outs, nout, n, flags, grid, st)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        best = max(best, 4.0 * (nin + nout) * n / (ms * 1e-3) / 1e9)
    del bufs
    torch.cuda.empty_cache()
    return best


def make_config(w, N):
    import lfm_hip as L
    mk = {"full_physics": L.full_physics, "gravity_em": L.gravity_em,
          "gravity_only": L.gravity_only}[w["preset"]]
    if not w["mask"]:
        return mk(N, boundary_type=L.BoundaryType.PERIODIC)
    return mk(N)


def gpu_ics(w, N, seed=42):
    """Seeded initial fields built on the GPU (same formulas as lfm_hip.fields)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    x = torch.arange(N, dtype=torch.float32, device=dev)
    X, Y, Z = x.view(N, 1, 1), x.view(1, N, 1), x.view(1, 1, N)

    def blob(pos, amp, sig):
        r2 = (X - pos[0]) ** 2 + (Y - pos[1]) ** 2 + (Z - pos[2]) ** 2
        return amp * torch.exp(-r2 / (2.0 * sig ** 2))

    c = N / 2.0
    if w["level"] == 0:
        E = blob((c, c, c), 6.0, 3.0)
        E += torch.from_numpy((rng.standard_normal((N, N, N)) * 1e-3).astype(np.float32)).to(dev)
        return E, None
    if w["level"] == 1:
        pos = rng.uniform(0, N, (8, 3))
        amp = rng.uniform(2, 6, 8)
        sig = rng.uniform(2, 4, 8)
        ph = rng.uniform(0, 2 * np.pi, 8)
        pr = torch.zeros((N, N, N), device=dev)
        pi = torch.zeros((N, N, N), device=dev)
        for s in range(8):
            env = blob(pos[s], float(amp[s]), float(sig[s]))
            pr += env * float(np.cos(ph[s]))
            pi += env * float(np.sin(ph[s]))
        return pr, pi
    pr = torch.zeros((3, N, N, N), device=dev)
    pi = torch.zeros((3, N, N, N), device=dev)
    ph = rng.uniform(0, 2 * np.pi, 3)
    for a in range(3):
        # uniform in the ball of radius 0.35 * N/2 about the centre
        while True:
            d = rng.uniform(-1, 1, 3)
            if d @ d <= 1:
                break
        pos = c + 0.35 * c * d
        env = blob(pos, 4.0, 3.0)
        pr[a] += env * float(np.cos(ph[a]))
        pi[a] += env * float(np.sin(ph[a]))
    return pr, pi


CPU_CASES = (("real64", 64), ("complex512", 128), ("color512", 64))


def host_info():
    """The host the CPU baseline ran on: core count, CPU model, NumPy version
    and the thread-count variables (NumPy's element-wise ops and pocketfft run
    on one thread whatever they say)."""
    import numpy as np
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                          "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS")}
    return {"nproc": os.cpu_count(), "cpu": model or platform.machine(), "numpy": np.__version__,
            "threads_env": env}


def cpu_case(name, N, budget_s=10.0):
    """The oracle's float32 NumPy step (one core) on the workload's seeded
    fields at size N, run for `budget_s` seconds after one untimed step."""
    import numpy as np
    from oracle import lfm_oracle as O
    w = WORKLOADS[name]
    cfg = make_config(w, N)
    pr_t, pi_t = gpu_ics(w, N)
    pr = pr_t.cpu().numpy()
    pi = pi_t.cpu().numpy() if pi_t is not None else None
    p = dict(dt2=cfg.dt ** 2, kappa=cfg.kappa, lambda_self=cfg.lambda_self, chi0=cfg.chi0,
             e0_sq=cfg.e0_sq, epsilon_w=cfg.epsilon_w, kappa_c=cfg.kappa_c,
             epsilon_cc=cfg.epsilon_cc, kappa_string=cfg.kappa_string,
             kappa_tube=cfg.kappa_tube, sa_gamma=cfg.sa_gamma, sa_d=cfg.sa_d)
    mask = O.boundary_mask(N, cfg.boundary_fraction) if w["mask"] else np.zeros(N ** 3, np.float32)
    st = O.OracleState(w["level"], N, pr, pi, np.full(N ** 3, cfg.chi0), mask, p, dtype=np.float32)
    st.step(1)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        st.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"case": f"{['REAL', 'COMPLEX', 'COLOR'][w['level']]} {w['preset']} {N}^3"
                    f"{'' if w['mask'] else ' periodic'}",
            "GLUPS": N ** 3 * steps / dt / 1e9, "MLUPS": N ** 3 * steps / dt / 1e6,
            "steps_per_s": steps / dt, "steps": steps, "seconds": dt}


def cpu_baseline(w, budget_s=10.0):
    """The oracle's float32 NumPy step on one core, timed on bounded samples:
    REAL gravity_only 64^3, COMPLEX gravity_em 128^3 (periodic) and COLOR
    full_physics 64^3, each for `budget_s` seconds.  `value` is the case of
    the measured workload's level."""
    cases = [cpu_case(name, N, budget_s) for name, N in CPU_CASES]
    mine = cases[w["level"]]
    return {"value": mine["GLUPS"], "unit": "GLUPS", "cores": 1, "kind": "port",
            "sample": f"oracle NumPy float32 {mine['case']}, {mine['steps']} steps in "
                      f"{mine['seconds']:.1f} s ({mine['steps_per_s']:.2f} steps/s, one thread)",
            "cases": cases, "host": host_info()}


def batch64_recipe(nlat=64, seed=42):
    """The batched-sweep workload's lattices: for each of `nlat`, N drawn from
    {32, 48, 64, 96, 128}, kappa = (1/63) U(0.5, 2), one centred soliton with
    amplitude U(2, 8) and width U(2, 4), plus N(0, 1e-3) noise on both parts;
    COMPLEX, default cos^2 mask.  Returns (configs, [(psi_r, psi_i)])."""
    import numpy as np
    import lfm_hip as L
    rng = np.random.default_rng(seed)
    cfgs, ics = [], []
    for _ in range(nlat):
        N = int(rng.choice([32, 48, 64, 96, 128]))
        kap = (1 / 63) * rng.uniform(0.5, 2)
        amp, sig = rng.uniform(2, 8), rng.uniform(2, 4)
        cfgs.append(L.SimulationConfig(grid_size=N, field_level=L.FieldLevel.COMPLEX, kappa=kap))
        r, i = L.fields.gaussian_soliton(N, (N // 2, N // 2, N // 2), amp, sig)
        nr, ni = L.fields.seed_noise(N, 1e-3, rng=rng)
        ics.append((r + nr, i + ni))
    return cfgs, ics


def run_batch(args):
    """Batched sweep workload: 64 independent COMPLEX lattices with uneven
    sizes from {32,48,64,96,128}^3, kappa_b = kappa * U(0.5, 2), one soliton
    (amplitude U(2,8), width U(2,4)) at the centre plus N(0, 1e-3) noise on both
    parts, default cos^2 mask, float32, 500 steps.  Batched: one launch per
    step for all lattices, replayed from a HIP graph, autotuned tile/chunk.
    Sequential: the same 64 lattices through the plain Evolver, one after
    another, with its HIP-graph replay of rotation cycles (and, reported
    beside it, with one eager launch per step as the reference's loop does);
    IC generation excluded in both."""
    import numpy as np
    import torch
    import lfm_hip as L
    cfgs, ics = batch64_recipe()
    nlat = len(cfgs)
    steps = args.steps
    sites = sum(c.grid_size ** 3 for c in cfgs)
    batch = L.LatticeBatch(cfgs, backend="hip", tile=0 if args.tile == "auto" else int(args.tile),
                           chunk=args.chunk, two_step=args.two_step)

    def reset_batch():
        for b, (r, i) in enumerate(ics):
            batch.set_psi_real(b, r)
            batch.set_psi_imag(b, i)
            batch.set_chi(b, np.full((cfgs[b].grid_size,) * 3, cfgs[b].chi0, np.float32))

    reset_batch()
    tuning = batch.autotune() if args.tile == "auto" else None
    # the tuner may pick the two-step kernel (or single steps) for this box
    two = batch.two_step
    # untimed runs that start at every slot-rotation phase capture the graphs
    # the timed run replays; then the same initial state again
    for n in WARMUP_CALLS[:3]:
        batch.evolve(n)
    reset_batch()
    batch.evolve(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.evolve(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    chi_b0 = batch.get_chi(0)
    # kernel-only time: events around eager launches on the launch stream
    evs = []
    for _ in range(12):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        batch._step2() if two else batch._step()
        e1.record()
        evs.append((e0, e1))
    torch.cuda.synchronize()
    # per step (a two-step launch advances every lattice twice)
    kms = sorted(a.elapsed_time(b) for a, b in evs[2:])[len(evs[2:]) // 2] / (2 if two else 1)
    ok = all(bool(torch.isfinite(t).all()) for t in batch._chi)
    del batch
    torch.cuda.empty_cache()
    # sequential plain runs: the Evolver as a user runs it (whole rotation
    # cycles replayed from HIP graphs), and with one eager launch per step as
    # the reference's loop does
    t_seq = {}
    for graphs in (True, False):
        evols = []
        for c, (r, i) in zip(cfgs, ics):
            ev = L.Evolver(c, backend="hip", graphs=graphs)
            ev.set_psi_real(r)
            ev.set_psi_imag(i)
            evols.append(ev)
        for ev in evols:
            ev.evolve(args.warmup)
            # the graphs the timed call replays from this phase (one cycle and
            # the multi-cycle graph), captured without running any step
            if graphs and ev._cycle_graph() is not None:
                ev._cycle_graph(ev.GRAPH_CYCLES)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for ev in evols:
            ev.evolve(steps)
        torch.cuda.synchronize()
        t_seq[graphs] = time.perf_counter() - t0
        if graphs:
            same = bool(np.array_equal(evols[0].get_chi(), chi_b0))
        del evols, ev
        torch.cuda.empty_cache()
    # COMPLEX with mask: 7 reads + 3 writes of float32 per step; two steps per
    # launch read the 7 once and write both steps' 3 (26 B per step)
    b_alg = 26 if two else 40
    achieved = b_alg * sites / (kms * 1e-3) / 1e9
    out = {"metric": METRIC, "value": nlat / t_batch, "unit": "lattices/s", "n_gpus": 1,
           "steps": steps, "warmup": args.warmup, "ms_per_step": t_batch * 1e3 / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "batched sweep: 64 COMPLEX lattices, N in {32,48,64,96,128}, "
                                  f"{steps} steps, mask on", "lattices": nlat, "sites": sites,
                      "steps_per_launch": 2 if two else 1},
           "batched": {"lattices_per_s": nlat / t_batch, "GLUPS": sites * steps / t_batch / 1e9,
                       "wall_s": t_batch, "tile": batch_tile(tuning)},
           "sequential_graphs": {"lattices_per_s": nlat / t_seq[True],
                                 "GLUPS": sites * steps / t_seq[True] / 1e9, "wall_s": t_seq[True],
                                 "launches": "HIP-graph replay of rotation cycles"},
           "sequential_eager": {"lattices_per_s": nlat / t_seq[False],
                                "GLUPS": sites * steps / t_seq[False] / 1e9,
                                "wall_s": t_seq[False], "launches": "one eager launch per step"},
           # against the faster of the two sequential launch modes
           "speedup_vs_sequential": min(t_seq.values()) / t_batch,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "step2_tiled_batch<COMPLEX>" if two else "step_tiled_batch<COMPLEX>",
                        "kernel_ms": kms,
                        "bytes_per_site": b_alg},
           "finite": ok, "batch_equals_sequential_lattice0": same}
    return out


def run_adjoint(args):
    """Differentiable workload: COMPLEX 128^3, 400 steps, float32, mask on.
    Loss sum((chi_T - chi_target)^2) with chi_target from a run at 1.05 kappa;
    gradients w.r.t. Psi_0 (two solitons, amplitude 5, width 3) and kappa.
    Cost ratio = time(forward with tape + adjoint backward) / time(plain
    forward of the same steps), for the keep-every-state tape and for the
    sqrt(T)-checkpoint tape."""
    import math
    import numpy as np
    import torch
    import lfm_hip as L
    from lfm_hip.adjoint import evolve_chi
    N = args.size or 128
    T = args.steps
    cfg = L.SimulationConfig(grid_size=N, field_level=L.FieldLevel.COMPLEX)
    rng = np.random.default_rng(42)
    pr = np.zeros((N,) * 3, np.float32)
    pi = np.zeros((N,) * 3, np.float32)
    for _ in range(2):
        r, i = L.fields.gaussian_soliton(N, tuple(N / 2 + rng.uniform(-0.2, 0.2, 3) * N), 5.0, 3.0,
                                         rng.uniform(0, 2 * np.pi))
        pr += r
        pi += i
    dev = torch.device("cuda")
    t = lambda a: torch.tensor(a, device=dev)
    with torch.no_grad():
        target = evolve_chi(t(pr), t(pi), 1.05 * cfg.kappa, cfg, T, stride=T)
    # plain forward: the Evolver, T steps
    ev = L.Evolver(cfg, backend="hip")
    for _ in range(3):
        # untimed runs from the same start capture the cycle graphs the timed
        # run replays
        ev.set_psi_real(pr)
        ev.set_psi_imag(pi)
        ev.evolve(T)
    ev.set_psi_real(pr)
    ev.set_psi_imag(pi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev.evolve(T)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    res = {}
    for name, stride in (("keep_all", 1), ("sqrtT", math.ceil(math.sqrt(T)))):
        for rep in range(2):    # first pass warms the allocator; the second is timed
            if rep:
                del pr_t, pi_t, k, chiT, loss
            pr_t, pi_t = t(pr).requires_grad_(True), t(pi).requires_grad_(True)
            k = torch.tensor(cfg.kappa, dtype=torch.float64, device=dev, requires_grad=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            chiT = evolve_chi(pr_t, pi_t, k, cfg, T, stride=stride)
            loss = ((chiT - target) ** 2).sum()
            loss.backward()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        res[name] = {"stride": stride, "wall_s": dt, "cost_ratio": dt / t_fwd,
                     "loss": float(loss.detach()), "grad_kappa": float(k.grad),
                     "grad_psi_norm": float(torch.sqrt((pr_t.grad ** 2).sum() + (pi_t.grad ** 2).sum())),
                     "finite": bool(torch.isfinite(pr_t.grad).all() and torch.isfinite(pi_t.grad).all())}
        del pr_t, pi_t, chiT, loss
        torch.cuda.empty_cache()
    best = min(r["cost_ratio"] for r in res.values())
    out = {"metric": METRIC, "value": best, "unit": "x forward (adjoint cost ratio)", "n_gpus": 1,
           "steps": T, "warmup": 3, "ms_per_step": t_fwd * 1e3 / T, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"adjoint COMPLEX {N}^3, {T} steps, mask on", "grid": N},
           "forward_s": t_fwd, "forward_GLUPS": N ** 3 * T / t_fwd / 1e9, "tapes": res}
    return out


def run_slab(args):
    """One GPU's share of an 8-way i-slab decomposition of the COLOR
    full_physics 512^3 workload: 64 interior planes plus one halo plane on
    each side (lfm_step_slab_f32), against the whole-lattice step on the same
    fields (lfm_step_color_f32 with the SCV term), both timed with HIP events
    over `steps` launches on the launch stream.  The share's interior equals
    the whole step's planes bit for bit (checked here on the last launch).
    The halo exchange between shares is not part of one GPU's work here."""
    import ctypes
    import torch
    import lfm_hip as L
    from lfm_hip import _abi
    lib = _abi.load()
    w = WORKLOADS["color512"]
    N, S, p0 = 512, 64, 192
    cfg = make_config(w, N)
    pr, pi = gpu_ics(w, N)
    g = torch.Generator(device="cuda").manual_seed(7)
    prp = pr + 1e-3 * torch.randn(pr.shape, generator=g, device="cuda")
    pip = pi + 1e-3 * torch.randn(pi.shape, generator=g, device="cuda")
    chi = 1e-2 * torch.randn((1, N, N, N), generator=g, device="cuda")     # chi - chi0
    chip = chi + 1e-4 * torch.randn((1, N, N, N), generator=g, device="cuda")
    scv = torch.rand((1, N, N, N), generator=g, device="cuda")
    mask = torch.from_numpy(L.fields.boundary_mask_host(N, cfg.boundary_fraction)).cuda().view(1, N, N, N)
    pr, pi = pr.view(3, N, N, N), pi.view(3, N, N, N)
    p = _abi.LfmParams(N, _abi.LFM_F_MASK | _abi.LFM_F_DAMP_PREV | _abi.LFM_F_CHI_DELTA | _abi.LFM_F_SA_SCV,
                       cfg.dt ** 2, cfg.kappa, cfg.lambda_self, cfg.chi0, cfg.e0_sq, cfg.epsilon_w,
                       cfg.kappa_c, cfg.epsilon_cc, cfg.kappa_string, cfg.kappa_tube,
                       2 if args.tile == "auto" else int(args.tile), args.chunk)
    st = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: t.data_ptr() if t is not None else None
    idx = torch.arange(p0 - 1, p0 + S + 1, device="cuda")
    cut = lambda t: t.index_select(1, idx).contiguous()
    sl = [cut(t) for t in (pr, prp, pi, pip, chi, chip, mask, scv)]
    o_w = [torch.empty_like(pr), torch.empty_like(pi), torch.empty_like(chi), torch.empty_like(pr)]
    o_s = [torch.empty_like(sl[0]), torch.empty_like(sl[0]), torch.empty_like(sl[4]),
           torch.empty_like(sl[0])]

    def whole():
        return lib.lfm_step_color_f32(ptr(pr), ptr(prp), ptr(pi), ptr(pip), ptr(chi), ptr(chip),
                                      ptr(mask), ptr(o_w[0]), None, ptr(o_w[1]), None, ptr(o_w[2]),
                                      None, ptr(scv), None, ptr(o_w[3]), ctypes.byref(p), st)

    def share():
        return lib.lfm_step_slab_f32(2, S + 2, *[ptr(t) for t in sl], *[ptr(t) for t in o_s],
                                     ctypes.byref(p), st)

    def timed(fn):
        for _ in range(args.warmup):
            _abi.check(fn(), "step")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    t_w, t_s = timed(whole), timed(share)
    same = all(bool(torch.equal(a[:, 1:-1], b.index_select(1, idx[1:-1])))
               for a, b in zip(o_s, o_w))
    sites = S * N * N
    achieved = w["b_alg"] * sites / (t_s * 1e-3) / 1e9
    return {"metric": METRIC, "value": sites / (t_s * 1e-3) / 1e9, "unit": "GLUPS", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_s,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"one 8-way i-slab share of COLOR full_physics {N}^3: {S} interior "
                                   f"+ 2 halo planes of {N}^2 (stencil kernel, SCV term given)",
                       "grid": N, "interior_planes": S, "halo_planes": 2},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "step_tiled<COLOR> (slab)", "kernel_ms": t_s,
                         "bytes_per_site": w["b_alg"]},
            "whole_lattice_kernel_ms": t_w, "ideal_share_ms": t_w * S / N,
            "share_efficiency": t_w * S / N / t_s, "interior_equals_whole_step": same}


def run_relax(args):
    """Eigenmode relaxation at the reference's default N = 64: the Gaussian
    seed (amplitude 14, sigma 4), one Poisson chi, then `steps` imaginary-time
    steps (each: stencil update with the max|E| fold and the renormalisation
    by amplitude / max, folded into the next update between checks) with a
    convergence check every 200 steps and an unreachable tolerance, so every
    step runs.  CPU baseline: the oracle's float32 NumPy relaxation step."""
    import numpy as np
    import torch
    from lfm_hip.eigenmode import _Relaxer, gaussian_seed
    from lfm_hip.evolver import get_backend
    N = args.size or 64
    steps = args.steps
    chi0 = 19.0
    dt = float(np.float32(0.8 / (chi0 * chi0 + 12.0)))
    R = _Relaxer(gaussian_seed(N, (N // 2,) * 3, 14.0, max(3.0, N / 16.0)), N, chi0, 1 / 63, 0.3,
                 get_backend("hip"))
    R.poisson()
    # an untimed run of the same length captures the six-step graphs the
    # timed run replays (same slot phases and buffer order)
    R.relax(14.0, dt, steps, 200, -1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    R.relax(14.0, dt, steps, 200, -1.0)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {"metric": METRIC, "value": N ** 3 * steps / t / 1e9, "unit": "GLUPS", "n_gpus": 1,
           "steps": steps, "warmup": steps, "ms_per_step": t * 1e3 / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"eigenmode relaxation {N}^3 (imaginary-time steps, "
                                  "max renormalisation, check every 200)", "grid": N},
           "finite": bool(torch.isfinite(R.E).all())}
    if not args.no_cpu_baseline:
        from oracle import lfm_oracle as O
        E = R.E.cpu().numpy().reshape(N, N, N).copy()
        chi2 = (R.chi2.cpu().numpy().reshape(N, N, N)).astype(np.float32)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0:
            En = E + np.float32(dt) * (O.laplacian_19pt(E) - chi2 * E)
            En *= np.float32(14.0 / float(np.max(np.abs(En))))
            E = En
            n += 1
        dt_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": N ** 3 * n / dt_cpu / 1e9, "unit": "GLUPS", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle NumPy float32 relaxation step {N}^3, {n} steps "
                                         f"in {dt_cpu:.1f} s (one thread)"}
    print(json.dumps(out), flush=True)


def tune_key(tile, chunk):
    """Autotuner entry label: tile/chunk, or "direct" for the one-thread-per-site kernel."""
    return "direct" if tile < 0 else f"{tile}/{chunk}"


def batch_tile(tuning):
    """LatticeBatch.autotune's timings; a negative tile is the two-step
    kernel at that tile shape."""
    if not tuning:
        return None
    best = min(tuning, key=tuning.get)
    label = lambda t, c: f"2step {-t}/{c}" if t < 0 else f"{t}/{c}"
    return {"best": list(best), "two_step": best[0] < 0,
            "ms": {label(t, c): round(v, 4) for (t, c), v in tuning.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 100; batch64 500, adjoint128 400, relax64 1000)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="color512",
                    choices=sorted(WORKLOADS) + ["copy", "batch64", "adjoint128", "relax64", "slab512"])
    ap.add_argument("--size", type=int, default=None, help="override the lattice size")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="COLOR line only: skip the batched and adjoint sub-results")
    ap.add_argument("--direct", action="store_true", help="use the untiled kernel")
    ap.add_argument("--tile", default="auto", help="tile shape 1..6, 0 default, or auto (tune)")
    ap.add_argument("--chunk", type=int, default=0, help="i planes per workgroup (0 auto)")
    ap.add_argument("--two-step", action="store_true",
                    help="REAL / COMPLEX and batch64: two steps per launch (lfm_step2_*, "
                         "lfm_batch_step2_f32)")
    ap.add_argument("--sa-transform", default="auto", choices=["auto", "fused", "separate"],
                    help="COLOR S_a: k-axis passes fused into the step kernel, or five separate passes")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"batch64": 500, "adjoint128": 400, "relax64": 1000}.get(args.workload, 100)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the leapfrog path here runs on a single MI355X "
              "(multi-GPU slab decomposition is design-only); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import lfm_hip as L

    if args.workload == "batch64":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "adjoint128":
        print(json.dumps(run_adjoint(args)), flush=True)
        return
    if args.workload == "relax64":
        run_relax(args)
        return
    if args.workload == "slab512":
        print(json.dumps(run_slab(args)), flush=True)
        return
    if args.workload == "copy":
        # Device-to-device copy of a known size: the HBM calibration for the
        # rocprof FETCH_SIZE / WRITE_SIZE counters and the measured copy rate.
        n = (args.size or 512) ** 3
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(args.warmup):
            y.copy_(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y.copy_(x)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        print(json.dumps({"workload": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                          "ms": ms, "GBps": 8 * n / (ms * 1e-3) / 1e9}), flush=True)
        return

    w = WORKLOADS[args.workload]
    N = args.size or w["N"]
    cfg = make_config(w, N)
    policy = "reference" if w["mask"] else "boundary_type"
    ev = L.Evolver(cfg, backend="hip", mask_policy=policy,
                   sa_transform=args.sa_transform if w["level"] == 2 else "auto",
                   two_step=args.two_step)
    pr, pi = gpu_ics(w, N)
    ev.set_psi_real(pr)
    if pi is not None:
        ev.set_psi_imag(pi)
    del pr, pi
    torch.cuda.synchronize()

    tuning = None
    if args.tile == "auto" and not args.direct:
        tuning = ev.autotune(chunks=(0, 1, 2, 4, 8, 16, 32, 64, 128) if args.chunk == 0 else (args.chunk,))
    elif not args.direct:
        ev.tile, ev.chunk = int(args.tile), args.chunk
    # REAL / COMPLEX advance two steps per launch (as Evolver.evolve does)
    pairs = getattr(ev, "two_step", False) and not args.direct and N <= 1024
    if args.direct:
        step = lambda: ev._step(direct=True)
    else:
        step = ev._step2 if pairs else ev._step
    per = 2 if pairs else 1
    # whole launches: with two steps per launch an odd warm-up would end on a
    # single step, off the two-step rotation's cycle, and the timed call
    # would then capture its graphs
    warm = max(per, -(-args.warmup // per) * per)
    if args.direct:
        for _ in range(warm):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    else:
        # Evolver.evolve as a user runs it: pairs of steps per launch with
        # --two-step, whole rotation cycles replayed from HIP graphs.
        # After the warm-up steps, the graphs the timed call replays from
        # the phase it starts at (one cycle, and the multi-cycle graph) are
        # captured; capture records the launches without running them, so
        # the state is unchanged and no capture falls in the timed call.
        ev.evolve(warm)
        if ev._cycle_graph() is not None:
            ev._cycle_graph(ev.GRAPH_CYCLES)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev.evolve(args.steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    # kernel-only timing: events on the launch stream around the stencil kernel
    ev.kernel_events, ev.kernel_event_steps = [], []
    nk = max(10, min(args.steps, 50)) // per
    for _ in range(nk):
        step()
    torch.cuda.synchronize()
    kms = (sum(a.elapsed_time(b) for a, b in ev.kernel_events) / sum(ev.kernel_event_steps))
    ev.kernel_events = None
    ok = bool(torch.isfinite(ev.native_chi()).all().item())

    n3 = N ** 3
    b_alg = w["b_alg2"] if pairs else w["b_alg"]
    achieved = b_alg * n3 / (kms * 1e-3) / 1e9
    # what the kernel itself must move: its extra outputs (COLOR: |Psi'|^2) in,
    # the S_a it no longer reads (COLOR on the SCV-term path) out
    scv_path = bool(getattr(ev, "_sa_scv", False)) and not args.direct
    b_kernel = b_alg + w.get("b_extra", 0) - (w.get("b_scv_saving", 0) if scv_path else 0)
    streams = w["streams2"] if pairs else (w["streams_scv"] if scv_path else w["streams"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    lvl = ['REAL', 'COMPLEX', 'COLOR'][w['level']]
    kname = (f"step_color_kfft<{N}>" if getattr(ev, "_fused", False)
             else f"step2_tiled<{lvl}>" if pairs else f"step_tiled<{lvl}>")
    traffic_cfg = None
    if os.path.exists(tfile) and args.size is None and not args.direct:
        # HBM bytes per launch from the committed counter runs: those of the
        # tile shape and chunk this run uses when they were profiled, else
        # those of the profiled configuration, named in traffic_tile_chunk
        rec = json.load(open(tfile))
        if rec.get("kernel") == kname:
            by = rec.get("by_tile_chunk") or {}
            traffic_cfg = f"{ev.tile}/{ev.chunk}"
            traffic = by.get(traffic_cfg)
            if traffic is None:
                traffic_cfg = rec.get("tile_chunk")
                traffic = rec.get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC, "value": n3 / (ms * 1e-3) / 1e9, "unit": "GLUPS", "n_gpus": 1,
        "steps": args.steps, "ms_per_step": ms,
        "warmup": warm,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "grid": N, "field_level": ["REAL", "COMPLEX", "COLOR"][w["level"]],
                   "preset": w["preset"], "mask": w["mask"], "kernel": "direct" if args.direct else "tiled"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_tile_chunk": traffic_cfg,
                     "kernel": kname,
                     "kernel_ms": kms, "bytes_per_site": b_alg,
                     "steps_per_launch": 2 if pairs else 1},
        "finite": ok,
        "tile": {"shape": ev.tile, "chunk": ev.chunk,
                 "tuned_ms": {tune_key(t, c): round(v, 4) for (t, c), v in tuning.items()} if tuning else None},
    }
    if b_kernel != b_alg:
        # the kernel's rate on the bytes it moves itself (COLOR: plus the
        # |Psi'|^2 it writes, minus the 8 B of S_a it does not read when the
        # transform hands it the SCV term)
        out["roofline"]["achieved_kernel_bytes"] = b_kernel * n3 / (kms * 1e-3) / 1e9
        out["roofline"]["kernel_bytes_per_site"] = b_kernel
        out["roofline"]["sa_input"] = "SCV term (4 B)" if scv_path else "S_a fields (12 B)"
    if "b_step" in w:
        # the whole step (stencil kernel + the five S_a transform passes) on
        # its algorithmic bytes per site
        ach = w["b_step"] * n3 / (ms * 1e-3) / 1e9
        out["step_roofline"] = {"bytes_per_site": w["b_step"], "achieved": ach, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "ms_per_step": ms}
    if args.size is None and not args.direct:
        # the same access pattern with nothing but streaming: the achievable rate
        ceil = stream_ceiling(*streams, n3)
        out["roofline"]["stream_ceiling"] = ceil
        out["roofline"]["frac_of_stream_ceiling"] = (b_kernel * n3 / (kms * 1e-3) / 1e9) / ceil
    del ev
    torch.cuda.empty_cache()
    if args.workload == "color512" and args.size is None and not args.direct and not args.quick:
        # the metric's other two clauses: batched sweep and adjoint speedups
        # against plain runs, at their own configurations
        out["batched"] = _summary_batch(run_batch(argparse.Namespace(
            steps=500, warmup=args.warmup, tile="auto", chunk=0, two_step=False)))
        torch.cuda.empty_cache()
        out["adjoint"] = _summary_adjoint(run_adjoint(argparse.Namespace(steps=400, size=None)))
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline and args.size is None:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


def _summary_batch(r):
    keep = ("value", "unit", "steps", "ms_per_step", "config", "batched", "sequential_graphs",
            "sequential_eager", "speedup_vs_sequential", "roofline", "finite",
            "batch_equals_sequential_lattice0")
    return {k: r[k] for k in keep if k in r}


def _summary_adjoint(r):
    out = {k: r[k] for k in ("value", "unit", "steps", "ms_per_step", "config", "forward_s",
                             "forward_GLUPS") if k in r}
    out["cost_ratio"] = {name: t["cost_ratio"] for name, t in r["tapes"].items()}
    out["tapes"] = r["tapes"]
    return out


if __name__ == "__main__":
    main()
