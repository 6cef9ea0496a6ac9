"""GPU tests of the slab-share step (lfm_step_slab_f32): one GPU's share of a
lattice sharded along i, its two halo planes copied from the neighbouring
shares.  The interior planes must equal the same planes of a whole-lattice
step bit for bit (the same per-site arithmetic), at every level, with the
mask, damped prev, chi deviation storage and, at COLOR, the SCV term and the
|Psi'|^2 output; at N = 64 for several slab positions and thicknesses, and at
the real shape of an 8-way share of the 512^3 benchmark lattice (64 interior
planes + 2 halo planes of 512^2).  A 2-way decomposition stepped with halo
copies between the shares reproduces the whole-lattice run over 12 steps.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _params(level, N):
    from lfm_hip import _abi
    p = _abi.LfmParams()
    p.N = N
    p.flags = _abi.LFM_F_MASK | _abi.LFM_F_DAMP_PREV | _abi.LFM_F_CHI_DELTA
    p.dt2 = np.float32(0.02 ** 2)
    p.kappa, p.lambda_self, p.chi0 = 1 / 63, 4 / 31, 19.0
    p.epsilon_w = 0.1
    if level == 2:
        p.kappa_c, p.epsilon_cc, p.kappa_string, p.kappa_tube = 1 / 189, 2 / 17, 1 / 189, 30 / 63
        p.flags |= _abi.LFM_F_SA_SCV
    p.tile = 2
    return p


def _case(level, N, p0, S, seed):
    import torch
    import lfm_hip as L
    from lfm_hip import _abi
    from lfm_hip.fields import boundary_mask_host
    lib = _abi.load()
    C = 3 if level == 2 else 1
    g = torch.Generator(device="cuda").manual_seed(seed)
    full = lambda c, s: s * torch.randn((c, N, N, N), generator=g, device="cuda")
    pr, prp = full(C, 0.3), full(C, 0.3)
    pi, pip = (full(C, 0.3), full(C, 0.3)) if level else (None, None)
    chi, chip = full(1, 0.01), full(1, 0.01)
    mask = torch.from_numpy(boundary_mask_host(N, 0.3)).cuda().view(1, N, N, N)
    scv = full(1, 1.0).abs() if level == 2 else None
    p = _params(level, N)
    st = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: t.data_ptr() if t is not None else None
    # whole lattice
    o_pr, o_chi = torch.empty_like(pr), torch.empty_like(chi)
    o_pi = torch.empty_like(pi) if level else None
    o_sq = torch.empty_like(pr) if level == 2 else None
    if level == 0:
        rc = lib.lfm_step_real_f32(ptr(pr), ptr(prp), ptr(chi), ptr(chip), ptr(mask), ptr(o_pr),
                                   None, ptr(o_chi), None, ctypes.byref(p), st)
    elif level == 1:
        rc = lib.lfm_step_complex_f32(ptr(pr), ptr(prp), ptr(pi), ptr(pip), ptr(chi), ptr(chip),
                                      ptr(mask), ptr(o_pr), None, ptr(o_pi), None, ptr(o_chi),
                                      None, ctypes.byref(p), st)
    else:
        rc = lib.lfm_step_color_f32(ptr(pr), ptr(prp), ptr(pi), ptr(pip), ptr(chi), ptr(chip),
                                    ptr(mask), ptr(o_pr), None, ptr(o_pi), None, ptr(o_chi), None,
                                    ptr(scv), None, ptr(o_sq), ctypes.byref(p), st)
    _abi.check(rc, "whole-lattice step")
    # the share: planes p0-1 .. p0+S of every field (periodic in i)
    idx = torch.arange(p0 - 1, p0 + S + 1, device="cuda") % N
    cut = lambda t: t.index_select(1, idx).contiguous() if t is not None else None
    s_in = [cut(t) for t in (pr, prp, pi, pip, chi, chip, mask, scv)]
    # outputs pre-filled, so halo planes left alone are seen as left alone
    s_pr = torch.full_like(s_in[0], 7.0)
    s_pi = torch.full_like(s_in[0], 7.0) if level else None
    s_chi = torch.full_like(s_in[4], 7.0)
    s_sq = torch.full_like(s_in[0], 7.0) if level == 2 else None
    rc = lib.lfm_step_slab_f32(level, S + 2, *[ptr(t) for t in s_in], ptr(s_pr), ptr(s_pi),
                               ptr(s_chi), ptr(s_sq), ctypes.byref(p), st)
    _abi.check(rc, "slab step")
    torch.cuda.synchronize()
    pairs = [("psi_r", s_pr, o_pr), ("chi", s_chi, o_chi)]
    if level:
        pairs.append(("psi_i", s_pi, o_pi))
    if level == 2:
        pairs.append(("psi_sq", s_sq, o_sq))
    for name, s, o in pairs:
        want = o.index_select(1, idx[1:-1])
        assert torch.equal(s[:, 1:-1], want), name
        assert bool((s[:, 0] == 7.0).all()) and bool((s[:, -1] == 7.0).all()), f"{name} halo"
    log(f"  level {level} N {N} planes {p0}..{p0 + S - 1}: interior equal to the whole-lattice "
        f"step, halo planes untouched")


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("p0,S", [(13, 20), (0, 9), (60, 7), (1, 62)])
def test_slab_equals_whole_lattice_64(gpu, level, p0, S):
    log(f"slab share vs whole lattice, level {level}, N 64, planes {p0}..{p0 + S - 1}")
    _case(level, 64, p0, S, seed=17 + level)


@pytest.mark.parametrize("level", [1, 2])
def test_slab_equals_whole_lattice_512_share(gpu, level):
    import torch
    log(f"slab share vs whole lattice, level {level}, an 8-way share of 512^3")
    _case(level, 512, 192, 64, seed=5)
    torch.cuda.empty_cache()


@pytest.mark.parametrize("level", [0, 1, 2])
def test_two_shares_with_halo_exchange_equal_whole_lattice(gpu, level):
    """A 2-way i-slab decomposition run on one GPU: each share keeps its own
    three-slot rotation of interior + halo planes, and before every step the
    current Psi and chi halo planes are copied from the neighbouring share
    (the exchange RCCL would do between ranks; prev fields are read only at
    interior sites and need none).  After 12 steps both shares' interiors
    equal the whole-lattice run bit for bit (COLOR without the S_a term,
    which needs the global transform)."""
    import torch
    from lfm_hip import _abi
    lib = _abi.load()
    N, T = 48, 12
    C = 3 if level == 2 else 1
    log(f"two shares with halo exchange vs whole lattice, level {level}, N {N}, {T} steps")
    p = _params(level, N)
    if level == 2:
        p.kappa_tube = 0.0
        p.flags &= ~_abi.LFM_F_SA_SCV
    g = torch.Generator(device="cuda").manual_seed(23 + level)
    from lfm_hip.fields import boundary_mask_host
    mask = torch.from_numpy(boundary_mask_host(N, 0.3)).cuda().view(1, N, N, N)
    psi0 = [0.3 * torch.randn((C, N, N, N), generator=g, device="cuda") for _ in range(2)]
    chi0 = 0.01 * torch.randn((1, N, N, N), generator=g, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: t.data_ptr() if t is not None else None
    fields = lambda: [torch.empty((C, N, N, N), device="cuda") for _ in range(2 if level else 1)] + \
        [torch.empty((1, N, N, N), device="cuda")]

    def step_whole(cur, prev, out, flags):
        p.flags = flags
        if level == 0:
            rc = lib.lfm_step_real_f32(ptr(cur[0]), ptr(prev[0]), ptr(cur[1]), ptr(prev[1]), ptr(mask),
                                       ptr(out[0]), None, ptr(out[1]), None, ctypes.byref(p), st)
        elif level == 1:
            rc = lib.lfm_step_complex_f32(ptr(cur[0]), ptr(prev[0]), ptr(cur[1]), ptr(prev[1]),
                                          ptr(cur[2]), ptr(prev[2]), ptr(mask), ptr(out[0]), None,
                                          ptr(out[1]), None, ptr(out[2]), None, ctypes.byref(p), st)
        else:
            rc = lib.lfm_step_color_f32(ptr(cur[0]), ptr(prev[0]), ptr(cur[1]), ptr(prev[1]),
                                        ptr(cur[2]), ptr(prev[2]), ptr(mask), ptr(out[0]), None,
                                        ptr(out[1]), None, ptr(out[2]), None, None, None, None,
                                        ctypes.byref(p), st)
        _abi.check(rc, "whole step")

    base = p.flags & ~_abi.LFM_F_DAMP_PREV
    damp = _abi.LFM_F_DAMP_PREV
    # whole lattice: three slots, the first step reads prev undamped
    init = [psi0[0], psi0[1]][:(2 if level else 1)] + [chi0]
    slots = [[t.clone() for t in init], [t.clone() for t in init], fields()]
    cur, prev, nxt = 0, 1, 2
    for s in range(T):
        step_whole(slots[cur], slots[prev], slots[nxt], base | (damp if s else 0))
        cur, prev, nxt = nxt, cur, prev
    whole = slots[cur]
    # two shares: interior planes [0, 20) and [20, 48), each stored with one
    # halo plane per side
    bounds = [(0, 20), (20, N)]
    shares = []
    for a, b in bounds:
        idx = torch.arange(a - 1, b + 1, device="cuda") % N
        cut = lambda t: t.index_select(1, idx).contiguous()
        sl = [[cut(t) for t in init], [cut(t) for t in init],
              [torch.empty_like(cut(t)) for t in init]]
        shares.append({"idx": idx, "slots": sl, "mask": cut(mask), "planes": b - a + 2})
    cur, prev, nxt = 0, 1, 2
    for s in range(T):
        # halo exchange of the current fields: plane 0 of a share is the last
        # interior plane of the share before it, its last plane the first
        # interior plane of the share after it
        for k, sh in enumerate(shares):
            lo, hi = shares[k - 1], shares[(k + 1) % 2]
            for f, t in enumerate(sh["slots"][cur]):
                t[:, 0].copy_(lo["slots"][cur][f][:, -2])
                t[:, -1].copy_(hi["slots"][cur][f][:, 1])
        for sh in shares:
            c, pv, o = sh["slots"][cur], sh["slots"][prev], sh["slots"][nxt]
            pi_c, pi_p, pi_o = (c[1], pv[1], o[1]) if level else (None, None, None)
            p.flags = base | (damp if s else 0)
            rc = lib.lfm_step_slab_f32(level, sh["planes"], ptr(c[0]), ptr(pv[0]), ptr(pi_c),
                                       ptr(pi_p), ptr(c[-1]), ptr(pv[-1]), ptr(sh["mask"]), None,
                                       ptr(o[0]), ptr(pi_o), ptr(o[-1]), None, ctypes.byref(p), st)
            _abi.check(rc, "slab step")
        cur, prev, nxt = nxt, cur, prev
    torch.cuda.synchronize()
    for (a, b), sh in zip(bounds, shares):
        for f, (t_sh, t_wh) in enumerate(zip(sh["slots"][cur], whole)):
            assert torch.equal(t_sh[:, 1:-1], t_wh[:, a:b]), (a, b, f)
    log(f"  both shares equal the whole lattice after {T} steps")
