"""The allocation-free energy reduction, lfm_energy_f64_noalloc, against the
entry point it varies, lfm_energy_f64 (which takes a caller-allocated
partials workspace), through the C ABI.

Both run the same kernels in the same summation order, so the four FP64
sums must be equal bit for bit.  Three calls in a row on decreasing sizes
(COLOR 48^3, COMPLEX 20^3, REAL 16^3 with the damped prev): each later call
finds the previous call's partial records in the shared static scratch and
must not be affected by them.  The sums are also checked against the FP64
oracle's energy_components on the same float32 arrays (1e-12 relative).
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_energy_noalloc_equals_energy(gpu):
    import torch
    from lfm_hip import _abi
    lib = _abi.load()
    st = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(8)
    dt, c, chi0 = 0.02, 1.0, 19.0
    for N, ncomp, has_pi, damp in ((48, 3, True, 1), (20, 1, True, 0), (16, 1, False, 1)):
        log(f"energy noalloc vs workspace: N={N} ncomp={ncomp} imag={has_pi} damp={damp}")
        n3 = N ** 3
        f = lambda n: rng.standard_normal(n).astype(np.float32)
        pr, prp = f(ncomp * n3), f(ncomp * n3)
        pi, pip = (f(ncomp * n3), f(ncomp * n3)) if has_pi else (None, None)
        chi = (0.05 * rng.standard_normal(n3)).astype(np.float32)     # deviation storage
        mask = O.boundary_mask(N, 0.3)
        d = lambda a: None if a is None else torch.from_numpy(a).cuda()
        t = {k: d(v) for k, v in dict(pr=pr, prp=prp, pi=pi, pip=pip, chi=chi, mask=mask).items()}
        ptr = lambda k: None if t[k] is None else t[k].data_ptr()
        args = (ptr("pr"), ptr("prp"), ptr("pi"), ptr("pip"), ptr("chi"), chi0, ptr("mask"), damp,
                ncomp, N, dt, c)
        part = torch.empty(int(lib.lfm_reduce_partials_len(N)), dtype=torch.float64, device="cuda")
        out_a = torch.full((4,), -1.0, dtype=torch.float64, device="cuda")
        out_b = torch.full((4,), -2.0, dtype=torch.float64, device="cuda")
        _abi.check(lib.lfm_energy_f64(*args, part.data_ptr(), out_a.data_ptr(), st), "lfm_energy_f64")
        _abi.check(lib.lfm_energy_f64_noalloc(*args, out_b.data_ptr(), st), "lfm_energy_f64_noalloc")
        torch.cuda.synchronize()
        a, b = out_a.cpu().numpy(), out_b.cpu().numpy()
        np.testing.assert_array_equal(b, a)
        sh = (ncomp, N, N, N) if ncomp > 1 else (N, N, N)
        af = (1.0 - mask) if damp else np.ones_like(mask)
        prev = lambda x: None if x is None else (x.reshape(ncomp, n3) * af).astype(np.float32).reshape(sh)
        T, G, V = O.energy_components(pr.reshape(sh), prev(prp), (chi0 + chi.astype(np.float64)).reshape(N, N, N),
                                      dt, c, None if pi is None else pi.reshape(sh), prev(pip))
        want = [T.sum(), G.sum(), V.sum()]
        for k in range(3):
            assert abs(b[k] - want[k]) <= 1e-12 * abs(want[k]), (k, b[k], want[k])
        log(f"  T {b[0]:.6e} G {b[1]:.6e} V {b[2]:.6e}: equal to the workspace entry point")
