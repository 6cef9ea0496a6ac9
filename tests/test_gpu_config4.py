"""The batched-sweep configuration as written: 64 COMPLEX lattices with sizes
drawn from {32, 48, 64, 96, 128}^3 (most sites sit in the 96^3 and 128^3
lattices), kappa = (1/63) U(0.5, 2), one centred soliton with amplitude
U(2, 8) and width U(2, 4), N(0, 1e-3) noise on both parts, cos^2 mask on,
float32 -- the seeded recipe bench.py's batch workload times
(bench.batch64_recipe).  24 steps through LatticeBatch (one launch per step
for all lattices, HIP-graph replay):

* every lattice equals its own plain Evolver run bit for bit;
* the first 32^3 and the first 128^3 lattice are within 1e-5 relative L2
  of the FP64 oracle (lfm/sweep.py:14-82 runs each value through the same
  step the oracle restates).
"""
import os
import sys

import numpy as np
import pytest

from oracle import lfm_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu
STEPS = 24


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_batch64_configuration(gpu):
    import torch
    import lfm_hip as L
    sys.path.insert(0, os.path.dirname(HERE))
    import bench
    cfgs, ics = bench.batch64_recipe()
    sizes = [c.grid_size for c in cfgs]
    log(f"batch64: {len(cfgs)} lattices, sizes {sorted(set(sizes))}, "
        f"{sum(n ** 3 for n in sizes)} sites, {STEPS} steps")
    assert len(cfgs) == 64 and set(sizes) == {32, 48, 64, 96, 128}
    batch = L.LatticeBatch(cfgs, backend="hip")
    for b, (r, i) in enumerate(ics):
        batch.set_psi_real(b, r)
        batch.set_psi_imag(b, i)
    batch.evolve(STEPS)
    torch.cuda.synchronize()
    for b, (c, (r, i)) in enumerate(zip(cfgs, ics)):
        ev = L.Evolver(c, backend="hip")
        ev.set_psi_real(r)
        ev.set_psi_imag(i)
        ev.evolve(STEPS)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(batch.get_psi_real(b), ev.get_psi_real(), err_msg=f"psi_r {b}")
        np.testing.assert_array_equal(batch.get_psi_imag(b), ev.get_psi_imag(), err_msg=f"psi_i {b}")
        np.testing.assert_array_equal(batch.get_chi(b), ev.get_chi(), err_msg=f"chi {b}")
        del ev
    log("  every lattice equals its Evolver run bit for bit")
    for N in (32, 128):
        b = sizes.index(N)
        c = cfgs[b]
        p = dict(dt2=c.dt ** 2, kappa=c.kappa, lambda_self=c.lambda_self, chi0=c.chi0,
                 e0_sq=c.e0_sq, epsilon_w=c.epsilon_w)
        st = O.OracleState(1, N, ics[b][0], ics[b][1], np.full(N ** 3, c.chi0),
                           O.boundary_mask(N, c.boundary_fraction), p)
        st.step(STEPS)
        cat = lambda *a: np.concatenate([np.asarray(x, np.float64).ravel() for x in a])
        e_psi = O.rel_l2(cat(batch.get_psi_real(b), batch.get_psi_imag(b)), cat(st.psi_r, st.psi_i))
        e_chi = O.rel_l2(batch.get_chi(b), st.chi)
        log(f"  lattice {b} ({N}^3, kappa {c.kappa:.5f}) vs FP64 oracle: psi {e_psi:.2e} chi {e_chi:.2e}")
        assert e_psi <= 1e-5 and e_chi <= 1e-5
