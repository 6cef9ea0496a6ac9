"""Energy drift at the COMPLEX and COLOR levels against the FP64 oracle.

The 64^3 COMPLEX gravity_em configuration (eight solitons, periodic,
mask-free) and the 64^3 COLOR full_physics configuration (three colour
blobs, cos^2 mask, S_a every step) run 500 graph-replayed steps on the GPU;
total energy with the true leapfrog prev (FP64 sums on the GPU) is compared
every 100 steps with the FP64 oracle's trace, which tests/golden/
make_oracle_traces.py computed from the same seeded fields (the oracle takes
~0.2 s per COLOR step at 64^3, too slow for a GPU test; the CPU test
tests/test_oracle_traces.py re-derives the early checkpoints from the oracle).
Every checkpoint within 1e-5 relative, and the drift E(t) - E(0) within
1e-5 of E(0) of the oracle's drift.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_oracle_traces as T  # noqa: E402

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("level,name", [(1, "complex"), (2, "color")])
def test_energy_trace_matches_fp64_oracle(gpu, level, name):
    import torch
    import lfm_hip as L
    g = np.load(os.path.join(HERE, "golden", "energy_trace_64.npz"), allow_pickle=False)
    want = g[f"{name}_energy"]
    steps = [int(s) for s in g["steps"]]
    log(f"{name} 64^3 energy trace, {steps[-1]} steps, checkpoints {steps}")
    cfg = T.trace_config(level)
    pr, pi = T.trace_ics(level)
    np.testing.assert_array_equal(T.ics_checksum(level), g[f"{name}_ics_checksum"])
    ev = L.Evolver(cfg, backend="hip", mask_policy="boundary_type" if level == 1 else "reference")
    assert ev.has_mask == (level == 2)
    ev.set_psi_real(pr)
    ev.set_psi_imag(pi)
    got = []
    for t in steps:
        ev.evolve(t - ev.step)
        torch.cuda.synchronize()
        got.append(ev.total_energy())
    got = np.array(got)
    rel = np.abs(got - want) / np.abs(want)
    drift_diff = np.abs((got - got[0]) - (want - want[0])) / abs(want[0])
    log(f"  E(0) {want[0]:.6e}; drift {(want[-1] - want[0]) / want[0]:+.3e} (oracle) "
        f"{(got[-1] - got[0]) / got[0]:+.3e} (GPU); worst rel {rel.max():.2e}, "
        f"worst drift diff {drift_diff.max():.2e}")
    assert rel.max() <= 1e-5 and drift_diff.max() <= 1e-5
