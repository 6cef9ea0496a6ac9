"""The committed FP64 oracle energy traces (golden/energy_trace_64.npz) still
describe the oracle and the initial fields: the trace script's seeded ICs
give the recorded checksums, and the oracle recomputes the early
checkpoints (COMPLEX to step 100, COLOR to step 20) to 1e-13.  The GPU
trace tests compare against the same file."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_oracle_traces as T  # noqa: E402


@pytest.mark.parametrize("level,name,upto", [(1, "complex", 100), (2, "color", 20)])
def test_trace_fixture_matches_oracle(level, name, upto):
    g = np.load(os.path.join(HERE, "golden", "energy_trace_64.npz"), allow_pickle=False)
    np.testing.assert_array_equal(T.ics_checksum(level), g[f"{name}_ics_checksum"])
    steps = [int(s) for s in g["steps"] if s <= upto]
    got = T.trace(level, steps, log=lambda *a: None)
    want = g[f"{name}_energy"][:len(steps)]
    np.testing.assert_allclose(got, want, rtol=1e-13)
    # the trace is not flat: the energy moved between the first checkpoints
    assert abs(want[-1] - want[0]) > 1e-9 * abs(want[0])
