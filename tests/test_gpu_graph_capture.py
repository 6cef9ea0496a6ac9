"""Graph captures and Python's garbage collector.

An Evolver whose captured graphs are reachable only through a reference cycle
(a Simulation and an experiment object that point at each other, say) is
freed by the cyclic collector.  Destroying a CUDAGraph while a stream is
capturing is not permitted, so a collection during another capture would
abort the process; every capture therefore holds the collector off while it
records, and dead cycles go at the collector's next run.
"""
import gc
import sys
import weakref

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_dead_evolver_graphs_freed_outside_captures(gpu):
    import torch
    import lfm_hip as L
    N = 32
    log("dead Evolver in a reference cycle, then a new capture")
    cfg = L.gravity_em(N)
    rng = np.random.default_rng(1)
    pr = (0.1 * rng.standard_normal((N, N, N))).astype(np.float32)
    ev = L.Evolver(cfg, backend="hip")
    ev.set_psi_real(pr)
    ev.evolve(40)
    torch.cuda.synchronize()
    assert any(v is not None for v in ev._graphs.values()), "no graph was captured"
    ev.cycle = ev                    # reachable only through itself once dropped
    ref = weakref.ref(ev)
    del ev
    assert ref() is not None
    assert gc.isenabled()
    ev2 = L.Evolver(cfg, backend="hip")
    ev2.set_psi_real(pr)
    # captures with the collector forced to run often: it runs only between
    # captures, never inside one
    old = gc.get_threshold()
    gc.set_threshold(1, 1, 1)
    try:
        ev2.evolve(40)
    finally:
        gc.set_threshold(*old)
    torch.cuda.synchronize()
    assert gc.isenabled()            # the collector is back on after the capture
    assert any(v is not None for v in ev2._graphs.values())
    gc.collect()
    assert ref() is None
