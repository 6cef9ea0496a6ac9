"""GPU eigenmode relaxation (lfm_hip.eigenmode.relax_eigenmode) against the
FP64 oracle's restatement of lfm/particles/solver.py's relaxation loop.

* Fixed step counts (tolerance unreachable): E, chi and the eigenvalue after
  three Poisson / relaxation cycles; E within 1e-5 of the oracle relative to
  its norm, chi relative L2 <= 1e-5, eigenvalue to 1e-5 relative.
* A converging run: same converged flag, same step count, same tolerances.
"""
import sys

import numpy as np
import pytest

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(N, amp, sigma, cycles, spc, check, tol, kappa=1 / 63, chi0=19.0, bf=0.3):
    import lfm_hip as L
    from lfm_hip.eigenmode import relax_eigenmode
    sol = relax_eigenmode(N=N, amplitude=amp, sigma=sigma, kappa=kappa, chi0=chi0,
                          max_cycles=cycles, steps_per_cycle=spc, check_interval=check,
                          tolerance=tol, boundary_fraction=bf, backend="hip")
    seed = O.relax_seed(N, (N // 2, N // 2, N // 2), amp, sigma)
    dt = float(np.float32(0.8 / (chi0 * chi0 + 12.0)))
    E, chi, conv, total, cyc = O.relax_eigenmode(seed, N, chi0, kappa, amp, dt, cycles, spc, check,
                                                 tol, bf)
    return sol, (E, chi, conv, total, O.relax_eigenvalue(E, chi))


@pytest.mark.parametrize("N,amp", [(32, 6.0), (48, 10.0)])
def test_relax_fixed_steps_matches_oracle(gpu, N, amp):
    log(f"eigenmode relaxation N={N} amplitude {amp}, 3 cycles x 150 steps")
    sol, (E, chi, conv, total, omega) = _run(N, amp, 3.0, 3, 150, 50, 1e-30)
    assert sol.cycles == total == 450 and sol.converged == conv == False  # noqa: E712
    e = float(np.linalg.norm(sol.psi_r.astype(np.float64) - E) / np.linalg.norm(E))
    ec = O.rel_l2(sol.chi, chi)
    eo = abs(sol.eigenvalue - omega) / omega
    log(f"  E rel {e:.2e}  chi rel {ec:.2e}  omega {sol.eigenvalue:.6f} vs {omega:.6f}  "
        f"chi_min {sol.chi_min:.4f}")
    assert e <= 1e-5 and ec <= 1e-5 and eo <= 1e-5
    assert abs(float(np.abs(sol.psi_r).max()) - amp) <= 1e-5 * amp


def test_relax_converging_run_matches_oracle(gpu):
    log("eigenmode relaxation N=32 to convergence")
    sol, (E, chi, conv, total, omega) = _run(32, 6.0, 3.0, 20, 400, 100, 1e-4)
    log(f"  GPU converged={sol.converged} steps={sol.cycles}; oracle converged={conv} "
        f"steps={total}; omega {sol.eigenvalue:.6f} vs {omega:.6f}")
    assert sol.converged == conv and sol.cycles == total
    e = float(np.linalg.norm(sol.psi_r.astype(np.float64) - E) / np.linalg.norm(E))
    assert e <= 1e-5 and abs(sol.eigenvalue - omega) / omega <= 1e-5


@pytest.mark.parametrize("steps,check", [(37, 10), (200, 50), (13, 200)])
def test_fused_relax_steps_equal_separate_rescale(gpu, steps, check):
    """Steps with the previous rescale folded into the update (eager and
    replayed from six-step graphs) give the same bits as one update launch
    and one rescale launch per step, including the convergence-check value."""
    import torch
    from lfm_hip import _abi
    from lfm_hip.backend import stream_handle
    from lfm_hip.eigenmode import _Relaxer, gaussian_seed
    from lfm_hip.evolver import get_backend
    N, amp = 40, 9.0
    dt = float(np.float32(0.8 / (19.0 ** 2 + 12.0)))
    log(f"fused vs separate relaxation steps={steps} check={check}")
    seed = gaussian_seed(N, (N // 2 - 1, N // 2, N // 2 + 2), amp, 3.5)
    runs = []
    for mode in ("graphs", "eager", "separate"):
        R = _Relaxer(seed, N, 19.0, 1 / 63, 0.3, get_backend("hip"))
        R.poisson()
        if mode != "separate":
            R.relax(amp, dt, steps, check, -1.0, graphs=(mode == "graphs"))
            diff = R.slot_max(3) if steps >= check else None
        else:
            lib, st = _abi.load(), stream_handle()
            R.slots.zero_()
            diff = None
            for n in range(1, steps + 1):
                k = n % 3
                _abi.check(lib.lfm_relax_step_f32(R.E.data_ptr(), R.chi2.data_ptr(),
                                                  R.E_alt.data_ptr(), R.slot(k), None,
                                                  R.slot((k + 1) % 3), float(amp), N, float(dt),
                                                  st), "step")
                R.E, R.E_alt = R.E_alt, R.E
                chk = n % check == 0
                if chk:
                    R.slots[3 * R.sw:].zero_()
                _abi.check(lib.lfm_relax_scale_f32(R.E.data_ptr(),
                                                   R.E_alt.data_ptr() if chk else None, R.slot(k),
                                                   None, R.slot(3) if chk else None, float(amp),
                                                   R.n3, st), "scale")
                if chk:
                    diff = R.slot_max(3)
        torch.cuda.synchronize()
        runs.append((R.E.cpu().numpy(), diff))
    for E, d in runs[:2]:
        np.testing.assert_array_equal(E, runs[2][0])
        assert d == runs[2][1]
