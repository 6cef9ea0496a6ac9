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
