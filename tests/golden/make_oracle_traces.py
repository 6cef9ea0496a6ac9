"""Write FP64 oracle energy traces for the 64^3 COMPLEX and COLOR
configurations (tests/golden/energy_trace_64.npz).

The oracle is too slow to run 500 COLOR steps inside a GPU test (about
0.2 s per float64 step at 64^3), so its trace is computed here once and
committed; tests/test_oracle_traces.py recomputes the early checkpoints on
the CPU to show the file still matches the oracle, and
tests/test_gpu_energy.py compares the GPU's trace with it.

Run:  python tests/golden/make_oracle_traces.py

Configurations (initial fields built on the host by ``trace_ics``):

* COMPLEX: gravity_em, periodic (mask-free), eight Gaussian solitons at
  uniform positions in [0, N)^3, amplitudes U(2, 6), widths U(2, 4), phases
  U(0, 2 pi); chi = chi0.
* COLOR: full_physics (cos^2 mask, S_a every step), three blobs, one colour
  each, centres uniform in the ball of radius 0.35 N/2 about the centre,
  amplitude 4, width 3, random phases; chi = chi0.

Energy: total_energy (T + G + V, lfm/analysis/energy.py:116-135) with the
true leapfrog prev (the damped prev the reference's step writes), in FP64.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for _p in (ROOT, os.path.join(ROOT, "lfm-physics_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

N = 64
STEPS = 500
CHECKPOINTS = (0, 20, 100, 200, 300, 400, 500)
OUT = os.path.join(HERE, "energy_trace_64.npz")


def trace_config(level):
    import lfm_hip as L
    if level == 1:
        return L.gravity_em(N, boundary_type=L.BoundaryType.PERIODIC)
    return L.full_physics(N)


def trace_ics(level, n=N, seed=42):
    """Seeded float32 initial fields of the trace configurations."""
    import lfm_hip as L
    rng = np.random.default_rng(seed)
    if level == 1:
        pos = rng.uniform(0, n, (8, 3))
        amp = rng.uniform(2, 6, 8)
        sig = rng.uniform(2, 4, 8)
        ph = rng.uniform(0, 2 * np.pi, 8)
        pr = np.zeros((n,) * 3, np.float32)
        pi = np.zeros((n,) * 3, np.float32)
        for s in range(8):
            r, i = L.fields.gaussian_soliton(n, tuple(pos[s]), float(amp[s]), float(sig[s]),
                                             float(ph[s]))
            pr += r
            pi += i
        return pr, pi
    c = n / 2.0
    ph = rng.uniform(0, 2 * np.pi, 3)
    pos = []
    for _ in range(3):
        while True:
            d = rng.uniform(-1, 1, 3)
            if d @ d <= 1:
                break
        pos.append(tuple(c + 0.35 * c * d))
    return L.fields.place_solitons(n, pos, 4.0, 3.0, phases=ph, colors=[0, 1, 2])


def oracle_params(cfg):
    return dict(dt2=cfg.dt ** 2, kappa=cfg.kappa, lambda_self=cfg.lambda_self, chi0=cfg.chi0,
                e0_sq=cfg.e0_sq, epsilon_w=cfg.epsilon_w, kappa_c=cfg.kappa_c,
                epsilon_cc=cfg.epsilon_cc, kappa_string=cfg.kappa_string,
                kappa_tube=cfg.kappa_tube, sa_gamma=cfg.sa_gamma, sa_d=cfg.sa_d)


def oracle_state(level):
    from oracle import lfm_oracle as O
    cfg = trace_config(level)
    pr, pi = trace_ics(level)
    mask = np.zeros(N ** 3) if level == 1 else O.boundary_mask(N, cfg.boundary_fraction)
    return O.OracleState(level, N, pr, pi, np.full(N ** 3, cfg.chi0), mask, oracle_params(cfg)), cfg


def oracle_energy(st, cfg):
    from oracle import lfm_oracle as O
    sh = (3, N, N, N) if st.level == 2 else (N, N, N)
    return O.total_energy(st.psi_r.reshape(sh), st.psi_r_prev.reshape(sh),
                          st.chi.reshape(N, N, N), cfg.dt, cfg.c,
                          st.psi_i.reshape(sh), st.psi_i_prev.reshape(sh))


def trace(level, checkpoints=CHECKPOINTS, log=print):
    st, cfg = oracle_state(level)
    out = []
    for t in checkpoints:
        st.step(t - st.steps)
        out.append(oracle_energy(st, cfg))
        log(f"level {level} step {t}: E = {out[-1]:.12e}")
    return np.array(out)


def ics_checksum(level):
    pr, pi = trace_ics(level)
    a = np.concatenate([np.asarray(pr, np.float64).ravel(), np.asarray(pi, np.float64).ravel()])
    return np.array([a.sum(), (a * a).sum()])


def main():
    out = {"steps": np.array(CHECKPOINTS, np.int64), "N": np.int64(N)}
    for level, name in ((1, "complex"), (2, "color")):
        out[f"{name}_energy"] = trace(level)
        out[f"{name}_ics_checksum"] = ics_checksum(level)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
