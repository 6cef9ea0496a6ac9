"""Write golden input/output vectors for the reference's ``step_color``.

Run in the build container only (the reference tree is not present on the
GPU host):  python tests/golden/make_golden.py

Why an import is used here: the reference's NumPy ``step_color``
(lfm/core/backends/numpy_backend.py:207-382) mixes float32 scratch with
float64 f_c and a float32-forward / float64-inverse Helmholtz FFT; no closed
form pins the full combination of all coupling terms, so its own output on
small seeded inputs is recorded here as data.  Every other oracle function
is pinned by closed forms and invariants (tests/test_oracle_pins.py).

Output: tests/golden/step_color_N8.npz (inputs + outputs after 1 and 3
steps, two parameter sets).
"""
import os
import sys

import numpy as np

REF = "/root/reference"


def main():
    sys.path.insert(0, REF)
    from lfm.core.backends.numpy_backend import NumpyBackend  # noqa: E402

    be = NumpyBackend()
    N = 8
    T = N ** 3
    rng = np.random.default_rng(1234)
    x = np.arange(N, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    pr = np.zeros((3, T), np.float32)
    pi = np.zeros((3, T), np.float32)
    for a, c in enumerate([(3.0, 4.0, 4.0), (4.5, 3.5, 4.0), (4.0, 4.5, 3.2)]):
        env = 2.5 * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * 1.5 ** 2))
        ph = rng.uniform(0, 2 * np.pi)
        pr[a] = (env * np.cos(ph)).ravel()
        pi[a] = (env * np.sin(ph)).ravel()
    pr += (rng.standard_normal((3, T)) * 1e-2).astype(np.float32)
    pi += (rng.standard_normal((3, T)) * 1e-2).astype(np.float32)
    prp = pr + (rng.standard_normal((3, T)) * 1e-3).astype(np.float32)
    pip = pi + (rng.standard_normal((3, T)) * 1e-3).astype(np.float32)
    chi = (19.0 - 0.3 * rng.random(T)).astype(np.float32)
    chip = (chi + (rng.standard_normal(T) * 1e-3)).astype(np.float32)
    mask = be.create_boundary_mask(N, 0.3)

    sets = {
        # full_physics preset constants (lfm/config_presets.py:59-79)
        "full": dict(kappa=1 / 63, lambda_self=4 / 31, chi0=19.0, e0_sq=0.0, epsilon_w=0.1,
                     kappa_c=1 / 189, epsilon_cc=2 / 17, kappa_string=1 / 189,
                     kappa_tube=30 / 63, sa_gamma=0.1, sa_d=4.9),
        # colour-variance only, as tests/test_gpu_parity.py:327 exercises
        "kc": dict(kappa=1 / 63, lambda_self=0.0, chi0=19.0, e0_sq=0.0, epsilon_w=0.1,
                   kappa_c=1 / 189, epsilon_cc=0.0, kappa_string=0.0, kappa_tube=0.0,
                   sa_gamma=0.1, sa_d=4.9),
    }
    out = dict(N=np.int32(N), dt=np.float64(0.02), pr=pr.ravel(), pi=pi.ravel(),
               prp=prp.ravel(), pip=pip.ravel(), chi=chi, chip=chip, mask=mask)
    for name, p in sets.items():
        A = dict(r=pr.ravel().copy(), rp=prp.ravel().copy(), i=pi.ravel().copy(),
                 ip=pip.ravel().copy(), c=chi.copy(), cp=chip.copy(),
                 sa=np.zeros(3 * T, np.float32))
        B = {k: np.zeros_like(v) for k, v in A.items()}
        for step in range(1, 4):
            be.step_color(A["r"], A["rp"], A["i"], A["ip"], A["c"], A["cp"], mask,
                          B["r"], B["rp"], B["i"], B["ip"], B["c"], B["cp"], N, 0.02 ** 2,
                          p["kappa"], p["lambda_self"], p["chi0"], p["e0_sq"], p["epsilon_w"],
                          p["kappa_c"], p["epsilon_cc"], kappa_string=p["kappa_string"],
                          kappa_tube=p["kappa_tube"], sa_fields_in=A["sa"],
                          sa_fields_out=B["sa"], sa_gamma=p["sa_gamma"], sa_d=p["sa_d"],
                          dt=0.02)
            A, B = B, A
            if step in (1, 3):
                for k in ("r", "rp", "i", "ip", "c", "cp", "sa"):
                    out[f"{name}_s{step}_{k}"] = A[k].copy()
        for k, v in p.items():
            out[f"{name}_param_{k}"] = np.float64(v)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "step_color_N8.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
