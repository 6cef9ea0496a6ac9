"""Write golden input/output vectors from the reference's own code.

Run in the build container only (the reference tree is not present on the
GPU host):  python tests/golden/make_golden.py

Why an import is used here, per fixture:

* ``step_color_N8.npz`` -- the reference's NumPy ``step_color``
  (lfm/core/backends/numpy_backend.py:207-382) mixes float32 scratch with
  float64 f_c and a float32-forward / float64-inverse Helmholtz FFT; no
  closed form pins the full combination of all coupling terms, so its output
  on small seeded inputs is recorded as data.
* ``step_real_complex.npz`` -- ``step_real`` / ``step_complex``
  (numpy_backend.py:80-205) with the cos^2 mask on and the damped prev the
  reference writes (psi_prev_out = (1-m) psi), three A/B steps, N = 8 and 16,
  in float32 and in float64 (both methods only use the input dtype and
  Python scalars, so the float64 run is a true FP64 vector).  The closed
  forms of the oracle's pins run with m = 0; these vectors pin the masked,
  damped-prev branches.  The masks of N = 16 (and SHA-256 digests of those
  of N = 64 and 128) come from ``create_boundary_mask``
  (numpy_backend.py:49-69).
* ``checkpoint_{complex,color}_N16.npz`` -- files written by the reference's
  ``Simulation.save_checkpoint`` (lfm/simulation.py:1474-1514): byte
  compatibility of the checkpoint format can only be shown on files the
  reference wrote.  ``checkpoint_expect.npz`` holds, for each, the
  reference's ``metrics()`` at save time and the fields after
  ``load_checkpoint`` (simulation.py:1516-1565) plus ``run(20)`` on its CPU
  backend.

Every other oracle function is pinned by closed forms and invariants
(tests/test_oracle_pins.py).
"""
import json
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _save(name, out):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def make_step_color(be):
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
    _save("step_color_N8.npz", out)


def make_step_real_complex(be):
    """Masked three-step A/B runs of step_real and step_complex."""
    out = {}
    # gravity_em constants with the Mexican hat and a non-zero E0^2, so every
    # term of both updates is exercised
    p = dict(kappa=1 / 63, lambda_self=4 / 31, chi0=19.0, e0_sq=0.05, epsilon_w=0.1)
    out.update({f"param_{k}": np.float64(v) for k, v in p.items()})
    out["dt2"] = np.float64(0.02 ** 2)
    for N in (8, 16):
        T = N ** 3
        rng = np.random.default_rng(100 + N)
        c = N / 2
        x = np.arange(N, dtype=np.float64)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        env = 3.0 * np.exp(-((X - c + 0.7) ** 2 + (Y - c) ** 2 + (Z - c - 0.4) ** 2) / (2 * 2.0 ** 2))
        ph = 0.9 * (X - c) / N + rng.uniform(0, 2 * np.pi)
        pr = (env * np.cos(ph) + rng.standard_normal((N, N, N)) * 1e-2).astype(np.float32).ravel()
        pi = (env * np.sin(ph) + rng.standard_normal((N, N, N)) * 1e-2).astype(np.float32).ravel()
        prp = (pr + rng.standard_normal(T) * 1e-3).astype(np.float32)
        pip = (pi + rng.standard_normal(T) * 1e-3).astype(np.float32)
        chi = (19.0 - 0.5 * env.ravel() / 3.0 + rng.standard_normal(T) * 1e-2).astype(np.float32)
        chip = (chi + rng.standard_normal(T) * 1e-3).astype(np.float32)
        mask = be.create_boundary_mask(N, 0.3)
        tag = f"N{N}"
        for k, v in dict(pr=pr, pi=pi, prp=prp, pip=pip, chi=chi, chip=chip, mask=mask).items():
            out[f"{tag}_{k}"] = v
        # N = 8: every output of steps 1 and 3 in both dtypes; N = 16 (larger,
        # incompressible noise): the float64 fields after step 3
        keep = (lambda step, k: True) if N == 8 else (lambda step, k: step == 3 and k in "ric")
        for dt_name, dt in ((("f32", np.float32), ("f64", np.float64)) if N == 8
                            else (("f64", np.float64),)):
            m = mask.astype(dt)
            # REAL: E = pr
            A = dict(r=pr.astype(dt), rp=prp.astype(dt), c=chi.astype(dt), cp=chip.astype(dt))
            B = {k: np.zeros_like(v) for k, v in A.items()}
            for step in range(1, 4):
                be.step_real(A["r"], A["rp"], A["c"], A["cp"], m, B["r"], B["rp"], B["c"], B["cp"],
                             N, 0.02 ** 2, p["kappa"], p["lambda_self"], p["chi0"], p["e0_sq"])
                A, B = B, A
                if step in (1, 3):
                    for k, v in A.items():
                        if keep(step, k):
                            out[f"{tag}_real_{dt_name}_s{step}_{k}"] = v.copy()
            # COMPLEX
            A = dict(r=pr.astype(dt), rp=prp.astype(dt), i=pi.astype(dt), ip=pip.astype(dt),
                     c=chi.astype(dt), cp=chip.astype(dt))
            B = {k: np.zeros_like(v) for k, v in A.items()}
            for step in range(1, 4):
                be.step_complex(A["r"], A["rp"], A["i"], A["ip"], A["c"], A["cp"], m,
                                B["r"], B["rp"], B["i"], B["ip"], B["c"], B["cp"], N, 0.02 ** 2,
                                p["kappa"], p["lambda_self"], p["chi0"], p["e0_sq"],
                                p["epsilon_w"])
                A, B = B, A
                if step in (1, 3):
                    for k, v in A.items():
                        if keep(step, k):
                            out[f"{tag}_complex_{dt_name}_s{step}_{k}"] = v.copy()
    import hashlib
    out["mask_N16"] = be.create_boundary_mask(16, 0.3)
    out["mask_N16_bf05"] = be.create_boundary_mask(16, 0.5)
    for N in (64, 128):
        m = np.ascontiguousarray(be.create_boundary_mask(N, 0.3), dtype=np.float32)
        out[f"mask_N{N}_sha256"] = np.array(hashlib.sha256(m.tobytes()).hexdigest())
    _save("step_real_complex.npz", out)


def make_checkpoints():
    """Reference-written checkpoints and what the reference computes from them."""
    import warnings

    from lfm import Simulation
    from lfm.config_presets import full_physics, gravity_em

    expect = {}
    for name, cfg in (("complex", gravity_em(16, report_interval=5)),
                      ("color", full_physics(16, report_interval=5))):
        sim = Simulation(cfg, backend="cpu")
        c = 8.0
        if name == "color":
            sim.place_solitons([(c - 2, c, c), (c + 2, c + 1, c), (c, c - 2, c + 1)],
                               amplitude=3.0, sigma=2.0, phases=[0.3, 1.7, 4.1],
                               colors=[0, 1, 2])
        else:
            sim.place_soliton((c, c - 1, c + 0.5), amplitude=4.0, sigma=2.5, phase=0.6)
        sim.equilibrate()
        sim.run(12)                      # history at steps 5 and 10; prev snapshot of step 10
        path = os.path.join(HERE, f"checkpoint_{name}_N16.npz")
        sim.save_checkpoint(path)
        print("wrote", path, os.path.getsize(path), "bytes")
        m = sim.metrics()
        expect[f"{name}_metrics_json"] = np.array(json.dumps(m))
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sim2 = Simulation.load_checkpoint(path, backend="cpu")
        expect[f"{name}_loaded_metrics_json"] = np.array(json.dumps(sim2.metrics()))
        sim2.run(20, record_metrics=False)
        expect[f"{name}_resume20_step"] = np.int64(sim2.step)
        expect[f"{name}_resume20_psi_real"] = sim2.psi_real.astype(np.float32)
        expect[f"{name}_resume20_psi_imag"] = sim2.psi_imag.astype(np.float32)
        expect[f"{name}_resume20_chi"] = sim2.chi.astype(np.float32)
    _save("checkpoint_expect.npz", expect)


def main():
    sys.path.insert(0, REF)
    from lfm.core.backends.numpy_backend import NumpyBackend  # noqa: E402

    be = NumpyBackend()
    make_step_color(be)
    make_step_real_complex(be)
    make_checkpoints()


if __name__ == "__main__":
    main()
