"""Pins for the FP64 autograd reference of the COMPLEX leapfrog gradient.

The torch forward must reproduce the NumPy oracle's trajectory, and its
autograd gradients must match central finite differences of the NumPy
oracle (FP64, relative step 1e-4) for kappa and for a random direction in
Psi_0.
"""
import numpy as np
import torch

from oracle import adjoint_ref as A
from oracle import lfm_oracle as O

N, STEPS, DT = 10, 12, 0.02


def setup():
    rng = np.random.default_rng(5)
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    env = 2.5 * np.exp(-((X - 4.6) ** 2 + (Y - 5.2) ** 2 + (Z - 4.9) ** 2) / (2 * 1.6 ** 2))
    pr = env * np.cos(0.7) + 0.05 * rng.standard_normal((N,) * 3)
    pi = env * np.sin(0.7) + 0.05 * rng.standard_normal((N,) * 3)
    mask = O.boundary_mask(N, 0.3).astype(np.float64).reshape((N,) * 3)
    chi = np.full((N,) * 3, 19.0)
    return pr, pi, chi, mask, rng


def oracle_chiT(pr, pi, chi, mask, kappa, lam):
    p = dict(dt2=DT * DT, kappa=kappa, lambda_self=lam, chi0=19.0, e0_sq=0.0, epsilon_w=0.1)
    st = O.OracleState(1, N, pr, pi, chi, mask.ravel(), p)
    st.step(STEPS)
    return st.chi.reshape((N,) * 3)


def test_torch_forward_matches_numpy_oracle():
    pr, pi, chi, mask, _ = setup()
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    _, _, chiT = A.forward_complex(t(pr), t(pi), t(chi), 1 / 63, t(mask), STEPS, lam=4 / 31)
    np.testing.assert_allclose(chiT.numpy(), oracle_chiT(pr, pi, chi, mask, 1 / 63, 4 / 31),
                               rtol=0, atol=1e-12)


def test_autograd_matches_central_differences():
    pr, pi, chi, mask, rng = setup()
    lam = 4 / 31
    kappa = 1 / 63
    target = oracle_chiT(pr, pi, chi, mask, 1.05 * kappa, lam)

    def loss(pr_, pi_, k):
        return float(((oracle_chiT(pr_, pi_, chi, mask, k, lam) - target) ** 2).sum())

    L0, gr, gi, gk = A.grads(pr, pi, chi, kappa, mask, STEPS, target, lam=lam)
    assert abs(L0 - loss(pr, pi, kappa)) <= 1e-12 * L0
    h = 1e-4 * kappa
    fd_k = (loss(pr, pi, kappa + h) - loss(pr, pi, kappa - h)) / (2 * h)
    assert abs(fd_k - gk) <= 1e-6 * abs(gk), (fd_k, gk)
    dr, di = rng.standard_normal((N,) * 3), rng.standard_normal((N,) * 3)
    eps = 1e-4
    fd_d = (loss(pr + eps * dr, pi + eps * di, kappa) - loss(pr - eps * dr, pi - eps * di, kappa)) / (2 * eps)
    ad_d = float((gr * dr).sum() + (gi * di).sum())
    assert abs(fd_d - ad_d) <= 1e-6 * abs(ad_d), (fd_d, ad_d)
