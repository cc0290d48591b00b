"""Pins of the torch FP64 tape reference (tests/tape_euler.py) used to check
the hand-written adjoint: its forward equals the C oracle's step on 1D and 2D
grids with each boundary kind, and its autograd gradient matches central
finite differences at a generic (differentiable) state."""
from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests import tape_euler as T  # noqa: E402


def _state(nd, rng):
    if nd == 1:
        n = 40
        U = np.zeros((n, 3))
        U[:, 0] = 1.0 + 0.2 * np.sin(np.linspace(0, 2 * np.pi, n, endpoint=False))
        U[:, 1] = 2.5 + 0.1 * rng.random(n)
        U[:, 2] = 0.3 + 0.05 * rng.random(n)
        return U
    ny, nx = 12, 14
    U = np.zeros((ny, nx, 4))
    U[..., 0] = 1.0 + 0.1 * rng.random((ny, nx))
    U[..., 1] = 2.5 + 0.1 * rng.random((ny, nx))
    U[..., 2] = 0.3 + 0.05 * rng.random((ny, nx))
    U[..., 3] = -0.2 + 0.05 * rng.random((ny, nx))
    return U


def _to_t(U):
    return [torch.tensor(np.ascontiguousarray(np.moveaxis(U, -1, 0)[k])) for k in range(U.shape[-1])]


def _to_np(Ut):
    return np.moveaxis(np.stack([u.detach().numpy() for u in Ut]), 0, -1)


@pytest.mark.parametrize("nd,bc", [(1, (2,)), (1, (1,)), (1, (0,)), (2, (2, 2)), (2, (0, 1)),
                                   (2, (1, 2))])
def test_tape_forward_matches_oracle(oracle, nd, bc):
    rng = np.random.default_rng(1)
    U = _state(nd, rng)
    s = oracle.scheme(gamma=1.4, bc=tuple(bc) + (0,) * (3 - len(bc)))
    dh = 1.0 / U.shape[-2]
    Uo, _, it, _ = oracle.simulate(U, s, 0.9, dh, max_iters=3)
    Ut = _to_t(U)
    g = torch.tensor(1.4, dtype=torch.float64)
    for _ in range(3):
        Ut = T.step(Ut, g, 0.9, dh, nd, bc)
    np.testing.assert_allclose(_to_np(Ut), Uo, rtol=1e-13, atol=1e-15)


def test_tape_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    U = _state(2, rng)
    dh = 1.0 / 14
    lam = rng.standard_normal(U.shape)

    def loss(Ut, g):
        for _ in range(2):
            Ut = T.step(Ut, g, 0.9, dh, 2, (2, 2))
        return sum((u * torch.tensor(np.ascontiguousarray(lam[..., k]))).sum()
                   for k, u in enumerate(Ut))
    Ut = [u.requires_grad_() for u in _to_t(U)]
    g = torch.tensor(1.4, dtype=torch.float64, requires_grad=True)
    L = loss(Ut, g)
    grads = torch.autograd.grad(L, Ut + [g])
    d = rng.standard_normal(U.shape)
    dg = 0.3
    ad = sum(float((gr * torch.tensor(np.ascontiguousarray(d[..., k]))).sum())
             for k, gr in enumerate(grads[:-1])) + float(grads[-1]) * dg
    # small enough that no branch (limiter ratio, wavespeed argmax) switches
    # between the two evaluations of this noisy state
    eps = 1e-7
    with torch.no_grad():
        lp = float(loss(_to_t(U + eps * d), torch.tensor(1.4 + eps * dg, dtype=torch.float64)))
        lm = float(loss(_to_t(U - eps * d), torch.tensor(1.4 - eps * dg, dtype=torch.float64)))
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - ad) <= 1e-6 * abs(fd)
