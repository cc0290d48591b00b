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


@pytest.mark.parametrize("flux,form", [("hllc", "cons"), ("force", "cons"), ("hllc", "prim"),
                                       ("force", "prim")])
@pytest.mark.parametrize("nd,bc", [(1, (2,)), (1, (1,)), (1, (0,)), (2, (2, 2)), (2, (0, 1)),
                                   (2, (1, 2))])
def test_tape_forward_matches_oracle(oracle, nd, bc, flux, form):
    rng = np.random.default_rng(1)
    U = _state(nd, rng)
    s = oracle.scheme(gamma=1.4, bc=tuple(bc) + (0,) * (3 - len(bc)),
                      flux=oracle.FLUX_FORCE if flux == "force" else oracle.FLUX_HLLC,
                      form=oracle.FORM_PRIM if form == "prim" else oracle.FORM_CONS)
    dh = 1.0 / U.shape[-2]
    Uo, _, it, _ = oracle.simulate(U, s, 0.9, dh, max_iters=3)
    Ut = _to_t(U)
    g = torch.tensor(1.4, dtype=torch.float64)
    for _ in range(3):
        Ut = T.step(Ut, g, 0.9, dh, nd, bc, flux_method=flux, form=form)
    np.testing.assert_allclose(_to_np(Ut), Uo, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("flux,form", [("hllc", "cons"), ("force", "cons"), ("hllc", "prim")])
def test_tape_gradient_matches_finite_differences(flux, form):
    rng = np.random.default_rng(2)
    U = _state(2, rng)
    dh = 1.0 / 14
    lam = rng.standard_normal(U.shape)

    def loss(Ut, g):
        for _ in range(2):
            Ut = T.step(Ut, g, 0.9, dh, 2, (2, 2), flux_method=flux, form=form)
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


def _shock_state():
    """A 2D state with smeared compressive pressure jumps along x and y whose
    flattening factors lie strictly inside (0, 1) at many cells."""
    ny, nx = 10, 18
    x = np.arange(nx)[None, :] + np.zeros((ny, 1))
    y = np.arange(ny)[:, None] + np.zeros((1, nx))
    p = 1.5 + 0.5 * np.tanh((x - 8.3) / 0.8) + 0.4 * np.tanh((y - 4.6) / 0.7)
    rho = 1.0 + 0.3 * np.tanh((x - 8.3) / 0.8)
    u = -0.4 * np.tanh((x - 8.3) / 0.8) + 0.02 * np.sin(y)
    v = -0.3 * np.tanh((y - 4.6) / 0.7)
    U = np.zeros((ny, nx, 4))
    U[..., 0] = rho
    U[..., 2] = rho * u
    U[..., 3] = rho * v
    U[..., 1] = p / 0.4 + 0.5 * rho * (u * u + v * v)
    # small noise: no two cells tie for the maximum wavespeed (a tie is a
    # kink of dt that finite differences would straddle)
    U *= 1.0 + 1e-3 * np.random.default_rng(8).standard_normal(U.shape)
    return U


def test_tape_chi_gradient_matches_finite_differences():
    """Where the flattening factor is strictly between 0 and 1 it carries a
    gradient: the tape's gradient through it matches central finite
    differences of its own forward (which equals the oracle's step)."""
    rng = np.random.default_rng(5)
    U = _shock_state()
    dh = 1.0 / 18
    Ut0 = _to_t(U)
    chi = T._chi_min(T.pad(Ut0, 2, (0, 0)), torch.tensor(1.4, dtype=torch.float64), 2)
    inside = ((chi > 0) & (chi < 1)).sum().item()
    assert inside >= 4, inside                      # the case this test is for
    lam = rng.standard_normal(U.shape)

    def loss(Ut, g):
        Ut = T.step(Ut, g, 0.9, dh, 2, (0, 0))
        return sum((u * torch.tensor(np.ascontiguousarray(lam[..., k]))).sum()
                   for k, u in enumerate(Ut))
    Ut = [u.requires_grad_() for u in _to_t(U)]
    g = torch.tensor(1.4, dtype=torch.float64, requires_grad=True)
    grads = torch.autograd.grad(loss(Ut, g), Ut + [g])
    for trial in range(3):
        d = rng.standard_normal(U.shape)
        dg = 0.2
        ad = sum(float((gr * torch.tensor(np.ascontiguousarray(d[..., k]))).sum())
                 for k, gr in enumerate(grads[:-1])) + float(grads[-1]) * dg
        # limiter and flattening switches lie within about 1e-7 of this state
        # along some directions: a smaller step stays on one side of them
        eps = 1e-8
        with torch.no_grad():
            lp = float(loss(_to_t(U + eps * d), torch.tensor(1.4 + eps * dg, dtype=torch.float64)))
            lm = float(loss(_to_t(U - eps * d), torch.tensor(1.4 - eps * dg, dtype=torch.float64)))
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - ad) <= 1e-6 * abs(fd), (trial, fd, ad)


def test_tape_chi_gradient_is_not_the_frozen_one():
    """Holding chi constant changes the gradient of the shock state (the
    previous test therefore checks the chi term)."""
    U = _shock_state()
    dh = 1.0 / 18
    lam = np.random.default_rng(6).standard_normal(U.shape)

    def grad(frozen):
        Ut = [u.requires_grad_() for u in _to_t(U)]
        g = torch.tensor(1.4, dtype=torch.float64)
        orig = T._chi_min
        if frozen:
            T._chi_min = lambda P, gamma, nd: orig([c.detach() for c in P], gamma, nd).detach()
        try:
            out = T.step(Ut, g, 0.9, dh, 2, (0, 0))
        finally:
            T._chi_min = orig
        L = sum((u * torch.tensor(np.ascontiguousarray(lam[..., k]))).sum() for k, u in enumerate(out))
        return np.stack([gr.numpy() for gr in torch.autograd.grad(L, Ut)])
    a, b = grad(False), grad(True)
    assert np.abs(a - b).max() > 1e-3 * np.abs(a).max()
