"""The hand-written adjoint on the GPU.

* One step's vector-Jacobian product (fl_split_step_vjp) against the torch
  FP64 tape of the same step, 1D and 2D, each boundary kind, HLLC and FORCE,
  conservative- and primitive-form limiting.
* The checkpointed multi-step adjoint (fl_adj_forward / fl_adj_backward)
  against the tape over 25 steps (HLLC and FORCE), and bit-for-bit
  invariance under the checkpoint interval.
* The forward run equals fl_simulate step for step.
* Central finite differences on 64 smooth random directions of
  (rho, rho u, rho v, E, gamma) on the 1024^2, 200-step workload with a
  smooth background flow (where the scheme is differentiable), relative
  error <= 1e-6.  On the literal static workload (u = v = 0, p = 1) every face
  sits on a branch point of HLLC (star speed exactly 0) and of the limiter
  (flat E and momentum), so only the rho and gamma directions, which keep
  the state on its static branch, are checked there.
"""
from __future__ import annotations

import numpy as np
import pytest

from fluidity_amd import native as N
from fluidity_amd import workloads as W
from tests import tape_euler as T

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def A():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from fluidity_amd import adjoint
    return adjoint


@pytest.fixture(scope="module")
def E():
    from fluidity_amd import euler
    return euler


def _soa(U):
    return np.ascontiguousarray(np.moveaxis(U, -1, 0))


def _dev(U):
    return torch.from_numpy(_soa(U)).to("cuda")


def _generic_state(nd, ny=20, nx=24, seed=3):
    rng = np.random.default_rng(seed)
    if nd == 1:
        x = (np.arange(nx) + 0.5) / nx
        W_ = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * x), 1.0 + 0.1 * np.cos(2 * np.pi * x),
                       0.4 + 0.1 * np.sin(4 * np.pi * x)], -1)
        return W.prim_to_cons(W_ * (1 + 0.01 * rng.random(W_.shape)), 1.4)
    pr = W.adjoint_problem(nx, flow=True)
    U = pr.U[:ny].copy()
    return U * (1 + 0.01 * rng.random(U.shape))


def _tape_vjp(U, lam, nd, bc, cfl, dh, steps=1, flux="hllc", form="cons"):
    Ut = [torch.tensor(np.ascontiguousarray(np.moveaxis(U, -1, 0)[k]), requires_grad=True)
          for k in range(U.shape[-1])]
    g = torch.tensor(1.4, dtype=torch.float64, requires_grad=True)
    X = Ut
    for _ in range(steps):
        X = T.step(X, g, cfl, dh, nd, bc, flux_method=flux, form=form)
    L = sum((x * torch.tensor(np.ascontiguousarray(lam[k]))).sum() for k, x in enumerate(X))
    grads = torch.autograd.grad(L, Ut + [g])
    return np.stack([gr.numpy() for gr in grads[:-1]]), float(grads[-1])


def test_binding_imported_before_torch_initialises(A):
    """The binding imported first, then torch's first GPU call, then a
    context: the library's runtime is torch's, so it finds the device (it
    found none while the library brought a second runtime)."""
    import subprocess
    import sys
    code = ("import fluidity_amd.native as N\n"
            "import torch\n"
            "torch.zeros(1, device='cuda')\n"
            "ctx = N.Context(0)\n"
            "print('context ok')\n")
    root = __import__("os").path.dirname(__import__("os").path.dirname(__file__))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=root,
                         timeout=120)
    assert out.returncode == 0 and "context ok" in out.stdout, out.stderr[-2000:]


def _scheme(E, bc, flux, form="cons"):
    return E.Scheme(gamma=1.4, bc=tuple(bc) + (0,) * (3 - len(bc)),
                    flux=N.FL_FLUX_FORCE if flux == "force" else N.FL_FLUX_HLLC,
                    form=N.FL_FORM_PRIMITIVE if form == "prim" else N.FL_FORM_CONSERVATIVE)


@pytest.mark.parametrize("nd,bc,flux,state", [(2, (2, 2), "hllc", "generic"),
                                              (2, (0, 1), "force", "generic"),
                                              (1, (1,), "hllc", "generic"),
                                              (2, (1, 2), "hllc", "shock"),
                                              (1, (0,), "force", "shock")])
def test_step_vjp_primitive_form_matches_tape(A, E, nd, bc, flux, state):
    """Primitive-form limiting (the limiter's ratio from differences of
    rho, p and vel): one step's VJP equals the tape's, on a generic state and
    on a shock state where the flattening factor lies inside (0, 1)."""
    U = _generic_state(nd) if state == "generic" else _shock_state(nd)
    rng = np.random.default_rng(10)
    lam = rng.standard_normal(_soa(U).shape)
    dh = 1.0 / U.shape[-2]
    adj = A.Adjoint(_scheme(E, bc, flux, "prim"), 0.9, dh)
    lam_g, dg = adj.step_vjp(_dev(U), torch.from_numpy(lam).to("cuda"))
    ref, dg_ref = _tape_vjp(U, lam, nd, bc, 0.9, dh, flux=flux, form="prim")
    got = lam_g.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-11 * scale, np.abs(got - ref).max() / scale
    assert abs(dg - dg_ref) <= 1e-11 * max(abs(dg_ref), scale)


@pytest.mark.parametrize("flux", ["hllc", "force"])
@pytest.mark.parametrize("nd,bc", [(2, (2, 2)), (2, (0, 1)), (2, (1, 2)), (1, (2,)), (1, (1,))])
def test_step_vjp_matches_tape(A, E, nd, bc, flux):
    U = _generic_state(nd)
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(_soa(U).shape)
    dh = 1.0 / U.shape[-2]
    adj = A.Adjoint(_scheme(E, bc, flux), 0.9, dh)
    lam_g, dg = adj.step_vjp(_dev(U), torch.from_numpy(lam).to("cuda"))
    ref, dg_ref = _tape_vjp(U, lam, nd, bc, 0.9, dh, flux=flux)
    got = lam_g.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-11 * scale
    assert abs(dg - dg_ref) <= 1e-11 * max(abs(dg_ref), scale)


def _shock_state(nd):
    """Smeared compressive pressure jumps (along x, and y in 2D) whose
    flattening factors lie strictly between 0 and 1 at many cells, with small
    noise so that no two cells tie for the maximum wavespeed."""
    ny, nx = (12, 20) if nd == 2 else (1, 40)
    x = np.arange(nx)[None, :] + np.zeros((ny, 1))
    y = np.arange(ny)[:, None] + np.zeros((1, nx))
    xc = 9.3 if nd == 2 else 19.3
    p = 1.5 + 0.5 * np.tanh((x - xc) / 0.8)
    rho = 1.0 + 0.3 * np.tanh((x - xc) / 0.8)
    u = -0.4 * np.tanh((x - xc) / 0.8)
    v = np.zeros_like(u)
    if nd == 2:
        p = p + 0.4 * np.tanh((y - 5.6) / 0.7)
        u = u + 0.02 * np.sin(y)
        v = -0.3 * np.tanh((y - 5.6) / 0.7)
    U = np.zeros((ny, nx, 2 + nd))
    U[..., 0] = rho
    U[..., 2] = rho * u
    if nd == 2:
        U[..., 3] = rho * v
    U[..., 1] = p / 0.4 + 0.5 * rho * (u * u + v * v)
    U *= 1.0 + 1e-3 * np.random.default_rng(8).standard_normal(U.shape)
    return U[0] if nd == 1 else U


@pytest.mark.parametrize("nd,bc,flux", [(2, (0, 0), "hllc"), (2, (1, 2), "hllc"),
                                        (2, (2, 1), "hllc"), (1, (0,), "hllc"),
                                        (1, (1,), "hllc"), (2, (1, 2), "force"),
                                        (1, (0,), "force")])
def test_step_vjp_through_flattening_matches_tape(A, E, nd, bc, flux):
    """Where the flattening factor chi lies strictly between 0 and 1 the
    adjoint differentiates through it (the factor the minimum selected, its
    pressure ratio, dp/dU and dp/dgamma of the cells it reads, through ghost
    cells too): one step's VJP equals the torch tape, whose chi gradient is
    pinned by finite differences on the CPU."""
    U = _shock_state(nd)
    Ut = [torch.tensor(np.ascontiguousarray(np.moveaxis(U, -1, 0)[k])) for k in range(U.shape[-1])]
    chi = T._chi_min(T.pad(Ut, nd, bc), torch.tensor(1.4, dtype=torch.float64), nd)
    assert int(((chi > 0) & (chi < 1)).sum()) >= 2     # the case this test is for
    rng = np.random.default_rng(9)
    lam = rng.standard_normal(_soa(U).shape)
    dh = 1.0 / U.shape[-2]
    adj = A.Adjoint(_scheme(E, bc, flux), 0.9, dh)
    lam_g, dg = adj.step_vjp(_dev(U), torch.from_numpy(lam).to("cuda"))
    ref, dg_ref = _tape_vjp(U, lam, nd, bc, 0.9, dh, flux=flux)
    got = lam_g.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-11 * scale, np.abs(got - ref).max() / scale
    assert abs(dg - dg_ref) <= 1e-11 * max(abs(dg_ref), scale)


def _adjoint_run(A, E, U, target, dh, steps, every, bc=(2, 2), flux="hllc"):
    adj = A.Adjoint(_scheme(E, bc, flux), 0.9, dh)
    Ud = _dev(U)
    run = adj.forward(Ud, steps, every)
    tgt = torch.from_numpy(target).to("cuda")
    grad = torch.zeros_like(Ud)
    grad[0] = 2.0 * dh * dh * (Ud[0] - tgt)
    loss = float((dh * dh * (Ud[0] - tgt) ** 2).sum())
    dg = adj.backward(grad, run)
    return loss, grad.cpu().numpy(), dg, Ud.cpu().numpy(), run


@pytest.mark.parametrize("flux", ["hllc", "force"])
def test_multistep_adjoint_matches_tape(A, E, flux):
    U = _generic_state(2)
    ny, nx = U.shape[:2]
    dh = 1.0 / nx
    target = 1.0 + 0.05 * np.random.default_rng(5).random((ny, nx))
    steps = 25
    loss, g, dg, _, _ = _adjoint_run(A, E, U, target, dh, steps, 7, flux=flux)
    lam = np.zeros((4, ny, nx))
    # tape: L = dh^2 sum (rho_N - target)^2
    Ut = [torch.tensor(np.ascontiguousarray(np.moveaxis(U, -1, 0)[k]), requires_grad=True)
          for k in range(4)]
    gam = torch.tensor(1.4, dtype=torch.float64, requires_grad=True)
    X = Ut
    for _ in range(steps):
        X = T.step(X, gam, 0.9, dh, 2, (2, 2), flux_method=flux)
    L = (dh * dh * (X[0] - torch.tensor(target)) ** 2).sum()
    grads = torch.autograd.grad(L, Ut + [gam])
    ref = np.stack([gr.numpy() for gr in grads[:-1]])
    Lv = float(L.detach())
    assert abs(loss - Lv) <= 1e-12 * abs(Lv)
    assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max()
    assert abs(dg - float(grads[-1])) <= 1e-10 * abs(float(grads[-1]))
    del lam


def test_checkpoint_interval_invariance(A, E):
    U = _generic_state(2, seed=6)
    ny, nx = U.shape[:2]
    dh = 1.0 / nx
    target = np.ones((ny, nx))
    res = [_adjoint_run(A, E, U, target, dh, 23, k) for k in (1, 5, 23, 40)]
    for r in res[1:]:
        assert r[0] == res[0][0]
        assert np.array_equal(r[1].view(np.uint64), res[0][1].view(np.uint64))
        assert r[2] == res[0][2]


def test_forward_equals_simulate(A, E):
    U = _generic_state(2, seed=7)
    dh = 1.0 / U.shape[1]
    adj = A.Adjoint(E.Scheme(gamma=1.4, bc=(2, 2, 0)), 0.9, dh)
    Ud = _dev(U)
    adj.forward(Ud, 17, 5)
    solver = E.SplitSolver(E.Scheme(gamma=1.4, bc=(2, 2, 0)), layout=1)
    Us = _dev(U)
    r = solver.simulate(Us, torch.empty_like(Us), 0.9, dh, max_iters=17)
    assert r.iters == 17
    assert torch.equal(Ud, Us)


def _loss_fd(A, E, U, target, gamma, dh, steps):
    adj = A.Adjoint(E.Scheme(gamma=gamma, bc=(2, 2, 0)), 0.9, dh)
    Ud = _dev(U)
    adj.forward(Ud, steps, steps)
    tgt = torch.from_numpy(target).to("cuda")
    return float((dh * dh * (Ud[0] - tgt) ** 2).sum())


def _fd_check(A, E, prob, dirs, steps, eps):
    loss, g, dg, _, _ = _adjoint_run(A, E, prob.U, prob.target, prob.dh, steps, 20)
    errs = []
    for d, dgam in dirs:
        dsoa = _soa(d)
        ad = float((g * dsoa).sum()) + dg * dgam
        lp = _loss_fd(A, E, prob.U + eps * d, prob.target, prob.gamma + eps * dgam, prob.dh, steps)
        lm = _loss_fd(A, E, prob.U - eps * d, prob.target, prob.gamma - eps * dgam, prob.dh, steps)
        fd = (lp - lm) / (2 * eps)
        errs.append(abs(fd - ad) / max(abs(fd), 1e-300))
    return errs


def test_finite_differences_with_flow(A, E):
    prob = W.adjoint_problem(1024, flow=True)
    dirs = W.smooth_directions(1024, 64, seed=1006)
    # The scheme is piecewise smooth (limiter and HLLC branches): the central
    # difference error grows linearly with the step while branches switch
    # inside it (measured max over the 64 directions: 9.6e-6 at 1e-6,
    # 1.6e-7 at 1e-7, 4.6e-8 at 1e-8), so the step is 1e-8.
    errs = _fd_check(A, E, prob, dirs, 200, 1e-8)
    print("FD vs adjoint, 64 directions: max rel err %.3g, median %.3g" %
          (max(errs), float(np.median(errs))))
    assert max(errs) <= 1e-6


def test_finite_differences_static_rho_gamma(A, E):
    """The literal workload (u = v = 0, p = 1): rho-only directions against
    central differences, and the gamma direction with an absolute bound.

    The state is an exact steady contact for every gamma: pressure
    (gamma - 1) E is the same in every cell and the velocity is zero, so each
    face flux difference cancels and the density never changes; the loss does
    not depend on gamma and dL/dgamma is 0 exactly.  The central difference
    is therefore exactly 0, and the adjoint's value may differ from 0 only by
    the rounding of its gamma terms (which cancel between faces), here held
    to 1e-12 of loss / gamma.  Directions in momentum and energy start a flow
    at a state where every limiter ratio and the HLLC star branches sit on
    their branch points (zero slopes, u_L = u_R = 0), where the scheme is not
    differentiable: those are checked on the flow workload
    (test_finite_differences_with_flow)."""
    prob = W.adjoint_problem(1024)
    full = W.smooth_directions(1024, 8, seed=1006)
    dirs = []
    for d, dgam in full[:4]:
        rho_only = np.zeros_like(d)
        rho_only[..., 0] = d[..., 0]
        dirs.append((rho_only, 0.0))
    errs = _fd_check(A, E, prob, dirs, 200, 1e-8)
    print("static workload, rho directions: rel errs", ["%.3g" % e for e in errs])
    assert max(errs) <= 1e-6
    loss, _, dg, _, _ = _adjoint_run(A, E, prob.U, prob.target, prob.dh, 200, 20)
    eps = 1e-8
    lp = _loss_fd(A, E, prob.U, prob.target, prob.gamma + eps, prob.dh, 200)
    lm = _loss_fd(A, E, prob.U, prob.target, prob.gamma - eps, prob.dh, 200)
    print(f"static workload: loss {loss:.6g}, dL/dgamma adjoint {dg:.3g}, "
          f"central difference {(lp - lm) / (2 * eps):.3g}")
    assert lp == lm == loss
    assert abs(dg) <= 1e-12 * abs(loss) / prob.gamma
