"""Adjoint of the split Euler step over the C ABI (fl_adj_*, fl_split_step_vjp),
with a thin torch.autograd.Function binding used to check the gradients.

The reference has no adjoint; this differentiates its update loop
(``SimulatorImpl.simulate``, simulator/simulator_impl.hpp:222-258, with
``SplitSolver.solve``, solver/split_solver.hpp:105-162): FP64, HLLC or FORCE,
either limiter and limiting form, dt = cfl dh / max wavespeed recomputed every step
and the flattening factor (their dependence on the state and on gamma
included).  Grids are FP64 device
tensors as in ``euler.grid_field``.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import native as N
from .euler import Scheme, _stream_handle, autotune_context, grid_field


def _bytes(n: int, device) -> torch.Tensor:
    return torch.empty(max(int(n), 1), dtype=torch.uint8, device=device)


class Adjoint:
    """Checkpointed forward run and its adjoint on one device."""

    def __init__(self, scheme: Scheme, cfl: float, dh: float, layout: int = N.FL_LAYOUT_SOA,
                 device: int = 0):
        self.scheme = scheme
        self.cfl = cfl
        self.dh = dh
        self.layout = layout
        self.ctx = N.Context(device)
        self._tuned = None

    def tune(self, U: torch.Tensor, stream=None):
        """Launch shapes for this grid shape (once per shape)."""
        if self._tuned != tuple(U.shape):
            autotune_context(self.ctx, U, torch.empty_like(U), self.scheme, self.layout,
                             stream=stream)
            self._tuned = tuple(U.shape)

    def forward(self, U: torch.Tensor, steps: int, checkpoint_every: int = 20, stream=None):
        """Advances U in place by `steps` steps; returns the checkpoints."""
        self.tune(U, stream)
        f = grid_field(U, self.layout)
        cb = N.lib.fl_adj_checkpoint_bytes(C.byref(f), steps, checkpoint_every)
        wb = N.lib.fl_adj_work_bytes(C.byref(f), checkpoint_every)
        ck = _bytes(cb, U.device)
        work = _bytes(wb, U.device)
        s = self.scheme.c()
        r = N.SimResult()
        N.check(N.lib.fl_adj_forward(self.ctx.handle, C.byref(f), C.byref(s), self.cfl, self.dh,
                                     steps, checkpoint_every, ck.data_ptr(), cb, work.data_ptr(),
                                     wb, C.byref(r), _stream_handle(stream)))
        return {"checkpoints": ck, "steps": steps, "every": checkpoint_every, "result": r}

    def backward(self, grad: torch.Tensor, run, stream=None) -> float:
        """grad: dLoss/d(final state), replaced in place by dLoss/d(initial
        state); returns dLoss/dgamma."""
        f = grid_field(grad, self.layout)
        wb = N.lib.fl_adj_work_bytes(C.byref(f), run["every"])
        work = _bytes(wb, grad.device)
        s = self.scheme.c()
        dg = C.c_double(0.0)
        ck = run["checkpoints"]
        N.check(N.lib.fl_adj_backward(self.ctx.handle, C.byref(f), C.byref(s), self.cfl, self.dh,
                                      run["steps"], run["every"], ck.data_ptr(), ck.numel(),
                                      work.data_ptr(), wb, C.byref(dg), _stream_handle(stream)))
        return dg.value

    def step_vjp(self, U: torch.Tensor, lam_next: torch.Tensor, stream=None):
        """(dU_{n+1}/dU_n)^T lam_next and lam_next . dU_{n+1}/dgamma."""
        self.tune(U, stream)
        fu = grid_field(U, self.layout)
        fn = grid_field(lam_next, self.layout)
        lam = torch.empty_like(lam_next)
        fl = grid_field(lam, self.layout)
        wb = N.lib.fl_adj_work_bytes(C.byref(fu), 1)
        work = _bytes(wb, U.device)
        s = self.scheme.c()
        dg = C.c_double(0.0)
        N.check(N.lib.fl_split_step_vjp(self.ctx.handle, C.byref(fu), C.byref(fn), C.byref(fl),
                                        C.byref(s), self.cfl, self.dh, C.byref(dg), work.data_ptr(),
                                        wb, _stream_handle(stream)))
        return lam, dg.value


class EulerRun(torch.autograd.Function):
    """U_N = run(U_0, gamma) for checking gradients with torch: the forward
    stores checkpoints, the backward runs the hand-written adjoint."""

    @staticmethod
    def forward(ctx, U0, gamma, steps, checkpoint_every, cfl, dh, limiter, bc):
        scheme = Scheme(limiter=limiter, gamma=float(gamma), bc=tuple(bc))
        adj = Adjoint(scheme, cfl, dh, N.FL_LAYOUT_SOA, U0.device.index or 0)
        U = U0.detach().clone().contiguous()
        run = adj.forward(U, steps, checkpoint_every)
        ctx.adj = adj
        ctx.run = run
        return U

    @staticmethod
    def backward(ctx, gU):
        g = gU.detach().clone().contiguous()
        dgamma = ctx.adj.backward(g, ctx.run)
        return (g, torch.tensor(dgamma, dtype=torch.float64), None, None, None, None, None, None)


def run(U0: torch.Tensor, gamma: torch.Tensor, steps: int, cfl: float, dh: float,
        checkpoint_every: int = 20, limiter: int = N.FL_LIMITER_VAN_LEER,
        bc=(N.FL_BC_PERIODIC, N.FL_BC_PERIODIC, N.FL_BC_PERIODIC)) -> torch.Tensor:
    """Differentiable (w.r.t. U0 in SoA layout and the 0-d tensor gamma) run of
    `steps` steps."""
    return EulerRun.apply(U0, gamma, steps, checkpoint_every, cfl, dh, limiter, bc)
