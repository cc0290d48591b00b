"""Host-side mirror of fluidity's split-solver API over the C ABI.

Names follow the reference: ``SplitSolver.solve`` (solver/split_solver.hpp:105-115),
``set_wavespeeds`` + ``max_element`` as ``max_wavespeed``
(simulator/wavespeed_initialization.hpp:34-53, algorithm/max_element.hpp:28-38),
and ``SimulatorImpl.simulate`` (simulator/simulator_impl.hpp:222-258).
Grids are torch tensors on a ROCm device holding interior cells only:
AoS ``(ny, nx, 2+D)`` / ``(nx, 3)`` as the reference stores ``State`` arrays,
or SoA planes ``(2+D, ny, nx)`` / ``(3, nx)``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import native as N

_DT = {torch.float64: N.FL_F64, torch.float32: N.FL_F32}


def _stream_handle(stream) -> int | None:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def grid_field(U: torch.Tensor, layout: int = N.FL_LAYOUT_AOS) -> N.Field:
    """Describes a device tensor as an fl_field."""
    if not U.is_cuda:
        raise ValueError("grid must live on a ROCm device")
    if not U.is_contiguous():
        raise ValueError("grid must be contiguous")
    if U.dtype not in _DT:
        raise ValueError(f"unsupported dtype {U.dtype}")
    if layout == N.FL_LAYOUT_AOS:
        S = U.shape[-1]
        nd = S - 2
        if U.dim() != nd + 1:
            raise ValueError(f"AoS grid shape {tuple(U.shape)} does not match {nd}D states")
        n = list(reversed(U.shape[:-1]))
    else:
        S = U.shape[0]
        nd = S - 2
        if U.dim() != nd + 1:
            raise ValueError(f"SoA grid shape {tuple(U.shape)} does not match {nd}D states")
        n = list(reversed(U.shape[1:]))
    return N.field(U.data_ptr(), nd, n, _DT[U.dtype], layout)


@dataclass
class Scheme:
    """The compile-time choices of a reference SplitSolver instantiation."""
    limiter: int = N.FL_LIMITER_VAN_LEER
    form: int = N.FL_FORM_CONSERVATIVE
    flux: int = N.FL_FLUX_HLLC
    eos: int = N.FL_EOS_IDEAL
    gamma: float = 1.4
    p_inf: float = 0.0
    bc: tuple = (N.FL_BC_TRANSMISSIVE, N.FL_BC_TRANSMISSIVE, N.FL_BC_TRANSMISSIVE)
    # FP64 arithmetic: N.FL_NUMERICS_EXACT (the reference's bits) or
    # N.FL_NUMERICS_FAST (within 1e-12 relative; an extension)
    numerics: int = N.FL_NUMERICS_EXACT

    def c(self) -> N.Scheme:
        return N.scheme(self.limiter, self.form, self.flux, self.eos, self.gamma, self.p_inf,
                        self.bc, self.numerics)


def autotune_context(ctx: N.Context, U, work, scheme: "Scheme", layout: int, reps: int = 2,
                     stream=None) -> N.Tuning:
    """Launch-shape autotuning of a context on the grid U (U is not
    modified; `work` is scratch of the same shape)."""
    fi, fo = grid_field(U, layout), grid_field(work, layout)
    s = scheme.c()
    t = N.Tuning()
    N.check(N.lib.fl_ctx_autotune(ctx.handle, C.byref(fi), C.byref(fo), C.byref(s), reps,
                                  C.byref(t), _stream_handle(stream)))
    return t


class SplitSolver:
    """Dimensionally split MUSCL-Hancock solver on one device."""

    def __init__(self, scheme: Scheme | None = None, device: int = 0,
                 layout: int = N.FL_LAYOUT_AOS):
        self.scheme = scheme or Scheme()
        self.layout = layout
        self.ctx = N.Context(device)

    def set_tuning(self, **kw):
        self.ctx.set_tuning(**kw)

    def set_graphs(self, enable: bool) -> None:
        """simulate() replays hipGraphs of steps (default; a 1D line of at
        most 2^16 cells runs the whole loop in one cooperative launch) or
        launches every kernel."""
        N.check(N.lib.fl_ctx_set_graphs(self.ctx.handle, 1 if enable else 0))

    def autotune(self, U, work, reps: int = 3, stream=None) -> N.Tuning:
        """Times the candidate launch shapes of both sweeps on this grid
        (U is not modified) and keeps the fastest."""
        return autotune_context(self.ctx, U, work, self.scheme, self.layout, reps, stream)

    def sweep(self, U_in, U_out, dim: int, dtdh: float, stream=None) -> None:
        fi, fo = grid_field(U_in, self.layout), grid_field(U_out, self.layout)
        N.check(N.lib.fl_split_sweep(self.ctx.handle, C.byref(fi), C.byref(fo), dim,
                                     C.byref(self.scheme.c()), dtdh, _stream_handle(stream)))

    def solve(self, U_in, U_out, dtdh: float, stream=None):
        """All sweeps; returns (in, out) after the reference's swap, out
        holding the result."""
        fi, fo = grid_field(U_in, self.layout), grid_field(U_out, self.layout)
        N.check(N.lib.fl_split_solve(self.ctx.handle, C.byref(fi), C.byref(fo),
                                     C.byref(self.scheme.c()), dtdh, _stream_handle(stream)))
        if fo.data == U_in.data_ptr():
            return U_out, U_in
        return U_in, U_out

    def step(self, U_in, U_out, dtdh: float, stream=None):
        """All sweeps with the CFL reduction fused into the last one; returns
        (in, out, max wavespeed of the new states) after the reference's swap."""
        fi, fo = grid_field(U_in, self.layout), grid_field(U_out, self.layout)
        ws = C.c_double()
        N.check(N.lib.fl_split_step(self.ctx.handle, C.byref(fi), C.byref(fo),
                                    C.byref(self.scheme.c()), dtdh, C.byref(ws),
                                    _stream_handle(stream)))
        if fo.data == U_in.data_ptr():
            return U_out, U_in, ws.value
        return U_in, U_out, ws.value

    def max_wavespeed(self, U, stream=None) -> float:
        f = grid_field(U, self.layout)
        out = C.c_double()
        N.check(N.lib.fl_max_wavespeed(self.ctx.handle, C.byref(f), C.byref(self.scheme.c()),
                                       C.byref(out), _stream_handle(stream)))
        return out.value

    def simulate(self, U, work, cfl: float, dh: float, sim_time: float = float("inf"),
                 max_iters: int = 1 << 40, stream=None, check_numeric: bool = True):
        """Runs the step loop in place on U; returns the fl_sim_result."""
        fs, fw = grid_field(U, self.layout), grid_field(work, self.layout)
        p = N.SimParams(cfl, dh, sim_time, max_iters)
        r = N.SimResult()
        st = N.lib.fl_simulate(self.ctx.handle, C.byref(fs), C.byref(fw),
                               C.byref(self.scheme.c()), C.byref(p), C.byref(r),
                               _stream_handle(stream))
        if st != N.FL_OK and (check_numeric or st != N.FL_ERR_NUMERIC):
            N.check(st)
        return r

    def conservation_sums(self, U, stream=None):
        f = grid_field(U, self.layout)
        S = f.ndim + 2
        out = (C.c_double * S)()
        N.check(N.lib.fl_conservation_sums(self.ctx.handle, C.byref(f), out,
                                           _stream_handle(stream)))
        return [out[i] for i in range(S)]
