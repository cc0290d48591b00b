"""Host-side API of the batched Euler step over the C ABI (fl_batch_*).

A batch advances many independent grids together, each with its own
material, CFL dt and stop condition, as a host loop of the reference's
``SimulatorImpl.simulate`` (simulator/simulator_impl.hpp:222-258) over the
problems would, but in one persistent launch per sweep phase with the steps
replayed from a hipGraph.  Grids are device tensors as in ``euler.grid_field``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import native as N
from .euler import Scheme, _stream_handle, grid_field


class Batch:
    def __init__(self, problems, layout: int = N.FL_LAYOUT_SOA, device: int = 0, stream=None):
        """problems: sequence of dicts with keys U (device tensor), scheme
        (euler.Scheme), cfl, dh and optionally sim_time (default inf) and
        max_iters (default 2^62)."""
        self.layout = layout
        self.ctx = N.Context(device)
        self.tensors = []
        arr = (N.BatchProblem * len(problems))()
        for i, p in enumerate(problems):
            U = p["U"]
            work = torch.empty_like(U)
            self.tensors.append((U, work))
            arr[i].state = grid_field(U, layout)
            arr[i].work = grid_field(work, layout)
            arr[i].scheme = p["scheme"].c()
            arr[i].params = N.SimParams(p["cfl"], p["dh"], p.get("sim_time", float("inf")),
                                        p.get("max_iters", 1 << 62))
        self.nprob = len(problems)
        self._arr = arr
        h = C.c_void_p()
        N.check(N.lib.fl_batch_create(self.ctx.handle, arr, self.nprob, C.byref(h),
                                      _stream_handle(stream)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            N.lib.fl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, x_block=0, x_rows=0, y_block=0, y_rows=0):
        t = N.Tuning(x_block, x_rows, y_block, y_rows)
        N.check(N.lib.fl_batch_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self) -> N.Tuning:
        t = N.Tuning()
        N.check(N.lib.fl_batch_get_tuning(self.handle, C.byref(t)))
        return t

    def autotune(self, reps: int = 3, stream=None) -> N.Tuning:
        t = N.Tuning()
        N.check(N.lib.fl_batch_autotune(self.handle, reps, C.byref(t), _stream_handle(stream)))
        return t

    def reset(self, stream=None):
        N.check(N.lib.fl_batch_reset(self.handle, _stream_handle(stream)))

    def step(self, steps: int, graph: bool = True, stream=None):
        N.check(N.lib.fl_batch_step(self.handle, steps, 1 if graph else 0,
                                    _stream_handle(stream)))

    def finish(self, allow_failed: bool = False, stream=None):
        """Results into the state tensors; returns the per-problem SimResults.
        A problem whose state went to NaN or vacuum stops after that step (as
        the reference loop does) and reports a NaN or non-positive
        max_wavespeed; that raises unless allow_failed."""
        res = (N.SimResult * self.nprob)()
        rc = N.lib.fl_batch_finish(self.handle, res, _stream_handle(stream))
        if not (allow_failed and rc == N.FL_ERR_NUMERIC):
            N.check(rc)
        return list(res)

    def conservation_sums(self, stream=None) -> np.ndarray:
        out = np.zeros((self.nprob, 4))
        N.check(N.lib.fl_batch_conservation_sums(
            self.handle, out.ctypes.data_as(C.POINTER(C.c_double)), _stream_handle(stream)))
        return out

    def state(self, i: int) -> torch.Tensor:
        return self.tensors[i][0]


def problems_from_workload(specs, dtype=torch.float64, layout: int = N.FL_LAYOUT_SOA,
                           cfl: float = 0.9, max_iters: int = 1 << 62, flux: int = N.FL_FLUX_HLLC,
                           device: str = "cuda", numerics: int = N.FL_NUMERICS_EXACT):
    """Device problems from workloads.batched() specs (AoS float64 arrays);
    numerics: the FP64 numerics of every problem (one per FP64 batch)."""
    out = []
    for sp in specs:
        U = sp.U
        if layout == N.FL_LAYOUT_SOA:
            U = np.ascontiguousarray(np.moveaxis(U, -1, 0))
        T = torch.from_numpy(U).to(device=device, dtype=dtype)
        out.append({"U": T, "scheme": Scheme(flux=flux, gamma=sp.gamma, numerics=numerics),
                    "cfl": cfl,
                    "dh": sp.dh, "max_iters": max_iters})
    return out
