"""Host-side mirror of fluidity's level-set and ghost-fluid API over the C ABI.

Names follow the reference: the interface predicates as ``masks``
(levelset/levelset.hpp:166-217), ``set_velocities``
(levelset/cuda/velocity_initialization.cuh:108-143), ``evolve`` with the
level-set evolver (scheme/cuda/evolve.cuh:51-157), ``project``
(levelset/cuda/levelset_projection.cuh:54-66), ``eikonal`` with the fast
iterative method (scheme/cuda/fast_iterative_method.cuh), ``load_ghost_cells``
for the Riemann ghost fluid (ghost_fluid/riemann_ghost_fluid.hpp:50-195,
ghost_fluid/ghost_fluid_extrapolator.hpp:54-202), ``combine_into_first``
(material/cuda/combine_materials.cuh:39-63) and
``MultimaterialSimulator.simulate_mm``
(simulator/multimaterial_simulator.hpp:227-303).

Level sets are float64 device tensors of shape (ny, nx) or (nx,); states are
float64 grids as in ``euler.grid_field`` (AoS or SoA), or float32 grids for
the FP32 bulk path (the level-set and ghost-fluid arithmetic stays FP64).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import native as N
from .euler import Scheme, _stream_handle, autotune_context, grid_field

INSIDE, OUTSIDE, ON_BOUNDARY, INSIDE_INTERFACIAL, OUTSIDE_INTERFACIAL = 1, 2, 4, 8, 16


def _shape(phi: torch.Tensor):
    if not phi.is_cuda or phi.dtype != torch.float64 or not phi.is_contiguous():
        raise ValueError("level sets are contiguous float64 device tensors")
    n = list(reversed(phi.shape))
    if len(n) not in (1, 2):
        raise ValueError(f"level set of shape {tuple(phi.shape)} is neither 1D nor 2D")
    return len(n), (C.c_int64 * 2)(*(n + [1])[:2])


def _work(nbytes: int, device) -> torch.Tensor:
    # torch's caching allocator returns 512-byte aligned blocks
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def masks(phi: torch.Tensor, dh: float, stream=None) -> torch.Tensor:
    """One byte per cell: INSIDE | OUTSIDE | ON_BOUNDARY | INSIDE_INTERFACIAL |
    OUTSIDE_INTERFACIAL bits."""
    nd, n = _shape(phi)
    out = torch.empty(phi.shape, dtype=torch.uint8, device=phi.device)
    N.check(N.lib.fl_ls_masks(phi.data_ptr(), nd, n, dh, out.data_ptr(), _stream_handle(stream)))
    return out


def set_velocities(phis, states, dh: float, target: int = -1, layout: int = N.FL_LAYOUT_AOS,
                   stream=None) -> torch.Tensor:
    """V = u . n; target < 0 gives the reference's field (normal of the present
    material's level set), target = m the normal of level set m."""
    nd, _ = _shape(phis[0])
    V = torch.empty_like(phis[0])
    ptrs = (C.c_void_p * len(phis))(*[p.data_ptr() for p in phis])
    fields = (N.Field * len(states))(*[grid_field(u, layout) for u in states])
    N.check(N.lib.fl_ls_velocity(len(phis), ptrs, fields, dh, target, V.data_ptr(),
                                 _stream_handle(stream)))
    return V


def evolve(phi: torch.Tensor, V: torch.Tensor, dt: float, dh: float, out=None, work=None,
           stream=None) -> torch.Tensor:
    """One SSP-RK3 step of phi_t + V |grad phi| = 0 (WENO-HJ5, Godunov)."""
    nd, n = _shape(phi)
    out = torch.empty_like(phi) if out is None else out
    work = torch.empty(2 * phi.numel(), dtype=torch.float64, device=phi.device) if work is None else work
    N.check(N.lib.fl_ls_advect(phi.data_ptr(), V.data_ptr(), nd, n, dt, dh, out.data_ptr(),
                               work.data_ptr(), _stream_handle(stream)))
    return out


def project(phi0: torch.Tensor, phi1: torch.Tensor, stream=None) -> None:
    """In place: avg = (phi0 + phi1) / 2; phi0 -= avg; phi1 = -phi0."""
    N.check(N.lib.fl_ls_project(phi0.data_ptr(), phi1.data_ptr(), phi0.numel(),
                                _stream_handle(stream)))


class _Ctx:
    _ctx = {}

    @classmethod
    def get(cls, device) -> N.Context:
        idx = torch.device(device).index or 0
        if idx not in cls._ctx:
            cls._ctx[idx] = N.Context(idx)
        return cls._ctx[idx]


def eikonal(phi: torch.Tensor, dh: float, band: int = 6, out=None, ctx: N.Context | None = None,
            stream=None):
    """Reinitialises phi to signed distance within `band` cells.  Returns
    (out, band_mask, iterations)."""
    nd, n = _shape(phi)
    ctx = ctx or _Ctx.get(phi.device)
    out = torch.empty_like(phi) if out is None else out
    bm = torch.empty(phi.shape, dtype=torch.uint8, device=phi.device)
    work = _work(N.lib.fl_ls_reinit_work_bytes(nd, n), phi.device)
    it = C.c_int64(0)
    N.check(N.lib.fl_ls_reinit(ctx.handle, phi.data_ptr(), nd, n, dh, band, out.data_ptr(),
                               bm.data_ptr(), work.data_ptr(), C.byref(it),
                               _stream_handle(stream)))
    return out, bm, it.value


def riemann_gfm_star(phi_m, phi_o, state_m, state_o, scheme_m: Scheme, scheme_o: Scheme,
                     dh: float, layout: int = N.FL_LAYOUT_AOS, stream=None):
    """Star states (AoS, cells x (2+D)) and the mask of cells that hold one."""
    S = state_m.shape[-1] if layout == N.FL_LAYOUT_AOS else state_m.shape[0]
    star = torch.zeros((phi_m.numel(), S), dtype=torch.float64, device=phi_m.device)
    has = torch.empty(phi_m.shape, dtype=torch.uint8, device=phi_m.device)
    fm, fo = grid_field(state_m, layout), grid_field(state_o, layout)
    sm, so = scheme_m.c(), scheme_o.c()
    N.check(N.lib.fl_gfm_star(phi_m.data_ptr(), phi_o.data_ptr(), C.byref(fm), C.byref(fo),
                              C.byref(sm), C.byref(so), dh, star.data_ptr(), has.data_ptr(),
                              _stream_handle(stream)))
    return star, has


def extrapolate(phi_m, state_m, star, has, dh: float, sweeps: int = 8,
                layout: int = N.FL_LAYOUT_AOS, stream=None) -> None:
    """Fills the ghost band of state_m in place."""
    nd, n = _shape(phi_m)
    work = _work(N.lib.fl_gfm_extrapolate_work_bytes(nd, n), phi_m.device)
    fm = grid_field(state_m, layout)
    N.check(N.lib.fl_gfm_extrapolate(phi_m.data_ptr(), C.byref(fm), star.data_ptr(),
                                     has.data_ptr(), dh, sweeps, work.data_ptr(),
                                     _stream_handle(stream)))


def load_ghost_cells(phis, states, schemes, dh: float, sweeps: int = 8,
                     layout: int = N.FL_LAYOUT_AOS, stream=None) -> None:
    """Riemann ghost fluid for two materials: both star states from the same
    data, then each material's extrapolation."""
    stars = [riemann_gfm_star(phis[m], phis[1 - m], states[m], states[1 - m], schemes[m],
                              schemes[1 - m], dh, layout, stream) for m in range(2)]
    for m in range(2):
        extrapolate(phis[m], states[m], stars[m][0], stars[m][1], dh, sweeps, layout, stream)


def combine_into_first(phi0, state0, state1, layout: int = N.FL_LAYOUT_AOS,
                       stream=None) -> torch.Tensor:
    out = torch.empty_like(state0)
    f0, f1, fo = grid_field(state0, layout), grid_field(state1, layout), grid_field(out, layout)
    N.check(N.lib.fl_mm_combine(phi0.data_ptr(), C.byref(f0), C.byref(f1), C.byref(fo),
                                _stream_handle(stream)))
    return out


class MultimaterialSimulator:
    """Riemann ghost-fluid simulator of 2 to 8 materials on one device."""

    def __init__(self, device: int = 0, layout: int = N.FL_LAYOUT_AOS):
        self.device = device
        self.layout = layout
        self.ctx = N.Context(device)
        self.materials = []
        self._tuned = False

    MAX_MATERIALS = 8

    def add_material(self, state: torch.Tensor, phi: torch.Tensor, scheme: Scheme):
        if len(self.materials) == self.MAX_MATERIALS:
            raise ValueError(f"at most {self.MAX_MATERIALS} materials are supported")
        self.materials.append((state, torch.empty_like(state), phi, scheme))

    def simulate_mm(self, cfl: float, dh: float, sim_time: float = float("inf"),
                    max_iters: int = 1, band: int = 6, extrap_sweeps: int = 8,
                    reinit_every: int = 1, stream=None) -> N.SimResult:
        nmat = len(self.materials)
        if nmat < 2:
            raise ValueError("add at least two materials first")
        mats = (N.Material * nmat)()
        for m, (state, work, phi, scheme) in enumerate(self.materials):
            mats[m].state = grid_field(state, self.layout)
            mats[m].work = grid_field(work, self.layout)
            mats[m].phi = phi.data_ptr()
            mats[m].scheme = scheme.c()
        if not self._tuned:
            state, work, _, scheme = self.materials[0]
            autotune_context(self.ctx, state, work, scheme, self.layout, stream=stream)
            self._tuned = True
        nd, n = _shape(self.materials[0][2])
        wbytes = N.lib.fl_mm_work_bytes_n(nmat, nd, n)
        work = _work(wbytes, self.materials[0][0].device)
        p = N.SimParams(cfl, dh, sim_time, max_iters)
        mp = N.MmParams(band, extrap_sweeps, reinit_every, 0)
        r = N.SimResult()
        N.check(N.lib.fl_mm_simulate(self.ctx.handle, mats, nmat, C.byref(p), C.byref(mp),
                                     work.data_ptr(), wbytes, C.byref(r), _stream_handle(stream)))
        return r
