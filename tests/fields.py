"""Seeded test grids that drive every branch of the per-cell numerics.

 - "rough": i.i.d. per-cell states, so neighbouring pressure ratios exceed the
   flattening threshold and converging velocities switch chi on, and the
   limiter ratio takes both signs;
 - "fast": large velocities, so HLLC takes its supersonic left/right branches;
 - "smooth": sinusoidal data (limiter mostly 0 < r < 1, chi = 1);
 - "riemann": the seeded Riemann workloads.
"""
from __future__ import annotations

import numpy as np

from fluidity_amd import workloads as W


def prim_field(kind: str, shape, seed: int, gamma: float = 1.4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nd = len(shape)
    Wp = np.empty(tuple(shape) + (2 + nd,))
    if kind == "rough":
        Wp[..., 0] = np.exp(rng.uniform(np.log(0.5), np.log(2.0), shape))
        Wp[..., 1] = np.exp(rng.uniform(np.log(0.5), np.log(10.0), shape))
        for d in range(nd):
            Wp[..., 2 + d] = rng.uniform(-0.5, 0.5, shape)
    elif kind == "fast":
        # supersonic flow towards the middle (+3 in one half, -3 in the other)
        # with mild perturbations: HLLC's upwind branches without vacuum
        # (use with transmissive boundaries; periodic ones separate the flow)
        Wp[..., 0] = np.exp(rng.uniform(np.log(0.8), np.log(1.25), shape))
        Wp[..., 1] = np.exp(rng.uniform(np.log(0.8), np.log(1.25), shape))
        # velocity component d converges along its own axis (no vacuum)
        idx = np.indices(shape)
        for d in range(nd):
            ax = nd - 1 - d  # x is the last array axis
            sign = np.where(idx[ax] < shape[ax] // 2, 1.0, -1.0)
            Wp[..., 2 + d] = 3.0 * sign + rng.uniform(-0.1, 0.1, shape)
    elif kind == "smooth":
        grids = np.meshgrid(*[np.linspace(0, 1, s, endpoint=False) for s in shape],
                            indexing="ij")
        ph = rng.uniform(0, 2 * np.pi, 4)
        arg = sum(g * (k + 1) for k, g in enumerate(grids)) * 2 * np.pi
        Wp[..., 0] = 1.0 + 0.3 * np.sin(arg + ph[0])
        Wp[..., 1] = 1.0 + 0.4 * np.cos(arg + ph[1])
        for d in range(nd):
            Wp[..., 2 + d] = 0.5 * np.sin(arg * (d + 1) + ph[2 + d % 2])
    else:
        raise ValueError(kind)
    return Wp


def cons_field(kind: str, shape, seed: int, gamma: float = 1.4) -> np.ndarray:
    """Conservative AoS grid; shape is (nx,) or (ny, nx)."""
    if kind == "riemann":
        if len(shape) == 1:
            return W.riemann_1d(n=shape[0], seed=seed, gamma=gamma).U
        return W.riemann_2d(n=shape[1], seed=seed, gamma=gamma, ny=shape[0]).U
    return np.ascontiguousarray(W.prim_to_cons(prim_field(kind, shape, seed, gamma), gamma))
