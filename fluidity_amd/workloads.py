"""Seeded synthetic initial data for the Euler-step workloads.

The reference ships no input generator; these follow the workload
definitions: Riemann data with rho log-uniform in [0.1, 10], u (and v)
uniform in [-2, 2], p log-uniform in [0.01, 1000], drawn from a SplitMix64
stream (u01 = (x >> 11) * 2^-53).  Draws whose neighbouring states would
create vacuum (Toro eq. 4.40) are rejected and redrawn from the same stream.

Primitive data is converted to the reference's conservative layout with the
reference's prim->cons expressions (state/state_impl.hpp:388-406): E =
rho * (1/2 |v|^2 + p / ((gamma-1) rho)), m = v * rho.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform01(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform(self, a: float, b: float) -> float:
        return a + self.uniform01() * (b - a)

    def log_uniform(self, a: float, b: float) -> float:
        la, lb = math.log(a), math.log(b)
        return math.exp(la + self.uniform01() * (lb - la))


def draw_prim(rng: SplitMix64, nd: int) -> list[float]:
    """One primitive state [rho, p, u, (v)] in the workload ranges."""
    rho = rng.log_uniform(0.1, 10.0)
    vel = [rng.uniform(-2.0, 2.0) for _ in range(nd)]
    p = rng.log_uniform(0.01, 1000.0)
    return [rho, p, *vel]


def sound_speed(w, gamma):
    return math.sqrt(gamma * w[1] / w[0])


def no_vacuum(wl, wr, gamma, dim=0) -> bool:
    """Pressure positivity condition (Toro eq. 4.40) in the normal direction."""
    du = wr[2 + dim] - wl[2 + dim]
    return (2.0 / (gamma - 1.0)) * (sound_speed(wl, gamma) + sound_speed(wr, gamma)) > du


def prim_to_cons(W: np.ndarray, gamma: float) -> np.ndarray:
    """Vectorised prim->cons with the reference's operation order
    (state/state_impl.hpp:135-142, 388-406); W[..., :] = [rho, p, v...]."""
    nd = W.shape[-1] - 2
    rho, p = W[..., 0], W[..., 1]
    vss = np.zeros_like(rho)
    for i in range(nd):
        v = W[..., 2 + i]
        vss = vss + v * v
    eos = p / ((gamma - 1.0) * rho)
    U = np.empty_like(W)
    U[..., 0] = rho
    U[..., 1] = rho * (0.5 * vss + eos)
    for i in range(nd):
        U[..., 2 + i] = W[..., 2 + i] * rho
    return U


# --------------------------------------------------------------------------
# exact Riemann solver (Toro, Riemann Solvers and Numerical Methods, ch. 4)
# --------------------------------------------------------------------------

def _f_k(p, rho, pk, ak, gamma):
    if p > pk:  # shock
        A = 2.0 / ((gamma + 1.0) * rho)
        B = (gamma - 1.0) / (gamma + 1.0) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    r = p / pk
    f = 2.0 * ak / (gamma - 1.0) * (r ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0)
    df = 1.0 / (rho * ak) * r ** (-(gamma + 1.0) / (2.0 * gamma))
    return f, df


def star_pressure(wl, wr, gamma, dim=0):
    rl, pl, ul = wl[0], wl[1], wl[2 + dim]
    rr, pr, ur = wr[0], wr[1], wr[2 + dim]
    al, ar = sound_speed(wl, gamma), sound_speed(wr, gamma)
    du = ur - ul
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * du * (rl + rr) * (al + ar))
    for _ in range(200):
        fl, dfl = _f_k(p, rl, pl, al, gamma)
        fr, dfr = _f_k(p, rr, pr, ar, gamma)
        pn = p - (fl + fr + du) / (dfl + dfr)
        pn = max(pn, 1e-12)
        if abs(pn - p) <= 1e-14 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f_k(p, rl, pl, al, gamma)
    fr, _ = _f_k(p, rr, pr, ar, gamma)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def max_wave_speed(wl, wr, gamma, dim=0) -> float:
    """Largest |wave speed| of the exact solution of the Riemann problem."""
    p, _ = star_pressure(wl, wr, gamma, dim)
    ul, ur = wl[2 + dim], wr[2 + dim]
    al, ar = sound_speed(wl, gamma), sound_speed(wr, gamma)
    q = (gamma + 1.0) / (2.0 * gamma)
    r = (gamma - 1.0) / (2.0 * gamma)
    sl = ul - al * math.sqrt(q * p / wl[1] + r) if p > wl[1] else ul - al
    sr = ur + ar * math.sqrt(q * p / wr[1] + r) if p > wr[1] else ur + ar
    return max(abs(sl), abs(sr))


def sample_exact(wl, wr, gamma, xi):
    """Exact primitive solution [rho, p, u] at similarity coordinates xi = x/t."""
    p_s, u_s = star_pressure(wl, wr, gamma)
    out = np.empty((len(xi), 3))
    rl, pl, ul = wl[0], wl[1], wl[2]
    rr, pr, ur = wr[0], wr[1], wr[2]
    al, ar = sound_speed(wl, gamma), sound_speed(wr, gamma)
    g = gamma
    for i, s in enumerate(xi):
        if s <= u_s:  # left of contact
            if p_s > pl:
                sl = ul - al * math.sqrt((g + 1) / (2 * g) * p_s / pl + (g - 1) / (2 * g))
                if s <= sl:
                    out[i] = (rl, pl, ul)
                else:
                    rs = rl * ((p_s / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p_s / pl + 1))
                    out[i] = (rs, p_s, u_s)
            else:
                shl = ul - al
                a_s = al * (p_s / pl) ** ((g - 1) / (2 * g))
                stl = u_s - a_s
                if s <= shl:
                    out[i] = (rl, pl, ul)
                elif s >= stl:
                    out[i] = (rl * (p_s / pl) ** (1 / g), p_s, u_s)
                else:
                    c = 2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)
                    out[i] = (rl * c ** (2 / (g - 1)), pl * c ** (2 * g / (g - 1)),
                              2 / (g + 1) * (al + (g - 1) / 2 * ul + s))
        else:  # right of contact
            if p_s > pr:
                sr = ur + ar * math.sqrt((g + 1) / (2 * g) * p_s / pr + (g - 1) / (2 * g))
                if s >= sr:
                    out[i] = (rr, pr, ur)
                else:
                    rs = rr * ((p_s / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p_s / pr + 1))
                    out[i] = (rs, p_s, u_s)
            else:
                shr = ur + ar
                a_s = ar * (p_s / pr) ** ((g - 1) / (2 * g))
                str_ = u_s + a_s
                if s >= shr:
                    out[i] = (rr, pr, ur)
                elif s <= str_:
                    out[i] = (rr * (p_s / pr) ** (1 / g), p_s, u_s)
                else:
                    c = 2 / (g + 1) - (g - 1) / ((g + 1) * ar) * (ur - s)
                    out[i] = (rr * c ** (2 / (g - 1)), pr * c ** (2 * g / (g - 1)),
                              2 / (g + 1) * (-ar + (g - 1) / 2 * ur + s))
    return out


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------

@dataclass
class Riemann1D:
    U: np.ndarray        # (n, 3) conservative AoS, float64
    wl: list
    wr: list
    gamma: float
    dh: float
    t_end: float         # time for the fastest wave to travel 0.4 (x = 0.5 -> 0.9)
    rejected: int


def riemann_1d(n: int = 4096, seed: int = 1000, gamma: float = 1.4,
               travel: float = 0.4) -> Riemann1D:
    """1D Riemann problem split at x = 0.5 (cell i is left if i/n < 0.5)."""
    rng = SplitMix64(seed)
    rejected = 0
    while True:
        wl = draw_prim(rng, 1)
        wr = draw_prim(rng, 1)
        if no_vacuum(wl, wr, gamma):
            break
        rejected += 1
    x = np.arange(n) / n
    W = np.where((x < 0.5)[:, None], np.array(wl)[None, :], np.array(wr)[None, :])
    U = prim_to_cons(W.astype(np.float64), gamma)
    smax = max_wave_speed(wl, wr, gamma)
    return Riemann1D(np.ascontiguousarray(U), wl, wr, gamma, 1.0 / n, travel / smax, rejected)


@dataclass
class Riemann2D:
    U: np.ndarray        # (ny, nx, 4) conservative AoS, float64
    quadrants: list      # primitive states [Q1 (x>=x0,y>=y0), Q2, Q3, Q4] anticlockwise
    split: tuple
    gamma: float
    dh: float
    rejected: int


def riemann_2d(n: int = 4096, seed: int = 1001, gamma: float = 1.4,
               ny: int | None = None) -> Riemann2D:
    """Four-quadrant Riemann problem on an n x n grid of [0,1]^2 (ny x n rows
    when ny is given, the split point in the same fractions of each side)
    with the split point uniform in [0.3, 0.7]^2."""
    rng = SplitMix64(seed)
    x0 = rng.uniform(0.3, 0.7)
    y0 = rng.uniform(0.3, 0.7)
    rejected = 0
    while True:
        q = [draw_prim(rng, 2) for _ in range(4)]
        ok = (no_vacuum(q[1], q[0], gamma, 0) and no_vacuum(q[2], q[3], gamma, 0)
              and no_vacuum(q[3], q[0], gamma, 1) and no_vacuum(q[2], q[1], gamma, 1))
        if ok:
            break
        rejected += 1
    U = _quadrant_grid(n, x0, y0, q, gamma, ny)
    return Riemann2D(U, q, (x0, y0), gamma, 1.0 / n, rejected)


def _quadrant_grid(n: int, x0: float, y0: float, q, gamma: float,
                   ny: int | None = None) -> np.ndarray:
    """Conservative ny x n grid (ny = n by default) of four quadrant states
    (Q1 x >= x0, y >= y0, then anticlockwise); cell i is right of the split if
    i/n >= x0, row j above it if j/ny >= y0."""
    ny = n if ny is None else ny
    right = np.arange(n) / n >= x0
    top = np.arange(ny) / ny >= y0
    W = np.empty((ny, n, 4))
    Q = [np.array(s) for s in q]
    for j_top in (False, True):
        for i_right in (False, True):
            qi = {(True, True): 0, (False, True): 1, (False, False): 2, (True, False): 3}[
                (i_right, j_top)]
            mask = (top[:, None] == j_top) & (right[None, :] == i_right)
            W[mask] = Q[qi]
    return np.ascontiguousarray(prim_to_cons(W, gamma))


def _draw_quadrants(rng: SplitMix64, gamma: float):
    rejected = 0
    while True:
        q = [draw_prim(rng, 2) for _ in range(4)]
        ok = (no_vacuum(q[1], q[0], gamma, 0) and no_vacuum(q[2], q[3], gamma, 0)
              and no_vacuum(q[3], q[0], gamma, 1) and no_vacuum(q[2], q[1], gamma, 1))
        if ok:
            return q, rejected
        rejected += 1


@dataclass
class BatchProblem:
    nd: int
    n: tuple             # (nx,) or (nx, ny)
    gamma: float
    dh: float
    U: np.ndarray        # conservative AoS, float64
    states: list         # primitive states: [wl, wr] (1D) or the four quadrants (2D)
    split: tuple         # (0.5,) or (x0, y0)


def batched(seed: int = 1003, n1d: int = 384, n2d: int = 128, len1d=(256, 65536),
            side2d=(64, 1024)) -> list:
    """Independent Riemann problems of uneven size: n1d 1D problems with N
    log-uniform in len1d (split at x = 0.5) and n2d 2D four-quadrant problems
    with side log-uniform in side2d (split point uniform in [0.3, 0.7]^2), each
    with gamma uniform in [1.2, 1.67] and dh = 1/N.  States are drawn as in the
    single-problem workloads from one SplitMix64 stream, vacuum-producing
    draws rejected."""
    rng = SplitMix64(seed)
    out = []
    for i in range(n1d + n2d):
        gamma = rng.uniform(1.2, 1.67)
        if i < n1d:
            n = int(round(rng.log_uniform(*len1d)))
            while True:
                wl = draw_prim(rng, 1)
                wr = draw_prim(rng, 1)
                if no_vacuum(wl, wr, gamma):
                    break
            x = np.arange(n) / n
            W = np.where((x < 0.5)[:, None], np.array(wl)[None, :], np.array(wr)[None, :])
            U = np.ascontiguousarray(prim_to_cons(W.astype(np.float64), gamma))
            out.append(BatchProblem(1, (n,), gamma, 1.0 / n, U, [wl, wr], (0.5,)))
        else:
            n = int(round(rng.log_uniform(*side2d)))
            x0 = rng.uniform(0.3, 0.7)
            y0 = rng.uniform(0.3, 0.7)
            q, _ = _draw_quadrants(rng, gamma)
            U = _quadrant_grid(n, x0, y0, q, gamma)
            out.append(BatchProblem(2, (n, n), gamma, 1.0 / n, U, q, (x0, y0)))
    return out


@dataclass
class ShockBubble:
    U0: np.ndarray       # (ny, nx, 4) material 0 (ambient gas) over the whole grid
    U1: np.ndarray       # (ny, nx, 4) material 1 (bubble) over the whole grid
    phi0: np.ndarray     # (ny, nx) signed distance, <= 0 in material 0
    phi1: np.ndarray     # -phi0
    gamma0: float
    gamma1: float
    p_inf1: float
    stiff: bool
    mach: float
    radius: float
    centre: tuple
    shock_x: float
    dh: float


def prim_to_cons_eos(W: np.ndarray, gamma: float, p_inf: float = 0.0,
                     stiff: bool = False) -> np.ndarray:
    """prim->cons with the stiffened-gas energy (p + gamma p_inf) / ((gamma-1) rho)
    when stiff (material/stiff_fluid.hpp:77-98), else the ideal gas one."""
    if not stiff:
        return prim_to_cons(W, gamma)
    nd = W.shape[-1] - 2
    rho, pr = W[..., 0], W[..., 1]
    vss = np.zeros_like(rho)
    for i in range(nd):
        vss = vss + W[..., 2 + i] * W[..., 2 + i]
    eos = (pr + gamma * p_inf) / ((gamma - 1.0) * rho)
    U = np.empty_like(W)
    U[..., 0] = rho
    U[..., 1] = rho * (0.5 * vss + eos)
    for i in range(nd):
        U[..., 2 + i] = W[..., 2 + i] * rho
    return U


def shock_bubble(nx: int = 2048, ny: int = 1024, seed: int = 1002,
                 stiff: bool = False) -> ShockBubble:
    """A planar shock in gas (gamma 1.4, rho 1, p 1 ahead of it) running in +x
    onto a bubble of a second material on [0, nx dh] x [0, 1], dh = 1/ny.
    Radius uniform in [0.1, 0.25], centre x in [0.5, 1.5], y in [r, 1 - r],
    shock Mach number uniform in [1.2, 3] with the post-shock state from the
    Rankine-Hugoniot relations, shock 0.05 ahead of the bubble.  Bubble:
    ideal gas gamma 1.67 at density ratio 0.14 in pressure equilibrium, or
    (stiff) a stiffened gas gamma 4.4, p_inf 6e8 at density 1000 in units
    where the ambient pressure is 1e5.  phi is the exact signed distance to the
    circle; each material's state is its own phase over the whole grid."""
    rng = SplitMix64(seed)
    r = rng.uniform(0.1, 0.25)
    cx = rng.uniform(0.5, 1.5)
    cy = rng.uniform(r, 1.0 - r)
    mach = rng.uniform(1.2, 3.0)
    g0 = 1.4
    rho_a, p_a = (1.0, 1.0) if not stiff else (1.0, 1e5)
    a_a = math.sqrt(g0 * p_a / rho_a)
    m2 = mach * mach
    rho_s = rho_a * (g0 + 1.0) * m2 / ((g0 - 1.0) * m2 + 2.0)
    p_s = p_a * (2.0 * g0 * m2 - (g0 - 1.0)) / (g0 + 1.0)
    u_s = mach * a_a * (1.0 - rho_a / rho_s)
    xs = cx - r - 0.05
    dh = 1.0 / ny
    x = (np.arange(nx) + 0.5) * dh
    y = (np.arange(ny) + 0.5) * dh
    X, Y = np.meshgrid(x, y)
    W0 = np.empty((ny, nx, 4))
    post = X < xs
    W0[..., 0] = np.where(post, rho_s, rho_a)
    W0[..., 1] = np.where(post, p_s, p_a)
    W0[..., 2] = np.where(post, u_s, 0.0)
    W0[..., 3] = 0.0
    if stiff:
        g1, pinf1, rho_b = 4.4, 6e8, 1000.0
    else:
        g1, pinf1, rho_b = 1.67, 0.0, 0.14 * rho_a
    W1 = np.empty((ny, nx, 4))
    W1[..., 0] = rho_b
    W1[..., 1] = p_a
    W1[..., 2] = 0.0
    W1[..., 3] = 0.0
    phi1 = np.hypot(X - cx, Y - cy) - r
    phi0 = -phi1
    U0 = prim_to_cons(W0, g0)
    U1 = prim_to_cons_eos(W1, g1, pinf1, stiff)
    return ShockBubble(np.ascontiguousarray(U0), np.ascontiguousarray(U1),
                       np.ascontiguousarray(phi0), np.ascontiguousarray(phi1), g0, g1, pinf1,
                       stiff, mach, r, (cx, cy), xs, dh)


@dataclass
class ShockBubbles:
    Us: list             # per material (ny, nx, 4) conservative AoS
    phis: list           # per material (ny, nx) level sets, phi <= 0 inside
    gammas: list
    mach: float
    dh: float


def shock_two_bubbles(nx: int = 512, ny: int = 256, seed: int = 1007) -> ShockBubbles:
    """Three materials: a planar shock (Mach uniform in [1.2, 3]) in gas
    (gamma 1.4, rho 1, p 1 ahead of it) running in +x onto two bubbles on
    [0, nx dh] x [0, 1], dh = 1/ny: a light one (gamma 1.67, density ratio
    0.14) and a heavy one (gamma 1.249, density ratio 3.15), both in pressure
    equilibrium with the gas, radii uniform in [0.08, 0.15], centres seeded
    and apart.  Level sets: each bubble's exact signed distance, and for the
    gas minus the smaller of the two, negative outside both bubbles.  Each
    material's state is its own phase over the whole grid."""
    rng = SplitMix64(seed)
    ra = rng.uniform(0.08, 0.15)
    rb = rng.uniform(0.08, 0.15)
    ca = (rng.uniform(0.6, 0.9), rng.uniform(0.2, 0.45))
    cb = (rng.uniform(1.1, 1.4), rng.uniform(0.55, 0.8))
    mach = rng.uniform(1.2, 3.0)
    g0 = 1.4
    rho_a, p_a = 1.0, 1.0
    a_a = math.sqrt(g0 * p_a / rho_a)
    m2 = mach * mach
    rho_s = rho_a * (g0 + 1.0) * m2 / ((g0 - 1.0) * m2 + 2.0)
    p_s = p_a * (2.0 * g0 * m2 - (g0 - 1.0)) / (g0 + 1.0)
    u_s = mach * a_a * (1.0 - rho_a / rho_s)
    xs = min(ca[0] - ra, cb[0] - rb) - 0.05
    dh = 1.0 / ny
    x = (np.arange(nx) + 0.5) * dh
    y = (np.arange(ny) + 0.5) * dh
    X, Y = np.meshgrid(x, y)
    W0 = np.empty((ny, nx, 4))
    post = X < xs
    W0[..., 0] = np.where(post, rho_s, rho_a)
    W0[..., 1] = np.where(post, p_s, p_a)
    W0[..., 2] = np.where(post, u_s, 0.0)
    W0[..., 3] = 0.0
    gammas = [g0, 1.67, 1.249]
    Us = [prim_to_cons(W0, g0)]
    for g, ratio in ((1.67, 0.14), (1.249, 3.15)):
        W = np.zeros((ny, nx, 4))
        W[..., 0] = ratio * rho_a
        W[..., 1] = p_a
        Us.append(prim_to_cons(W, g))
    phia = np.hypot(X - ca[0], Y - ca[1]) - ra
    phib = np.hypot(X - cb[0], Y - cb[1]) - rb
    phis = [-np.minimum(phia, phib), phia, phib]
    return ShockBubbles([np.ascontiguousarray(U) for U in Us],
                        [np.ascontiguousarray(p) for p in phis], gammas, mach, dh)


def aos_to_soa(U: np.ndarray) -> np.ndarray:
    """(.., S) AoS -> (S, ..) SoA planes."""
    return np.ascontiguousarray(np.moveaxis(U, -1, 0))


def soa_to_aos(P: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(P, 0, -1))


def _bump_sum(rng: SplitMix64, n: int, count: int = 16, amp=(-0.3, 0.3), width=(0.02, 0.08)):
    """sum of `count` Gaussian bumps A exp(-|x - c|^2 / w^2) on the periodic
    unit square (minimum-image distance), centres uniform, n x n cells."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    f = np.zeros((n, n))
    for _ in range(count):
        a = rng.uniform(*amp)
        w = rng.uniform(*width)
        cx = rng.uniform(0.0, 1.0)
        cy = rng.uniform(0.0, 1.0)
        dx = X - cx
        dx -= np.round(dx)
        dy = Y - cy
        dy -= np.round(dy)
        f += a * np.exp(-(dx * dx + dy * dy) / (w * w))
    return f


@dataclass
class AdjointProblem:
    U: np.ndarray        # (n, n, 4) conservative AoS initial state
    target: np.ndarray   # (n, n) target density
    gamma: float
    dh: float


def adjoint_problem(n: int = 1024, seed: int = 1004, target_seed: int = 1005,
                    gamma: float = 1.4, flow: bool = False) -> AdjointProblem:
    """Density 1 + 16 Gaussian bumps (seed), u = v = 0, p = 1 on the periodic
    unit square; E = p / (gamma - 1) exactly, so the state is an exact
    steady contact.  The target density is a second bump set (target_seed).
    flow=True adds the smooth background flow u = 0.5 + 0.1 sin(2 pi y),
    v = 0.3 + 0.1 sin(2 pi x) and p = 1 + 0.1 cos(2 pi x) cos(2 pi y): the
    same bumps, but a state at which the scheme is differentiable (no face on
    a branch point), for finite-difference checks."""
    rho = 1.0 + _bump_sum(SplitMix64(seed), n)
    target = 1.0 + _bump_sum(SplitMix64(target_seed), n)
    U = np.zeros((n, n, 4))
    U[..., 0] = rho
    if not flow:
        U[..., 1] = 1.0 / (gamma - 1.0)
    else:
        x = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(x, x)
        W = np.empty((n, n, 4))
        W[..., 0] = rho
        W[..., 1] = 1.0 + 0.1 * np.cos(2 * np.pi * X) * np.cos(2 * np.pi * Y)
        W[..., 2] = 0.5 + 0.1 * np.sin(2 * np.pi * Y)
        W[..., 3] = 0.3 + 0.1 * np.sin(2 * np.pi * X)
        U = prim_to_cons(W, gamma)
    return AdjointProblem(np.ascontiguousarray(U), np.ascontiguousarray(target), gamma, 1.0 / n)


def smooth_directions(n: int, count: int = 64, seed: int = 1006):
    """`count` random smooth perturbation directions of (rho, rho u, rho v, E)
    on an n x n periodic grid (each field a sum of 4 Gaussian bumps of width
    0.05..0.15) plus a gamma component, each normalised to unit RMS."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        d = np.stack([_bump_sum(rng, n, 4, (-1.0, 1.0), (0.05, 0.15)) for _ in range(4)], -1)
        dg = rng.uniform(-1.0, 1.0)
        scale = np.sqrt((d * d).mean() + dg * dg / d.size)
        out.append((np.ascontiguousarray(d / scale), dg / scale))
    return out
