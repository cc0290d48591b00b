"""Literal pure-Python restatement of the reference's split-sweep call graph.

TEST INFRASTRUCTURE ONLY.  It exists to pin the structure of the C oracle
(fl_oracle.c), which computes each cell's face states once and each face flux
once: this module instead walks the reference's call graph as written, per
cell, recomputing everything the way the templates do:

  SplitSolver::invoke (solver/split_solver.hpp:138-162)
    *out = *patch + dtdh * FaceFlux::flux_delta(patch, dim)
  FaceFlux::flux_delta = backward - forward (solver/flux_solver.hpp:66-72)
  backward = flux(recon.backward_left, recon.backward_right)   (:148-155)
  forward  = flux(recon.forward_left,  recon.forward_right)    (:92-98)
  MHReconstructor::input_* -> evolve(it +- 1, .., face)
    (reconstruction/muscl_reconstructor.hpp:196-263), each evolve recomputing
    chi over all dimensions (:106-114) and the limiter (:118)

on a ghost-padded copy of the grid (4 cells per side, mirror ghosts through
BoundarySetter, solver/boundary_loader.hpp:42-199).  Python floats are IEEE
doubles and every expression keeps the reference's operation order, so the
result must equal the C oracle bit for bit.  Only for tiny grids.
"""
from __future__ import annotations

import math

PAD = 4


def sqrt(x):
    """IEEE sqrt: NaN for negative arguments instead of an exception."""
    return math.sqrt(x) if x >= 0.0 else math.nan


class Material:
    def __init__(self, gamma=1.4, p_inf=0.0, stiff=False):
        self.gamma, self.p_inf, self.stiff = gamma, p_inf, stiff

    def adiabatic(self):
        return self.gamma

    def sound_speed(self, st):
        if self.stiff:
            return sqrt(self.gamma * (st.pressure(self) + self.p_inf) / st.density())
        return sqrt(self.gamma * st.pressure(self) / st.density())


class State:
    """Conservative state [rho, E, m_0, ..] (state/state.hpp)."""

    def __init__(self, v):
        self.v = list(v)
        self.nd = len(self.v) - 2

    def __getitem__(self, i):
        return self.v[i]

    def density(self):
        return self.v[0]

    def velocity(self, d):
        return self.v[2 + d] / self.v[0]

    def v_squared_sum(self):
        s = 0.0
        for i in range(self.nd):
            v = self.v[2 + i] / self.v[0]
            s += v * v
        return s

    def pressure(self, mat):
        return (mat.adiabatic() - 1.0) * (self.v[1] - 0.5 * self.density() * self.v_squared_sum())

    def energy(self, mat):
        return self.v[1]

    def flux(self, mat, dim):
        v = self.velocity(dim)
        p = self.pressure(mat)
        f = [0.0] * len(self.v)
        f[0] = self.v[2 + dim]
        f[2 + dim] = self.v[2 + dim] * v + p
        f[1] = v * (self.energy(mat) + p)
        for i in range(self.nd - 1):
            idx = 2 + i + (1 if i >= dim else 0)
            f[idx] = self.v[idx] * v
        return f

    def primitive(self, mat):
        return [self.v[0], self.pressure(mat)] + [self.v[2 + i] / self.density()
                                                  for i in range(self.nd)]

    def set_velocity(self, value, d):
        value *= self.v[0]
        self.v[2 + d] = value


def add(a, b):
    return [x + y for x, y in zip(a, b)]


def sub(a, b):
    return [x - y for x, y in zip(a, b)]


def scale(s, a):
    return [x * s for x in a]


def std_min(a, b):
    return b if b < a else a


def std_max(a, b):
    return b if a < b else a


def signum(x):
    return (1 if 0.0 < x else 0) - (1 if x < 0.0 else 0)


class Iterator:
    """MultidimIterator over a padded grid: offset() moves in one dimension."""

    def __init__(self, grid, pos):
        self.grid, self.pos = grid, tuple(pos)

    def offset(self, amount, dim):
        p = list(self.pos)
        p[dim] += amount
        return Iterator(self.grid, p)

    def __call__(self):
        return self.grid[self.pos]

    def backward_diff(self, dim):
        return sub(self().v, self.offset(-1, dim)().v)

    def forward_diff(self, dim):
        return sub(self.offset(1, dim)().v, self().v)


# ---- limiters (limiting/limiter.hpp, van_leer_limiter.hpp, superbee_limiter.hpp)

def van_leer(left, right):
    if right == 0.0:  # the reference's (r <= 0 || right == 0) test
        return 0.0
    r = left / right
    return 0.0 if r <= 0.0 else 2.0 * std_min(r, 1.0) / (1.0 + r)


def superbee(left, right):
    if right == 0.0:
        return 0.0
    r = left / right
    if r <= 0.0:
        return 0.0
    elif r <= 0.5:
        return 2.0 * r
    elif r <= 1.0:
        return 1.0
    Er = 2.0 / (1.0 + r)
    return std_min(Er, std_min(r, 2.0))


class Limiter:
    def __init__(self, single, primitive_form):
        self.single, self.primitive_form = single, primitive_form

    def limit(self, it, mat, dim):
        if self.primitive_form:
            fwrd = sub(it.offset(1, dim)().primitive(mat), it().primitive(mat))
            back = sub(it().primitive(mat), it.offset(-1, dim)().primitive(mat))
        else:
            fwrd = it.forward_diff(dim)
            back = it.backward_diff(dim)
        limited = scale(0.5, add(it.backward_diff(dim), it.forward_diff(dim)))
        for i in range(len(limited)):
            limited[i] *= self.single(back[i], fwrd[i])
        return limited


# ---- MUSCL-Hancock (reconstruction/muscl_reconstructor.hpp)

def chi(it, mat, dim):
    def eita(z):
        temp = 1.0 - ((z - 0.75) / (0.85 - 0.75))
        return std_min(1.0, std_max(0.0, temp))

    p_next = it.offset(1, dim)().pressure(mat)
    p_prev = it.offset(-1, dim)().pressure(mat)
    p_nnext = it.offset(2, dim)().pressure(mat)
    p_pprev = it.offset(-2, dim)().pressure(mat)
    u_factor = it.offset(-1, dim)().velocity(dim) - it.offset(1, dim)().velocity(dim)
    flatten = u_factor > 0.0 and (abs(p_next - p_prev) / std_min(p_next, p_prev)) > 0.33
    if not flatten:
        return 1.0
    den = abs(p_nnext - p_pprev)
    z = abs(p_next - p_prev) / den if den != 0.0 else math.inf
    return eita(z)


def evolve(it, mat, dtdh, dim, face, limiter):
    nd = it().nd
    chi_v = 1.7976931348623157e308
    for d in range(nd):
        sign = signum(it.offset(1, d)().pressure(mat) - it.offset(-1, d)().pressure(mat))
        chi_v = std_min(chi_v, std_min(chi(it, mat, d), chi(it.offset(sign, d), mat, d)))
    delta = scale(chi_v * 0.5, limiter.limit(it, mat, dim))
    u = it().v
    if face == +1:
        bev = State(add(u, delta))
        return State(add(bev.v, scale(0.5 * dtdh, sub(State(sub(u, delta)).flux(mat, dim),
                                                        bev.flux(mat, dim)))))
    bev = State(sub(u, delta))
    return State(add(bev.v, scale(0.5 * dtdh, sub(bev.flux(mat, dim),
                                                  State(add(u, delta)).flux(mat, dim)))))


# ---- flux methods (flux_method/*.hpp)

def hllc(mat, dtdh, dim):
    def wavespeed(st, pstar, sound, adi):
        p = st.pressure(mat)
        if pstar <= p:
            return st.velocity(dim) + sound
        return st.velocity(dim) + sound * sqrt(1.0 + adi * (pstar / p - 1.0))

    def star_speed(sl, sr, wl, wr):
        vl, vr = sl.velocity(dim), sr.velocity(dim)
        fl = sl.density() * (wl - vl)
        fr = sr.density() * (wr - vr)
        return (sr.pressure(mat) - sl.pressure(mat) + vl * fl - vr * fr) / (fl - fr)

    def star_state(st, s, s_star):
        sf = s - st.velocity(dim)
        scl = sf / (s - s_star)
        u = scale(scl, st.v)
        u[2 + dim] = s_star * st.density() * scl
        u[1] = scl * (st.energy(mat) + st.density() * (s_star - st.velocity(dim)) *
                      (s_star + st.pressure(mat) / (st.density() * sf)))
        return u

    def f(ul, ur):
        al, ar = mat.sound_speed(ul), mat.sound_speed(ur)
        adi = (mat.adiabatic() + 1.0) / (2.0 * mat.adiabatic())
        p_pvrs = 0.5 * ((ul.pressure(mat) + ur.pressure(mat)) -
                        (0.25 * (ur.velocity(dim) - ul.velocity(dim)) *
                         (ur.density() + ul.density()) * (al + ar)))
        p_star = std_max(0.0, p_pvrs)
        wsl = wavespeed(ul, p_star, -al, adi)
        fluxl = ul.flux(mat, dim)
        if 0.0 <= wsl:
            return fluxl
        wsr = wavespeed(ur, p_star, ar, adi)
        fluxr = ur.flux(mat, dim)
        if 0.0 >= wsr:
            return fluxr
        ws = star_speed(ul, ur, wsl, wsr)
        if 0.0 <= ws:
            return add(fluxl, scale(wsl, sub(star_state(ul, wsl, ws), ul.v)))
        return add(fluxr, scale(wsr, sub(star_state(ur, wsr, ws), ur.v)))
    return f


def force(mat, dtdh, dim):
    dhdt = 1.0 / dtdh

    def lf(ul, ur):
        return scale(0.5, add(add(ul.flux(mat, dim), ur.flux(mat, dim)),
                              scale(dhdt, sub(ul.v, ur.v))))

    def ri(ul, ur):
        u = State(scale(0.5, add(add(ul.v, ur.v),
                                 scale(dtdh, sub(ul.flux(mat, dim), ur.flux(mat, dim))))))
        return u.flux(mat, dim)

    def f(ul, ur):
        return scale(0.5, add(lf(ul, ur), ri(ul, ur)))
    return f


# ---- sweep ---------------------------------------------------------------------

def _ghost(i, n, kind):
    if 0 <= i < n:
        return i, False
    if kind == 2:
        return i % n, False
    return (-i - 1 if i < 0 else 2 * n - 1 - i), kind == 1


def make_grid(U, nd, n, bc):
    grid = {}
    if nd == 1:
        for x in range(-PAD, n[0] + PAD):
            xi, fx = _ghost(x, n[0], bc[0])
            st = State(U[xi])
            if fx:
                st.set_velocity(-st.velocity(0), 0)
            grid[(x,)] = st
    else:
        for y in range(-PAD, n[1] + PAD):
            yi, fy = _ghost(y, n[1], bc[1])
            for x in range(-PAD, n[0] + PAD):
                xi, fx = _ghost(x, n[0], bc[0])
                st = State(U[yi][xi])
                if fx:
                    st.set_velocity(-st.velocity(0), 0)
                if fy:
                    st.set_velocity(-st.velocity(1), 1)
                grid[(x, y)] = st
    return grid


def sweep(U, nd, n, dim, dtdh, gamma=1.4, flux="hllc", limiter="van_leer", form="cons",
          bc=(0, 0)):
    """One sweep; U is a nested list [y][x][k] (2D) or [x][k] (1D)."""
    mat = Material(gamma)
    lim = Limiter(van_leer if limiter == "van_leer" else superbee, form == "prim")
    fm = (hllc if flux == "hllc" else force)(mat, dtdh, dim)
    grid = make_grid(U, nd, n, bc)
    out = [[None] * n[0] for _ in range(n[1])] if nd == 2 else [None] * n[0]
    cells = [(x,) for x in range(n[0])] if nd == 1 else [(x, y) for y in range(n[1])
                                                         for x in range(n[0])]
    for pos in cells:
        it = Iterator(grid, pos)
        back = fm(evolve(it.offset(-1, dim), mat, dtdh, dim, +1, lim),
                  evolve(it, mat, dtdh, dim, -1, lim))
        fwrd = fm(evolve(it, mat, dtdh, dim, +1, lim),
                  evolve(it.offset(1, dim), mat, dtdh, dim, -1, lim))
        res = add(it().v, scale(dtdh, sub(back, fwrd)))
        if nd == 1:
            out[pos[0]] = res
        else:
            out[pos[1]][pos[0]] = res
    return out
