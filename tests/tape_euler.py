"""FP64 tape reference of the split MUSCL-Hancock + HLLC / FORCE step in torch (CPU).

TEST INFRASTRUCTURE ONLY: the reference gradient ("tape") the hand-written
adjoint is checked against.  The forward restates the oracle's step (the
reference's solver/split_solver.hpp:138-162 with
reconstruction/muscl_reconstructor.hpp:59-172, limiting/van_leer_limiter.hpp,
flux_method/flux_hllc.hpp:56-245 or flux_method/flux_force.hpp:73-84, dt from
simulator/parameters.hpp:60-63)
with vectorised torch operations in the same operation order, so that its
values equal the oracle's; torch.autograd then differentiates it.  Branches
are torch.where selections (the derivative of the branch taken, with the
unselected branch fed safe operands), the flattening factor chi included
(its flattening test and the choice of the neighbour and of the minimum are
branches; the clamped ratio carries the gradient).

Grids are lists of component tensors [rho, E, m_x, (m_y)], each (ny, nx) in
2D or (nx,) in 1D; bc is one of 0 transmissive, 1 reflective, 2 periodic per
dimension.
"""
from __future__ import annotations

import torch

PAD = 4


def _index_map(n, bc):
    idx = torch.arange(-PAD, n + PAD)
    flip = torch.zeros(n + 2 * PAD, dtype=torch.bool)
    if bc == 2:
        return torch.remainder(idx, n), flip
    lo, hi = idx < 0, idx >= n
    m = torch.where(lo, -idx - 1, torch.where(hi, 2 * n - 1 - idx, idx))
    if bc == 1:
        flip = lo | hi
    return m, flip


def _reflect(U, d, mask, nd):
    """m_d <- (-(m_d / rho)) * rho where mask (the reflective ghost rule)."""
    out = list(U)
    v = -(U[2 + d] / U[0])
    v = v * U[0]
    out[2 + d] = torch.where(mask, v, U[2 + d])
    return out


def pad(U, nd, bcs):
    """Pads every component by PAD ghost cells in every dimension."""
    out = list(U)
    for d in range(nd):
        axis = out[0].dim() - 1 - d
        n = out[0].shape[axis]
        m, flip = _index_map(n, bcs[d])
        out = [torch.index_select(c, axis, m) for c in out]
        shape = [1] * out[0].dim()
        shape[axis] = flip.numel()
        out = _reflect(out, d, flip.view(shape), nd)
    return out


def prim(U, gamma, nd):
    rho = U[0]
    vel = [U[2 + i] / rho for i in range(nd)]
    s = vel[0] * vel[0]
    for i in range(1, nd):
        s = s + vel[i] * vel[i]
    p = (gamma - 1.0) * (U[1] - 0.5 * rho * s)
    return vel, p


def sound(p, rho, gamma):
    return torch.sqrt(gamma * p / rho)


def flux(U, vel, p, D, nd):
    v = vel[D]
    F = [None] * (2 + nd)
    F[0] = U[2 + D]
    F[2 + D] = U[2 + D] * v + p
    F[1] = v * (U[1] + p)
    for o in range(nd):
        if o != D:
            F[2 + o] = U[2 + o] * v
    return F


def van_leer(L, R):
    zero = (R == 0)
    Rs = torch.where(zero, torch.ones_like(R), R)
    r = L / Rs
    clip = (r <= 0) | zero
    rs = torch.where(clip, torch.ones_like(r), r)
    val = (2.0 * torch.where(1.0 < rs, torch.ones_like(rs), rs)) / (1.0 + rs)
    return torch.where(clip, torch.zeros_like(val), val)


def _chi_min(P, gamma, nd):
    """chi_v of MHReconstructor::evolve for every cell of the grid padded by
    one ghost layer in each dimension; returns a tensor shaped like the grid
    padded by 1 everywhere."""
    vel, p = prim(P, gamma, nd)
    full = p.shape                                     # padded by PAD everywhere

    def at(t, offs):
        """t at offset offs (tuple per dim, x first) for every cell of the
        grid padded by 1."""
        sl = t
        for d in range(nd):
            axis = t.dim() - 1 - d
            n = full[axis] - 2 * PAD
            sl = sl.narrow(axis, PAD - 1 + offs[d], n + 2)
        return sl

    def chi_at(offs, d):
        e = [0] * nd
        e[d] = 1

        def o(k):
            return tuple(offs[i] + k * e[i] for i in range(nd))
        p_next, p_prev = at(p, o(1)), at(p, o(-1))
        p_nnext, p_pprev = at(p, o(2)), at(p, o(-2))
        u_factor = at(vel[d], o(-1)) - at(vel[d], o(1))
        dp = torch.abs(p_next - p_prev)
        pmin = torch.where(p_prev < p_next, p_prev, p_next)
        with torch.no_grad():
            flatten = (u_factor > 0) & ((dp / pmin) > 0.33)
        # the ratio only where it is used and finite (elsewhere 0 / 1, and
        # +inf where the reference divides by a zero difference): no
        # infinite or NaN operand whose zero gradient would turn into NaN
        den = torch.abs(p_nnext - p_pprev)
        ok = flatten & (den > 0)
        z = torch.where(ok, torch.where(ok, dp, torch.zeros_like(dp))
                        / torch.where(ok, den, torch.ones_like(den)),
                        torch.where(flatten, torch.full_like(dp, float("inf")),
                                    torch.zeros_like(dp)))
        temp = 1.0 - ((z - 0.75) / (0.85 - 0.75))
        lo = torch.where(0.0 < temp, temp, torch.zeros_like(temp))
        c = torch.where(lo < 1.0, lo, torch.ones_like(lo))
        return torch.where(flatten, c, torch.ones_like(c))

    zero = tuple([0] * nd)
    chi = torch.full(at(p, zero).shape, torch.finfo(torch.float64).max, dtype=torch.float64)
    for d in range(nd):
        e = [0] * nd
        e[d] = 1
        plus = tuple(e)
        minus = tuple(-x for x in e)
        sign = torch.sign(at(p, plus) - at(p, minus))
        c0 = chi_at(zero, d)
        cp = chi_at(plus, d)
        cm = chi_at(minus, d)
        cn = torch.where(sign > 0, cp, torch.where(sign < 0, cm, c0))
        m = torch.where(cn < c0, cn, c0)                 # smin(chi(c), chi(c+s))
        chi = torch.where(m < chi, m, chi)               # smin(chi_v, .)
    return chi


def hllc(ul, ur, gamma, D, nd):
    wl_v, pl = prim(ul, gamma, nd)
    wr_v, pr = prim(ur, gamma, nd)
    al = sound(pl, ul[0], gamma)
    ar = sound(pr, ur[0], gamma)
    vl, vr = wl_v[D], wr_v[D]
    q = (gamma + 1.0) / (2.0 * gamma)
    p_pvrs = 0.5 * ((pl + pr) - (0.25 * (vr - vl) * (ur[0] + ul[0]) * (al + ar)))
    p_star = torch.where(0.0 < p_pvrs, p_pvrs, torch.zeros_like(p_pvrs))

    def wavespeed(u, p, s):
        shock = ~(p_star <= p)
        t = torch.where(shock, p_star / p, torch.ones_like(p))
        return torch.where(shock, u + s * torch.sqrt(1.0 + q * (t - 1.0)), u + s)
    wsl = wavespeed(vl, pl, -al)
    wsr = wavespeed(vr, pr, ar)
    m_l = 0.0 <= wsl
    m_r = (~m_l) & (0.0 >= wsr)
    star = ~(m_l | m_r)
    Fl = flux(ul, wl_v, pl, D, nd)
    Fr = flux(ur, wr_v, pr, D, nd)
    factorl = ul[0] * (wsl - vl)
    factorr = ur[0] * (wsr - vr)
    den = factorl - factorr
    den = torch.where(star, den, torch.ones_like(den))
    ws_star = (pr - pl + vl * factorl - vr * factorr) / den
    left = 0.0 <= ws_star
    S = 2 + nd
    st = [torch.where(left, ul[k], ur[k]) for k in range(S)]
    v = [torch.where(left, wl_v[i], wr_v[i]) for i in range(nd)]
    p = torch.where(left, pl, pr)
    ws = torch.where(left, wsl, wsr)
    f = flux(st, v, p, D, nd)
    sf = ws - v[D]
    dd = torch.where(star, ws - ws_star, torch.ones_like(ws))
    scale = sf / dd
    us = [st[k] * scale for k in range(S)]
    us[2 + D] = ws_star * st[0] * scale
    den2 = torch.where(star, st[0] * sf, torch.ones_like(sf))
    us[1] = scale * (st[1] + st[0] * (ws_star - v[D]) * (ws_star + p / den2))
    Fs = [f[k] + (us[k] - st[k]) * ws for k in range(S)]
    return [torch.where(m_l, Fl[k], torch.where(m_r, Fr[k], Fs[k])) for k in range(S)]


def force(ul, ur, gamma, D, nd, dtdh):
    """FORCE flux (flux_method/flux_force.hpp:73-84): the mean of the
    Lax-Friedrichs flux (flux_lax_friedrichs.hpp:62-78) and the Richtmyer flux
    (flux_richtmyer.hpp:78-88), dh/dt = 1 / (dt/dh)."""
    dhdt = 1.0 / dtdh
    S = 2 + nd
    wl_v, pl = prim(ul, gamma, nd)
    wr_v, pr = prim(ur, gamma, nd)
    fl = flux(ul, wl_v, pl, D, nd)
    fr = flux(ur, wr_v, pr, D, nd)
    uri = [(ul[k] + ur[k] + (fl[k] - fr[k]) * dtdh) * 0.5 for k in range(S)]
    wi_v, pi = prim(uri, gamma, nd)
    fri = flux(uri, wi_v, pi, D, nd)
    return [((fl[k] + fr[k] + (ul[k] - ur[k]) * dhdt) * 0.5 + fri[k]) * 0.5 for k in range(S)]


def sweep(U, gamma, dtdh, D, nd, bcs, limiter=0, flux_method="hllc", form="cons"):
    """One split sweep in dimension D (van Leer; form "cons" limits on the
    conservative differences, "prim" takes the limiter's ratio from the
    differences of (rho, p, vel), limiting/limiting_impl.hpp:77-81, 142-146;
    flux_method "hllc" or "force")."""
    assert limiter == 0
    P = pad(U, nd, bcs)
    chi = _chi_min(P, gamma, nd)                     # grid padded by 1 everywhere
    axis = U[0].dim() - 1 - D
    S = 2 + nd

    def cells(t, k):
        """t at offset k along D for cells -1..n along D, interior in the
        other dimension (t padded by PAD everywhere)."""
        out = t
        for d in range(nd):
            ax = t.dim() - 1 - d
            n = t.shape[ax] - 2 * PAD
            if d == D:
                out = out.narrow(ax, PAD - 1 + k, n + 2)
            else:
                out = out.narrow(ax, PAD, n)
        return out

    def chi_cells():
        out = chi
        for d in range(nd):
            ax = chi.dim() - 1 - d
            n = chi.shape[ax] - 2
            if d != D:
                out = out.narrow(ax, 1, n)
        return out
    um = [cells(c, -1) for c in P]
    u0 = [cells(c, 0) for c in P]
    up = [cells(c, 1) for c in P]
    cscale = chi_cells() * 0.5
    half = 0.5 * dtdh
    if form == "prim":
        vm_, pm_ = prim(um, gamma, nd)
        v0_, p0_ = prim(u0, gamma, nd)
        vp_, pp_ = prim(up, gamma, nd)
        bd = [u0[0] - um[0], p0_ - pm_] + [v0_[i] - vm_[i] for i in range(nd)]
        fd = [up[0] - u0[0], pp_ - p0_] + [vp_[i] - v0_[i] for i in range(nd)]
    plus, minus = [], []
    for k in range(S):
        back = u0[k] - um[k]
        fwrd = up[k] - u0[k]
        lim = (back + fwrd) * 0.5
        if form == "prim":
            lim = lim * van_leer(bd[k], fd[k])
        else:
            lim = lim * van_leer(back, fwrd)
        delta = lim * cscale
        plus.append(u0[k] + delta)
        minus.append(u0[k] - delta)
    vm, pm = prim(minus, gamma, nd)
    vp, pp = prim(plus, gamma, nd)
    fm = flux(minus, vm, pm, D, nd)
    fp = flux(plus, vp, pp, D, nd)
    R, L = [], []
    for k in range(S):
        dflux = (fm[k] - fp[k]) * half
        R.append(plus[k] + dflux)
        L.append(minus[k] + dflux)
    n = U[0].shape[axis]
    Rf = [r.narrow(axis, 0, n + 1) for r in R]       # R_i, i = -1..n-1
    Lf = [l.narrow(axis, 1, n + 1) for l in L]       # L_{i+1}
    if flux_method == "force":                       # faces -1/2 .. n-1/2
        G = force(Rf, Lf, gamma, D, nd, dtdh)
    else:
        G = hllc(Rf, Lf, gamma, D, nd)
    out = []
    for k in range(S):
        gm = G[k].narrow(axis, 0, n)
        gp = G[k].narrow(axis, 1, n)
        out.append(U[k] + (gm - gp) * dtdh)
    return out


def max_wavespeed(U, gamma, nd):
    vel, p = prim(U, gamma, nd)
    mv = torch.abs(vel[0])
    for i in range(1, nd):
        a = torch.abs(vel[i])
        mv = torch.where(mv < a, a, mv)
    return (mv + sound(p, U[0], gamma)).max()


def step(U, gamma, cfl, dh, nd, bcs, flux_method="hllc", form="cons"):
    maxws = max_wavespeed(U, gamma, nd)
    dt = cfl * dh / maxws
    dtdh = dt / dh
    X = sweep(U, gamma, dtdh, 0, nd, bcs, flux_method=flux_method, form=form)
    if nd == 1:
        return X
    return sweep(X, gamma, dtdh, 1, nd, bcs, flux_method=flux_method, form=form)


def split(Usoa):
    """(S, ...) tensor -> list of components."""
    return [Usoa[k] for k in range(Usoa.shape[0])]


def join(U):
    return torch.stack(U, 0)
