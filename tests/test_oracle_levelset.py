"""Pins of the level-set / ghost-fluid oracle by what the mathematics fixes.

The reference ships no vectors for this path and its GPU code is undefined in
places, so every function of the oracle is checked against closed forms:
the interface predicates on hand-picked values (levelset/levelset.hpp:166-217),
WENO-HJ5 + Godunov + SSP-RK3 on linear data (exact) and on a smooth
translating profile (high-order convergence to the exact solution), the
eikonal solve on planes (exact distances) and circles (first-order accuracy),
the projection identities, the speed u.n on planes, the ghost-fluid star
state on contacts and uniform states (where HLLC is exact), the extrapolation
of constants and of the 1D star value, and the two-material step on uniform
and translating data."""
from __future__ import annotations

import math

import numpy as np
import pytest

from fluidity_amd import workloads as W


def cons(rho, p, vel, gamma, p_inf=0.0):
    vel = np.atleast_1d(np.asarray(vel, dtype=np.float64))
    e = (p + gamma * p_inf) / ((gamma - 1.0) * rho)
    E = rho * (0.5 * float(np.dot(vel, vel)) + e)
    return np.concatenate([[rho, E], rho * vel])


# ---- predicates ----------------------------------------------------------------

@pytest.mark.parametrize("phi,bits", [
    (-0.5, 1),               # inside, deep
    (-0.1, 1 | 8),           # inside interfacial (phi >= -dh)
    (float(np.nextafter(-0.1, -1.0)), 1),  # one ulp past -dh
    (-0.10000001, 1),        # just past -dh
    (0.0, 1 | 4 | 8),        # on the boundary: inside and inside-interfacial
    (-0.0, 1 | 4 | 8),
    (0.05, 2 | 16),          # outside interfacial (phi < |dh|)
    (0.1, 2),                # phi == dh is not outside-interfacial (strict)
    (0.7, 2),
])
def test_mask_bits(oracle, phi, bits):
    assert oracle.ls_mask_value(phi, 0.1) == bits


def test_masks_grid_matches_scalar(oracle):
    rng = np.random.default_rng(3)
    phi = rng.choice([-0.2, -0.1, -0.05, 0.0, 0.05, 0.1, 0.2], size=(7, 9))
    m = oracle.ls_masks(phi, 0.1)
    ref = np.vectorize(lambda p: oracle.ls_mask_value(p, 0.1))(phi)
    assert np.array_equal(m, ref)


# ---- WENO-HJ5 + Godunov + RK3 ------------------------------------------------------

@pytest.mark.parametrize("slope", [1.3, -0.7])
@pytest.mark.parametrize("v", [0.8, -0.6])
def test_advect_linear_is_exact(oracle, slope, v):
    """phi = a x: every WENO candidate equals a, so |grad phi| = |a| and each
    RK3 stage subtracts dt v |a|: out = phi - dt v |a| (away from the mirrored
    ends, which the three stages reach by 9 cells)."""
    n, dh, dt = 64, 1.0 / 64, 0.4 / 64
    x = (np.arange(n) + 0.5) * dh
    phi = slope * (x - 0.41)
    out = oracle.ls_advect(phi, np.full(n, v), dt, dh)
    ref = phi - dt * v * abs(slope)
    np.testing.assert_allclose(out[12:-12], ref[12:-12], rtol=0, atol=1e-13)


def _translate_error(oracle, n, v, sign):
    """phi(x) = s (x + 0.1 sin 2 pi x) is monotone, so phi_t + v |phi_x| = 0 is
    solved by translation along the direction that v sign(phi_x) picks."""
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    f = lambda z: sign * (z + 0.1 * np.sin(2 * np.pi * z))
    phi = f(x)
    dt = 0.25 * dh
    steps = 4
    V = np.full(n, v)
    for _ in range(steps):
        phi = oracle.ls_advect(phi, V, dt, dh)
    shift = v * steps * dt * sign
    ref = f(x - shift)
    lo, hi = int(0.3 * n), int(0.7 * n)
    return np.max(np.abs(phi[lo:hi] - ref[lo:hi]))


@pytest.mark.parametrize("v,sign", [(0.9, 1.0), (-0.9, 1.0), (0.9, -1.0)])
def test_advect_smooth_profile_converges(oracle, v, sign):
    e1 = _translate_error(oracle, 128, v, sign)
    e2 = _translate_error(oracle, 256, v, sign)
    assert e1 < 2e-7
    assert e1 / e2 > 6.0  # at least third order (RK3 with dt ~ dh)


def test_advect_2d_plane(oracle):
    """A plane phi = n.x - c with |n| = 1 moves with speed v along n:
    phi -> phi - dt v per RK3 step, exact in the interior."""
    n, dh, dt, v = 48, 1.0 / 48, 0.3 / 48, 0.7
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    nx_, ny_ = 0.6, 0.8
    phi = nx_ * X + ny_ * Y - 0.5
    out = oracle.ls_advect(phi, np.full((n, n), v), dt, dh)
    np.testing.assert_allclose(out[12:-12, 12:-12], (phi - dt * v)[12:-12, 12:-12], atol=1e-13)


def test_advect_zero_speed_is_identity(oracle):
    rng = np.random.default_rng(5)
    phi = rng.normal(size=(16, 20))
    out = oracle.ls_advect(phi, np.zeros_like(phi), 0.01, 0.05)
    # 0.75 phi + 0.25 phi and (1/3) phi + (2/3) phi round to within an ulp
    np.testing.assert_allclose(out, phi, rtol=4e-16, atol=0)


# ---- projection --------------------------------------------------------------------

def test_project_identities(oracle):
    rng = np.random.default_rng(7)
    a = rng.normal(size=50)
    b = rng.normal(size=50)
    p0, p1 = oracle.ls_project2(a, b)
    np.testing.assert_allclose(p0, 0.5 * (a - b), atol=1e-15)
    assert np.array_equal(p1, -p0)
    q0, q1 = oracle.ls_project2(a, -a)           # consistent pair: unchanged bitwise
    assert np.array_equal(q0, a) and np.array_equal(q1, -a)
    r0, r1 = oracle.ls_project2(p0, p1)          # idempotent
    assert np.array_equal(r0, p0) and np.array_equal(r1, p1)


# ---- eikonal ---------------------------------------------------------------------------

@pytest.mark.parametrize("band", [3, 6])
@pytest.mark.parametrize("slope", [1.0, -1.0, 2.5])
def test_reinit_1d_distance(oracle, band, slope):
    """Sources (the smaller cell of each sign change) keep |phi|; every other
    in-band cell is the signed source value plus whole cells, so a linear phi
    of unit slope comes back as its exact distance."""
    n, dh = 64, 1.0 / 64
    x = (np.arange(n) + 0.5) * dh
    x0 = 0.4137
    phi = slope * (x - x0)
    out, bm, it = oracle.ls_reinit(phi, dh, band)
    k0 = int(np.floor(x0 / dh - 0.5))
    s_idx = k0 if abs(phi[k0]) < abs(phi[k0 + 1]) else k0 + 1
    signed_src = phi[s_idx]
    # distance along x from the source, measured in its own sign frame
    ref = np.sign(phi) * (np.sign(phi) * signed_src + np.abs(np.arange(n) - s_idx) * dh *
                          np.where(np.sign(phi) == np.sign(signed_src), 1.0, 1.0))
    cross = np.sign(phi) != np.sign(signed_src)
    ref[cross] = np.sign(phi[cross]) * (-abs(signed_src) + np.abs(np.arange(n) - s_idx)[cross] * dh)
    inband = np.abs(ref) <= band * dh
    np.testing.assert_allclose(out[inband], ref[inband], rtol=0, atol=1e-15)
    assert np.all(np.abs(out[~inband]) == (band + 1) * dh)
    assert np.array_equal(bm.astype(bool), inband)
    assert np.all(np.sign(out) == np.sign(phi))
    assert it <= band + 3
    if abs(slope) == 1.0:
        np.testing.assert_allclose(out[inband], phi[inband], rtol=0, atol=1e-15)


def test_reinit_2d_plane_exact(oracle):
    """On linear distance the 2D Godunov update is exact,
    (d - |n_x| h - q)^2 + (d - |n_y| h - q)^2 = h^2 at q = d, and cross-interface
    neighbours enter with their sign, so cells whose upwind chain stays off the
    grid edges get the exact distance."""
    n, dh = 64, 1.0 / 64
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    dist = 0.6 * X + 0.8 * Y - 0.7
    out, bm, _ = oracle.ls_reinit(dist, dh, 6)
    core = np.zeros_like(bm, dtype=bool)
    core[10:-10, 10:-10] = True
    sel = core & (np.abs(dist) <= 5.5 * dh)
    assert sel.sum() > 300
    np.testing.assert_allclose(out[sel], dist[sel], rtol=0, atol=1e-15)


def test_reinit_2d_axis_plane_equals_1d(oracle):
    """Constant in y: the y neighbours equal the cell, |a - b| >= h, and the 2D
    update reduces to the 1D one bit for bit."""
    n, dh = 48, 1.0 / 48
    x = (np.arange(n) + 0.5) * dh
    row = 1.8 * (x - 0.537)
    phi = np.broadcast_to(row, (20, n)).copy()
    out2, bm2, _ = oracle.ls_reinit(phi, dh, 5)
    out1, bm1, _ = oracle.ls_reinit(row, dh, 5)
    assert np.array_equal(out2, np.broadcast_to(out1, out2.shape))
    assert np.array_equal(bm2, np.broadcast_to(bm1, bm2.shape))


def _godunov(a, b, h):
    amb = a - b
    with np.errstate(invalid="ignore", over="ignore"):
        return np.where(np.abs(amb) >= h, np.minimum(a, b) + h,
                        0.5 * (a + b + np.sqrt(2.0 * h * h - amb * amb)))


@pytest.mark.parametrize("shape", ["plane", "circle"])
def test_reinit_2d_fixed_point_and_accuracy(oracle, shape):
    """Jacobi stops at the exact fixed point: every in-band cell that is not a
    source equals the Godunov update of its final neighbours, and the result
    is within a cell of the true distance (first order, unsigned neighbours)."""
    n, dh, band = 96, 1.0 / 96, 6
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    if shape == "plane":
        dist = 0.6 * X + 0.8 * Y - 0.7
    else:
        dist = np.hypot(X - 0.51, Y - 0.47) - 0.23
    out, bm, it = oracle.ls_reinit(dist, dh, band)
    d = np.abs(out)
    inb = bm.astype(bool)
    big = np.finfo(np.float64).max
    sgn = np.sign(dist)
    dd = np.where(inb, d, big)
    pad = np.pad(dd, 1, constant_values=big)
    spad = np.pad(sgn, 1, constant_values=0)

    def nb(sl):
        v = pad[sl]
        return np.where((spad[sl] == sgn) | (v == big), v, -v)
    a = np.minimum(nb((slice(1, -1), slice(0, -2))), nb((slice(1, -1), slice(2, None))))
    b = np.minimum(nb((slice(0, -2), slice(1, -1))), nb((slice(2, None), slice(1, -1))))
    q = _godunov(a, b, dh)
    src = np.zeros_like(inb)
    for sl_c, sl_o in [((slice(None), slice(0, -1)), (slice(None), slice(1, None))),
                       ((slice(0, -1), slice(None)), (slice(1, None), slice(None)))]:
        c, o = dist[sl_c], dist[sl_o]
        src[sl_c] |= (np.sign(c) != np.sign(o)) & (np.abs(c) < np.abs(o))
        src[sl_o] |= (np.sign(c) != np.sign(o)) & (np.abs(o) < np.abs(c))
    assert np.array_equal(d[src], np.abs(dist[src]))
    rest = inb & ~src
    assert np.array_equal(q[rest], d[rest])
    assert np.all(np.sign(out) == sgn)
    core = np.zeros_like(inb)
    core[12:-12, 12:-12] = True
    err = np.max(np.abs(out - dist)[inb & core])
    assert err < (1e-15 if shape == "plane" else 0.2 * dh)
    assert np.all(np.abs(out[~inb]) == (band + 1) * dh)


def test_reinit_far_field_clamped(oracle):
    phi = np.full((12, 12), 5.0)
    out, bm, it = oracle.ls_reinit(phi, 0.1, 4)
    assert np.all(out == 5 * 0.1) and not bm.any() and it == 1


# ---- level-set speed -------------------------------------------------------------------

def test_velocity_plane(oracle):
    n, dh = 32, 1.0 / 32
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    nrm = np.array([0.6, -0.8])
    phi0 = nrm[0] * X + nrm[1] * Y + 0.2
    phi1 = -phi0
    u0 = cons(1.2, 0.9, [0.3, 0.5], 1.4)
    u1 = cons(0.2, 0.9, [-0.4, 0.1], 1.67)
    U0 = np.broadcast_to(u0, (n, n, 4)).copy()
    U1 = np.broadcast_to(u1, (n, n, 4)).copy()
    V = oracle.ls_velocity([phi0, phi1], [U0, U1], dh)
    core = (slice(2, -2), slice(2, -2))
    ref = np.where(phi0 <= 0, 0.3 * 0.6 + 0.5 * -0.8, -(-0.4 * 0.6 + 0.1 * -0.8))
    np.testing.assert_allclose(V[core], ref[core], atol=1e-14)


def test_velocity_flat_levelset_is_zero(oracle):
    phi = np.full((8, 8), -1.0)
    U = np.broadcast_to(cons(1.0, 1.0, [0.5, 0.5], 1.4), (8, 8, 4)).copy()
    V = oracle.ls_velocity([phi, -phi], [U, U], 0.1)
    assert np.all(V == 0.0)


# ---- ghost fluid -----------------------------------------------------------------------

def _line_phi(n, x0):
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    return x - x0, dh


def test_gfm_star_contact_is_left_state(oracle):
    """Equal pressure and velocity on both sides: HLLC resolves the contact, so
    the star state is the material's own (rho, u, p)."""
    n = 40
    phi, dh = _line_phi(n, 0.5013)
    s0 = oracle.scheme(gamma=1.4)
    s1 = oracle.scheme(gamma=1.67)
    U0 = np.broadcast_to(cons(1.0, 2.0, [0.3], 1.4), (n, 3)).copy()
    U1 = np.broadcast_to(cons(0.14, 2.0, [0.3], 1.67), (n, 3)).copy()
    star, has = oracle.gfm_star(phi, -phi, U0, U1, s0, s1, dh)
    idx = np.nonzero(has)[0]
    assert list(idx) == [int(np.floor(0.5013 * n - 0.5))]
    np.testing.assert_allclose(star[idx[0]], U0[0], rtol=1e-14)
    # the other side sees the mirror contact
    star1, has1 = oracle.gfm_star(-phi, phi, U1, U0, s1, s0, dh)
    i1 = np.nonzero(has1)[0]
    assert list(i1) == [idx[0] + 1]
    np.testing.assert_allclose(star1[i1[0]], U1[0], rtol=1e-14)


def test_gfm_star_2d_keeps_tangential_velocity(oracle):
    n = 32
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    nrm = np.array([0.8, 0.6])
    phi = nrm[0] * X + nrm[1] * Y - 0.61
    vel = 0.25 * nrm + 0.7 * np.array([-0.6, 0.8])   # normal 0.25, tangential 0.7
    s = oracle.scheme(gamma=1.4)
    U = np.broadcast_to(cons(1.1, 1.5, vel, 1.4), (n, n, 4)).copy()
    star, has = oracle.gfm_star(phi, -phi, U, U, s, s, dh)
    sel = has.astype(bool)
    sel[:4, :] = sel[-4:, :] = sel[:, :4] = sel[:, -4:] = False
    assert sel.sum() > 10
    np.testing.assert_allclose(star[sel], np.broadcast_to(U[0, 0], star[sel].shape), rtol=1e-13)


def test_gfm_star_against_exact_riemann(oracle):
    """A mild shock tube across the interface: the star pressure of HLLC (via
    the star state) lies near the exact Riemann star pressure."""
    n = 64
    phi, dh = _line_phi(n, 0.5)
    g = 1.4
    s = oracle.scheme(gamma=g)
    wl, wr = (1.0, 1.0, 0.0), (0.8, 0.7, 0.0)   # [rho, p, u]
    U0 = np.broadcast_to(cons(wl[0], wl[1], [wl[2]], g), (n, 3)).copy()
    U1 = np.broadcast_to(cons(wr[0], wr[1], [wr[2]], g), (n, 3)).copy()
    star, has = oracle.gfm_star(phi, -phi, U0, U1, s, s, dh)
    i = int(np.nonzero(has)[0][0])
    Ws = oracle.cons_to_prim(star[i], s)
    p_exact, u_exact = W.star_pressure(wl, wr, g)
    assert abs(Ws[1] - p_exact) / p_exact < 0.02
    assert abs(Ws[2] - u_exact) < 0.02


def test_extrapolate_constant_field(oracle):
    n = 40
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    phi = np.hypot(X - 0.5, Y - 0.45) - 0.2
    u = cons(0.9, 1.3, [0.2, -0.1], 1.4)
    U = np.broadcast_to(u, (n, n, 4)).copy()
    has = np.zeros((n, n), dtype=np.uint8)
    out = oracle.gfm_extrapolate(phi, U, np.zeros_like(U), has, dh, 8)
    np.testing.assert_allclose(out, U, rtol=2e-16 * 4)


def test_extrapolate_1d_carries_star(oracle):
    n = 48
    phi, dh = _line_phi(n, 0.5)
    U = np.zeros((n, 3))
    U[:] = cons(1.0, 1.0, [0.0], 1.4)
    star = np.zeros_like(U)
    has = np.zeros(n, dtype=np.uint8)
    k = int(np.floor(0.5 * n - 0.5))
    has[k] = 1
    star[k] = cons(2.0, 3.0, [0.5], 1.4)
    out = oracle.gfm_extrapolate(phi, U, star, has, dh, 8)
    band = (phi > 0) & (phi <= 6 * dh / math.sqrt(2))
    assert band.sum() == 4
    assert np.array_equal(out[band], np.broadcast_to(star[k], out[band].shape))
    assert np.array_equal(out[~band], U[~band])   # real cells and far field untouched


# ---- two-material step ---------------------------------------------------------------------

def test_mm_uniform_state_stays_uniform(oracle):
    n = 32
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    X, Y = np.meshgrid(x, x)
    phi0 = np.hypot(X - 0.5, Y - 0.5) - 0.2
    s = oracle.scheme(gamma=1.4)
    u = cons(1.0, 1.0, [0.0, 0.0], 1.4)
    U = np.broadcast_to(u, (n, n, 4)).copy()
    U0, U1, p0, p1, t, it, dts = oracle.mm_simulate(U, U, phi0, -phi0, s, s, 0.5, dh,
                                                    max_iters=3)
    assert it == 3 and t > 0
    np.testing.assert_allclose(U0, U, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(U1, U, rtol=1e-14, atol=1e-14)
    assert np.array_equal(p1, -p0)
    assert np.all(np.sign(p0[np.abs(phi0) > dh]) == np.sign(phi0[np.abs(phi0) > dh]))


def test_mm_interface_moves_with_flow(oracle):
    """Uniform flow u through a 1D interface: the zero of phi moves by u t."""
    n = 128
    phi, dh = _line_phi(n, 0.4)
    g = 1.4
    s = oracle.scheme(gamma=g)
    uvel = 0.5
    U = np.broadcast_to(cons(1.0, 1.0, [uvel], g), (n, 3)).copy()
    U0, U1, p0, p1, t, it, dts = oracle.mm_simulate(U, U, phi, -phi, s, s, 0.5, dh,
                                                    max_iters=40)
    i = int(np.nonzero((p0[:-1] <= 0) & (p0[1:] > 0))[0][0])
    x = (np.arange(n) + 0.5) * dh
    xz = x[i] - p0[i] * dh / (p0[i + 1] - p0[i])
    assert abs(xz - (0.4 + uvel * t)) < 0.05 * dh
    np.testing.assert_allclose(U0[20:100], U[20:100], rtol=1e-12)


# ---- n > 2 materials (decision L9) ------------------------------------------

def test_project_n_no_overlap_no_void(oracle):
    """The n-level-set projection (levelset_projection.cuh:87-118): after
    subtracting the average of the two smallest values, the smallest is <= 0
    and every other >= 0 at each cell (no overlap, no void), the order of the
    level sets at a cell is kept, their differences are kept to rounding, and
    cells already projected (two smallest opposite, 0 their average) are
    unchanged bit for bit."""
    rng = np.random.default_rng(5)
    phis = [rng.normal(size=4000) for _ in range(4)]
    out = oracle.ls_project_n(phis)
    P = np.stack(phis)
    Q = np.stack(out)
    order = np.argsort(P, axis=0, kind="stable")
    srt = np.take_along_axis(Q, order, axis=0)
    assert (srt[0] <= 0).all() and (srt[1:] >= 0).all()
    assert np.array_equal(np.argsort(Q, axis=0, kind="stable"), order)
    np.testing.assert_allclose(Q - Q[0], P - P[0], atol=1e-14)
    # projected data: the two smallest are -a and +a, so their average is 0
    a = np.abs(rng.normal(size=500)) + 0.1
    proj = [-a, a, a + 1.0, a + 2.5]
    again = oracle.ls_project_n(proj)
    for p, q in zip(proj, again):
        assert np.array_equal(p, q)


def test_project_n_hand_values(oracle):
    """Hand-worked cells: (3, -1, 2) -> the two smallest are -1 and 2, their
    average 0.5, so (2.5, -1.5, 1.5); (-4, -2, 7, 0) -> average -3, so
    (-1, 1, 10, 3); a tie (1, 1, 5) -> average 1, so (0, 0, 4)."""
    cells = [([3.0, -1.0, 2.0], [2.5, -1.5, 1.5]),
             ([-4.0, -2.0, 7.0, 0.0], [-1.0, 1.0, 10.0, 3.0]),
             ([1.0, 1.0, 5.0], [0.0, 0.0, 4.0])]
    for vals, want in cells:
        out = oracle.ls_project_n([np.array([v]) for v in vals])
        assert [float(q[0]) for q in out] == want


def test_gfm_star_n_two_materials_is_pairwise(oracle):
    """With two materials the n-material star states equal the pairwise
    ones bit for bit (decision L9 reduces to L5)."""
    sb = W.shock_bubble(128, 64, seed=1002)
    s0, s1 = oracle.scheme(gamma=sb.gamma0), oracle.scheme(gamma=sb.gamma1)
    for m, (pm, po, Um, Uo, sm, so) in enumerate([(sb.phi0, sb.phi1, sb.U0, sb.U1, s0, s1),
                                                  (sb.phi1, sb.phi0, sb.U1, sb.U0, s1, s0)]):
        ref, has = oracle.gfm_star(pm, po, Um, Uo, sm, so, sb.dh)
        got, has_n = oracle.gfm_star_n(m, [sb.phi0, sb.phi1], [sb.U0, sb.U1], [s0, s1], sb.dh)
        assert np.array_equal(has, has_n)
        assert np.array_equal(ref[has == 1].view(np.uint64), got[has_n == 1].view(np.uint64))


def test_gfm_star_n_last_other_material_wins(oracle):
    """Three materials where two others qualify at a cell: the star state is
    the pairwise one against the later material (the reference's loop over
    the other materials overwrites)."""
    sbs = W.shock_two_bubbles(96, 48, seed=1007)
    ss = [oracle.scheme(gamma=g) for g in sbs.gammas]
    dh = sbs.dh
    # material 1's level set, with both others outside-interfacial wherever
    # material 1 is inside-interfacial
    phi1 = sbs.phis[1]
    phis = [np.abs(phi1) * 0 + 0.5 * dh, phi1, np.full_like(phi1, 0.5 * dh)]
    star, has = oracle.gfm_star_n(1, phis, sbs.Us, ss, dh)
    ref2, has2 = oracle.gfm_star(phis[1], phis[2], sbs.Us[1], sbs.Us[2], ss[1], ss[2], dh)
    ref0, has0 = oracle.gfm_star(phis[1], phis[0], sbs.Us[1], sbs.Us[0], ss[1], ss[0], dh)
    assert has.any()
    assert np.array_equal(has, has2)
    assert np.array_equal(star[has == 1].view(np.uint64), ref2[has2 == 1].view(np.uint64))
    assert not np.array_equal(star[has == 1].view(np.uint64), ref0[has0 == 1].view(np.uint64))


def test_mm_simulate_n_two_materials_equals_pairwise(oracle):
    """The n-material loop with two materials computes exactly the
    two-material loop (same states, level sets, run time, dts)."""
    sb = W.shock_bubble(96, 48, seed=1002)
    s0, s1 = oracle.scheme(gamma=sb.gamma0), oracle.scheme(gamma=sb.gamma1)
    R = oracle.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, s0, s1, 0.5, sb.dh, max_iters=6,
                           reinit_every=3)
    Us, phis, rt, it, dts = oracle.mm_simulate_n([sb.U0, sb.U1], [sb.phi0, sb.phi1], [s0, s1],
                                                 0.5, sb.dh, max_iters=6, reinit_every=3)
    assert it == R[5] == 6 and rt == R[4]
    assert np.array_equal(dts, R[6])
    for a, b in zip(Us + phis, R[:4]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_mm_simulate_n_three_materials(oracle):
    """Three materials (shock onto two bubbles of different gases): the loop
    runs, states stay finite and positive in density where each material is
    present, the level sets stay projected (exactly one non-positive at each
    cell away from exact ties), and each bubble keeps its area to 5%."""
    sbs = W.shock_two_bubbles(128, 64, seed=1007)
    ss = [oracle.scheme(gamma=g) for g in sbs.gammas]
    Us, phis, rt, it, dts = oracle.mm_simulate_n(sbs.Us, sbs.phis, ss, 0.5, sbs.dh, max_iters=12,
                                                 reinit_every=5)
    assert it == 12 and rt > 0
    for U in Us:
        assert np.isfinite(U).all()
    P = np.stack(phis)
    inside = (P <= 0).sum(axis=0)
    assert (inside <= 1).all()
    assert (inside == 1).mean() > 0.99
    for m in (1, 2):
        a0 = (sbs.phis[m] <= 0).sum()
        assert abs((phis[m] <= 0).sum() - a0) / a0 < 0.05
