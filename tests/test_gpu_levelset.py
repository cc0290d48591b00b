"""Parity of the gfx950 level-set / ghost-fluid path (through the C ABI) with
the CPU oracle: masks, band masks and sweep counts exactly, FP64 fields bit
for bit.  Inputs are the seeded shock-bubble workload at reduced and full
size, and random fields for the single kernels."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from fluidity_amd import workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def L():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from fluidity_amd import levelset
    return levelset


@pytest.fixture(scope="module")
def E():
    from fluidity_amd import euler
    return euler


def _dev(a, layout=0):
    if layout == 1:
        a = np.moveaxis(a, -1, 0)
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _host(t, layout=0):
    a = t.cpu().numpy()
    if layout == 1:
        a = np.moveaxis(a, 0, -1)
    return np.ascontiguousarray(a)


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.dtype == np.float64:
        na, nb = np.isnan(a), np.isnan(b)
        assert np.array_equal(na, nb)
        bad = ~na & (a.view(np.uint64) != b.view(np.uint64))
        assert not bad.any(), f"{bad.sum()} of {a.size} differ; first at {np.argwhere(bad)[0]}"
    else:
        assert np.array_equal(a, b)


def _bubble(nx, ny, stiff=False):
    return W.shock_bubble(nx, ny, seed=1002, stiff=stiff)


def _eos1(sb):
    """Equation of state of the bubble: 0 ideal, 1 the reference's stiffened
    gas, 2 the corrected stiffened gas (W.shock_bubble(..., stiff=True) with
    .corrected set)."""
    if not sb.stiff:
        return 0
    return 2 if getattr(sb, "corrected", False) else 1


def _schemes(O, E, sb):
    eos1 = _eos1(sb)
    os0 = O.scheme(gamma=sb.gamma0)
    os1 = O.scheme(gamma=sb.gamma1, eos=eos1, p_inf=sb.p_inf1)
    es0 = E.Scheme(gamma=sb.gamma0)
    es1 = E.Scheme(gamma=sb.gamma1, eos=eos1, p_inf=sb.p_inf1)
    return os0, os1, es0, es1


def _noisy_phi(shape, seed):
    rng = np.random.default_rng(seed)
    ny, nx = shape if len(shape) == 2 else (1, shape[0])
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / max(ny, 1)
    X, Y = np.meshgrid(x, y)
    phi = 1.3 * (np.hypot(X - 0.45, Y - 0.55) - 0.27) + 0.01 * rng.normal(size=X.shape)
    phi[phi.shape[0] // 2, 3] = 0.0                        # exact zeros
    phi[phi.shape[0] // 3, 5] = -0.0
    return np.ascontiguousarray(phi.reshape(shape))


# ---- single kernels ------------------------------------------------------------

@pytest.mark.parametrize("shape", [(257,), (67, 131)])
def test_masks(L, oracle, shape):
    phi = _noisy_phi(shape, 1)
    dh = 1.0 / shape[-1]
    phi.flat[7] = dh
    phi.flat[11] = -dh
    _same(_host(L.masks(_dev(phi), dh)), oracle.ls_masks(phi, dh))


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("target", [-1, 0, 1])
def test_velocity(L, oracle, layout, target):
    sb = _bubble(192, 96)
    V = L.set_velocities([_dev(sb.phi0), _dev(sb.phi1)],
                         [_dev(sb.U0, layout), _dev(sb.U1, layout)], sb.dh, target,
                         layout=layout)
    _same(_host(V), oracle.ls_velocity([sb.phi0, sb.phi1], [sb.U0, sb.U1], sb.dh, target))


@pytest.mark.parametrize("shape", [(300,), (70, 90)])
def test_advect(L, oracle, shape):
    phi = _noisy_phi(shape, 2)
    rng = np.random.default_rng(3)
    V = rng.uniform(-1.0, 1.0, size=shape)
    dh = 1.0 / shape[-1]
    dt = 0.3 * dh
    out = L.evolve(_dev(phi), _dev(V), dt, dh)
    _same(_host(out), oracle.ls_advect(phi, V, dt, dh))


def test_project(L, oracle):
    rng = np.random.default_rng(4)
    a, b = rng.normal(size=(40, 50)), rng.normal(size=(40, 50))
    ta, tb = _dev(a), _dev(b)
    L.project(ta, tb)
    ra, rb = oracle.ls_project2(a, b)
    _same(_host(ta), ra)
    _same(_host(tb), rb)


@pytest.mark.parametrize("band", [3, 6])
@pytest.mark.parametrize("shape", [(300,), (80, 96)])
def test_reinit(L, oracle, shape, band):
    phi = _noisy_phi(shape, 5)
    dh = 1.0 / shape[-1]
    out, bm, it = L.eikonal(_dev(phi), dh, band)
    ro, rbm, rit = oracle.ls_reinit(phi, dh, band)
    assert it == rit
    _same(_host(out), ro)
    _same(_host(bm), rbm)


def test_reinit_band_covers_grid(L, oracle):
    """A band wider than the grid: every cell reached, many sweeps (several
    convergence tests on the host)."""
    phi = _noisy_phi((48, 40), 6)
    dh = 1.0 / 40
    out, bm, it = L.eikonal(_dev(phi), dh, 200)
    ro, rbm, rit = oracle.ls_reinit(phi, dh, 200)
    assert it == rit and it > 16
    assert rbm.all()
    _same(_host(out), ro)


@pytest.mark.parametrize("stiff", [False, True])
@pytest.mark.parametrize("layout", [0, 1])
def test_gfm_star_and_extrapolate(L, E, oracle, stiff, layout):
    sb = _bubble(192, 96, stiff)
    os0, os1, es0, es1 = _schemes(oracle, E, sb)
    U0, U1 = _dev(sb.U0, layout), _dev(sb.U1, layout)
    p0, p1 = _dev(sb.phi0), _dev(sb.phi1)
    for (pm, po, Um, Uo, em, eo, npm, npo, nUm, nUo, om, oo) in [
            (p0, p1, U0, U1, es0, es1, sb.phi0, sb.phi1, sb.U0, sb.U1, os0, os1),
            (p1, p0, U1, U0, es1, es0, sb.phi1, sb.phi0, sb.U1, sb.U0, os1, os0)]:
        star, has = L.riemann_gfm_star(pm, po, Um, Uo, em, eo, sb.dh, layout)
        rstar, rhas = oracle.gfm_star(npm, npo, nUm, nUo, om, oo, sb.dh)
        _same(_host(has), rhas)
        assert rhas.sum() > 50
        _same(_host(star).reshape(rstar.shape), rstar)
        L.extrapolate(pm, Um, star, has, sb.dh, 8, layout)
        ref = oracle.gfm_extrapolate(npm, nUm, rstar, rhas, sb.dh, 8)
        _same(_host(Um, layout), ref)


def test_gfm_1d(L, E, oracle):
    n = 200
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    phi0 = np.ascontiguousarray(x - 0.4137)
    phi1 = -phi0
    U0 = W.prim_to_cons(np.broadcast_to(np.array([1.0, 2.0, 0.3]), (n, 3)).copy(), 1.4)
    U1 = W.prim_to_cons(np.broadcast_to(np.array([0.2, 0.5, -0.2]), (n, 3)).copy(), 1.67)
    os0, os1 = oracle.scheme(gamma=1.4), oracle.scheme(gamma=1.67)
    es0, es1 = E.Scheme(gamma=1.4), E.Scheme(gamma=1.67)
    star, has = L.riemann_gfm_star(_dev(phi0), _dev(phi1), _dev(U0), _dev(U1), es0, es1, dh)
    rstar, rhas = oracle.gfm_star(phi0, phi1, U0, U1, os0, os1, dh)
    _same(_host(has), rhas)
    assert rhas.sum() == 1
    _same(_host(star).reshape(rstar.shape), rstar)


def test_combine(L, oracle):
    sb = _bubble(160, 80)
    out = L.combine_into_first(_dev(sb.phi0), _dev(sb.U0), _dev(sb.U1))
    _same(_host(out), oracle.mm_combine(sb.phi0, sb.U0, sb.U1))


# ---- the two-material step ---------------------------------------------------------

def _run_mm(L, E, sb, steps, layout=0, band=6, sweeps=8, every=1, dtype=None, numerics=0):
    es0, es1 = _schemes_e(E, sb)
    es0.numerics = es1.numerics = numerics
    sim = L.MultimaterialSimulator(layout=layout)
    U0, U1 = _dev(sb.U0, layout), _dev(sb.U1, layout)
    if dtype is not None:
        U0, U1 = U0.to(dtype), U1.to(dtype)
    p0, p1 = _dev(sb.phi0), _dev(sb.phi1)
    sim.add_material(U0, p0, es0)
    sim.add_material(U1, p1, es1)
    r = sim.simulate_mm(0.5, sb.dh, max_iters=steps, band=band, extrap_sweeps=sweeps,
                        reinit_every=every)
    return r, _host(U0, layout), _host(U1, layout), _host(p0), _host(p1)


def _schemes_e(E, sb):
    eos1 = _eos1(sb)
    return E.Scheme(gamma=sb.gamma0), E.Scheme(gamma=sb.gamma1, eos=eos1, p_inf=sb.p_inf1)


@pytest.mark.parametrize("stiff,layout,every", [(False, 0, 1), (False, 1, 1), (True, 0, 1),
                                               (False, 0, 5)])
def test_mm_simulate_bitwise(L, E, oracle, stiff, layout, every):
    sb = _bubble(256, 128, stiff)
    os0, os1, _, _ = _schemes(oracle, E, sb)
    steps = 12
    r, U0, U1, p0, p1 = _run_mm(L, E, sb, steps, layout, every=every)
    R = oracle.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, os0, os1, 0.5, sb.dh,
                           max_iters=steps, reinit_every=every)
    assert r.iters == R[5] == steps
    assert r.run_time == R[4]
    assert r.last_dt == R[6][-1]
    _same(U0, R[0])
    _same(U1, R[1])
    _same(p0, R[2])
    _same(p1, R[3])


@pytest.mark.parametrize("flux1,lim1,form1", [(1, 0, 0), (0, 1, 1), (0, 1, 0)],
                         ids=["force-fallback", "superbee-prim", "superbee"])
def test_mm_simulate_mixed_schemes(L, E, oracle, flux1, lim1, form1):
    """Materials with different schemes: with equal flux methods both grids
    share each sweep launch (each with its own limiter, form and EOS); with
    different ones each grid is swept on its own.  Bit for bit either way."""
    sb = _bubble(192, 96)
    os0 = oracle.scheme(gamma=sb.gamma0)
    os1 = oracle.scheme(gamma=sb.gamma1, flux=flux1, limiter=lim1, form=form1)
    es0 = E.Scheme(gamma=sb.gamma0)
    es1 = E.Scheme(gamma=sb.gamma1, flux=flux1, limiter=lim1, form=form1)
    steps = 6
    sim = L.MultimaterialSimulator(layout=1)
    U0, U1 = _dev(sb.U0, 1), _dev(sb.U1, 1)
    p0, p1 = _dev(sb.phi0), _dev(sb.phi1)
    sim.add_material(U0, p0, es0)
    sim.add_material(U1, p1, es1)
    r = sim.simulate_mm(0.5, sb.dh, max_iters=steps, reinit_every=2)
    R = oracle.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, os0, os1, 0.5, sb.dh,
                           max_iters=steps, reinit_every=2)
    assert r.iters == R[5] == steps
    _same(_host(U0, 1), R[0])
    _same(_host(U1, 1), R[1])
    _same(_host(p0), R[2])
    _same(_host(p1), R[3])


def test_mm_simulate_mixed_layouts(L, E, oracle):
    """fl_mm_simulate through the C ABI with material 0 in AoS and material 1
    in SoA (each grid keeps its own strides in every joint launch)."""
    import ctypes as C
    from fluidity_amd import native as N
    from fluidity_amd.euler import grid_field
    sb = _bubble(160, 96)
    os0, os1, es0, es1 = _schemes(oracle, E, sb)
    steps = 5
    U0, U1 = _dev(sb.U0, 0), _dev(sb.U1, 1)
    W0, W1 = torch.empty_like(U0), torch.empty_like(U1)
    p0, p1 = _dev(sb.phi0), _dev(sb.phi1)
    mats = (N.Material * 2)()
    for m, (U, Wk, p, sc, lay) in enumerate([(U0, W0, p0, es0, N.FL_LAYOUT_AOS),
                                            (U1, W1, p1, es1, N.FL_LAYOUT_SOA)]):
        mats[m].state = grid_field(U, lay)
        mats[m].work = grid_field(Wk, lay)
        mats[m].phi = p.data_ptr()
        mats[m].scheme = sc.c()
    n = (C.c_int64 * 2)(sb.phi0.shape[1], sb.phi0.shape[0])
    wbytes = N.lib.fl_mm_work_bytes(2, n)
    work = torch.empty(wbytes, dtype=torch.uint8, device="cuda")
    ctx = N.Context(0)
    p = N.SimParams(0.5, sb.dh, float("inf"), steps)
    mp = N.MmParams(6, 8, 2, 0)
    r = N.SimResult()
    N.check(N.lib.fl_mm_simulate(ctx.handle, mats, 2, C.byref(p), C.byref(mp), work.data_ptr(),
                                 wbytes, C.byref(r), None))
    torch.cuda.synchronize()
    R = oracle.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, os0, os1, 0.5, sb.dh,
                           max_iters=steps, reinit_every=2)
    assert r.iters == R[5] == steps
    _same(_host(U0, 0), R[0])
    _same(_host(U1, 1), R[1])
    _same(_host(p0), R[2])
    _same(_host(p1), R[3])


@pytest.mark.parametrize("layout", [0, 1])
def test_mm_simulate_1d_bitwise(L, E, oracle, layout):
    """The two-material loop on a 1D shock tube (every pass in its 1D form:
    joint sweeps and maximum, star states, band seeds, RK3, eikonal), 12
    steps with reinitialisation every 3, bit for bit against the oracle."""
    n = 300
    dh = 1.0 / n
    x = (np.arange(n) + 0.5) * dh
    phi0 = np.ascontiguousarray(x - 0.4137)
    phi1 = -phi0
    U0 = W.prim_to_cons(np.broadcast_to(np.array([1.0, 2.0, 0.3]), (n, 3)).copy(), 1.4)
    U1 = W.prim_to_cons(np.broadcast_to(np.array([0.2, 0.5, -0.2]), (n, 3)).copy(), 1.67)
    os0, os1 = oracle.scheme(gamma=1.4), oracle.scheme(gamma=1.67)
    es0, es1 = E.Scheme(gamma=1.4), E.Scheme(gamma=1.67)
    sim = L.MultimaterialSimulator(layout=layout)
    d0, d1 = _dev(U0, layout), _dev(U1, layout)
    p0, p1 = _dev(phi0), _dev(phi1)
    sim.add_material(d0, p0, es0)
    sim.add_material(d1, p1, es1)
    steps = 12
    r = sim.simulate_mm(0.5, dh, max_iters=steps, reinit_every=3)
    R = oracle.mm_simulate(U0, U1, phi0, phi1, os0, os1, 0.5, dh, max_iters=steps,
                           reinit_every=3)
    assert r.iters == R[5] == steps
    assert r.run_time == R[4]
    _same(_host(d0, layout), R[0])
    _same(_host(d1, layout), R[1])
    _same(_host(p0), R[2])
    _same(_host(p1), R[3])


def test_mm_simulate_full_size(L, E, oracle):
    """The 2048 x 1024 shock-bubble workload: six steps bit for bit against
    the oracle, then more steps checked through size-independent properties
    (finite states, phi1 == -phi0 exactly after projection + reinitialisation,
    the bubble's area conserved to first order)."""
    sb = _bubble(2048, 1024)
    os0, os1, _, _ = _schemes(oracle, E, sb)
    r, U0, U1, p0, p1 = _run_mm(L, E, sb, 6)
    R = oracle.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, os0, os1, 0.5, sb.dh,
                           max_iters=6, nthreads=16)
    assert r.iters == R[5] == 6
    _same(U0, R[0])
    _same(U1, R[1])
    _same(p0, R[2])
    _same(p1, R[3])
    r, U0, U1, p0, p1 = _run_mm(L, E, sb, 20)
    assert r.iters == 20
    assert np.isfinite(U0).all() and np.isfinite(U1).all()
    assert np.array_equal(p1, -p0)
    area0 = (sb.phi1 <= 0).sum()
    area = (p1 <= 0).sum()
    assert abs(area - area0) / area0 < 0.05


@pytest.mark.parametrize("layout", [0, 1])
def test_mm_simulate_corrected_stiff_bubble(L, E, oracle, layout):
    """The shock-bubble workload's stiffened-gas variant (gamma 4.4,
    p_inf 6e8, water-like bubble at pressure equilibrium with the ambient
    gas) with the corrected stiffened gas: the whole two-material loop
    (reinitialisation every 5 steps) bit-identical to the oracle's corrected
    loop, and the bubble's pressure stays at the ambient 1e5 where the
    reference's pressure would read p + gamma p_inf."""
    import dataclasses
    sb = dataclasses.replace(_bubble(256, 128, True))
    sb.corrected = True
    os0, os1, _, _ = _schemes(oracle, E, sb)
    assert os1.eos == 2
    steps = 10
    r, U0, U1, p0, p1 = _run_mm(L, E, sb, steps, layout, every=5)
    R = oracle.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, os0, os1, 0.5, sb.dh,
                           max_iters=steps, reinit_every=5)
    assert r.iters == R[5] == steps
    assert r.run_time == R[4]
    _same(U0, R[0])
    _same(U1, R[1])
    _same(p0, R[2])
    _same(p1, R[3])
    inside = sb.phi1 < -4 * sb.dh
    p_bubble = np.array([oracle.pressure(u, os1) for u in sb.U1[inside][:50]])
    np.testing.assert_allclose(p_bubble, 1e5, rtol=1e-6)


@pytest.mark.parametrize("nx,ny,steps", [(256, 128, 12), (2048, 1024, 20)])
def test_mm_simulate_fp32_bulk(L, E, nx, ny, steps):
    """FP32 bulk storage on the two-material loop (FP32 sweeps; dt, the
    wavespeed maximum, level sets, ghost-fluid states and the eikonal solve in
    FP64) against the FP64 loop, which equals the oracle bit for bit
    (test_mm_simulate_bitwise, test_mm_simulate_full_size): the relative L1
    and L-inf error of every conserved field of both materials (momentum
    components relative to the momentum magnitude) and of both level sets is
    printed; the bound is 1e-5 relative L1 on density.  Steps reinitialise
    every 5 steps, as the shock-bubble workload does."""
    sb = _bubble(nx, ny)
    ref = _run_mm(L, E, sb, steps, layout=1, every=5)
    got = _run_mm(L, E, sb, steps, layout=1, every=5, dtype=torch.float32)
    assert got[0].iters == ref[0].iters == steps
    names = ["rho", "E", "m_x", "m_y"]
    worst = 0.0
    for m, (g, r) in enumerate([(got[1], ref[1]), (got[2], ref[2])]):
        g = np.asarray(g, dtype=np.float64)
        # momentum components relative to the momentum magnitude: m_y is zero
        # bit for bit in most of the planar-shock domain
        mom = np.hypot(r[..., 2], r[..., 3])
        for k in range(4):
            d = np.abs(g[..., k] - r[..., k])
            scale = np.abs(r[..., k]) if k < 2 else mom
            l1 = float(d.sum() / scale.sum())
            linf = float(d.max() / scale.max())
            print(f"fp32 {nx}x{ny} material {m} {names[k]}: rel L1 {l1:.3e}, rel Linf {linf:.3e}")
            if k == 0:
                worst = max(worst, l1)
    for m, (g, r) in enumerate([(got[3], ref[3]), (got[4], ref[4])]):
        d = np.abs(g - r)
        print(f"fp32 {nx}x{ny} phi{m}: rel L1 {d.sum() / np.abs(r).sum():.3e}, "
              f"max |diff| / dh {d.max() / sb.dh:.3e}, sign changes {(np.sign(g) != np.sign(r)).sum()}")
    assert np.isfinite(got[1]).all() and np.isfinite(got[2]).all()
    assert abs(got[0].run_time - ref[0].run_time) <= 1e-5 * ref[0].run_time
    assert worst <= 1e-5, worst


@pytest.mark.parametrize("nx,ny,steps,stiff", [(256, 128, 12, False), (256, 128, 12, True),
                                               (2048, 1024, 20, False)])
def test_mm_simulate_fast_numerics(L, E, nx, ny, steps, stiff):
    """FL_NUMERICS_FAST on the two-material loop (both materials' sweeps in
    their fast forms, refined reciprocals and square root in the level sets'
    WENO-HJ5 / Godunov right-hand side; star states, extrapolation, dt and the
    eikonal solve keep the exact numerics) against the exact loop, which
    equals the oracle bit for bit (test_mm_simulate_bitwise,
    test_mm_simulate_full_size): every conserved field of both materials and
    both level sets within 1e-12 relative L-inf, the same step count, the run
    time within 1e-12 and no level-set sign change.  Relative L-inf is
    max |diff| / max |exact| over both materials' grids, for the momentum
    components max |diff| / the largest momentum magnitude: the bubble is at
    rest until the shock reaches it and the planar shock has no y momentum,
    so those fields hold only rounding-level values of either sign.
    Reinitialisation every 5 steps, as the shock-bubble workload."""
    sb = _bubble(nx, ny, stiff)
    ref = _run_mm(L, E, sb, steps, layout=1, every=5)
    got = _run_mm(L, E, sb, steps, layout=1, every=5, numerics=1)
    assert got[0].iters == ref[0].iters == steps
    assert abs(got[0].run_time - ref[0].run_time) <= 1e-12 * ref[0].run_time
    rs = [np.asarray(ref[1]), np.asarray(ref[2])]
    mom = max(float(np.hypot(r[..., 2], r[..., 3]).max()) for r in rs)
    scale = [max(float(np.abs(r[..., 0]).max()) for r in rs),
             max(float(np.abs(r[..., 1]).max()) for r in rs), mom, mom]
    errs = {}
    for m in (1, 2):
        g, r = np.asarray(got[m]), rs[m - 1]
        for k, name in enumerate(["rho", "E", "m_x", "m_y"]):
            errs[f"U{m - 1} {name}"] = float(np.abs(g[..., k] - r[..., k]).max() / scale[k])
    for m in (3, 4):
        g, r = np.asarray(got[m]), np.asarray(ref[m])
        errs[f"phi{m - 3}"] = float(np.abs(g - r).max() / np.abs(r).max())
        assert (np.sign(g) == np.sign(r)).all(), f"phi{m - 3} changes sign"
    print(f"fast {nx}x{ny} stiff={stiff}: rel Linf {errs}", file=sys.stderr)
    assert max(errs.values()) <= 1e-12, errs


@pytest.mark.parametrize("layout,every", [(0, 1), (1, 5)])
def test_mm_simulate_three_materials_bitwise(L, E, oracle, layout, every):
    """Three materials (a shock onto two bubbles of different gases, seed
    1007): the n-material loop (star states from the last qualifying other
    material, n-level-set projection) bit-identical to the oracle's
    n-material loop, states, level sets, run time and step count."""
    sbs = W.shock_two_bubbles(192, 96, seed=1007)
    ss = [oracle.scheme(gamma=g) for g in sbs.gammas]
    steps = 10
    Us, phis, rt, it, dts = oracle.mm_simulate_n(sbs.Us, sbs.phis, ss, 0.5, sbs.dh,
                                                 max_iters=steps, reinit_every=every)
    sim = L.MultimaterialSimulator(layout=layout)
    dev_U = [_dev(U, layout) for U in sbs.Us]
    dev_p = [_dev(p) for p in sbs.phis]
    for U, p, g in zip(dev_U, dev_p, sbs.gammas):
        sim.add_material(U, p, E.Scheme(gamma=g))
    r = sim.simulate_mm(0.5, sbs.dh, max_iters=steps, reinit_every=every)
    assert r.iters == it == steps
    assert r.run_time == rt
    for m in range(3):
        _same(_host(dev_U[m], layout), Us[m])
        _same(_host(dev_p[m]), phis[m])


def test_mm_simulate_three_materials_fast_numerics(L, E, oracle):
    """The n-material loop in FL_NUMERICS_FAST (three materials, seed 1007)
    against the oracle's exact n-material loop: every field of every
    material and every level set within 1e-12 relative L-inf (momentum
    against the largest momentum magnitude, as in the two-material test), the
    same step count, run time within 1e-12, no level-set sign change."""
    sbs = W.shock_two_bubbles(192, 96, seed=1007)
    ss = [oracle.scheme(gamma=g) for g in sbs.gammas]
    steps = 10
    Us, phis, rt, it, _ = oracle.mm_simulate_n(sbs.Us, sbs.phis, ss, 0.5, sbs.dh,
                                               max_iters=steps, reinit_every=5)
    sim = L.MultimaterialSimulator(layout=1)
    dev_U = [_dev(U, 1) for U in sbs.Us]
    dev_p = [_dev(p) for p in sbs.phis]
    for U, p, g in zip(dev_U, dev_p, sbs.gammas):
        sim.add_material(U, p, E.Scheme(gamma=g, numerics=1))
    r = sim.simulate_mm(0.5, sbs.dh, max_iters=steps, reinit_every=5)
    assert r.iters == it == steps
    assert abs(r.run_time - rt) <= 1e-12 * rt
    mom = max(float(np.hypot(U[..., 2], U[..., 3]).max()) for U in Us)
    scale = [max(float(np.abs(U[..., 0]).max()) for U in Us),
             max(float(np.abs(U[..., 1]).max()) for U in Us), mom, mom]
    worst = 0.0
    for m in range(3):
        g = _host(dev_U[m], 1)
        for k in range(4):
            worst = max(worst, float(np.abs(g[..., k] - Us[m][..., k]).max()) / scale[k])
        gp = _host(dev_p[m])
        worst = max(worst, float(np.abs(gp - phis[m]).max() / np.abs(phis[m]).max()))
        assert (np.sign(gp) == np.sign(phis[m])).all()
    print(f"fast three materials: worst rel Linf {worst:.3e}", file=sys.stderr)
    assert worst <= 1e-12, worst
