"""Parity of the gfx950 path (through the C ABI) with the CPU oracle.

FP64 results must be bit-identical to the oracle (stricter than the 1e-12
relative L-inf bound); FP32 results are checked against the FP64 oracle with
the tolerance written in each test.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from tests.fields import cons_field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def E():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from fluidity_amd import euler
    return euler


@pytest.fixture(scope="module")
def N():
    from fluidity_amd import native
    return native


def _dev(U, dtype=torch.float64, layout=0):
    if layout == 1:
        U = np.ascontiguousarray(np.moveaxis(U, -1, 0))
    return torch.from_numpy(np.ascontiguousarray(U)).to(device="cuda", dtype=dtype)


def _host(T, layout=0):
    a = T.double().cpu().numpy()
    if layout == 1:
        a = np.ascontiguousarray(np.moveaxis(a, 0, -1))
    return a


def _oscheme(O, lim, form, flux, bc, gamma=1.4, eos=0, p_inf=0.0):
    return O.scheme(limiter=lim, form=form, flux=flux, eos=eos, gamma=gamma, p_inf=p_inf, bc=bc)


def _escheme(E, lim, form, flux, bc, gamma=1.4, eos=0, p_inf=0.0):
    return E.Scheme(limiter=lim, form=form, flux=flux, eos=eos, gamma=gamma, p_inf=p_inf,
                    bc=tuple(bc))


def _bitwise(a, b):
    """Bit-identical, except that NaNs only need to sit at the same places
    (x86 and gfx950 produce different NaN payloads)."""
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64))


def _report(a, b):
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return f"NaN positions differ: gpu {na.sum()} NaNs, oracle {nb.sum()} NaNs"
    d = np.where(na, 0.0, np.abs(a - b))
    i = np.unravel_index(np.argmax(d), d.shape)
    nbad = int((a[~na].view(np.uint64) != b[~nb].view(np.uint64)).sum())
    return f"{nbad} values differ; max |diff| {d.max():.3e} at {i}: gpu {a[i]!r} oracle {b[i]!r}"


CASES_1D = [
    # kind, n, limiter, form, flux, bc, layout
    ("rough", 1000, 0, 0, 0, (0,), 0),
    ("rough", 1000, 1, 0, 1, (1,), 1),
    ("fast", 777, 0, 1, 0, (0,), 0),
    ("rough", 640, 1, 1, 0, (2,), 1),
    ("smooth", 300, 1, 1, 1, (0,), 1),
    ("riemann", 4096, 0, 0, 0, (0,), 0),
    ("rough", 900, 0, 0, 2, (0,), 1),     # Lax-Friedrichs
    ("fast", 700, 1, 1, 3, (1,), 0),      # Richtmyer
]


@pytest.mark.parametrize("case", CASES_1D, ids=lambda c: "-".join(map(str, c[:3])))
def test_sweep_1d_bitwise(E, oracle, case):
    kind, n, lim, form, flux, bc, layout = case
    U = cons_field(kind, (n,), seed=11 + n)
    os_ = _oscheme(oracle, lim, form, flux, bc)
    dtdh = 0.4 / oracle.max_wavespeed_grid(U, os_)
    ref = oracle.sweep(U, os_, dtdh, 0)
    solver = E.SplitSolver(_escheme(E, lim, form, flux, bc), layout=layout)
    for xb in (64, 256):
        solver.set_tuning(x_block=xb)
        a = _dev(U, layout=layout)
        b = torch.empty_like(a)
        solver.sweep(a, b, 0, dtdh)
        torch.cuda.synchronize()
        got = _host(b, layout)
        assert _bitwise(got, ref), _report(got, ref)


CASES_2D = [
    # kind, (ny, nx), limiter, form, flux, bc, layout
    ("rough", (45, 70), 0, 0, 0, (0, 0), 0),
    ("rough", (37, 130), 1, 1, 1, (1, 0), 1),
    ("fast", (64, 64), 0, 0, 0, (0, 1), 1),
    ("fast", (50, 64), 1, 0, 0, (0, 0), 0),
    ("smooth", (33, 90), 0, 1, 1, (2, 2), 0),
    ("rough", (20, 300), 0, 0, 0, (1, 1), 0),
    ("rough", (41, 66), 0, 0, 2, (0, 1), 1),    # Lax-Friedrichs
    ("fast", (40, 72), 1, 1, 3, (2, 0), 0),     # Richtmyer
]

TUNINGS = [(256, 32, 256, 64), (64, 1, 64, 1), (128, 7, 192, 5)]


@pytest.mark.parametrize("case", CASES_2D, ids=lambda c: "-".join(map(str, c[:3])))
@pytest.mark.parametrize("dim", [0, 1])
def test_sweep_2d_bitwise(E, oracle, case, dim):
    kind, shape, lim, form, flux, bc, layout = case
    U = cons_field(kind, shape, seed=7 + shape[0] * shape[1])
    os_ = _oscheme(oracle, lim, form, flux, bc)
    dtdh = 0.4 / oracle.max_wavespeed_grid(U, os_)
    ref = oracle.sweep(U, os_, dtdh, dim)
    solver = E.SplitSolver(_escheme(E, lim, form, flux, bc), layout=layout)
    for xb, xr, yb, yr in TUNINGS:
        solver.set_tuning(x_block=xb, x_rows=xr, y_block=yb, y_rows=yr)
        a = _dev(U, layout=layout)
        b = torch.empty_like(a)
        solver.sweep(a, b, dim, dtdh)
        torch.cuda.synchronize()
        got = _host(b, layout)
        assert _bitwise(got, ref), f"tuning {(xb, xr, yb, yr)}: " + _report(got, ref)


def test_max_wavespeed_bitwise(E, oracle):
    for shape in [(1000,), (45, 70)]:
        U = cons_field("rough", shape, seed=3)
        os_ = _oscheme(oracle, 0, 0, 0, (0, 0))
        ref = oracle.max_wavespeed_grid(U, os_)
        solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0, 0)))
        got = solver.max_wavespeed(_dev(U))
        assert got == ref


def test_simulate_2d_bitwise(E, oracle):
    """The device-resident step loop (fused CFL, device dt) against the
    oracle's simulator loop: final state, run time and step count."""
    U = cons_field("riemann", (96, 96), seed=1001)
    for flux in (0, 1):
        os_ = _oscheme(oracle, 0, 0, flux, (0, 0))
        ref, rt, it, _ = oracle.simulate(U, os_, 0.9, 1.0 / 96, max_iters=40)
        solver = E.SplitSolver(_escheme(E, 0, 0, flux, (0, 0)), layout=1)
        a = _dev(U, layout=1)
        w = torch.empty_like(a)
        r = solver.simulate(a, w, 0.9, 1.0 / 96, max_iters=40)
        got = _host(a, 1)
        assert r.iters == it
        assert r.run_time == rt
        assert _bitwise(got, ref), _report(got, ref)


@pytest.mark.parametrize("k", [-900, -400, 400, 900])
def test_simulate_2d_scaled_states_bitwise(E, oracle, k):
    """Density, momenta and energy scaled by 2^k (exact: velocities and
    wavespeeds unchanged, pressures scaled by 2^k): the step loop stays
    bit-identical to the oracle, so the FP64 division and square root agree
    with IEEE ones for operands far from 1 as well."""
    U = cons_field("riemann", (64, 64), seed=1001) * (2.0 ** k)
    for flux in (0, 1):
        os_ = _oscheme(oracle, 0, 0, flux, (0, 0))
        ref, rt, it, _ = oracle.simulate(U, os_, 0.9, 1.0 / 64, max_iters=20)
        solver = E.SplitSolver(_escheme(E, 0, 0, flux, (0, 0)), layout=1)
        a = _dev(U, layout=1)
        r = solver.simulate(a, torch.empty_like(a), 0.9, 1.0 / 64, max_iters=20)
        got = _host(a, 1)
        assert r.iters == it
        assert r.run_time == rt
        assert _bitwise(got, ref), _report(got, ref)


def test_division_window_and_edges(E):
    """The exact FP64 division (fl_numerics.hpp dv: v_rcp_f64, two Newton
    steps, a residual correction and v_div_fixup, without the IEEE
    sequence's operand rescaling) through the primitive-planes kernel, whose
    velocity plane is one quotient m / rho per cell: bit-identical to IEEE
    m / rho for divisors from 2^-900 to 2^900 with normal quotients and for
    divisors of 2^1023 and 1.5 * 2^1023, and the recorded outcome outside
    the window: a denormal divisor (2^-1060, 2^-1030; its reciprocal
    overflows) gives +inf where IEEE gives 0.5.  No state of a physical run
    comes near these magnitudes."""
    from fluidity_amd import simulator as SM
    rng = np.random.default_rng(11)
    inside = np.concatenate([2.0 ** rng.uniform(-900, 900, 4000), [2.0 ** -900, 2.0 ** 900,
                                                                  1.0, 3.0, 0.1, 7.0]])
    quot = rng.uniform(0.25, 4.0, inside.size) * np.where(rng.random(inside.size) < 0.5, -1, 1)
    edge_rho = np.array([2.0 ** -1060, 2.0 ** -1030, 2.0 ** 1023, 1.5 * 2.0 ** 1023])
    rho = np.concatenate([inside, edge_rho])
    m = np.concatenate([quot * inside, 0.5 * edge_rho])
    U = np.stack([rho, np.ones_like(rho), m], axis=-1)   # [rho, E, m]: 1D AoS
    prim = SM.primitive_planes(_dev(U), E.Scheme(), 0)
    vel = prim[2].cpu().numpy()
    ieee = m / rho
    k = inside.size
    assert _bitwise(vel[:k], ieee[:k]), _report(vel[:k], ieee[:k])
    got = vel[k:]
    print("dv outside the window (rho, GPU, IEEE):",
          [(float(r), float(g), float(q)) for r, g, q in zip(edge_rho, got, ieee[k:])],
          file=sys.stderr)
    assert np.all(ieee[k:] == 0.5)
    assert np.all(got[:2] == np.inf)
    assert np.all(got[2:] == 0.5)


def test_simulate_1d_config0_bitwise(E, oracle):
    """The 1D Riemann workload (4096 cells, run until the fastest wave has
    travelled 0.4) reproduces the oracle bit for bit, including the step
    count and the accumulated run time (no final-dt clip)."""
    from fluidity_amd import workloads as W
    r1 = W.riemann_1d()
    os_ = _oscheme(oracle, 0, 0, 0, (0,))
    ref, rt, it, _ = oracle.simulate(r1.U, os_, 0.9, r1.dh, sim_time=r1.t_end, max_iters=100000)
    solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0,)))
    a = _dev(r1.U)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.9, r1.dh, sim_time=r1.t_end)
    got = _host(a)
    assert r.iters == it and r.run_time == rt
    assert _bitwise(got, ref), _report(got, ref)


def test_solve_swap_semantics(E, oracle):
    """fl_split_solve leaves the result in `out` after the reference's swap."""
    U = cons_field("rough", (40, 50), seed=5)
    os_ = _oscheme(oracle, 0, 0, 0, (0, 0))
    dtdh = 0.3 / oracle.max_wavespeed_grid(U, os_)
    ref = oracle.sweep(oracle.sweep(U, os_, dtdh, 0), os_, dtdh, 1)
    solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0, 0)))
    a = _dev(U)
    b = torch.empty_like(a)
    new_in, new_out = solver.solve(a, b, dtdh)
    torch.cuda.synchronize()
    assert new_out.data_ptr() == a.data_ptr()
    assert _bitwise(_host(new_out), ref)


@pytest.mark.parametrize("flux", [0, 1], ids=["hllc", "force"])
def test_fp32_against_fp64_oracle(E, oracle, flux):
    """FP32 bulk sweeps, FP64 dt: relative L1 density error <= 1e-5 after
    100 steps on a 2D Riemann workload, with HLLC and with FORCE (the 4096^2
    workload's two flux variants)."""
    U = cons_field("riemann", (128, 128), seed=1001)
    os_ = _oscheme(oracle, 0, 0, flux, (0, 0))
    ref, rt, it, _ = oracle.simulate(U, os_, 0.9, 1.0 / 128, max_iters=100)
    solver = E.SplitSolver(_escheme(E, 0, 0, flux, (0, 0)), layout=1)
    a = _dev(U, dtype=torch.float32, layout=1)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.9, 1.0 / 128, max_iters=100)
    got = _host(a, 1)
    assert r.iters == it
    rel_l1 = np.abs(got[..., 0] - ref[..., 0]).sum() / np.abs(ref[..., 0]).sum()
    print(f"fp32 rel L1 density {rel_l1:.3e}", file=sys.stderr)
    assert rel_l1 <= 1e-5


def test_vacuum_reports_numeric_error(E, N):
    """Colliding-away states that drive the pressure negative make the fused
    wavespeed NaN; fl_simulate reports FL_ERR_NUMERIC instead of silently
    continuing."""
    n = 64
    Wp = np.zeros((n, 3))
    Wp[:, 0] = 1.0
    Wp[:, 1] = 1e-6
    Wp[: n // 2, 2] = -50.0
    Wp[n // 2:, 2] = 50.0
    from fluidity_amd import workloads as W
    U = W.prim_to_cons(Wp, 1.4)
    solver = E.SplitSolver(E.Scheme())
    a = _dev(U)
    w = torch.empty_like(a)
    with pytest.raises(N.FluidityError) as ei:
        solver.simulate(a, w, 0.9, 1.0 / n, max_iters=50)
    assert ei.value.status == N.FL_ERR_NUMERIC


@pytest.mark.parametrize("nd,steps", [(1, 37), (2, 13), (2, 45), (1, 300)])
def test_simulate_graph_replay_equals_launches(E, nd, steps):
    """The hipGraph replay (16-step runs while a batch has 16 steps left, then
    step pairs) gives the same state, step count and run time as launching
    every kernel: odd counts, counts that mix both graphs, and a count past
    one 256-step batch."""
    U = cons_field("smooth", (97,) if nd == 1 else (40, 52), 0)
    res = []
    for graphs in (True, False):
        solver = E.SplitSolver(E.Scheme(), layout=1)
        solver.set_graphs(graphs)
        a = _dev(U, layout=1)
        r = solver.simulate(a, torch.empty_like(a), 0.9, 1.0 / U.shape[-2], max_iters=steps)
        res.append((_host(a, layout=1), r.iters, r.run_time, r.last_dt))
    assert res[0][1] == res[1][1] == steps
    assert res[0][2] == res[1][2] and res[0][3] == res[1][3]
    assert np.array_equal(res[0][0].view(np.uint64), res[1][0].view(np.uint64))


@pytest.mark.parametrize("case", [
    # n, dtype, flux, bc, layout, max_iters, sim_time
    (4096, torch.float64, 0, 0, 0, 301, float("inf")),
    (700, torch.float64, 1, 1, 1, 40, float("inf")),
    (1000, torch.float64, 0, 2, 0, 10 ** 6, 0.05),
    (62, torch.float64, 0, 0, 1, 9, float("inf")),
    (65536, torch.float64, 0, 0, 1, 12, float("inf")),
    (65537, torch.float64, 0, 0, 1, 12, float("inf")),   # past the persistent loop's limit
    (3000, torch.float32, 0, 0, 1, 50, float("inf")),
    (2000, torch.float64, 0, 0, 0, 0, float("inf")),
])
def test_simulate_1d_persistent_equals_launches(E, case):
    """fl_simulate on a 1D line of at most 2^16 cells runs the whole step loop
    in one cooperative launch (a grid barrier per step, dt computed by every
    workgroup from the same control words); with graphs off every step is
    launched from the host.  Both give the same state bit for bit, the same
    step count, run time, last dt and final maximum wavespeed: odd and even
    step counts, a run stopped by sim_time, every boundary kind, FP64 and
    FP32, AoS and SoA, zero steps."""
    n, dtype, flux, bc, layout, steps, sim_time = case
    U = cons_field("riemann", (n,), seed=1000)
    res = []
    for graphs in (True, False):
        solver = E.SplitSolver(E.Scheme(flux=flux, bc=(bc, 0, 0)), layout=layout)
        solver.set_graphs(graphs)
        a = _dev(U, dtype=dtype, layout=layout)
        r = solver.simulate(a, torch.empty_like(a), 0.9, 1.0 / n, sim_time=sim_time,
                            max_iters=steps)
        res.append((_host(a, layout=layout), r.iters, r.run_time, r.last_dt, r.max_wavespeed))
    assert res[0][1:] == res[1][1:]
    if sim_time == float("inf"):
        assert res[0][1] == steps
    assert _bitwise(res[0][0], res[1][0]), _report(res[0][0], res[1][0])


def test_headline_workload_full_size_bitwise(E, oracle):
    """The 4096^2 FP64 HLLC four-quadrant workload itself (seed 1001), after
    the launch-shape autotuner: 12 steps of the device loop equal 12 steps of
    the oracle (16 host threads) bit for bit, with the same run time."""
    import os
    from fluidity_amd import workloads as W
    wl = W.riemann_2d(n=4096, seed=1001)
    os_ = _oscheme(oracle, 0, 0, 0, (0, 0))
    steps = 12
    ref, rt, it, _ = oracle.simulate(wl.U, os_, 0.9, wl.dh, max_iters=steps,
                                     nthreads=min(os.cpu_count() or 1, 16))
    solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0, 0)), layout=1)
    a = _dev(wl.U, layout=1)
    w = torch.empty_like(a)
    solver.autotune(a, w, reps=1)
    r = solver.simulate(a, w, 0.9, wl.dh, max_iters=steps)
    got = _host(a, 1)
    del a, w
    torch.cuda.empty_cache()
    assert r.iters == it == steps and r.run_time == rt
    assert _bitwise(got, ref), _report(got, ref)


@pytest.mark.parametrize("flux", [0, 1], ids=["hllc", "force"])
def test_fp32_full_size_against_fp64(E, flux):
    """FP32 bulk on the 4096^2 workload, 100 steps, against the FP64 device
    loop (bit-identical to the oracle): relative L1 density error <= 1e-5;
    the relative L1 and L-inf error of every field is printed."""
    from fluidity_amd import workloads as W
    wl = W.riemann_2d(n=4096, seed=1001)
    out = {}
    for dt in (torch.float64, torch.float32):
        solver = E.SplitSolver(_escheme(E, 0, 0, flux, (0, 0)), layout=1)
        a = _dev(wl.U, dtype=dt, layout=1)
        w = torch.empty_like(a)
        r = solver.simulate(a, w, 0.9, wl.dh, max_iters=100)
        assert r.iters == 100
        out[dt] = a.double()
        del a, w
    ref, got = out[torch.float64], out[torch.float32]
    names = ["rho", "E", "m_x", "m_y"]
    for k in range(4):
        d = (got[k] - ref[k]).abs()
        l1 = float(d.sum() / ref[k].abs().sum())
        linf = float(d.max() / ref[k].abs().max())
        print(f"4096^2 fp32 {names[k]}: rel L1 {l1:.3e}, rel Linf {linf:.3e}", file=sys.stderr)
    rel_l1 = float((got[0] - ref[0]).abs().sum() / ref[0].abs().sum())
    assert rel_l1 <= 1e-5
    del out
    torch.cuda.empty_cache()


def test_large_grid_constant_in_y_equals_1d(E, oracle):
    """Maximum-size indexing: a 65536 x 8192 FP64 SoA grid (537 M cells, each
    component plane 4.3 GB, element offsets beyond 2^31) holding a 1D Riemann
    problem constant in y with v = 0.  The y fluxes cancel exactly and the
    flattening factors across y are 1, so after 2 steps every row equals the
    oracle's 1D run of the same line bit for bit, with the same run time."""
    from fluidity_amd import workloads as W
    if torch.cuda.get_device_properties(0).total_memory < 64 * 2**30:
        pytest.skip("needs 64 GiB of device memory")
    nx, ny, steps = 65536, 8192, 2
    r1 = W.riemann_1d(n=nx, seed=1000)
    os1 = _oscheme(oracle, 0, 0, 0, (0,))
    ref, rt, it, _ = oracle.simulate(r1.U, os1, 0.9, r1.dh, max_iters=steps)
    row = torch.zeros((4, 1, nx), dtype=torch.float64, device="cuda")
    row[:3, 0, :] = torch.from_numpy(np.ascontiguousarray(r1.U.T)).to("cuda")
    a = row.expand(4, ny, nx).contiguous()
    w = torch.empty_like(a)
    solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0, 0)), layout=1)
    r = solver.simulate(a, w, 0.9, r1.dh, max_iters=steps)
    del w
    assert r.iters == it == steps and r.run_time == rt
    assert torch.equal(a, a[:, :1, :].expand(4, ny, nx))
    got = a[:, 0, :].cpu().numpy()
    del a, row
    torch.cuda.empty_cache()
    assert np.all(got[3] == 0.0)
    assert _bitwise(np.ascontiguousarray(got[:3].T), ref), _report(got[:3].T, ref)


@pytest.mark.parametrize("flux", [2, 3])
def test_simulate_lf_richtmyer_bitwise(E, oracle, flux):
    """The step loop with the standalone Lax-Friedrichs and Richtmyer
    fluxes against the oracle's loop.  Richtmyer (two-step Lax-Wendroff) has
    no dissipation at discontinuities and reaches negative pressure on the
    four-quadrant data within two steps, so it runs on smooth data."""
    U = cons_field("riemann" if flux == 2 else "smooth", (64, 64), seed=1001)
    os_ = _oscheme(oracle, 0, 0, flux, (0, 0))
    ref, rt, it, _ = oracle.simulate(U, os_, 0.5, 1.0 / 64, max_iters=15)
    solver = E.SplitSolver(_escheme(E, 0, 0, flux, (0, 0)), layout=1)
    a = _dev(U, layout=1)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.5, 1.0 / 64, max_iters=15)
    got = _host(a, 1)
    assert r.iters == it and r.run_time == rt
    assert _bitwise(got, ref), _report(got, ref)


@pytest.mark.parametrize("shape", [(1000,), (45, 70)])
def test_split_step_fused_max(E, oracle, shape):
    """fl_split_step equals fl_split_solve and its fused maximum equals the
    oracle's maximum over the new states."""
    U = cons_field("rough", shape, seed=17)
    nd = len(shape)
    os_ = _oscheme(oracle, 0, 0, 0, (0,) * nd)
    dtdh = 0.4 / oracle.max_wavespeed_grid(U, os_)
    solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0,) * nd))
    a, b = _dev(U), torch.empty(U.shape, dtype=torch.float64, device="cuda")
    _, out, ws = solver.step(a, b, dtdh)
    got = _host(out)
    c, d = _dev(U), torch.empty_like(b)
    _, out2 = solver.solve(c, d, dtdh)
    assert _bitwise(got, _host(out2))
    ref = U
    for dim in range(nd):
        ref = oracle.sweep(ref, os_, dtdh, dim)
    assert _bitwise(got, ref), _report(got, ref)
    assert ws == oracle.max_wavespeed_grid(ref, os_)


@pytest.mark.parametrize("shape", [(1001,), (40, 52)])
def test_simulate_zero_steps_is_defined(E, oracle, shape):
    """max_iters = 0 runs no step: a fresh context and one that has run
    before both report t = 0, no iterations and the initial maximum
    wavespeed, and leave the state untouched (1D too, whose odd step counts
    end in the work buffer)."""
    U = cons_field("rough", shape, seed=23)
    nd = len(shape)
    os_ = _oscheme(oracle, 0, 0, 0, (0,) * nd)
    ws0 = oracle.max_wavespeed_grid(U, os_)
    solver = E.SplitSolver(_escheme(E, 0, 0, 0, (0,) * nd))
    a = _dev(U)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.5, 0.01, max_iters=0)
    assert r.iters == 0 and r.run_time == 0.0 and r.max_wavespeed == ws0
    assert _bitwise(_host(a), U)
    r = solver.simulate(a, w, 0.5, 0.01, max_iters=3)   # odd count: 1D ends in `work`
    assert r.iters == 3
    after = _host(a)
    r = solver.simulate(a, w, 0.5, 0.01, max_iters=0)
    assert r.iters == 0 and r.run_time == 0.0
    assert r.max_wavespeed == oracle.max_wavespeed_grid(after, os_)
    assert _bitwise(_host(a), after)


def test_simulator_continues_after_iteration_limit(E):
    """Simulator.simulate() after its iteration limit was reached runs no
    step: iterations, run time and the state stay as they were."""
    from fluidity_amd.simulator import Simulator
    sim = Simulator(dimensions=1)
    sim.configure_resolution(1.0 / 256)
    sim.configure_dimension(0, 0.0, 1.0)
    sim.configure_sim_time(1.0)      # the reference's default sim_time is 0: no step
    sim.configure_max_iterations(5)
    sim.fill_data([("rho", lambda p: np.where(p[:, 0] < 0.5, 1.0, 0.125)),
                   ("p", lambda p: np.where(p[:, 0] < 0.5, 1.0, 0.1))])
    sim.simulate()
    it, t = sim.params.iters, sim.params.run_time
    state = sim.state.clone()
    assert it == 5 and t > 0
    r = sim.simulate()
    assert r.iters == 0 and r.run_time == 0.0
    assert sim.params.iters == it and sim.params.run_time == t
    assert torch.equal(sim.state, state)


@pytest.mark.parametrize("flux", [0, 1])
@pytest.mark.parametrize("shape", [(900,), (52, 60)])
def test_simulate_corrected_stiff_bitwise(E, oracle, flux, shape):
    """The corrected stiffened gas (pressure of the stiffened-gas energy,
    ratios of p + p_inf) through the step loop: bit-identical to the
    oracle's corrected loop, and different from the reference's pressure."""
    g, pinf = 4.4, 6.0
    nd = len(shape)
    rng = np.random.default_rng(41)
    n = int(np.prod(shape))
    Wp = np.zeros((n, 2 + nd))
    Wp[:, 0] = np.where(rng.uniform(size=n) < 0.5, 1.0, 3.0)
    Wp[:, 1] = rng.uniform(0.5, 20.0, n)
    Wp[:, 2:] = rng.uniform(-1.0, 1.0, (n, nd))
    from fluidity_amd import workloads as W
    U = W.prim_to_cons_eos(Wp, g, pinf, True).reshape(shape + (2 + nd,))
    bc = (0,) * nd
    os_ = _oscheme(oracle, 0, 0, flux, bc, gamma=g, eos=2, p_inf=pinf)
    ref, rt, it, _ = oracle.simulate(U, os_, 0.8, 1.0 / shape[-1], max_iters=10)
    solver = E.SplitSolver(_escheme(E, 0, 0, flux, bc, gamma=g, eos=2, p_inf=pinf))
    a = _dev(U)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.8, 1.0 / shape[-1], max_iters=10)
    got = _host(a)
    assert r.iters == it and r.run_time == rt
    assert _bitwise(got, ref), _report(got, ref)
    quirk, _, _, _ = oracle.simulate(U, _oscheme(oracle, 0, 0, flux, bc, gamma=g, eos=1,
                                                 p_inf=pinf), 0.8, 1.0 / shape[-1], max_iters=10)
    assert np.abs(quirk - ref).max() > 1e-6



@pytest.mark.parametrize("prec", ["f64", "f64fast", "f32"])
def test_step_loop_independent_of_x_tile_width(E, prec):
    """x-sweep tiles of 64, 128 and 256 lanes (one, two and four waves per
    workgroup) and several heights give the same bits over a full step loop
    (x and y sweeps, fused CFL), FP64 exact and fast and FP32."""
    U = cons_field("riemann", (150, 200), seed=77)
    dt = torch.float32 if prec == "f32" else torch.float64
    outs = []
    for xb, xr in ((64, 5), (64, 33), (128, 7), (256, 16)):
        sch = _escheme(E, 0, 0, 0, (0, 0))
        sch.numerics = 1 if prec == "f64fast" else 0
        solver = E.SplitSolver(sch, layout=1)
        solver.set_tuning(x_block=xb, x_rows=xr, y_block=64, y_rows=16)
        a = _dev(U, dtype=dt, layout=1)
        w = torch.empty_like(a)
        r = solver.simulate(a, w, 0.9, 1.0 / 200, max_iters=25)
        assert r.iters == 25
        outs.append((_host(a, 1), r.run_time))
    for got, rt in outs[1:]:
        assert rt == outs[0][1]
        assert _bitwise(got, outs[0][0])


# ---- fast FP64 numerics ---------------------------------------------------------------------

def _rel_linf(got, ref):
    """Per conserved field, max |got - ref| / max |ref| (fields last)."""
    S = ref.shape[-1]
    g = got.reshape(-1, S)
    r = ref.reshape(-1, S)
    return [float(np.abs(g[:, k] - r[:, k]).max() / np.abs(r[:, k]).max()) for k in range(S)]


FAST_CASES = [
    # kind, shape, limiter, form, flux, bc, layout
    ("riemann", (160, 192), 0, 0, 0, (0, 0), 1),
    ("riemann", (160, 192), 0, 0, 1, (0, 0), 1),
    ("rough", (96, 80), 1, 0, 0, (1, 2), 0),
    ("rough", (96, 80), 0, 1, 0, (2, 1), 1),
    ("rough", (96, 80), 1, 1, 1, (0, 0), 0),
    ("rough", (2000,), 0, 0, 0, (0,), 0),
    ("rough", (2000,), 1, 1, 1, (2,), 1),
]


@pytest.mark.parametrize("case", FAST_CASES)
def test_fast_numerics_against_oracle(E, oracle, case):
    """FL_NUMERICS_FAST (one reciprocal per density, closed-form van Leer
    slope, contracted multiply-adds) against the oracle's exact loop: every
    field within 1e-12 relative L-inf after 100 steps, same step count and
    run time within 1e-12."""
    kind, shape, lim, form, flux, bc, layout = case
    U = cons_field(kind, shape, seed=1001)
    n = shape[-1]
    ref, rt, it, _ = oracle.simulate(U, _oscheme(oracle, lim, form, flux, bc), 0.8, 1.0 / n,
                                     max_iters=100)
    sch = _escheme(E, lim, form, flux, bc)
    sch.numerics = 1
    solver = E.SplitSolver(sch, layout=layout)
    a = _dev(U, layout=layout)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.8, 1.0 / n, max_iters=100)
    got = _host(a, layout)
    assert r.iters == it == 100
    assert abs(r.run_time - rt) <= 1e-12 * rt
    err = _rel_linf(got, ref)
    print(f"fast {case}: rel Linf per field {err}", file=sys.stderr)
    assert max(err) <= 1e-12, err


def test_fast_numerics_1024_against_oracle(E, oracle):
    """The four-quadrant workload at 1024^2, 100 steps, fast numerics against
    the oracle (16 host threads): every field within 1e-12 relative L-inf."""
    from fluidity_amd import workloads as W
    wl = W.riemann_2d(n=1024, seed=1001)
    ref, rt, it, _ = oracle.simulate(wl.U, _oscheme(oracle, 0, 0, 0, (0, 0)), 0.9, wl.dh,
                                     max_iters=100, nthreads=16)
    sch = _escheme(E, 0, 0, 0, (0, 0))
    sch.numerics = 1
    solver = E.SplitSolver(sch, layout=1)
    a = _dev(wl.U, layout=1)
    w = torch.empty_like(a)
    r = solver.simulate(a, w, 0.9, wl.dh, max_iters=100)
    got = _host(a, 1)
    assert r.iters == it == 100
    err = _rel_linf(got, ref)
    print(f"fast 1024^2 rel Linf per field {err}", file=sys.stderr)
    assert max(err) <= 1e-12, err


@pytest.mark.parametrize("flux", [0, 1], ids=["hllc", "force"])
def test_fast_numerics_full_size(E, flux):
    """The 4096^2 workload, 100 steps: fast numerics against the exact device
    loop (bit-identical to the oracle, test_headline_workload_full_size_bitwise):
    every field within 1e-12 relative L-inf, and the same run time within
    1e-12."""
    from fluidity_amd import workloads as W
    wl = W.riemann_2d(n=4096, seed=1001)
    out = {}
    for numerics in (0, 1):
        sch = _escheme(E, 0, 0, flux, (0, 0))
        sch.numerics = numerics
        solver = E.SplitSolver(sch, layout=1)
        a = _dev(wl.U, layout=1)
        w = torch.empty_like(a)
        r = solver.simulate(a, w, 0.9, wl.dh, max_iters=100)
        assert r.iters == 100
        out[numerics] = (a, r.run_time)
        del w
    ref, got = out[0][0], out[1][0]
    err = [float((got[k] - ref[k]).abs().max() / ref[k].abs().max()) for k in range(4)]
    print(f"fast 4096^2 rel Linf per field {err}", file=sys.stderr)
    assert abs(out[1][1] - out[0][1]) <= 1e-12 * out[0][1]
    assert max(err) <= 1e-12, err
    del out, ref, got
    torch.cuda.empty_cache()


def test_fast_numerics_rejected_where_unsupported(E, N):
    """The batched, two-material and adjoint paths run the exact numerics
    and refuse the fast flag with FL_ERR_UNSUPPORTED."""
    from fluidity_amd import adjoint as AJ
    sch = _escheme(E, 0, 0, 0, (2, 2))
    sch.numerics = 1
    adj = AJ.Adjoint(sch, 0.5, 1.0 / 32)
    U = _dev(cons_field("smooth", (32, 32), seed=3), layout=1)
    with pytest.raises(N.FluidityError) as e:
        adj.forward(U, steps=2, checkpoint_every=1)
    assert e.value.status == N.FL_ERR_UNSUPPORTED
