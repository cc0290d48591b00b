"""fl_simulate_noalloc, the step loop with its step-control block in a
file-scope device array (no context, no allocation), against fl_simulate on
the same inputs: bit-identical states, iteration counts, run times and final
wavespeeds, on two calls in a row (the second on a smaller grid), in 1D and
2D, FP64 and FP32, AoS and SoA."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.fields import cons_field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def N():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from fluidity_amd import native
    return native


@pytest.fixture(scope="module")
def E():
    from fluidity_amd import euler
    return euler


def _dev(U, layout, dtype=torch.float64):
    a = np.moveaxis(U, -1, 0) if layout == 1 else U
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype)


def _noalloc(N, E, U, scheme, layout, params, tuning=None):
    w = torch.empty_like(U)
    fs, fw = E.grid_field(U, layout), E.grid_field(w, layout)
    r = N.SimResult()
    t = C.byref(tuning) if tuning is not None else None
    st = N.lib.fl_simulate_noalloc(t, C.byref(fs), C.byref(fw), C.byref(scheme.c()),
                                   C.byref(params), C.byref(r), None)
    return st, r


def _reference(N, E, U, scheme, layout, params, tuning=None):
    solver = E.SplitSolver(scheme, layout=layout)
    if tuning is not None:
        solver.set_tuning(x_block=tuning.x_block, x_rows=tuning.x_rows,
                          y_block=tuning.y_block, y_rows=tuning.y_rows)
    w = torch.empty_like(U)
    fs, fw = E.grid_field(U, layout), E.grid_field(w, layout)
    r = N.SimResult()
    st = N.lib.fl_simulate(solver.ctx.handle, C.byref(fs), C.byref(fw), C.byref(scheme.c()),
                           C.byref(params), C.byref(r), None)
    return st, r


CASES = [
    # shape of the first call, of the second (smaller) call, dtype, layout, flux, sim_time, max_iters
    ((96, 112), (40, 48), torch.float64, 1, 0, float("inf"), 23),
    ((96, 112), (40, 48), torch.float64, 0, 1, float("inf"), 17),
    ((3000,), (700,), torch.float64, 0, 0, 0.05, 10 ** 6),
    ((2500,), (600,), torch.float64, 1, 1, float("inf"), 301),
    ((64, 80), (32, 32), torch.float32, 1, 0, float("inf"), 30),
]


@pytest.mark.parametrize("case", CASES)
def test_noalloc_matches_simulate(N, E, case):
    shape1, shape2, dtype, layout, flux, sim_time, max_iters = case
    scheme = E.Scheme(flux=flux)
    tuning = N.Tuning(128, 16, 64, 32)
    for shape in (shape1, shape2):
        U0 = cons_field("riemann", shape, seed=1001)
        dh = 1.0 / shape[-1]
        p = N.SimParams(0.9, dh, sim_time, max_iters)
        a, b = _dev(U0, layout, dtype), _dev(U0, layout, dtype)
        st_a, ra = _noalloc(N, E, a, scheme, layout, p, tuning)
        st_b, rb = _reference(N, E, b, scheme, layout, p, tuning)
        assert st_a == st_b == N.FL_OK, N.lib.fl_last_error()
        assert ra.iters == rb.iters > 0
        assert ra.run_time == rb.run_time
        assert ra.last_dt == rb.last_dt
        assert ra.max_wavespeed == rb.max_wavespeed
        assert torch.equal(a, b)


def test_noalloc_default_tuning_and_zero_steps(N, E):
    """NULL tuning selects the built-in launch shapes; max_iters = 0 runs no
    step and reports the initial maximum wavespeed."""
    scheme = E.Scheme()
    U0 = cons_field("riemann", (64, 64), seed=7)
    for iters in (5, 0):
        p = N.SimParams(0.9, 1.0 / 64, float("inf"), iters)
        a, b = _dev(U0, 1), _dev(U0, 1)
        st_a, ra = _noalloc(N, E, a, scheme, 1, p)
        st_b, rb = _reference(N, E, b, scheme, 1, p)
        assert st_a == st_b == N.FL_OK
        assert (ra.iters, ra.run_time, ra.max_wavespeed) == (rb.iters, rb.run_time, rb.max_wavespeed)
        assert ra.iters == iters
        assert torch.equal(a, b)


def test_noalloc_rejects_bad_arguments(N, E):
    scheme = E.Scheme()
    U = _dev(cons_field("riemann", (32, 32), seed=3), 1)
    p = N.SimParams(0.9, 1.0 / 32, float("inf"), 3)
    bad = N.Tuning(100, 16, 64, 32)          # block width not a multiple of 64
    st, _ = _noalloc(N, E, U, scheme, 1, p, bad)
    assert st == N.FL_ERR_ARGUMENT
    p0 = N.SimParams(0.0, 1.0 / 32, float("inf"), 3)
    st, _ = _noalloc(N, E, U, scheme, 1, p0)
    assert st == N.FL_ERR_ARGUMENT
