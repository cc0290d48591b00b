"""The reference-named C++ template API (include/fluidity_amd/fluidity_amd.hpp)
driven like the reference's simulator, checked bit for bit against the oracle."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from tests.fields import cons_field

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "shim_driver")


@pytest.fixture(scope="module")
def driver():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/shim_driver is not built (run __graft_entry__.build())")
    return DRIVER


def _run(driver, tmp_path, mode, combo, U, arg):
    nd = U.shape[-1] - 2
    ny, nx = (1, U.shape[0]) if nd == 1 else U.shape[:2]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(U, dtype=np.float64).tofile(fin)
    res = subprocess.run([driver, mode, str(combo), str(nd), str(nx), str(ny), repr(arg),
                          str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    return np.fromfile(fout, dtype=np.float64).reshape(U.shape), res.stdout


@pytest.mark.parametrize("combo", [0, 1])
@pytest.mark.parametrize("shape", [(500,), (40, 56)])
def test_shim_solve_matches_oracle(driver, oracle, tmp_path, combo, shape):
    U = cons_field("rough", shape, seed=17)
    s = oracle.scheme(limiter=combo, form=combo, flux=combo, bc=(0, 0))
    dtdh = 0.35 / oracle.max_wavespeed_grid(U, s)
    ref = oracle.sweep(U, s, dtdh, 0)
    if len(shape) == 2:
        ref = oracle.sweep(ref, s, dtdh, 1)
    got, _ = _run(driver, tmp_path, "solve", combo, U, dtdh)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_shim_simulator_loop(driver, oracle, tmp_path):
    U = cons_field("riemann", (64, 64), seed=1001)
    s = oracle.scheme(flux=0)
    ref, rt, it, _ = oracle.simulate(U, s, 0.9, 1.0 / 64, max_iters=12)
    got, out = _run(driver, tmp_path, "simulate", 0, U, 12)
    assert f"iters {it}" in out
    assert float(out.split()[1]) == rt
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("combo", [0, 1])
@pytest.mark.parametrize("shape", [(300,), (48, 40)])
def test_shim_reference_simulate_loop(driver, oracle, tmp_path, combo, shape):
    """SimulatorImpl::simulate's loop lines (set_wavespeeds, max_element,
    update_time_delta, solve with the reference's in/out swap, the
    SimulationData swap) compiled unchanged against the shim: bit-identical
    to the oracle's loop, with the same run time and iteration count."""
    U = cons_field("riemann" if len(shape) == 2 else "rough", shape, seed=1001)
    nx = shape[-1] if len(shape) == 2 else shape[0]
    s = oracle.scheme(limiter=combo, form=combo, flux=combo, bc=(0, 0))
    ref, rt, it, _ = oracle.simulate(U, s, 0.9, 1.0 / nx, max_iters=9)
    got, out = _run(driver, tmp_path, "impl", combo, U, 9)
    assert f"iters {it}" in out
    assert float(out.split()[1]) == rt
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_shim_simulator_continues_across_calls(driver, oracle, tmp_path):
    """Two simulate() calls with the iteration limit raised in between run
    the steps of one call of the full length (run time and iterations
    accumulate, simulator/parameters.hpp:43-76)."""
    U = cons_field("riemann", (40, 48), seed=1001)
    s = oracle.scheme(flux=0)
    ref, rt, it, _ = oracle.simulate(U, s, 0.9, 1.0 / 48, max_iters=12)
    got, out = _run(driver, tmp_path, "twice", 0, U, 12)
    f = out.split()
    assert int(f[3]) == it == 12 and int(f[5]) == 6
    assert float(f[1]) == rt
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
