"""The C ABI library on the CPU: it loads, exports every entry point the
header declares, and rejects bad arguments before touching a device."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluidity_amd.h")
LIB = os.path.join(ROOT, "fluidity_amd", "libfluidity_amd.so")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fl_[a-z_0-9]+)\s*\(", src)))


@pytest.fixture(scope="module")
def N():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    from fluidity_amd import native
    return native


def test_library_exports_every_declared_symbol(N):
    names = declared_functions()
    assert len(names) >= 10
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    exported = set(re.findall(r"\bT (fl_[a-z_0-9]+)\b", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for n in names:
        assert hasattr(N.lib, n)
    assert set(N.EXPORTED) == set(names)


_MAPPED_RUNTIMES = """
import sys
sys.path.insert(0, {root!r})
import fluidity_amd.native
paths = set()
for line in open('/proc/self/maps'):
    parts = line.split()
    if len(parts) >= 6 and ('libamdhip64' in parts[-1] or 'libhsa-runtime64' in parts[-1]):
        paths.add(parts[-1])
print(len([p for p in paths if 'libamdhip64' in p]), len([p for p in paths if 'libhsa' in p]))
"""


def test_one_hip_runtime_per_process():
    """Importing the binding before torch leaves one HIP and one HSA runtime
    in the process (torch's, which the library's libamdhip64.so.7 dependency
    binds to).  Loading the library first used to map /opt/rocm's runtime
    beside torch's, and the runtime initialised second found no device."""
    out = subprocess.check_output(["python", "-c", _MAPPED_RUNTIMES.format(root=ROOT)],
                                  text=True, cwd=ROOT)
    assert out.split() == ["1", "1"], out


def test_status_strings(N):
    assert N.lib.fl_status_string(0) == b"ok"
    assert N.lib.fl_status_string(N.FL_ERR_ARGUMENT) == b"invalid argument"
    assert N.lib.fl_version().startswith(b"fluidity_amd")


def test_null_context_rejected(N):
    f = N.field(0x1000, 2, (8, 8))
    g = N.field(0x2000, 2, (8, 8))
    s = N.scheme()
    rc = N.lib.fl_split_sweep(None, C.byref(f), C.byref(g), 0, C.byref(s), 0.1, None)
    assert rc == N.FL_ERR_ARGUMENT
    assert b"context" in N.lib.fl_last_error()


def test_ctx_create_null_out(N):
    assert N.lib.fl_ctx_create(0, None) == N.FL_ERR_ARGUMENT


def _fake_ctx(N):
    """A context-shaped handle is never dereferenced before argument checks;
    only fl_ctx_create touches the HIP runtime."""
    return C.c_void_p(0xdead0000)


@pytest.mark.parametrize("mutate,needle", [
    (lambda f, g, s: setattr(f, "data", None), b"in->data"),
    (lambda f, g, s: f.n.__setitem__(0, 3), b"at least 4"),
    (lambda f, g, s: setattr(f, "dtype", 7), b"dtype"),
    (lambda f, g, s: setattr(f, "layout", 9), b"layout"),
    (lambda f, g, s: setattr(g, "data", f.data), b"alias"),
    (lambda f, g, s: setattr(s, "gamma", 1.0), b"gamma"),
    (lambda f, g, s: setattr(s, "limiter", 5), b"limiter"),
    (lambda f, g, s: setattr(s, "flux", 5), b"flux"),
    (lambda f, g, s: s.bc.__setitem__(1, 7), b"bc[1]"),
    (lambda f, g, s: setattr(f, "ndim", 3), b"3D"),
])
def test_sweep_argument_errors(N, mutate, needle):
    f = N.field(0x1000, 2, (16, 16))
    g = N.field(0x2000, 2, (16, 16))
    s = N.scheme()
    mutate(f, g, s)
    rc = N.lib.fl_split_sweep(_fake_ctx(N), C.byref(f), C.byref(g), 0, C.byref(s), 0.1, None)
    assert rc in (N.FL_ERR_ARGUMENT, N.FL_ERR_UNSUPPORTED)
    assert needle in N.lib.fl_last_error()


@pytest.mark.parametrize("dtdh", [0.0, -1.0, float("nan"), float("inf")])
def test_sweep_rejects_bad_dtdh(N, dtdh):
    f = N.field(0x1000, 2, (16, 16))
    g = N.field(0x2000, 2, (16, 16))
    s = N.scheme()
    rc = N.lib.fl_split_sweep(_fake_ctx(N), C.byref(f), C.byref(g), 0, C.byref(s), dtdh, None)
    assert rc == N.FL_ERR_ARGUMENT


def test_sweep_rejects_bad_dim(N):
    f = N.field(0x1000, 1, (16,))
    g = N.field(0x2000, 1, (16,))
    s = N.scheme()
    rc = N.lib.fl_split_sweep(_fake_ctx(N), C.byref(f), C.byref(g), 1, C.byref(s), 0.1, None)
    assert rc == N.FL_ERR_ARGUMENT


def test_simulate_rejects_bad_params(N):
    f = N.field(0x1000, 2, (16, 16))
    g = N.field(0x2000, 2, (16, 16))
    s = N.scheme()
    for p in [N.SimParams(0.0, 0.1, 1.0, 10), N.SimParams(0.9, -1.0, 1.0, 10),
              N.SimParams(0.9, 0.1, float("nan"), 10), N.SimParams(0.9, 0.1, 1.0, -1)]:
        r = N.SimResult()
        rc = N.lib.fl_simulate(_fake_ctx(N), C.byref(f), C.byref(g), C.byref(s), C.byref(p),
                               C.byref(r), None)
        assert rc == N.FL_ERR_ARGUMENT
    rc = N.lib.fl_simulate(_fake_ctx(N), C.byref(f), C.byref(g), C.byref(s), None, None, None)
    assert rc == N.FL_ERR_ARGUMENT


def test_tuning_validation(N):
    t = N.Tuning(100, 0, 0, 0)
    assert N.lib.fl_ctx_set_tuning(None, C.byref(t)) == N.FL_ERR_ARGUMENT


def test_io_entry_points_reject_bad_arguments(N):
    """The initial-condition and output entry points fail with
    FL_ERR_ARGUMENT (and a message) before any device work."""
    f = N.field(0x1000, 2, (8, 8))
    s = N.scheme()
    lib = N.lib
    assert lib.fl_primitive_planes(None, C.byref(f), C.byref(s), 0x2000, None) == N.FL_ERR_ARGUMENT
    assert lib.fl_from_primitive(None, 0x2000, C.byref(s), C.byref(f), None) == N.FL_ERR_ARGUMENT
    w = (C.c_double * 4)(1.0, 1.0, 0.0, 0.0)
    assert lib.fl_set_states(None, 0x2000, w, C.byref(s), C.byref(f), None) == N.FL_ERR_ARGUMENT
    assert lib.fl_write_results(None, C.byref(f), C.byref(s), None, 0.0, 0.1, b"x.dat",
                                None) == N.FL_ERR_ARGUMENT
    assert lib.fl_write_results_separate_raw(None, C.byref(f), C.byref(s), 0.0, b"x",
                                             None) == N.FL_ERR_ARGUMENT
    assert b"ctx" in lib.fl_last_error()


def test_level_set_grid_size_limits(N):
    """Level-set grids hold cell indices as int in their band lists: a grid of
    2^31 cells or more is rejected by the workspace queries (host-only calls),
    the largest square grid below that is accepted."""
    import ctypes as C
    big = (C.c_int64 * 3)(46341, 46341, 1)       # 2 147 488 281 cells >= 2^31
    ok = (C.c_int64 * 3)(46340, 46340, 1)        # 2 147 395 600 cells
    for fn in (N.lib.fl_ls_reinit_work_bytes, N.lib.fl_mm_work_bytes,
               N.lib.fl_gfm_extrapolate_work_bytes):
        assert fn(2, big) == -1
        assert fn(2, ok) > 0
    long1d = (C.c_int64 * 3)((1 << 30) + 1, 1, 1)  # a side above 2^30
    assert N.lib.fl_ls_reinit_work_bytes(1, long1d) == -1


def _batch_problems(N, numerics):
    probs = (N.BatchProblem * len(numerics))()
    for i, num in enumerate(numerics):
        probs[i].state = N.field(0x100000 * (2 * i + 1), 2, (16, 16))
        probs[i].work = N.field(0x100000 * (2 * i + 2), 2, (16, 16))
        probs[i].scheme = N.scheme(numerics=num)
        probs[i].params = N.SimParams(0.9, 1.0 / 16, 1.0, 10)
    return probs


def test_batch_rejects_mixed_fp64_numerics(N):
    """One FP64 numerics per batch: a batch mixing exact and fast problems is
    refused before any device work."""
    probs = _batch_problems(N, [N.FL_NUMERICS_EXACT, N.FL_NUMERICS_FAST])
    out = C.c_void_p()
    rc = N.lib.fl_batch_create(_fake_ctx(N), probs, 2, C.byref(out), None)
    assert rc == N.FL_ERR_UNSUPPORTED
    assert b"numerics" in N.lib.fl_last_error()


def test_mm_simulate_rejects_mixed_numerics(N):
    """Every material of a two-material step takes the same numerics."""
    mats = (N.Material * 2)()
    for m in range(2):
        mats[m].state = N.field(0x100000 * (3 * m + 1), 2, (16, 16))
        mats[m].work = N.field(0x100000 * (3 * m + 2), 2, (16, 16))
        mats[m].phi = 0x100000 * (3 * m + 3)
        mats[m].scheme = N.scheme(numerics=N.FL_NUMERICS_FAST if m else N.FL_NUMERICS_EXACT)
    p = N.SimParams(0.5, 1.0 / 16, 1.0, 10)
    r = N.SimResult()
    rc = N.lib.fl_mm_simulate(_fake_ctx(N), mats, 2, C.byref(p), None, C.c_void_p(0x7000000),
                              1 << 30, C.byref(r), None)
    assert rc == N.FL_ERR_ARGUMENT
    assert b"same numerics" in N.lib.fl_last_error()


@pytest.mark.parametrize("mutate,status,needle", [
    (lambda s: setattr(s, "flux", 2), 2, b"HLLC and FORCE"),
    (lambda s: setattr(s, "flux", 3), 2, b"HLLC and FORCE"),
    (lambda s: setattr(s, "numerics", 1), 2, b"exact"),
])
def test_adjoint_scheme_limits(N, mutate, status, needle):
    """The adjoint runs HLLC and FORCE (either limiter and form) in the exact
    numerics; the standalone Lax-Friedrichs and Richtmyer fluxes and the fast
    numerics are refused before any device work."""
    f = N.field(0x1000, 2, (16, 16))
    s = N.scheme()
    mutate(s)
    dg = C.c_double()
    rc = N.lib.fl_split_step_vjp(_fake_ctx(N), C.byref(f), C.byref(f), C.byref(f), C.byref(s),
                                 0.9, 1.0 / 16, C.byref(dg), C.c_void_p(0x9000000), 1 << 30, None)
    assert rc == status
    assert needle in N.lib.fl_last_error()

