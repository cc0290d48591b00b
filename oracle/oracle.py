"""ctypes binding of the CPU oracle (liborc.so built from fl_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the timed
CPU baseline.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this module; the product path never does.

Arrays are float64 numpy arrays in the reference's AoS conservative layout
(state/state.hpp:88-98): shape (ny, nx, 2+nd) for 2D, (nx, 3) for 1D.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

LIM_VANLEER, LIM_SUPERBEE = 0, 1
FORM_CONS, FORM_PRIM = 0, 1
FLUX_HLLC, FLUX_FORCE, FLUX_LF, FLUX_RICHTMYER = 0, 1, 2, 3
EOS_IDEAL, EOS_STIFF = 0, 1
BC_TRANSMISSIVE, BC_REFLECTIVE, BC_PERIODIC = 0, 1, 2


class Scheme(C.Structure):
    _fields_ = [
        ("limiter", C.c_int32),
        ("form", C.c_int32),
        ("flux", C.c_int32),
        ("eos", C.c_int32),
        ("gamma", C.c_double),
        ("p_inf", C.c_double),
        ("bc", C.c_int32 * 3),
        ("pad", C.c_int32),
    ]


def scheme(limiter=LIM_VANLEER, form=FORM_CONS, flux=FLUX_HLLC, eos=EOS_IDEAL,
           gamma=1.4, p_inf=0.0, bc=(0, 0, 0)) -> Scheme:
    s = Scheme()
    s.limiter, s.form, s.flux, s.eos = limiter, form, flux, eos
    s.gamma, s.p_inf = gamma, p_inf
    for i in range(3):
        s.bc[i] = bc[i] if i < len(bc) else 0
    return s


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        sp = C.POINTER(Scheme)
        i64p = C.POINTER(C.c_int64)
        L.orc_pressure.argtypes = [dp, C.c_int, C.c_int, sp]
        L.orc_pressure.restype = C.c_double
        L.orc_sound_speed.argtypes = [dp, C.c_int, C.c_int, sp]
        L.orc_sound_speed.restype = C.c_double
        L.orc_max_wavespeed.argtypes = [dp, C.c_int, C.c_int, sp]
        L.orc_max_wavespeed.restype = C.c_double
        for f in (L.orc_flux, L.orc_flux_prim):
            f.argtypes = [dp, C.c_int, C.c_int, sp, C.c_int, dp]
            f.restype = None
        for f in (L.orc_cons_to_prim, L.orc_prim_to_cons):
            f.argtypes = [dp, C.c_int, C.c_int, sp, dp]
            f.restype = None
        L.orc_apply_setter.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.orc_apply_setter.restype = None
        L.orc_limit_single.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_limit_single.restype = C.c_double
        L.orc_numerical_flux.argtypes = [dp, dp, C.c_int, sp, C.c_double, C.c_int, dp]
        L.orc_numerical_flux.restype = None
        L.orc_sweep.argtypes = [dp, dp, i64p, C.c_int, sp, C.c_double, C.c_int, C.c_int]
        L.orc_sweep.restype = C.c_int
        L.orc_max_wavespeed_grid.argtypes = [dp, i64p, C.c_int, sp, C.c_int]
        L.orc_max_wavespeed_grid.restype = C.c_double
        L.orc_flux_delta_at.argtypes = [dp, i64p, C.c_int, sp, C.c_double, C.c_int,
                                        C.c_int64, C.c_int64, dp]
        L.orc_flux_delta_at.restype = None
        L.orc_face_states_at.argtypes = [dp, i64p, C.c_int, sp, C.c_double, C.c_int,
                                         C.c_int64, C.c_int64, dp, dp, dp]
        L.orc_face_states_at.restype = None
        L.orc_simulate.argtypes = [dp, dp, i64p, C.c_int, sp, C.c_double, C.c_double,
                                   C.c_double, C.c_int64, C.c_int, dp, i64p, dp]
        L.orc_simulate.restype = C.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _dims(U: np.ndarray):
    nd = U.shape[-1] - 2
    if nd == 1:
        n = np.array([U.shape[0], 1], dtype=np.int64)
    else:
        n = np.array([U.shape[1], U.shape[0]], dtype=np.int64)
    return nd, n


def _n_ptr(n: np.ndarray):
    return n.ctypes.data_as(C.POINTER(C.c_int64))


# ---- single-state functions ------------------------------------------------

def pressure(U, s, nd=None, na=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd = len(U) - 2 - na if nd is None else nd
    return lib().orc_pressure(_dp(U), nd, na, C.byref(s))


def sound_speed(U, s, nd=None, na=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd = len(U) - 2 - na if nd is None else nd
    return lib().orc_sound_speed(_dp(U), nd, na, C.byref(s))


def max_wavespeed(U, s, nd=None, na=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd = len(U) - 2 - na if nd is None else nd
    return lib().orc_max_wavespeed(_dp(U), nd, na, C.byref(s))


def flux(U, s, dim, nd=None, na=0, primitive=False):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd = len(U) - 2 - na if nd is None else nd
    F = np.zeros_like(U)
    f = lib().orc_flux_prim if primitive else lib().orc_flux
    f(_dp(U), nd, na, C.byref(s), dim, _dp(F))
    return F


def cons_to_prim(U, s, nd=None, na=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd = len(U) - 2 - na if nd is None else nd
    W = np.zeros_like(U)
    lib().orc_cons_to_prim(_dp(U), nd, na, C.byref(s), _dp(W))
    return W


def prim_to_cons(W, s, nd=None, na=0):
    W = np.ascontiguousarray(W, dtype=np.float64)
    nd = len(W) - 2 - na if nd is None else nd
    U = np.zeros_like(W)
    lib().orc_prim_to_cons(_dp(W), nd, na, C.byref(s), _dp(U))
    return U


def apply_setter(U, dim, kind):
    U = np.array(U, dtype=np.float64)
    lib().orc_apply_setter(_dp(U), len(U) - 2, dim, kind)
    return U


def limit_single(limiter, left, right):
    return lib().orc_limit_single(limiter, left, right)


def numerical_flux(ul, ur, s, dtdh, dim):
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    F = np.zeros_like(ul)
    lib().orc_numerical_flux(_dp(ul), _dp(ur), len(ul) - 2, C.byref(s), dtdh, dim, _dp(F))
    return F


# ---- grid functions ----------------------------------------------------------

def sweep(U, s, dtdh, dim, nthreads=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd, n = _dims(U)
    out = np.empty_like(U)
    rc = lib().orc_sweep(_dp(U), _dp(out), _n_ptr(n), nd, C.byref(s), dtdh, dim, nthreads)
    if rc:
        raise RuntimeError(f"orc_sweep failed: {rc}")
    return out


def max_wavespeed_grid(U, s, nthreads=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd, n = _dims(U)
    return lib().orc_max_wavespeed_grid(_dp(U), _n_ptr(n), nd, C.byref(s), nthreads)


def flux_delta_at(U, s, dtdh, dim, x, y=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd, n = _dims(U)
    out = np.zeros(2 + nd)
    lib().orc_flux_delta_at(_dp(U), _n_ptr(n), nd, C.byref(s), dtdh, dim, x, y, _dp(out))
    return out


def face_states_at(U, s, dtdh, dim, x, y=0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd, n = _dims(U)
    R = np.zeros(2 + nd)
    L = np.zeros(2 + nd)
    chi = np.zeros(1)
    lib().orc_face_states_at(_dp(U), _n_ptr(n), nd, C.byref(s), dtdh, dim, x, y,
                             _dp(R), _dp(L), _dp(chi))
    return R, L, float(chi[0])


def simulate(U, s, cfl, dh, sim_time=float("inf"), max_iters=1 << 62, nthreads=0,
             record_dt=False):
    """Returns (U_final, run_time, iters, dts)."""
    U = np.array(U, dtype=np.float64, order="C")
    nd, n = _dims(U)
    work = np.empty_like(U)
    rt = C.c_double(0.0)
    it = C.c_int64(0)
    dts = None
    dtp = None
    if record_dt:
        if max_iters > 10_000_000:
            raise ValueError("record_dt needs a bounded max_iters")
        dts = np.zeros(int(max_iters))
        dtp = _dp(dts)
    rc = lib().orc_simulate(_dp(U), _dp(work), _n_ptr(n), nd, C.byref(s), cfl, dh,
                            sim_time, int(max_iters), nthreads, C.byref(rt), C.byref(it), dtp)
    if rc:
        raise RuntimeError(f"orc_simulate failed: {rc}")
    if dts is not None:
        dts = dts[: it.value]
    return U, rt.value, it.value, dts
