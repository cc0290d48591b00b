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
EOS_IDEAL, EOS_STIFF, EOS_STIFF_CORRECTED = 0, 1, 2
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
        u8p = C.POINTER(C.c_uint8)
        L.orc_ls_mask_value.argtypes = [C.c_double, C.c_double]
        L.orc_ls_mask_value.restype = C.c_uint8
        L.orc_ls_masks.argtypes = [dp, i64p, C.c_int, C.c_double, u8p]
        L.orc_ls_masks.restype = None
        L.orc_ls_velocity.argtypes = [C.c_int, C.POINTER(dp), C.POINTER(dp), i64p, C.c_int,
                                      C.c_double, C.c_int, dp]
        L.orc_ls_velocity.restype = None
        L.orc_ls_advect.argtypes = [dp, dp, i64p, C.c_int, C.c_double, C.c_double, dp]
        L.orc_ls_advect.restype = None
        L.orc_ls_weno.argtypes = [dp, C.c_int64, C.c_double, C.c_int64, dp, dp]
        L.orc_ls_weno.restype = None
        L.orc_ls_project2.argtypes = [dp, dp, C.c_int64]
        L.orc_ls_project2.restype = None
        L.orc_ls_reinit.argtypes = [dp, i64p, C.c_int, C.c_double, C.c_int, dp, u8p, i64p]
        L.orc_ls_reinit.restype = C.c_int
        L.orc_gfm_star.argtypes = [dp, dp, dp, dp, sp, sp, i64p, C.c_int, C.c_double, dp, u8p]
        L.orc_gfm_star.restype = None
        L.orc_gfm_extrapolate.argtypes = [dp, dp, dp, u8p, i64p, C.c_int, C.c_double, C.c_int]
        L.orc_gfm_extrapolate.restype = None
        L.orc_mm_combine.argtypes = [dp, dp, dp, i64p, C.c_int, dp]
        L.orc_mm_combine.restype = None
        L.orc_mm_simulate.argtypes = [dp, dp, dp, dp, i64p, C.c_int, sp, sp, C.c_double,
                                      C.c_double, C.c_double, C.c_int64, C.c_int, C.c_int,
                                      C.c_int, C.c_int, dp, i64p, dp]
        L.orc_mm_simulate.restype = C.c_int
        L.orc_mm_simulate_n.argtypes = [C.c_int, C.POINTER(dp), C.POINTER(dp), i64p, C.c_int,
                                        C.POINTER(sp), C.c_double, C.c_double, C.c_double,
                                        C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_double), C.POINTER(C.c_int64), dp]
        L.orc_mm_simulate_n.restype = C.c_int
        L.orc_ls_project_n.argtypes = [C.c_int, C.POINTER(dp), C.c_int64]
        L.orc_ls_project_n.restype = None
        L.orc_gfm_star_n.argtypes = [C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp),
                                     C.POINTER(sp), i64p, C.c_int, C.c_double, dp, u8p]
        L.orc_gfm_star_n.restype = None
        L.orc_write_ascii.argtypes = [dp, i64p, C.c_int, sp, dp, C.c_double, C.c_double,
                                      C.c_char_p]
        L.orc_write_ascii.restype = C.c_int
        L.orc_write_separate_raw.argtypes = [dp, i64p, C.c_int, sp, C.c_double, C.c_char_p]
        L.orc_write_separate_raw.restype = C.c_int
        L.orc_set_states.argtypes = [dp, dp, sp, dp, C.c_int64, C.c_int]
        L.orc_set_states.restype = None
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


# ---- level set / ghost fluid ---------------------------------------------------
# Level-set fields are float64 arrays of shape (ny, nx) in 2D, (nx,) in 1D.

def _u8p(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _ls_dims(phi: np.ndarray):
    if phi.ndim == 1:
        return 1, np.array([phi.shape[0], 1], dtype=np.int64)
    return 2, np.array([phi.shape[1], phi.shape[0]], dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ls_mask_value(phi, dh):
    return int(lib().orc_ls_mask_value(phi, dh))


def ls_masks(phi, dh):
    phi = _f64(phi)
    nd, n = _ls_dims(phi)
    m = np.zeros(phi.shape, dtype=np.uint8)
    lib().orc_ls_masks(_dp(phi), _n_ptr(n), nd, dh, _u8p(m))
    return m


def ls_velocity(phis, states, dh, target=-1):
    phis = [_f64(p) for p in phis]
    states = [_f64(u) for u in states]
    nd, n = _ls_dims(phis[0])
    dp = C.POINTER(C.c_double)
    pa = (dp * len(phis))(*[_dp(p) for p in phis])
    ua = (dp * len(states))(*[_dp(u) for u in states])
    V = np.zeros(phis[0].shape)
    lib().orc_ls_velocity(len(phis), pa, ua, _n_ptr(n), nd, dh, target, _dp(V))
    return V


def ls_advect(phi, V, dt, dh):
    phi, V = _f64(phi), _f64(V)
    nd, n = _ls_dims(phi)
    out = np.empty_like(phi)
    lib().orc_ls_advect(_dp(phi), _dp(V), _n_ptr(n), nd, dt, dh, _dp(out))
    return out


def ls_weno(phi, dh, i):
    """(D-phi, D+phi) of WENO-HJ5 at cell i of the 1D line phi (mirrored ghosts)."""
    phi = _f64(phi)
    dm = C.c_double(0.0)
    dp_ = C.c_double(0.0)
    lib().orc_ls_weno(_dp(phi), phi.shape[0], dh, int(i), C.byref(dm), C.byref(dp_))
    return dm.value, dp_.value


def ls_project2(phi0, phi1):
    a = np.array(phi0, dtype=np.float64, order="C")
    b = np.array(phi1, dtype=np.float64, order="C")
    lib().orc_ls_project2(_dp(a), _dp(b), a.size)
    return a, b


def ls_reinit(phi, dh, band):
    """Returns (phi_out, band_mask, iterations)."""
    phi = _f64(phi)
    nd, n = _ls_dims(phi)
    out = np.empty_like(phi)
    bm = np.zeros(phi.shape, dtype=np.uint8)
    it = C.c_int64(0)
    rc = lib().orc_ls_reinit(_dp(phi), _n_ptr(n), nd, dh, band, _dp(out), _u8p(bm), C.byref(it))
    if rc:
        raise RuntimeError(f"orc_ls_reinit failed: {rc}")
    return out, bm, it.value


def gfm_star(phi_m, phi_o, U_m, U_o, s_m, s_o, dh):
    """Returns (star, has_star): star states (AoS, zero where has_star is 0)."""
    phi_m, phi_o, U_m, U_o = _f64(phi_m), _f64(phi_o), _f64(U_m), _f64(U_o)
    nd, n = _ls_dims(phi_m)
    star = np.zeros_like(U_m)
    has = np.zeros(phi_m.shape, dtype=np.uint8)
    lib().orc_gfm_star(_dp(phi_m), _dp(phi_o), _dp(U_m), _dp(U_o), C.byref(s_m), C.byref(s_o),
                       _n_ptr(n), nd, dh, _dp(star), _u8p(has))
    return star, has


def gfm_extrapolate(phi_m, U_m, star, has_star, dh, sweeps):
    phi_m, star = _f64(phi_m), _f64(star)
    U = np.array(U_m, dtype=np.float64, order="C")
    has = np.ascontiguousarray(has_star, dtype=np.uint8)
    nd, n = _ls_dims(phi_m)
    lib().orc_gfm_extrapolate(_dp(phi_m), _dp(U), _dp(star), _u8p(has), _n_ptr(n), nd, dh, sweeps)
    return U


def mm_combine(phi0, U0, U1):
    phi0, U0, U1 = _f64(phi0), _f64(U0), _f64(U1)
    nd, n = _ls_dims(phi0)
    out = np.empty_like(U0)
    lib().orc_mm_combine(_dp(phi0), _dp(U0), _dp(U1), _n_ptr(n), nd, _dp(out))
    return out


def mm_simulate(U0, U1, phi0, phi1, s0, s1, cfl, dh, sim_time=float("inf"), max_iters=1,
                band=6, extrap_sweeps=8, reinit_every=1, nthreads=0):
    """Two-material ghost-fluid run; returns (U0, U1, phi0, phi1, run_time, iters, dts)."""
    U0 = np.array(U0, dtype=np.float64, order="C")
    U1 = np.array(U1, dtype=np.float64, order="C")
    phi0 = np.array(phi0, dtype=np.float64, order="C")
    phi1 = np.array(phi1, dtype=np.float64, order="C")
    nd, n = _ls_dims(phi0)
    if max_iters > 10_000_000:
        raise ValueError("mm_simulate needs a bounded max_iters")
    dts = np.zeros(int(max_iters))
    rt = C.c_double(0.0)
    it = C.c_int64(0)
    rc = lib().orc_mm_simulate(_dp(U0), _dp(U1), _dp(phi0), _dp(phi1), _n_ptr(n), nd,
                               C.byref(s0), C.byref(s1), cfl, dh, sim_time, int(max_iters),
                               band, extrap_sweeps, reinit_every, nthreads, C.byref(rt),
                               C.byref(it), _dp(dts))
    if rc:
        raise RuntimeError(f"orc_mm_simulate failed: {rc}")
    return U0, U1, phi0, phi1, rt.value, it.value, dts[: it.value]


def mm_simulate_n(Us, phis, schemes, cfl, dh, sim_time=float("inf"), max_iters=1, band=6,
                  extrap_sweeps=8, reinit_every=1, nthreads=0):
    """n-material ghost-fluid run (decision L9); returns (Us, phis, run_time,
    iters, dts) with copies of the inputs advanced."""
    Us = [np.array(U, dtype=np.float64, order="C") for U in Us]
    phis = [np.array(p, dtype=np.float64, order="C") for p in phis]
    nmat = len(Us)
    nd, n = _ls_dims(phis[0])
    if max_iters > 10_000_000:
        raise ValueError("mm_simulate_n needs a bounded max_iters")
    dts = np.zeros(int(max_iters))
    rt = C.c_double(0.0)
    it = C.c_int64(0)
    dp = C.POINTER(C.c_double)
    ua = (dp * nmat)(*[_dp(U) for U in Us])
    pa = (dp * nmat)(*[_dp(p) for p in phis])
    sa = (C.POINTER(Scheme) * nmat)(*[C.pointer(s) for s in schemes])
    rc = lib().orc_mm_simulate_n(nmat, ua, pa, _n_ptr(n), nd, sa, cfl, dh, sim_time,
                                 int(max_iters), band, extrap_sweeps, reinit_every, nthreads,
                                 C.byref(rt), C.byref(it), _dp(dts))
    if rc:
        raise RuntimeError(f"orc_mm_simulate_n failed: {rc}")
    return Us, phis, rt.value, it.value, dts[: it.value]


def ls_project_n(phis):
    """The n-level-set projection, on copies."""
    phis = [np.array(p, dtype=np.float64, order="C") for p in phis]
    dp = C.POINTER(C.c_double)
    pa = (dp * len(phis))(*[_dp(p) for p in phis])
    lib().orc_ls_project_n(len(phis), pa, phis[0].size)
    return phis


def gfm_star_n(m, phis, Us, schemes, dh):
    """Material m's star states among n materials: (star, has_star)."""
    phis = [np.ascontiguousarray(p, dtype=np.float64) for p in phis]
    Us = [np.ascontiguousarray(U, dtype=np.float64) for U in Us]
    nmat = len(phis)
    nd, n = _ls_dims(phis[0])
    star = np.zeros(Us[m].shape)
    has = np.zeros(phis[0].shape, dtype=np.uint8)
    dp = C.POINTER(C.c_double)
    pa = (dp * nmat)(*[_dp(p) for p in phis])
    ua = (dp * nmat)(*[_dp(U) for U in Us])
    sa = (C.POINTER(Scheme) * nmat)(*[C.pointer(s) for s in schemes])
    lib().orc_gfm_star_n(m, nmat, pa, ua, sa, _n_ptr(n), nd, dh, _dp(star),
                         has.ctypes.data_as(C.POINTER(C.c_uint8)))
    return star, has


# ---- initial conditions and result files (fl_oracle_io.cpp) ---------------

def fill_positions(n):
    """The positions SimulatorImpl::fill_data hands its fillers
    (simulator/simulator_impl.hpp:318-326): pos[d] = float(index_d) / size_d in
    single precision, index_d = (i / stride_d) % size_d
    (dimension/dimension_info.hpp flattened_index).  n = sizes, x first;
    returns an (cells, 3) float32 array, x fastest, unused dimensions 0."""
    n = [int(v) for v in n]
    cells = int(np.prod(n))
    pos = np.zeros((cells, 3), dtype=np.float32)
    stride = 1
    i = np.arange(cells, dtype=np.int64)
    for d, size in enumerate(n):
        idx = (i // stride) % size
        pos[:, d] = idx.astype(np.float32) / np.float32(size)
        stride *= size
    return pos


def fill_data(n, fillers, s):
    """SimulatorImpl::fill_data (simulator/simulator_impl.hpp:292-340): each
    filler (name, f) maps the float32 positions to float32 values of the
    primitive element `name`; every cell becomes conservative(fill_state).
    Elements no filler names keep the default 0 (the fill state is
    value-initialised once and reused).  Returns the AoS conservative grid."""
    n = [int(v) for v in n]
    nd = len(n)
    names = ["rho", "p", "v_x", "v_y"][:2 + nd]
    pos = fill_positions(n)
    W = np.zeros((pos.shape[0], 2 + nd), dtype=np.float64)
    for name, f in fillers:
        if name not in names:
            raise RuntimeError("State does not have a data element named: \n\t" + name)
        W[:, names.index(name)] = np.asarray(f(pos), dtype=np.float32)
    U = np.zeros_like(W)
    for i in range(W.shape[0]):
        lib().orc_prim_to_cons(_dp(W[i]), nd, 0, C.byref(s), _dp(U[i]))
    return U.reshape((n[0], 2 + nd) if nd == 1 else (n[1], n[0], 2 + nd))


def write_ascii(U, s, path, run_time, resolution, phi=None):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd, n = _dims(U)
    ph = None if phi is None else _dp(np.ascontiguousarray(phi, dtype=np.float64))
    if lib().orc_write_ascii(_dp(U), _n_ptr(n), nd, C.byref(s), ph, run_time, resolution,
                             str(path).encode()):
        raise OSError(f"cannot write {path}")


def write_separate_raw(U, s, sim_time, prefix):
    U = np.ascontiguousarray(U, dtype=np.float64)
    nd, n = _dims(U)
    if lib().orc_write_separate_raw(_dp(U), _n_ptr(n), nd, C.byref(s), sim_time,
                                    str(prefix).encode()):
        raise OSError(f"cannot write {prefix}_*")


def set_states(phi, W, s, U):
    """In place on the AoS grid U: cells with phi <= 0 become conservative(W)."""
    nd = U.shape[-1] - 2
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64)
    assert U.dtype == np.float64 and U.flags.c_contiguous
    lib().orc_set_states(_dp(phi), _dp(W), C.byref(s), _dp(U), phi.size, nd)
