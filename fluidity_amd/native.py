"""ctypes binding of libfluidity_amd.so (the C ABI in include/fluidity_amd.h).

Device buffers are torch tensors on a ROCm device; torch is only the
allocator and stream provider here.  Every call goes through the HIP
library: if the library is missing this module raises at import time instead
of falling back to any other implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FLUIDITY_AMD_LIB selects another build of the same library (tuning builds);
# by default the in-tree library is loaded.
LIB_PATH = os.environ.get("FLUIDITY_AMD_LIB") or os.path.join(HERE, "libfluidity_amd.so")

FL_OK, FL_ERR_ARGUMENT, FL_ERR_UNSUPPORTED, FL_ERR_HIP, FL_ERR_NUMERIC = range(5)
FL_F64, FL_F32 = 0, 1
FL_LAYOUT_AOS, FL_LAYOUT_SOA = 0, 1
FL_LIMITER_VAN_LEER, FL_LIMITER_SUPERBEE = 0, 1
FL_FORM_CONSERVATIVE, FL_FORM_PRIMITIVE = 0, 1
FL_FLUX_HLLC, FL_FLUX_FORCE, FL_FLUX_LAX_FRIEDRICHS, FL_FLUX_RICHTMYER = 0, 1, 2, 3
FL_EOS_IDEAL, FL_EOS_STIFF, FL_EOS_STIFF_CORRECTED = 0, 1, 2
FL_NUMERICS_EXACT, FL_NUMERICS_FAST = 0, 1
FL_BC_TRANSMISSIVE, FL_BC_REFLECTIVE, FL_BC_PERIODIC = 0, 1, 2

EXPORTED = [
    "fl_status_string", "fl_last_error", "fl_version", "fl_ctx_create", "fl_ctx_destroy",
    "fl_ctx_set_tuning", "fl_ctx_get_tuning", "fl_ctx_autotune", "fl_ctx_set_graphs",
    "fl_split_sweep",
    "fl_split_solve", "fl_split_step",
    "fl_max_wavespeed", "fl_set_wavespeeds", "fl_max_element", "fl_simulate",
    "fl_simulate_noalloc",
    "fl_conservation_sums",
    "fl_ls_masks", "fl_ls_velocity", "fl_ls_advect", "fl_ls_project",
    "fl_ls_reinit_work_bytes", "fl_ls_reinit", "fl_gfm_star", "fl_gfm_extrapolate_work_bytes",
    "fl_gfm_extrapolate",
    "fl_mm_combine", "fl_mm_work_bytes", "fl_mm_work_bytes_n", "fl_mm_simulate",
    "fl_batch_create", "fl_batch_destroy", "fl_batch_set_tuning", "fl_batch_get_tuning",
    "fl_batch_autotune", "fl_batch_reset", "fl_batch_step", "fl_batch_finish",
    "fl_batch_conservation_sums",
    "fl_adj_checkpoint_bytes", "fl_adj_work_bytes", "fl_adj_forward", "fl_adj_backward",
    "fl_split_step_vjp",
    "fl_primitive_planes", "fl_from_primitive", "fl_set_states", "fl_write_results",
    "fl_write_results_separate_raw",
]

# entry points returning a size rather than a status
_SIZE_RETURNING = ("fl_ls_reinit_work_bytes", "fl_mm_work_bytes", "fl_mm_work_bytes_n",
                   "fl_adj_checkpoint_bytes",
                   "fl_gfm_extrapolate_work_bytes",
                   "fl_adj_work_bytes")


class Field(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("dtype", C.c_int32),
        ("layout", C.c_int32),
        ("ndim", C.c_int32),
        ("reserved", C.c_int32),
        ("n", C.c_int64 * 3),
    ]


class Scheme(C.Structure):
    _fields_ = [
        ("limiter", C.c_int32),
        ("form", C.c_int32),
        ("flux", C.c_int32),
        ("eos", C.c_int32),
        ("gamma", C.c_double),
        ("p_inf", C.c_double),
        ("bc", C.c_int32 * 3),
        ("numerics", C.c_int32),
    ]


class SimParams(C.Structure):
    _fields_ = [
        ("cfl", C.c_double),
        ("dh", C.c_double),
        ("sim_time", C.c_double),
        ("max_iters", C.c_int64),
    ]


class SimResult(C.Structure):
    _fields_ = [
        ("run_time", C.c_double),
        ("iters", C.c_int64),
        ("last_dt", C.c_double),
        ("max_wavespeed", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("x_block", C.c_int32),
        ("x_rows", C.c_int32),
        ("y_block", C.c_int32),
        ("y_rows", C.c_int32),
    ]


class Material(C.Structure):
    _fields_ = [
        ("state", Field),
        ("work", Field),
        ("phi", C.c_void_p),
        ("scheme", Scheme),
    ]


class MmParams(C.Structure):
    _fields_ = [
        ("band", C.c_int32),
        ("extrap_sweeps", C.c_int32),
        ("reinit_every", C.c_int32),
        ("reserved", C.c_int32),
    ]


class BatchProblem(C.Structure):
    _fields_ = [
        ("state", Field),
        ("work", Field),
        ("scheme", Scheme),
        ("params", SimParams),
    ]


class FluidityError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m fluidity_amd.build` "
            "(there is no fallback implementation)")
    # torch ships its own HIP and HSA runtimes (SONAMEs libamdhip64.so.7 and
    # libhsa-runtime64.so.1, requested by torch's libraries under the
    # unversioned names).  With torch loaded first, this library's
    # libamdhip64.so.7 dependency binds to that runtime and the process holds
    # one.  Loaded first, this library would bring in /opt/rocm's runtime and
    # torch a second one; whichever of the two initialises later then finds
    # no device.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    fp = C.POINTER(Field)
    sp = C.POINTER(Scheme)
    L.fl_status_string.argtypes = [C.c_int]
    L.fl_status_string.restype = C.c_char_p
    L.fl_last_error.argtypes = []
    L.fl_last_error.restype = C.c_char_p
    L.fl_version.argtypes = []
    L.fl_version.restype = C.c_char_p
    L.fl_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.fl_ctx_destroy.argtypes = [vp]
    L.fl_ctx_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.fl_ctx_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.fl_ctx_autotune.argtypes = [vp, fp, fp, sp, C.c_int, C.POINTER(Tuning), vp]
    L.fl_ctx_set_graphs.argtypes = [vp, C.c_int]
    L.fl_split_sweep.argtypes = [vp, fp, fp, C.c_int, sp, C.c_double, vp]
    L.fl_split_solve.argtypes = [vp, fp, fp, sp, C.c_double, vp]
    L.fl_split_step.argtypes = [vp, fp, fp, sp, C.c_double, C.POINTER(C.c_double), vp]
    L.fl_max_wavespeed.argtypes = [vp, fp, sp, C.POINTER(C.c_double), vp]
    L.fl_simulate.argtypes = [vp, fp, fp, sp, C.POINTER(SimParams), C.POINTER(SimResult), vp]
    L.fl_simulate_noalloc.argtypes = [C.POINTER(Tuning), fp, fp, sp, C.POINTER(SimParams),
                                      C.POINTER(SimResult), vp]
    L.fl_conservation_sums.argtypes = [vp, fp, C.POINTER(C.c_double), vp]
    i64p = C.POINTER(C.c_int64)
    L.fl_ls_masks.argtypes = [vp, C.c_int, i64p, C.c_double, vp, vp]
    L.fl_ls_velocity.argtypes = [C.c_int, C.POINTER(vp), fp, C.c_double, C.c_int, vp, vp]
    L.fl_ls_advect.argtypes = [vp, vp, C.c_int, i64p, C.c_double, C.c_double, vp, vp, vp]
    L.fl_ls_project.argtypes = [vp, vp, C.c_int64, vp]
    L.fl_ls_reinit_work_bytes.argtypes = [C.c_int, i64p]
    L.fl_ls_reinit.argtypes = [vp, vp, C.c_int, i64p, C.c_double, C.c_int, vp, vp, vp, i64p, vp]
    L.fl_gfm_star.argtypes = [vp, vp, fp, fp, sp, sp, C.c_double, vp, vp, vp]
    L.fl_gfm_extrapolate_work_bytes.argtypes = [C.c_int, i64p]
    L.fl_gfm_extrapolate.argtypes = [vp, fp, vp, vp, C.c_double, C.c_int, vp, vp]
    L.fl_mm_combine.argtypes = [vp, fp, fp, fp, vp]
    L.fl_mm_work_bytes.argtypes = [C.c_int, i64p]
    L.fl_mm_work_bytes_n.argtypes = [C.c_int, C.c_int, i64p]
    L.fl_mm_simulate.argtypes = [vp, C.POINTER(Material), C.c_int, C.POINTER(SimParams),
                                 C.POINTER(MmParams), vp, C.c_int64, C.POINTER(SimResult), vp]
    L.fl_batch_create.argtypes = [vp, C.POINTER(BatchProblem), C.c_int, C.POINTER(vp), vp]
    L.fl_batch_destroy.argtypes = [vp]
    L.fl_batch_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.fl_batch_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.fl_batch_autotune.argtypes = [vp, C.c_int, C.POINTER(Tuning), vp]
    L.fl_batch_reset.argtypes = [vp, vp]
    L.fl_batch_step.argtypes = [vp, C.c_int64, C.c_int, vp]
    L.fl_batch_finish.argtypes = [vp, C.POINTER(SimResult), vp]
    L.fl_batch_conservation_sums.argtypes = [vp, C.POINTER(C.c_double), vp]
    L.fl_adj_checkpoint_bytes.argtypes = [fp, C.c_int64, C.c_int]
    L.fl_adj_work_bytes.argtypes = [fp, C.c_int]
    L.fl_adj_forward.argtypes = [vp, fp, sp, C.c_double, C.c_double, C.c_int64, C.c_int, vp,
                                 C.c_int64, vp, C.c_int64, C.POINTER(SimResult), vp]
    L.fl_adj_backward.argtypes = [vp, fp, sp, C.c_double, C.c_double, C.c_int64, C.c_int, vp,
                                  C.c_int64, vp, C.c_int64, C.POINTER(C.c_double), vp]
    L.fl_split_step_vjp.argtypes = [vp, fp, fp, fp, sp, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), vp, C.c_int64, vp]
    L.fl_primitive_planes.argtypes = [vp, fp, sp, vp, vp]
    L.fl_from_primitive.argtypes = [vp, vp, sp, fp, vp]
    L.fl_set_states.argtypes = [vp, vp, C.POINTER(C.c_double), sp, fp, vp]
    L.fl_write_results.argtypes = [vp, fp, sp, vp, C.c_double, C.c_double, C.c_char_p, vp]
    L.fl_write_results_separate_raw.argtypes = [vp, fp, sp, C.c_double, C.c_char_p, vp]
    for name in EXPORTED:
        if name in _SIZE_RETURNING:
            getattr(L, name).restype = C.c_int64
        elif name not in ("fl_status_string", "fl_last_error", "fl_version"):
            getattr(L, name).restype = C.c_int
    return L


lib = _load()


def check(status: int) -> None:
    if status != FL_OK:
        raise FluidityError(status, lib.fl_last_error().decode())


def scheme(limiter=FL_LIMITER_VAN_LEER, form=FL_FORM_CONSERVATIVE, flux=FL_FLUX_HLLC,
           eos=FL_EOS_IDEAL, gamma=1.4, p_inf=0.0, bc=(0, 0, 0),
           numerics=FL_NUMERICS_EXACT) -> Scheme:
    s = Scheme()
    s.limiter, s.form, s.flux, s.eos = limiter, form, flux, eos
    s.numerics = numerics
    s.gamma, s.p_inf = gamma, p_inf
    for i in range(3):
        s.bc[i] = bc[i] if i < len(bc) else 0
    return s


def field(data_ptr: int, ndim: int, n, dtype: int = FL_F64,
          layout: int = FL_LAYOUT_AOS) -> Field:
    f = Field()
    f.data = data_ptr
    f.dtype = dtype
    f.layout = layout
    f.ndim = ndim
    for i in range(3):
        f.n[i] = n[i] if i < len(n) else 1
    return f


class Context:
    """Owns an fl_context on one device."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib.fl_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib.fl_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, x_block=0, x_rows=0, y_block=0, y_rows=0):
        t = Tuning(x_block, x_rows, y_block, y_rows)
        check(lib.fl_ctx_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self) -> Tuning:
        t = Tuning()
        check(lib.fl_ctx_get_tuning(self.handle, C.byref(t)))
        return t
