"""ctypes binding of libhsmooth.so (the C ABI declared in include/hsmooth.h).

The library is the only compute path: if it is missing this module raises at
import time instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhsmooth.so")

# status codes (hsmooth.h)
HS_OK = 0
HS_EINVAL = -1
HS_ESINGULAR = -2
HS_EDEVICE = -3
HS_ENOMEM = -4
HS_ENODEVICE = -5
HS_EUNSUPPORTED = -6
HS_EAGAIN = -7

FMT_CSC, FMT_CSR = 0, 1
SWEEP_FORWARD, SWEEP_BACKWARD, SWEEP_SYMMETRIC = 0, 1, 2
CACHE_PACKED, CACHE_VIEW, CACHE_PACKED_SYM = 0, 1, 2
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
LOC_HOST, LOC_DEVICE = 0, 1
FLAG_ASYNC = 1
GET_D_DL1, GET_SL, GET_SU = 0, 1, 2
PCG_NONE, PCG_L1GS, PCG_CHEB = 0, 1, 2

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python hybridsmoothers.jl_amd/build.py` "
        "(hipcc, gfx950) -- there is no fallback compute path")

# One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64 with the
# same SONAME (libamdhip64.so.7) as /opt/rocm's; loading torch first lets the
# dynamic loader resolve libhsmooth's dependency to that already-loaded copy,
# so device pointers and streams from torch are valid in this library.
try:
    import torch  # noqa: F401
except ImportError:  # torch is optional (plain host vectors need no torch)
    torch = None

lib = C.CDLL(LIB_PATH)

vp = C.c_void_p
i64 = C.c_int64
i64p = C.POINTER(C.c_int64)
dp = C.c_void_p


class PCGStats(C.Structure):
    _fields_ = [("iters", C.c_int64), ("status", C.c_int64), ("rnorm0", C.c_double),
                ("rnorm", C.c_double), ("true_relres", C.c_double), ("seconds", C.c_double)]


class BatchConfig(C.Structure):
    """hs_batch_config"""
    _fields_ = [("arena_bytes", C.c_int64), ("max_systems", C.c_int64),
                ("staging_bytes", C.c_int64), ("partsize", C.c_int64),
                ("precision", C.c_int), ("passes", C.c_int), ("grid", C.c_int),
                ("eta", C.c_double), ("triangles", C.c_int), ("smoother", C.c_int),
                ("degree", C.c_int), ("power_iters", C.c_int)]


class BatchResult(C.Structure):
    """hs_batch_result"""
    _fields_ = [("id", C.c_int64), ("n", C.c_int64), ("iters", C.c_int64),
                ("status", C.c_int64), ("rnorm0", C.c_double), ("rnorm", C.c_double),
                ("true_relres", C.c_double), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double)]


def _sig(name, args, res=C.c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("hs_init", [C.c_int])
_sig("hs_last_error", [], C.c_char_p)
_sig("hs_last_error_aux", [], C.c_int64)
_sig("hs_device_bytes", [], C.c_int64)
_sig("hs_l1gs_build", [i64, vp, vp, vp, C.c_int, C.c_int, i64, C.c_double, C.c_int, C.c_int,
                       C.c_int, i64, i64, vp, C.POINTER(vp)])
_sig("hs_l1gs_apply", [vp, vp, vp, C.c_int, vp, C.c_int])
_sig("hs_l1gs_info", [vp, vp])
_sig("hs_l1gs_get", [vp, C.c_int, vp])
_sig("hs_l1gs_destroy", [vp])
_sig("hs_l1gs_set_variant", [vp, C.c_int])
_sig("hs_l1gs_save", [vp, C.c_char_p])
_sig("hs_l1gs_load", [C.c_char_p, vp, C.POINTER(vp)])

EXPORTED = [
    "hs_init", "hs_last_error", "hs_last_error_aux", "hs_device_bytes",
    "hs_l1gs_build", "hs_l1gs_apply", "hs_l1gs_info", "hs_l1gs_get", "hs_l1gs_destroy",
    "hs_l1gs_set_variant", "hs_l1gs_save", "hs_l1gs_load",
    "hs_csr_create", "hs_csr_spmv", "hs_csr_info", "hs_csr_get", "hs_csr_destroy",
    "hs_csr_column_bytes",
    "hs_cheb_build", "hs_cheb_smooth", "hs_cheb_apply", "hs_cheb_info", "hs_cheb_get",
    "hs_cheb_matrix", "hs_cheb_destroy", "hs_pcg_solve", "hs_pcg_solve_true", "hs_l1gs_smooth", "hs_l1gs_smooth_noalloc",
    "hs_problem_lognormal",
    "hs_batch_create", "hs_batch_footprint", "hs_batch_submit", "hs_batch_step",
    "hs_batch_collect", "hs_batch_info", "hs_batch_clock", "hs_batch_peek", "hs_batch_busy", "hs_batch_destroy",
]


def bind_optional():
    """Bind the entry points beyond L1GS (present once their sources are built)."""
    if hasattr(lib, "hs_csr_create"):
        _sig("hs_csr_create", [i64, vp, vp, vp, C.c_int, vp, C.POINTER(vp)])
        _sig("hs_csr_spmv", [vp, vp, vp, C.c_int, vp, C.c_int])
        _sig("hs_csr_info", [vp, vp])
        _sig("hs_csr_get", [vp, vp, vp, vp])
    if hasattr(lib, "hs_csr_column_bytes"):
        _sig("hs_csr_column_bytes", [vp, vp])
        _sig("hs_csr_destroy", [vp])
    if hasattr(lib, "hs_cheb_build"):
        _sig("hs_cheb_build", [i64, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                               C.POINTER(vp)])
        _sig("hs_cheb_smooth", [vp, vp, vp, C.c_int, vp, C.c_int])
        _sig("hs_cheb_apply", [vp, vp, vp, C.c_int, vp, C.c_int])
        _sig("hs_cheb_info", [vp, vp])
        _sig("hs_cheb_get", [vp, C.c_int, vp])
        _sig("hs_cheb_matrix", [vp], vp)
        _sig("hs_cheb_destroy", [vp])
    if hasattr(lib, "hs_pcg_solve"):
        _sig("hs_pcg_solve", [vp, C.c_int, vp, vp, vp, C.c_double, C.c_double, i64, C.c_int, vp,
                              C.POINTER(PCGStats)])
    if hasattr(lib, "hs_pcg_solve_true"):
        _sig("hs_pcg_solve_true", [vp, C.c_int, vp, vp, vp, C.c_double, C.c_double, i64,
                                   C.c_double, C.c_int, vp, C.POINTER(PCGStats)])
    if hasattr(lib, "hs_l1gs_smooth"):
        _sig("hs_l1gs_smooth", [vp, vp, vp, vp, C.c_int, C.c_int, vp, C.c_int])
    if hasattr(lib, "hs_l1gs_smooth_noalloc"):
        _sig("hs_l1gs_smooth_noalloc", [vp, vp, vp, vp, C.c_int, vp, C.c_int])
    if hasattr(lib, "hs_problem_lognormal"):
        _sig("hs_problem_lognormal", [i64, i64, i64, C.c_double, C.c_uint64, vp, vp, vp, vp])
    if hasattr(lib, "hs_batch_create"):
        _sig("hs_batch_create", [C.POINTER(BatchConfig), C.POINTER(vp)])
        _sig("hs_batch_footprint", [vp, i64, vp, vp])
        _sig("hs_batch_submit", [vp, i64, vp, vp, vp, vp, vp, C.c_double, C.c_double, i64, vp])
        _sig("hs_batch_step", [vp])
        _sig("hs_batch_collect", [vp, vp, i64, vp])
        _sig("hs_batch_info", [vp, vp])
        _sig("hs_batch_clock", [vp], C.c_double)
        _sig("hs_batch_peek", [vp, i64, C.c_int, vp])
        _sig("hs_batch_busy", [vp])
        _sig("hs_batch_destroy", [vp])


bind_optional()


class HSError(RuntimeError):
    """ErrorException of the reference: any failure other than a singular diagonal."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class SingularException(HSError):
    """LinearAlgebra.SingularException(col) (src/l1_gauss_seidel.jl:429-431)."""

    def __init__(self, col, msg=""):
        super().__init__(HS_ESINGULAR, msg or f"SingularException({col})")
        self.col = col


def check(status):
    if status == HS_OK:
        return
    msg = lib.hs_last_error().decode(errors="replace")
    if status == HS_ESINGULAR:
        raise SingularException(int(lib.hs_last_error_aux()), msg)
    raise HSError(status, msg)


_initialized = {}


def init(device: int = 0):
    if device not in _initialized:
        check(lib.hs_init(device))
        _initialized[device] = True


# ---------------------------------------------------------------------------
# vectors: numpy (host) or torch CUDA tensors (device)
# ---------------------------------------------------------------------------
def vec_ptr(v, n):
    """(pointer, loc) of a float64 vector of length n; numpy -> host, torch cuda -> device."""
    if isinstance(v, np.ndarray):
        if v.dtype != np.float64 or not v.flags.c_contiguous or v.shape != (n,):
            raise HSError(HS_EINVAL, f"expected a contiguous float64 vector of length {n}")
        return v.ctypes.data, LOC_HOST
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(v, torch.Tensor):
        if v.dtype != torch.float64 or not v.is_contiguous() or tuple(v.shape) != (n,):
            raise HSError(HS_EINVAL, f"expected a contiguous float64 tensor of length {n}")
        if not v.is_cuda:
            return v.data_ptr(), LOC_HOST
        return v.data_ptr(), LOC_DEVICE
    raise HSError(HS_EINVAL, f"unsupported vector type {type(v)!r}")


def current_stream():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.current_stream().cuda_stream
    except Exception:  # pragma: no cover
        pass
    return None
