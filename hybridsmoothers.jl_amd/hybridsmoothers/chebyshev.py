"""Chebyshev smoothers: host-side mirror of src/chebyshev.jl.

    s = _build_smoother(ChebyshevFourth(degree=3, backend=ROCBackend()), A)   # :141-156
    smooth_(x, s, b)                                                          # :158-183
    ldiv_(z, s, r)     # z = 0; smooth!(z, s, r): the smoother as a left preconditioner

Fields the reference's tests read are properties downloaded on access:
s.A (a device CSR, materialised as scipy CSR), s.l1inv, s.ρ / s.rho, s.a, s.r,
s.z, s.tmp, s.degree, s.backend.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .l1gs import ROCBackend, functional


class ChebyshevFourth:
    """ChebyshevFourth(; degree=4, backend) (:115-119)."""

    kind = 4

    def __init__(self, degree: int = 4, backend=None, precision: str = "fp64"):
        self.degree = int(degree)
        self.backend = backend if backend is not None else ROCBackend()
        self.precision = precision


class ChebyshevFirst:
    """ChebyshevFirst(; degree=4, backend) (:218-222)."""

    kind = 1

    def __init__(self, degree: int = 4, backend=None, precision: str = "fp64"):
        self.degree = int(degree)
        self.backend = backend if backend is not None else ROCBackend()
        self.precision = precision


_PREC = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32}


class ChebyshevSmoother:
    """ChebyshevFourthSmoother / ChebyshevFirstSmoother (:121-129, :224-234)."""

    def __init__(self, handle, config):
        self._h = handle
        self.config = config
        self.degree = config.degree
        self.backend = config.backend
        self.kind = config.kind
        info = np.zeros(5)
        L.check(L.lib.hs_cheb_info(self._h, info.ctypes.data))
        self.rho = float(info[0])
        self.a = float(info[1])
        self.n = int(info[4])

    @property
    def ρ(self):  # noqa: N802 - the reference's field name
        return self.rho

    def _get(self, what):
        out = np.zeros(self.n)
        L.check(L.lib.hs_cheb_get(self._h, what, out.ctypes.data))
        return out

    @property
    def l1inv(self):
        return self._get(0)

    @property
    def r(self):
        return self._get(1)

    @property
    def z(self):
        return self._get(2)

    @property
    def tmp(self):
        return self._get(3)

    @property
    def A(self):
        """The smoother's CSR operator, downloaded as a scipy CSR matrix."""
        return csr_download(L.lib.hs_cheb_matrix(self._h))

    @property
    def matrix_handle(self):
        return L.lib.hs_cheb_matrix(self._h)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.hs_cheb_destroy(h)
            self._h = None


def csr_download(h):
    info = np.zeros(2, np.int64)
    L.check(L.lib.hs_csr_info(h, info.ctypes.data))
    n, nnz = int(info[0]), int(info[1])
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(max(nnz, 1), np.int64)
    cv = np.zeros(max(nnz, 1))
    L.check(L.lib.hs_csr_get(h, rp.ctypes.data, ci.ctypes.data, cv.ctypes.data))
    return sp.csr_matrix((cv[:nnz], ci[:nnz] - 1, rp - 1), shape=(n, n))


def _build_smoother(config, A, power_iters: int = 50, stream=None) -> ChebyshevSmoother:
    """_build_smoother(cfg, A::SparseMatrixCSC) (:141-156, :258-276) on the device."""
    if not functional(config.backend):
        raise L.HSError(L.HS_ENODEVICE,
                        f" {config.backend} is not functional, please check your backend.")
    A = sp.csc_matrix(A)
    if not A.has_sorted_indices:
        A = A.sorted_indices()
    n = A.shape[0]
    ptr = A.indptr.astype(np.int64) + 1
    ind = A.indices.astype(np.int64) + 1
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    h = C.c_void_p()
    st = stream if stream is not None else L.current_stream()
    L.check(L.lib.hs_cheb_build(n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data,
                                config.kind, config.degree, power_iters,
                                _PREC[config.precision], st, C.byref(h)))
    return ChebyshevSmoother(h, config)


build_smoother = _build_smoother


def _call(fn, s, x, b, stream, sync):
    px, lx = L.vec_ptr(x, s.n)
    pb, lb = L.vec_ptr(b, s.n)
    if lx != lb:
        raise L.HSError(L.HS_EINVAL, "x and b must both be host or both be device vectors")
    st = stream if stream is not None else (L.current_stream() if lx == L.LOC_DEVICE else None)
    L.check(fn(s.handle, px, pb, lx, st, 0 if sync else L.FLAG_ASYNC))


def smooth_(x, s: ChebyshevSmoother, b, stream=None, sync=True):
    """smooth!(x, s, b): x is updated in place (:158-183, :278-322)."""
    _call(L.lib.hs_cheb_smooth, s, x, b, stream, sync)


def ldiv_(z, s: ChebyshevSmoother, r, stream=None, sync=True):
    """z = 0; smooth!(z, s, r) -- the smoother applied as a left preconditioner."""
    _call(L.lib.hs_cheb_apply, s, z, r, stream, sync)
