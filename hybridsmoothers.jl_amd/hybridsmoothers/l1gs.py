"""L1 Gauss-Seidel preconditioner: host-side mirror of HybridSmoothers.jl's API.

Reference: src/l1_gauss_seidel.jl (v0.1.4).  Names, argument meaning, defaults
and errors follow the Julia package:

    builder = L1GSPrecBuilder(ROCBackend(), threads=256, blocks=20)   # :258-262
    P = builder(A, 16, isSymA=True, sweep=SymmetricSweep())           # :287-296
    ldiv_(y, P, x)        # LinearSolve.ldiv!(y, P, x)                 :367-384
    y = backslash(P, x)   # P \\ x                                      :398-407

All computation runs in libhsmooth.so on the GPU.  The Julia KA.CPU backend is
not part of this library (it is the reference's own CPU path); asking for it
raises instead of silently running anything on the host.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib as L


# ---------------------------------------------------------------------------
# sweeps, caches, backends, configs (:11-100, :198-224)
# ---------------------------------------------------------------------------
class AbstractSweep:
    code = -1

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def __repr__(self):
        return f"{type(self).__name__}()"


class ForwardSweep(AbstractSweep):
    """M = D + L (:18)."""
    code = L.SWEEP_FORWARD


class BackwardSweep(AbstractSweep):
    """M = D + U (:25)."""
    code = L.SWEEP_BACKWARD


class SymmetricSweep(AbstractSweep):
    """M = (D + L) D^-1 (D + U) (:83)."""
    code = L.SWEEP_SYMMETRIC


class AbstractCacheStrategy:
    code = -1

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def __repr__(self):
        return f"{type(self).__name__}()"


class PackedBufferCache(AbstractCacheStrategy):
    """Packed strict triangles (:93); the GPU default (:614)."""
    code = L.CACHE_PACKED


class MatrixViewCache(AbstractCacheStrategy):
    """Read the in-partition entries from the matrix (:100)."""
    code = L.CACHE_VIEW


class PackedSymmetricCache(AbstractCacheStrategy):
    """Packed lower triangle only, U = L^T inside each partition: for a
    symmetric A (isSymA=True) with SymmetricSweep and partsize <= 64.  Not in
    the reference, whose PackedBufferCache stores both triangles (:788-799);
    the sweep gives the same results from half the packed bytes."""
    code = L.CACHE_PACKED_SYM


class Backend:
    def __repr__(self):
        return f"{type(self).__name__}()"


class CPU(Backend):
    """KernelAbstractions.CPU(); present for API shape only (see module docstring)."""


class ROCBackend(Backend):
    """The MI355X device (AMDGPU.jl's ROCBackend)."""

    def __init__(self, device: int = 0):
        self.device = device

    def __eq__(self, other):
        return isinstance(other, ROCBackend) and other.device == self.device

    def __hash__(self):
        return hash(("roc", self.device))


GPU = ROCBackend


def functional(backend) -> bool:
    if not isinstance(backend, ROCBackend):
        return False
    try:
        L.init(backend.device)
        return True
    except L.HSError:
        return False


class DeviceConfig:
    pass


class CPUConfig(DeviceConfig):
    def __init__(self, chunks: int):
        if chunks <= 0:
            raise L.HSError(L.HS_EINVAL, "`chunks` must be greater than 0")
        self.chunks = int(chunks)


class GPUConfig(DeviceConfig):
    def __init__(self, threads: int, blocks: int):
        if threads <= 0:
            raise L.HSError(L.HS_EINVAL, "`threads` must be greater than 0")
        if blocks <= 0:
            raise L.HSError(L.HS_EINVAL, "`blocks` must be greater than 0")
        self.threads = int(threads)
        self.blocks = int(blocks)


class Symmetric:
    """LinearAlgebra.Symmetric(A, uplo): only the `uplo` triangle of A is read."""

    def __init__(self, A, uplo: str = "U"):
        self.data = A
        self.uplo = uplo

    def full(self):
        """get_data (:604-610): tri + tri' - Diagonal(tri), same format as the input."""
        A = self.data
        tri = sp.triu(A) if self.uplo == "U" else sp.tril(A)
        full = tri + tri.T - sp.diags(tri.diagonal())
        return full.tocsr() if sp.isspmatrix_csr(A) else full.tocsc()


# ---------------------------------------------------------------------------
# builder (:240-338)
# ---------------------------------------------------------------------------
class L1GSPrecBuilder:
    def __init__(self, backend, device_config=None, *, chunks=None, threads=None, blocks=None):
        if device_config is None:
            if isinstance(backend, CPU):
                device_config = CPUConfig(chunks if chunks is not None else 1)
            else:
                if threads is None or blocks is None:
                    raise TypeError("L1GSPrecBuilder(::GPU; threads, blocks) needs both keywords")
                device_config = GPUConfig(threads, blocks)
        if isinstance(backend, CPU):
            raise L.HSError(L.HS_EUNSUPPORTED,
                            "KA.CPU() is the reference's host path; this library runs the "
                            "preconditioner on the MI355X (use ROCBackend())")
        if not isinstance(device_config, GPUConfig):
            raise TypeError("a GPU backend takes a GPUConfig")
        if not functional(backend):
            raise L.HSError(L.HS_ENODEVICE,
                            f" {backend} is not functional, please check your backend.")
        self.backend = backend
        self.device_config = device_config

    def default_partsize(self):
        return self.device_config.blocks  # _default_partsize(::GPUConfig) (:265-266)

    def __call__(self, A, partsize=None, *, isSymA=False, eta=1.5, η=None,
                 sweep=None, cache_strategy=None, precision="fp64", stream=None):
        if η is not None:
            eta = η
        if partsize is None:
            partsize = self.default_partsize()
        if isinstance(A, Symmetric):
            A = A.full()
            isSymA = True
        if sweep is None:
            sweep = SymmetricSweep()
        if cache_strategy is None:
            cache_strategy = PackedBufferCache()  # _choose_default_device_storage(::GPU)
        return build_l1prec(self, A, int(partsize), bool(isSymA), float(eta), sweep,
                            cache_strategy, precision, stream)


@dataclass
class BlockPartitioning:
    """(:118-124)"""
    partsize: int
    nparts: int
    nchunks: int
    chunksize: int
    backend: Backend


_PREC = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32, "bf16": L.PREC_BF16}


def sparse_arrays(A):
    """1-based Int64 (ptr, ind) + Float64 values of a scipy CSC/CSR matrix, and its format."""
    if sp.isspmatrix_csr(A) or isinstance(A, sp.csr_array):
        fmt = L.FMT_CSR
    else:
        if not (sp.isspmatrix_csc(A) or isinstance(A, sp.csc_array)):
            A = sp.csc_matrix(A)
        fmt = L.FMT_CSC
    if not A.has_sorted_indices:
        A = A.sorted_indices()
    n, m = A.shape
    if n != m:
        raise L.HSError(L.HS_EINVAL, "matrix must be square")
    ptr = A.indptr.astype(np.int64) + 1
    ind = A.indices.astype(np.int64) + 1
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    return n, ptr, ind, val, fmt


def build_l1prec(builder, A, partsize, isSymA, eta, sweep, cache_strategy, precision="fp64",
                 stream=None):
    """build_l1prec (:317-338) on the device."""
    if partsize == 0:
        raise L.HSError(L.HS_EINVAL, "partsize must be greater than 0")
    n, ptr, ind, val, fmt = sparse_arrays(A)
    cfg = builder.device_config
    h = C.c_void_p()
    st = stream if stream is not None else L.current_stream()
    L.check(L.lib.hs_l1gs_build(n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data, fmt,
                                int(isSymA), partsize, eta, sweep.code, cache_strategy.code,
                                _PREC[precision], cfg.threads, cfg.blocks, st, C.byref(h)))
    return L1GSPreconditioner(h, builder.backend, sweep, cache_strategy, precision)


# ---------------------------------------------------------------------------
# preconditioner object and its lazily downloaded fields (:180-183, :487-537)
# ---------------------------------------------------------------------------
class _Packed:
    def __init__(self, P, what):
        self._P, self._what = P, what

    @property
    def SLbuffer(self):
        if self._what != L.GET_SL:
            raise AttributeError("SLbuffer")
        return self._P._get(L.GET_SL)

    @property
    def SUbuffer(self):
        if self._what != L.GET_SU:
            raise AttributeError("SUbuffer")
        return self._P._get(L.GET_SU)


class _Op:
    def __init__(self, P, what):
        self._P, self._what = P, what

    @property
    def D_DL1(self):
        return self._P._get(L.GET_D_DL1)

    @property
    def L(self):
        if self._what != L.GET_SL:
            raise AttributeError("L")
        return _Packed(self._P, L.GET_SL)

    @property
    def U(self):
        if self._what != L.GET_SU:
            raise AttributeError("U")
        return _Packed(self._P, L.GET_SU)


class _SweepPlan:
    def __init__(self, P, sweep):
        self._P = P
        self._sweep = sweep

    @property
    def op(self):
        if isinstance(self._sweep, SymmetricSweep):
            raise AttributeError("op (a SymmetricL1GSSweep has lop and uop)")
        return _Op(self._P, L.GET_SL if isinstance(self._sweep, ForwardSweep) else L.GET_SU)

    @property
    def lop(self):
        if not isinstance(self._sweep, SymmetricSweep):
            raise AttributeError("lop")
        return _Op(self._P, L.GET_SL)

    @property
    def uop(self):
        if not isinstance(self._sweep, SymmetricSweep):
            raise AttributeError("uop")
        return _Op(self._P, L.GET_SU)


class L1GSPreconditioner:
    def __init__(self, handle, backend, sweep, cache_strategy, precision):
        self._h = handle
        self.backend = backend
        self.cache_strategy = cache_strategy
        self.precision = precision
        info = np.zeros(8, np.int64)
        L.check(L.lib.hs_l1gs_info(self._h, info.ctypes.data))
        self.partitioning = BlockPartitioning(int(info[0]), int(info[1]), int(info[2]),
                                              int(info[3]), backend)
        self.n = int(info[4])
        self.packed_size = int(info[5])
        self.sweep = _SweepPlan(self, sweep)
        self.sweep_kind = sweep

    @property
    def handle(self):
        return self._h

    def set_variant(self, variant: int):
        """Selects the kernel variant of the partsize-64 symmetric sweep (0-2)."""
        L.check(L.lib.hs_l1gs_set_variant(self._h, int(variant)))

    def save(self, path):
        """Writes the built preconditioner to `path` (see load)."""
        L.check(L.lib.hs_l1gs_save(self._h, os.fsencode(path)))

    def _get(self, what):
        size = self.n if what == L.GET_D_DL1 else self.packed_size
        out = np.zeros(size)
        L.check(L.lib.hs_l1gs_get(self._h, what, out.ctypes.data))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.hs_l1gs_destroy(h)
            self._h = None

    def __repr__(self):
        return (f"L1GSPreconditioner({type(self.sweep_kind).__name__[:-5]}, ROCBackend, "
                f"nparts={self.partitioning.nparts})")


_SWEEPS = {L.SWEEP_FORWARD: ForwardSweep, L.SWEEP_BACKWARD: BackwardSweep,
           L.SWEEP_SYMMETRIC: SymmetricSweep}
_CACHES = {L.CACHE_PACKED: PackedBufferCache, L.CACHE_VIEW: MatrixViewCache,
           L.CACHE_PACKED_SYM: PackedSymmetricCache}


def load(path, backend=None, stream=None) -> L1GSPreconditioner:
    """Recreates a preconditioner written by L1GSPreconditioner.save on the
    device, without the matrix and without rebuilding it."""
    backend = backend or ROCBackend()
    if not functional(backend):
        raise L.HSError(L.HS_ENODEVICE, f" {backend} is not functional, please check your backend.")
    h = C.c_void_p()
    st = stream if stream is not None else L.current_stream()
    L.check(L.lib.hs_l1gs_load(os.fsencode(path), st, C.byref(h)))
    info = np.zeros(8, np.int64)
    L.check(L.lib.hs_l1gs_info(h, info.ctypes.data))
    with open(path, "rb") as f:  # header: magic, 6 x int64, then sweep, cache, precision
        fields = struct.unpack("<8s6q3i", f.read(8 + 48 + 12))
    precision = {v: k for k, v in _PREC.items()}[fields[9]]
    return L1GSPreconditioner(h, backend, _SWEEPS[int(info[6])](), _CACHES[int(info[7])](),
                              precision)


def ldiv_(y, P: L1GSPreconditioner, x, stream=None, sync=True):
    """LinearSolve.ldiv!(y, P, x): y = M^-1 x, x unchanged, returns None (:367-384)."""
    py, ly = L.vec_ptr(y, P.n)
    px, lx = L.vec_ptr(x, P.n)
    if ly != lx:
        raise L.HSError(L.HS_EINVAL, "x and y must both be host or both be device vectors")
    st = stream if stream is not None else (L.current_stream() if ly == L.LOC_DEVICE else None)
    L.check(L.lib.hs_l1gs_apply(P.handle, py, px, ly, st, 0 if sync else L.FLAG_ASYNC))
    return None


def backslash(P: L1GSPreconditioner, x):
    """P \\ x (:398-407): allocates y like x."""
    try:
        import torch
        if isinstance(x, torch.Tensor):
            y = torch.empty_like(x)
            ldiv_(y, P, x)
            return y
    except ImportError:  # pragma: no cover
        pass
    y = np.empty_like(np.asarray(x, dtype=np.float64))
    ldiv_(y, P, np.ascontiguousarray(x, dtype=np.float64))
    return y
