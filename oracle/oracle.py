"""ctypes front-end of the CPU oracle (oracle/hs_oracle.c).

TEST INFRASTRUCTURE ONLY: the test-suite, the smoke check and the CPU-baseline
leg of the benchmark use this module as the checker.  The product library
(hybridsmoothers) never imports it.

Matrices are scipy.sparse CSC/CSR matrices; they are converted here to the
reference's 1-based Int64 arrays before the call, exactly what Julia's
SparseMatrixCSC{Float64,Int64} holds.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libhs_oracle.so")
_lib = None

I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i64 = C.c_int64


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.hs_or_partition_cpu.argtypes = [i64, i64, i64, I64P]
        L.hs_or_partition_gpu.argtypes = [i64, i64, i64, i64, I64P]
        L.hs_or_diag_indices.argtypes = [i64, I64P, I64P, F64P, I64P]
        L.hs_or_diag_indices.restype = i64
        L.hs_or_d_dl1.argtypes = [i64, I64P, I64P, F64P, C.c_int, C.c_int, C.c_void_p,
                                  i64, i64, i64, i64, C.c_double, F64P, C.c_int]
        L.hs_or_d_dl1_nonsym_fast.argtypes = [i64, I64P, I64P, F64P, I64P, i64, i64, i64, i64,
                                              C.c_double, F64P, C.c_int]
        L.hs_or_csc_to_csr.argtypes = [i64, i64, I64P, I64P, F64P, I64P, I64P, F64P]
        L.hs_or_packed_size.argtypes = [i64, i64, i64]
        L.hs_or_packed_size.restype = i64
        L.hs_or_pack.argtypes = [i64, I64P, I64P, F64P, C.c_int, C.c_int, i64, i64, i64, i64,
                                 C.c_int, F64P, C.c_int]
        L.hs_or_sweep_packed.argtypes = [i64, F64P, F64P, C.c_int, F64P, i64, i64, i64, i64, C.c_int]
        L.hs_or_sweep_view.argtypes = [i64, F64P, I64P, I64P, F64P, C.c_int, C.c_int, C.c_void_p,
                                       C.c_int, F64P, i64, i64, i64, i64, C.c_int]
        L.hs_or_scale.argtypes = [i64, F64P, F64P, C.c_int]
        L.hs_or_l1_rowsum.argtypes = [i64, i64, I64P, I64P, F64P, F64P]
        L.hs_or_spmv.argtypes = [i64, I64P, I64P, F64P, F64P, F64P, C.c_int]
        L.hs_or_spmv_sub.argtypes = [i64, I64P, I64P, F64P, F64P, F64P, C.c_int]
        L.hs_or_spectral_radius.argtypes = [i64, I64P, I64P, F64P, F64P, C.c_int, C.c_int]
        L.hs_or_spectral_radius.restype = C.c_double
        L.hs_or_cheb1_astar.argtypes = [C.c_int]
        L.hs_or_cheb1_astar.restype = C.c_double
        L.hs_or_cheb4_smooth.argtypes = [i64, I64P, I64P, F64P, F64P, C.c_double, C.c_int,
                                         F64P, F64P, F64P, F64P, C.c_int]
        L.hs_or_cheb1_smooth.argtypes = [i64, I64P, I64P, F64P, F64P, C.c_double, C.c_double,
                                         C.c_int, F64P, F64P, F64P, F64P, F64P, C.c_int]
        L.hs_or_l1gs_ldiv.argtypes = [i64, F64P, F64P, F64P, F64P, F64P, C.c_int, i64, i64, i64,
                                      i64, C.c_int]
        L.hs_or_pcg_l1gs.argtypes = [i64, I64P, I64P, F64P, F64P, F64P, C.c_double, C.c_double,
                                     i64, F64P, F64P, F64P, C.c_int, i64, i64, i64, i64, C.c_int,
                                     F64P]
        L.hs_or_pcg_cheb.argtypes = [i64, I64P, I64P, F64P, F64P, F64P, C.c_double, C.c_double,
                                     i64, F64P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                     F64P]
        L.hs_or_pcg_none.argtypes = [i64, I64P, I64P, F64P, F64P, F64P, C.c_double, C.c_double,
                                     i64, C.c_int, F64P]
        L.hs_or_max_threads.restype = C.c_int
        _lib = L
    return _lib


def nthreads_default() -> int:
    return int(lib().hs_or_max_threads())


# ---------------------------------------------------------------------------
# Sparse matrices in the reference's 1-based Int64 form
# ---------------------------------------------------------------------------
@dataclass
class Sparse1:
    n: int
    ptr: np.ndarray  # int64, 1-based
    ind: np.ndarray  # int64, 1-based
    val: np.ndarray  # float64
    fmt: int  # 0 CSC, 1 CSR


def to1(A) -> Sparse1:
    """Convert a scipy CSC or CSR matrix (sorted indices, no duplicates) to 1-based Int64."""
    if sp.isspmatrix_csr(A) or isinstance(A, sp.csr_array):
        fmt = 1
    else:
        A = A.tocsc() if not (sp.isspmatrix_csc(A) or isinstance(A, sp.csc_array)) else A
        fmt = 0
    A = A.copy()
    A.sum_duplicates()
    A.sort_indices()
    return Sparse1(A.shape[0], A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1,
                   np.ascontiguousarray(A.data, dtype=np.float64), fmt)


def partition_cpu(n, partsize, chunks):
    out = np.zeros(4, np.int64)
    lib().hs_or_partition_cpu(n, partsize, chunks, out)
    return tuple(int(v) for v in out)


def partition_gpu(n, partsize, threads, blocks):
    out = np.zeros(4, np.int64)
    lib().hs_or_partition_gpu(n, partsize, threads, blocks, out)
    return tuple(int(v) for v in out)


class SingularException(Exception):
    def __init__(self, col):
        super().__init__(f"SingularException({col})")
        self.col = col


def diag_indices(S: Sparse1) -> np.ndarray:
    d = np.zeros(S.n, np.int64)
    col = lib().hs_or_diag_indices(S.n, S.ptr, S.ind, S.val, d)
    if col:
        raise SingularException(int(col))
    return d


def packed_size(n, partsize, nparts):
    return int(lib().hs_or_packed_size(n, partsize, nparts))


@dataclass
class L1GS:
    """Oracle L1GS plan, mirroring L1GSPreconditioner's fields."""
    n: int
    part: tuple  # (partsize, nparts, nchunks, chunksize)
    sweep: str  # "forward" | "backward" | "symmetric"
    cache: str  # "packed" | "view"
    D: np.ndarray
    SL: np.ndarray | None
    SU: np.ndarray | None
    S: Sparse1
    sym: bool
    diag: np.ndarray | None
    nthreads: int


def l1gs_build(A, partsize, *, is_sym=False, eta=1.5, sweep="symmetric", cache="packed",
               chunks=None, gpu=None, fast_nonsym=None, nthreads=None) -> L1GS:
    """build_l1prec (l1_gauss_seidel.jl:317-338) + _make_sweep_plan (:626-843).

    chunks: KA.CPU chunk count; gpu=(threads, blocks) selects the GPU partitioning formula.
    fast_nonsym: use the O(nnz) D_Dl1 for non-symmetric CSC (default: when n > 4096).
    """
    if partsize == 0:
        raise ValueError("partsize must be greater than 0")
    S = to1(A)
    n = S.n
    nthreads = nthreads or nthreads_default()
    if gpu is not None:
        part = partition_gpu(n, partsize, gpu[0], gpu[1])
    else:
        part = partition_cpu(n, partsize, chunks or nthreads)
    P, nparts, nchunks, chunksize = part
    sym = bool(is_sym)
    diag = None
    if S.fmt == 0 and not sym:
        diag = diag_indices(S)
    D = np.zeros(n)
    if diag is not None and (fast_nonsym if fast_nonsym is not None else n > 4096):
        lib().hs_or_d_dl1_nonsym_fast(n, S.ptr, S.ind, S.val, diag, P, nparts, nchunks, chunksize,
                                      float(eta), D, nthreads)
    else:
        dptr = diag.ctypes.data_as(C.c_void_p) if diag is not None else None
        lib().hs_or_d_dl1(n, S.ptr, S.ind, S.val, S.fmt, int(sym), dptr, P, nparts, nchunks,
                          chunksize, float(eta), D, nthreads)
    SL = SU = None
    if cache == "packed":
        size = packed_size(n, P, nparts)
        if sweep in ("forward", "symmetric"):
            SL = np.zeros(size)
            lib().hs_or_pack(n, S.ptr, S.ind, S.val, S.fmt, int(sym), P, nparts, nchunks, chunksize,
                             0, SL, nthreads)
        if sweep in ("backward", "symmetric"):
            SU = np.zeros(size)
            lib().hs_or_pack(n, S.ptr, S.ind, S.val, S.fmt, int(sym), P, nparts, nchunks, chunksize,
                             1, SU, nthreads)
    return L1GS(n, part, sweep, cache, D, SL, SU, S, sym, diag, nthreads)


def _sweep(p: L1GS, y, upper):
    P, nparts, nchunks, chunksize = p.part
    if p.cache == "packed":
        buf = p.SU if upper else p.SL
        lib().hs_or_sweep_packed(p.n, y, buf, int(upper), p.D, P, nparts, nchunks, chunksize,
                                 p.nthreads)
    else:
        S = p.S
        dptr = p.diag.ctypes.data_as(C.c_void_p) if p.diag is not None else None
        lib().hs_or_sweep_view(p.n, y, S.ptr, S.ind, S.val, S.fmt, int(p.sym), dptr, int(upper),
                               p.D, P, nparts, nchunks, chunksize, p.nthreads)


def l1gs_ldiv(p: L1GS, x) -> np.ndarray:
    """ldiv!(y, P, x) (l1_gauss_seidel.jl:386-396 -> :854-906)."""
    y = np.array(x, dtype=np.float64, copy=True)
    if p.sweep in ("forward", "symmetric"):
        _sweep(p, y, False)
    if p.sweep == "symmetric":
        lib().hs_or_scale(p.n, y, p.D, p.nthreads)
    if p.sweep in ("backward", "symmetric"):
        _sweep(p, y, True)
    return y


def l1gs_sweep_inplace(p: L1GS, y, upper: bool):
    """_apply_sweep! of a single pass on y (in place), as used by the GS sanity checks."""
    _sweep(p, y, upper)


# ---------------------------------------------------------------------------
# Chebyshev
# ---------------------------------------------------------------------------
def l1_rowsum(A) -> np.ndarray:
    S = to1(sp.csc_matrix(A))
    out = np.zeros(S.n)
    lib().hs_or_l1_rowsum(S.n, S.n, S.ptr, S.ind, S.val, out)
    return out


def csr1(A) -> Sparse1:
    """SparseMatrixCSR(A_csc) (chebyshev.jl:144): 1-based CSR with ascending columns."""
    S = to1(sp.csc_matrix(A))
    n = S.n
    nnz = int(S.ptr[-1] - 1)
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(max(nnz, 1), np.int64)
    cv = np.zeros(max(nnz, 1))
    lib().hs_or_csc_to_csr(n, n, S.ptr, S.ind, S.val, rp, ci, cv)
    return Sparse1(n, rp, ci[:nnz].copy(), cv[:nnz].copy(), 1)


def spmv(R: Sparse1, x, nthreads=None):
    y = np.zeros(R.n)
    lib().hs_or_spmv(R.n, R.ptr, R.ind, R.val, np.ascontiguousarray(x, np.float64), y,
                     nthreads or nthreads_default())
    return y


def spectral_radius(R: Sparse1, l1inv, power_iters=50, nthreads=None):
    return float(lib().hs_or_spectral_radius(R.n, R.ptr, R.ind, R.val,
                                              np.ascontiguousarray(l1inv), power_iters,
                                              nthreads or nthreads_default()))


def cheb1_astar(k):
    return float(lib().hs_or_cheb1_astar(int(k)))


@dataclass
class Cheb:
    kind: int  # 4 or 1
    degree: int
    R: Sparse1
    l1inv: np.ndarray
    rho: float
    a: float
    nthreads: int


def cheb_build(A, kind, degree, power_iters=50, nthreads=None) -> Cheb:
    """_build_smoother (chebyshev.jl:141-156, :258-276)."""
    nthreads = nthreads or nthreads_default()
    l1inv = 1.0 / l1_rowsum(A)
    R = csr1(A)
    rho = spectral_radius(R, l1inv, power_iters, nthreads)
    a = cheb1_astar(degree) if kind == 1 else 0.0
    return Cheb(kind, degree, R, l1inv, rho, a, nthreads)


def cheb_smooth(s: Cheb, x, b):
    """smooth!(x, s, b) in place on a copy; returns (x, r, z[, tmp])."""
    n = s.R.n
    x = np.array(x, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    r = np.zeros(n)
    z = np.zeros(n)
    if s.kind == 4:
        lib().hs_or_cheb4_smooth(n, s.R.ptr, s.R.ind, s.R.val, s.l1inv, s.rho, s.degree, x, b, r, z,
                                 s.nthreads)
        return x, r, z
    tmp = np.zeros(n)
    lib().hs_or_cheb1_smooth(n, s.R.ptr, s.R.ind, s.R.val, s.l1inv, s.rho, s.a, s.degree, x, b, r,
                             z, tmp, s.nthreads)
    return x, r, z


# ---------------------------------------------------------------------------
# PCG (Krylov.jl-style preconditioned-norm stopping rule)
# ---------------------------------------------------------------------------
@dataclass
class PCGStats:
    iters: int
    rnorm: float
    rnorm0: float
    status: int


def pcg(A, b, prec=None, rtol=1e-8, atol=0.0, itmax=None, nthreads=None):
    """CG with x0 = 0 and left preconditioner `prec` (an L1GS or Cheb oracle plan, or None)."""
    R = prec.R if isinstance(prec, Cheb) else csr1(A)
    n = R.n
    nthreads = nthreads or nthreads_default()
    itmax = itmax if itmax is not None else 2 * n
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    st = np.zeros(4)
    if prec is None:
        lib().hs_or_pcg_none(n, R.ptr, R.ind, R.val, b, x, rtol, atol, itmax, nthreads, st)
    elif isinstance(prec, L1GS):
        assert prec.cache == "packed"
        sw = {"forward": 0, "backward": 1, "symmetric": 2}[prec.sweep]
        SL = prec.SL if prec.SL is not None else np.zeros(1)
        SU = prec.SU if prec.SU is not None else np.zeros(1)
        P, nparts, nchunks, chunksize = prec.part
        lib().hs_or_pcg_l1gs(n, R.ptr, R.ind, R.val, b, x, rtol, atol, itmax, prec.D, SL, SU, sw,
                             P, nparts, nchunks, chunksize, nthreads, st)
    else:
        lib().hs_or_pcg_cheb(n, R.ptr, R.ind, R.val, b, x, rtol, atol, itmax, prec.l1inv, prec.rho,
                             prec.a, prec.kind, prec.degree, nthreads, st)
    return x, PCGStats(int(st[0]), float(st[1]), float(st[2]), int(st[3]))
