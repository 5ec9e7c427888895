"""Preconditioned CG on the device, mirroring how the reference is used:

    sol = solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P)      # README.md:27-29

Here: ``sol = solve(LinearProblem(A, b), KrylovJL_CG(), Pl=P, reltol=1e-8)``.
`P` is an L1GSPreconditioner or a Chebyshev smoother (used as z = 0;
smooth!(z, S, r)).  The stopping rule is Krylov.jl `cg`'s:
sqrt(r'z) <= abstol + reltol*sqrt(r0'z0).  Every solve also reports the FP64
true relative residual ||b - A x|| / ||b|| computed on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .chebyshev import ChebyshevSmoother
from .l1gs import L1GSPreconditioner, sparse_arrays


class DeviceMatrix:
    """A device CSR copy of A (the operator CG multiplies by)."""

    def __init__(self, A, stream=None):
        n, ptr, ind, val, fmt = sparse_arrays(A)
        self.n = n
        self.nnz = int(A.nnz)
        h = C.c_void_p()
        L.check(L.lib.hs_csr_create(n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data, fmt,
                                    stream, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def column_bytes(self) -> int:
        """Bytes the products read per stored column: 2 (16-bit offsets from the
        row, every column within 32,767 of it) or 4."""
        out = C.c_int64(0)
        L.check(L.lib.hs_csr_column_bytes(self._h, C.byref(out)))
        return int(out.value)

    def mul(self, y, x, stream=None, sync=True):
        py, ly = L.vec_ptr(y, self.n)
        px, lx = L.vec_ptr(x, self.n)
        if ly != lx:
            raise L.HSError(L.HS_EINVAL, "x and y must both be host or both be device vectors")
        if stream is None and ly == L.LOC_DEVICE:
            stream = L.current_stream()
        L.check(L.lib.hs_csr_spmv(self._h, py, px, ly, stream, 0 if sync else L.FLAG_ASYNC))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.hs_csr_destroy(h)
            self._h = None


@dataclass
class LinearProblem:
    A: object
    b: object


class KrylovJL_CG:  # noqa: N801 - LinearSolve's name
    """CG from Krylov.jl as LinearSolve names it."""


@dataclass
class Solution:
    u: object
    iters: int
    status: str
    resid: float          # final sqrt(r'z)
    resid0: float         # sqrt(r0'z0)
    true_relres: float    # ||b - A x|| / ||b||
    seconds: float


_STATUS = {0: "solved", 1: "maximum number of iterations exceeded", 2: "breakdown"}


def pcg_solve(A, b, Pl=None, reltol=1e-8, abstol=0.0, maxiters=None, stream=None,
              matrix: DeviceMatrix | None = None, true_reltol: float | None = None) -> Solution:
    """FP64 PCG with x0 = 0.  `b` is a numpy (host) or torch CUDA (device) vector.

    true_reltol: also meet ||b - A x|| / ||b|| <= true_reltol -- at each stop of
    Krylov.jl's rule the true residual is computed on the device and the same CG
    recurrence continues, with a proportionally tighter threshold, while it is
    above (hs_pcg_solve_true).  None: Krylov.jl's rule alone."""
    M = matrix if matrix is not None else (A if isinstance(A, DeviceMatrix) else DeviceMatrix(A))
    n = M.n
    if isinstance(Pl, L1GSPreconditioner):
        kind, ph = L.PCG_L1GS, Pl.handle
    elif isinstance(Pl, ChebyshevSmoother):
        kind, ph = L.PCG_CHEB, Pl.handle
    elif Pl is None:
        kind, ph = L.PCG_NONE, None
    else:
        raise TypeError(f"unsupported preconditioner {type(Pl)!r}")
    try:
        import torch
        is_dev = isinstance(b, torch.Tensor) and b.is_cuda
    except ImportError:  # pragma: no cover
        is_dev = False
    if is_dev:
        import torch
        x = torch.empty_like(b)
    else:
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n)
    pb, lb = L.vec_ptr(b, n)
    px, _ = L.vec_ptr(x, n)
    st = stream if stream is not None else (L.current_stream() if lb == L.LOC_DEVICE else None)
    stats = L.PCGStats()
    itmax = -1 if maxiters is None else int(maxiters)
    if true_reltol is None:
        L.check(L.lib.hs_pcg_solve(M.handle, kind, ph, pb, px, float(reltol), float(abstol),
                                   itmax, lb, st, C.byref(stats)))
    else:
        L.check(L.lib.hs_pcg_solve_true(M.handle, kind, ph, pb, px, float(reltol), float(abstol),
                                        itmax, float(true_reltol), lb, st, C.byref(stats)))
    return Solution(x, int(stats.iters), _STATUS.get(int(stats.status), "unknown"),
                    float(stats.rnorm), float(stats.rnorm0), float(stats.true_relres),
                    float(stats.seconds))


def solve(prob: LinearProblem, alg=None, Pl=None, reltol=1e-8, abstol=0.0, maxiters=None,
          **kw) -> Solution:
    """LinearSolve.solve with KrylovJL_CG (the only algorithm of this path)."""
    if alg is not None and not isinstance(alg, KrylovJL_CG):
        raise TypeError("only KrylovJL_CG is provided")
    return pcg_solve(prob.A, prob.b, Pl=Pl, reltol=reltol, abstol=abstol, maxiters=maxiters, **kw)
