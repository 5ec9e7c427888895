"""AlgebraicMultigrid smoother glue on the device.

Mirrors ext/HybridSmoothersAMGExt.jl (HybridSmoothers.jl v0.1.4):

    s = setup_smoother(config, A, symmetry)   # AMG.setup_smoother   (:15-27, :37-43)
    smooth_(x, s, b)                          # AMG.smooth!          (:29-35, :45-51)

For an L1GS builder the smoother is two rounds of x -= P^-1 (A x - b) with a
symmetric sweep and isSymA = false (:23), the residual and correction kept as
device workspace with the matrix; a Chebyshev configuration gives the
Chebyshev smoother itself, whose smooth! is used unchanged.  The reference
only defines the L1GS form for its CPU backend, where partsize =
n / Threads.nthreads() (:22); here the builder's default partsize (its GPU
`blocks`, :265-266) is used unless one is given.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib as L
from .chebyshev import ChebyshevFirst, ChebyshevFourth, ChebyshevSmoother, _build_smoother
from .chebyshev import smooth_ as _cheb_smooth
from .l1gs import L1GSPrecBuilder, L1GSPreconditioner, SymmetricSweep
from .pcg import DeviceMatrix


@dataclass
class HybridSmoother:
    """HybridSmoother{AT, ST, VT} (:8-13): the matrix, the preconditioner and (on
    the device, owned by the matrix) the residual and correction vectors."""
    A: DeviceMatrix
    s: L1GSPreconditioner
    iters: int = 2


def setup_smoother(config, A, symmetry=None, partsize=None):
    if isinstance(config, (ChebyshevFourth, ChebyshevFirst)):
        return _build_smoother(config, A)
    if not isinstance(config, L1GSPrecBuilder):
        raise TypeError(f"unsupported smoother configuration {type(config)!r}")
    P = config(A, partsize if partsize is not None else config.default_partsize(),
               sweep=SymmetricSweep(), isSymA=False)
    return HybridSmoother(DeviceMatrix(A), P)


def smooth_(x, s, b, stream=None, sync=True):
    """AMG.smooth!(x, s, b): x updated in place (numpy = host, torch CUDA = device).
    With packed triangles and partsize <= 64 each round is one fused launch
    (residual, sweep and update); sync=False leaves device work queued."""
    if isinstance(s, ChebyshevSmoother):
        return _cheb_smooth(x, s, b)
    n = s.A.n
    px, lx = L.vec_ptr(x, n)
    pb, lb = L.vec_ptr(b, n)
    if lx != lb:
        raise L.HSError(L.HS_EINVAL, "x and b must both be host or both be device vectors")
    st = stream if stream is not None else (L.current_stream() if lx == L.LOC_DEVICE else None)
    L.check(L.lib.hs_l1gs_smooth(s.s.handle, s.A.handle, px, pb, s.iters, lx, st,
                                 0 if sync else L.FLAG_ASYNC))
