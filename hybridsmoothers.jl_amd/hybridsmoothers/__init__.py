"""hybridsmoothers -- MI355X-native smoother hot path of HybridSmoothers.jl.

Host-side mirror of the Julia package's API over the C ABI of libhsmooth.so
(include/hsmooth.h).  Exports follow src/HybridSmoothers.jl:41-44, plus the
LinearSolve-style solve used by the reference's README.
"""
from ._lib import HSError, SingularException, init  # noqa: F401
from .chebyshev import (  # noqa: F401
    ChebyshevFirst, ChebyshevFourth, ChebyshevSmoother, _build_smoother, build_smoother, smooth_,
)
from .chebyshev import ldiv_ as cheb_ldiv_  # noqa: F401
from .l1gs import (  # noqa: F401
    CPU, GPU, BackwardSweep, BlockPartitioning, CPUConfig, ForwardSweep, GPUConfig,
    L1GSPrecBuilder, L1GSPreconditioner, MatrixViewCache, PackedBufferCache,
    PackedSymmetricCache, ROCBackend,
    Symmetric, SymmetricSweep, backslash, functional,
)
from .l1gs import ldiv_ as _l1gs_ldiv_
from .l1gs import load as load_preconditioner  # noqa: F401
from .pcg import DeviceMatrix, KrylovJL_CG, LinearProblem, Solution, pcg_solve, solve  # noqa: F401
from .batch import BatchSolution, BatchSolver  # noqa: F401

__all__ = [
    "L1GSPrecBuilder", "ForwardSweep", "BackwardSweep", "SymmetricSweep",
    "PackedBufferCache", "MatrixViewCache", "ChebyshevFourth", "ChebyshevFirst",
]


def ldiv_(y, P, x, **kw):
    """LinearSolve.ldiv!(y, P, x) for an L1GS preconditioner or a Chebyshev smoother."""
    if isinstance(P, ChebyshevSmoother):
        return cheb_ldiv_(y, P, x, **kw)
    return _l1gs_ldiv_(y, P, x, **kw)
