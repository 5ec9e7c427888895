"""hybridsmoothers -- MI355X-native smoother hot path of HybridSmoothers.jl.

Host-side mirror of the Julia package's API over the C ABI of libhsmooth.so
(include/hsmooth.h).  Exports follow src/HybridSmoothers.jl:41-44.
"""
from ._lib import HSError, SingularException, init  # noqa: F401
from .l1gs import (  # noqa: F401
    CPU, GPU, BackwardSweep, BlockPartitioning, CPUConfig, ForwardSweep, GPUConfig,
    L1GSPrecBuilder, L1GSPreconditioner, MatrixViewCache, PackedBufferCache, ROCBackend,
    Symmetric, SymmetricSweep, backslash, functional, ldiv_,
)

__all__ = [
    "L1GSPrecBuilder", "ForwardSweep", "BackwardSweep", "SymmetricSweep",
    "PackedBufferCache", "MatrixViewCache", "ChebyshevFourth", "ChebyshevFirst",
]

try:  # components whose native sources are built alongside
    from .chebyshev import (  # noqa: F401
        ChebyshevFirst, ChebyshevFourth, build_smoother, smooth_, _build_smoother,
    )
    from .pcg import pcg_solve  # noqa: F401
except ImportError:  # pragma: no cover - only while a component is being added
    pass
