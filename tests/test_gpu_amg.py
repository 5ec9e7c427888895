"""AMG smoother glue (ext/HybridSmoothersAMGExt.jl) through hs_l1gs_smooth.

Parity bar: bit-identical to the same steps composed from the C ABI calls
(SpMV, numpy subtraction, L1GS apply, subtraction -- each rounded as the
reference's mul!, .-=, ldiv!, .-= round); within 1e-12 of the oracle's
composition (symmetric sweeps agree to that tolerance); and it smooths: the
error of a Poisson problem shrinks, high frequencies fastest.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _compose(M, P, x, b, hs, iters=2):
    x = x.copy()
    r = np.empty_like(x)
    for _ in range(iters):
        M.mul(r, x)
        r = r - b
        x = x - hs.backslash(P, r)
    return x


@pytest.mark.parametrize("case", ["poisson3d_P16", "lognormal2d_P32", "convdiff_P8_nonsym"])
def test_l1gs_amg_smoother(gpu, case):
    import oracle as O
    from hybridsmoothers import amg, problems as pb
    hs = gpu
    if case == "poisson3d_P16":
        A, P = pb.poisson3d(14), 16
    elif case == "lognormal2d_P32":
        A, P = pb.lognormal2d(40, seed=6), 32
    else:
        A, P = pb.convdiff2d(20), 8
    n = A.shape[0]
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=P)
    s = amg.setup_smoother(builder, A, None)
    assert s.s.partitioning.partsize == P and s.iters == 2
    b = pb.uniform_rhs(n, 9)
    x0 = pb.uniform_rhs(n, 10)
    x = x0.copy()
    amg.smooth_(x, s, b)
    np.testing.assert_array_equal(x, _compose(s.A, s.s, x0, b, hs))
    # the oracle's composition (isSymA = false, symmetric sweep)
    ref = O.l1gs_build(A, P, is_sym=False, sweep="symmetric", gpu=(256, P))
    R = O.csr1(A)
    xo = x0.copy()
    for _ in range(2):
        xo = xo - O.l1gs_ldiv(ref, O.spmv(R, xo) - b)
    assert relerr(x, xo) < 1e-12
    # device vectors, in place, same result
    import torch
    xd = torch.tensor(x0, device="cuda")
    amg.smooth_(xd, s, torch.tensor(b, device="cuda"))
    np.testing.assert_array_equal(xd.cpu().numpy(), x)


def test_smoother_reduces_error(gpu):
    from hybridsmoothers import amg, problems as pb
    hs = gpu
    A = pb.poisson3d(16)
    n = A.shape[0]
    xs = pb.uniform_rhs(n, 1)
    b = A @ xs
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=32)
    s = amg.setup_smoother(builder, A, None)
    x = np.zeros(n)
    errs = [np.linalg.norm(x - xs)]
    for _ in range(5):
        amg.smooth_(x, s, b)
        errs.append(np.linalg.norm(x - xs))
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:]))
    # a checkerboard (highest frequency) error is damped by orders of magnitude
    g = np.indices((16, 16, 16)).sum(axis=0).ravel() % 2 * 2.0 - 1.0
    x = xs + g
    amg.smooth_(x, s, b)
    assert np.linalg.norm(x - xs) < 0.1 * np.linalg.norm(g)


def test_chebyshev_amg_smoother(gpu):
    from hybridsmoothers import amg, problems as pb
    hs = gpu
    A = pb.poisson3d(10)
    s = amg.setup_smoother(hs.ChebyshevFourth(degree=3, backend=hs.ROCBackend()), A, None)
    assert isinstance(s, hs.ChebyshevSmoother)
    b = pb.uniform_rhs(A.shape[0], 2)
    x = np.zeros_like(b)
    y = np.zeros_like(b)
    amg.smooth_(x, s, b)
    hs.smooth_(y, s, b)
    np.testing.assert_array_equal(x, y)


_ROUNDS = [(sw, p, c, it) for sw in ("forward", "backward", "symmetric")
           for p, c, it in [(3, "packed", 1), (32, "packed", 2), (64, "packed", 3),
                            (64, "view", 2), (100, "packed", 2), (128, "view", 3)]]
_ROUNDS.append(("symmetric", 16, "sym1", 2))  # the one-triangle layout is symmetric-only


@pytest.mark.parametrize("sweep,partsize,cache,iters", _ROUNDS)
def test_fused_smoother_rounds(gpu, sweep, partsize, cache, iters):
    """Each round of hs_l1gs_smooth is one fused launch for packed layouts with
    partsize <= 64 (residual in the sweep prologue, x - z stored to the other
    of two buffers, a final copy after an odd count); the view cache and larger
    partsizes take the three-launch path.  Every case equals the composition of
    separate calls bit for bit."""
    import torch
    from hybridsmoothers import amg, problems as pb
    hs = gpu
    A = pb.lognormal3d(14, seed=12)
    n = A.shape[0]
    sw = {"forward": hs.ForwardSweep(), "backward": hs.BackwardSweep(),
          "symmetric": hs.SymmetricSweep()}[sweep]
    strategy = {"packed": hs.PackedBufferCache(), "view": hs.MatrixViewCache(),
                "sym1": hs.PackedSymmetricCache()}[cache]
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=16)(
        A, partsize, isSymA=True, sweep=sw, cache_strategy=strategy)
    s = amg.HybridSmoother(hs.DeviceMatrix(A), P, iters=iters)
    b = pb.uniform_rhs(n, 21)
    x0 = pb.uniform_rhs(n, 22)
    x = x0.copy()
    amg.smooth_(x, s, b)
    np.testing.assert_array_equal(x, _compose(s.A, P, x0, b, hs, iters=iters))
    xd = torch.tensor(x0, device="cuda")
    amg.smooth_(xd, s, torch.tensor(b, device="cuda"), sync=False)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(xd.cpu().numpy(), x)
