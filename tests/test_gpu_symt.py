"""GPU parity of the one-triangle packed cache (PackedSymmetricCache: the lower
triangle alone, U = L^T read back from an LDS copy) through the C ABI.

For a symmetric matrix the sweep performs the same loads, products and sums
as the two-triangle PackedBufferCache, so results are compared BIT FOR BIT
with it in every precision, and the FP64 ones with the oracle within 1e-12.
The packed getters rebuild the reference layout (SU from L) and equal the
oracle's buffers exactly.
"""
import numpy as np
import pytest

from conftest import relerr
from test_gpu_l1gs import _random_sym

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("partsize", [1, 2, 3, 5, 8, 13, 16, 17, 31, 32, 33, 63, 64])
def test_one_triangle_equals_two(gpu, partsize, precision):
    import oracle as O
    hs = gpu
    n = 64 * 37 + 19 if partsize == 64 else 517  # ragged last partition
    A = _random_sym(n, 6.0 / n, 100 + partsize)
    x = np.random.default_rng(partsize).uniform(-1, 1, n)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=7)
    P2 = b(A, partsize, isSymA=True, precision=precision)
    P1 = b(A, partsize, isSymA=True, precision=precision,
           cache_strategy=hs.PackedSymmetricCache())
    assert P1.packed_size == P2.packed_size
    y1 = hs.backslash(P1, x)
    y2 = hs.backslash(P2, x)
    np.testing.assert_array_equal(y1, y2)
    # in place (x and y alias) gives the same bits
    z = x.copy()
    hs.ldiv_(z, P1, z)
    np.testing.assert_array_equal(z, y1)
    if precision == "fp64":
        ref = O.l1gs_build(A, partsize, is_sym=True, sweep="symmetric", gpu=(64, 7))
        assert relerr(y1, O.l1gs_ldiv(ref, x)) < TOL
        np.testing.assert_array_equal(P1.sweep.lop.D_DL1, ref.D)
        np.testing.assert_array_equal(P1.sweep.lop.L.SLbuffer, ref.SL)
        np.testing.assert_array_equal(P1.sweep.uop.U.SUbuffer, ref.SU)


def test_one_triangle_poisson128(gpu):
    """128^3 Poisson, partsize 64: the one-triangle sweep equals the oracle's
    symmetric sweep exactly (one in-partition neighbour per side)."""
    import torch
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(128)
    n = A.shape[0]
    x = pb.uniform_rhs(n, 0)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 1024))
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)(
        A, 64, isSymA=True, cache_strategy=hs.PackedSymmetricCache())
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    hs.ldiv_(yd, P, xd)
    np.testing.assert_array_equal(yd.cpu().numpy(), O.l1gs_ldiv(ref, x))


@pytest.mark.parametrize("partsize,precision", [(64, "fp64"), (64, "fp32"), (64, "bf16"),
                                                (32, "fp64"), (16, "fp32")])
def test_one_triangle_pcg(gpu, partsize, precision):
    """PCG with the one-triangle cache: the same iteration count as with both
    triangles.  At partsize 64 both forms reduce r'z per partition in the same
    order, so the solutions are bit-identical; for narrower partitions the
    workgroup partials group partitions differently (last-bits differences)."""
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal3d(40, seed=3)
    n = A.shape[0]
    rhs = pb.uniform_rhs(n, 9)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)
    out = []
    for cache in (hs.PackedBufferCache(), hs.PackedSymmetricCache()):
        P = b(A, partsize, isSymA=True, precision=precision, cache_strategy=cache)
        sol = hs.solve(hs.LinearProblem(A, rhs), hs.KrylovJL_CG(), Pl=P, reltol=1e-8)
        out.append(sol)
    s2, s1 = out
    assert s1.iters == s2.iters and s1.status == s2.status == "solved"
    if partsize == 64:
        np.testing.assert_array_equal(s1.u, s2.u)
    else:
        assert relerr(s1.u, s2.u) < 1e-9
    r = rhs - A @ s1.u
    assert np.linalg.norm(r) / np.linalg.norm(rhs) < 1.5e-8


def test_one_triangle_save_load(gpu, tmp_path):
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal2d(50, seed=6)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)(
        A, 64, isSymA=True, precision="fp32", cache_strategy=hs.PackedSymmetricCache())
    path = tmp_path / "sym1.hsl1gs"
    P.save(path)
    Q = hs.load_preconditioner(path)
    assert isinstance(Q.cache_strategy, hs.PackedSymmetricCache)
    x = pb.uniform_rhs(A.shape[0], 2)
    np.testing.assert_array_equal(hs.backslash(Q, x), hs.backslash(P, x))
    np.testing.assert_array_equal(Q._get(2), P._get(2))


def test_one_triangle_amg_smoother(gpu):
    """x -= M^-1 (A x - b) with the one-triangle cache equals the two-triangle
    smoother bit for bit."""
    import torch
    from hybridsmoothers import amg, problems as pb
    hs = gpu
    A = pb.lognormal3d(20, seed=8)
    n = A.shape[0]
    b = torch.tensor(pb.uniform_rhs(n, 4), device="cuda")
    xs = []
    for cache in (hs.PackedBufferCache(), hs.PackedSymmetricCache()):
        P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(
            A, 32, isSymA=True, cache_strategy=cache)
        s = amg.HybridSmoother(hs.DeviceMatrix(A), P, iters=3)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        amg.smooth_(x, s, b)
        xs.append(x.cpu().numpy())
    np.testing.assert_array_equal(xs[0], xs[1])
