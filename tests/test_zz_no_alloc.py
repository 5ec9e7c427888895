"""hs_l1gs_smooth_noalloc: the AMG smoother round on device vectors with its
residual and correction in the library's fixed device array (no allocation).

Called through the C ABI and compared with hs_l1gs_smooth (the same round with
the matrix-owned workspace) bit for bit, on two calls in a row, the second
with a smaller system, for the fused one-launch round (packed triangles,
partsize 64) and the composed round (MatrixViewCache: residual, sweep and
update as three launches).  The library's device-memory counter does not move
during the allocation-free calls; a system past the array's 2,097,152 rows is
refused before any device work.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call_noalloc(L, P, M, x, b, iters):
    import torch
    st = L.lib.hs_l1gs_smooth_noalloc(P.handle, M.handle, C.c_void_p(x.data_ptr()),
                                      C.c_void_p(b.data_ptr()), iters,
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream), 0)
    return st


def _call_alloc(L, P, M, x, b, iters):
    import torch
    return L.lib.hs_l1gs_smooth(P.handle, M.handle, C.c_void_p(x.data_ptr()),
                                C.c_void_p(b.data_ptr()), iters, L.LOC_DEVICE,
                                C.c_void_p(torch.cuda.current_stream().cuda_stream), 0)


@pytest.mark.parametrize("cache", ["packed", "view"])
def test_noalloc_smoother_equals_smoother(gpu, cache):
    import torch
    import hybridsmoothers._lib as L
    from hybridsmoothers import problems as pb
    hs = gpu
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)
    cs = hs.PackedBufferCache() if cache == "packed" else hs.MatrixViewCache()
    # two calls in a row, the second on a smaller system (the array keeps the
    # first call's contents; nothing may depend on them)
    for grid, iters in ((48, 3), (20, 2)):
        A = pb.lognormal3d(grid, seed=grid)
        n = A.shape[0]
        print(f"[noalloc] cache={cache} n={n} iters={iters}", flush=True)
        P = builder(A, 64, isSymA=False, sweep=hs.SymmetricSweep(), cache_strategy=cs)
        M = hs.DeviceMatrix(A)
        b = torch.tensor(pb.uniform_rhs(n, 3), device="cuda")
        x0 = torch.tensor(pb.uniform_rhs(n, 4), device="cuda")
        xa, xn = x0.clone(), x0.clone()
        torch.cuda.synchronize()
        before = L.lib.hs_device_bytes()
        assert _call_noalloc(L, P, M, xn, b, iters) == L.HS_OK, L.lib.hs_last_error()
        torch.cuda.synchronize()
        assert L.lib.hs_device_bytes() == before  # nothing allocated
        assert _call_alloc(L, P, M, xa, b, iters) == L.HS_OK, L.lib.hs_last_error()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(xn.cpu().numpy(), xa.cpu().numpy())
        assert not np.array_equal(xn.cpu().numpy(), x0.cpu().numpy())
        # zero rounds: x unchanged
        xz = x0.clone()
        assert _call_noalloc(L, P, M, xz, b, 0) == L.HS_OK
        np.testing.assert_array_equal(xz.cpu().numpy(), x0.cpu().numpy())


def test_noalloc_smoother_rejects_large_systems(gpu):
    import torch
    import hybridsmoothers._lib as L
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.laplace1d((1 << 21) + 1)
    n = A.shape[0]
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(
        A, 64, isSymA=True, sweep=hs.SymmetricSweep())
    M = hs.DeviceMatrix(A)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    assert _call_noalloc(L, P, M, x, b, 2) == L.HS_EINVAL
    assert b"2,097,152 rows" in L.lib.hs_last_error()
    assert float(x.abs().max()) == 0.0
