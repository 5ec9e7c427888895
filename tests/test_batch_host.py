"""Host-side parts of the batched / serving path (no GPU needed).

* the native workload generator against its numpy statement (same sparsity,
  values within libm-vs-numpy rounding), and the properties that make it a
  valid CG workload (symmetric, positive definite, row sums >= 0);
* the Poisson arrival process and the producer's ordering;
* argument checking of hs_batch_create (rejected before any device call).
"""
import ctypes as C

import numpy as np
import pytest


@pytest.mark.parametrize("dims", [(7, 5, 1), (6, 4, 3), (1, 9, 1), (5, 1, 1), (3, 3, 3),
                                  (40, 30, 1), (12, 11, 10), (2, 1, 1)])
def test_generator_matches_numpy_statement(dims):
    from hybridsmoothers import problems as pb
    A = pb.lognormal_grid_numpy(*dims, sigma=1.0, seed=11).tocsr()
    rp, ci, cv = pb.lognormal_grid(*dims, sigma=1.0, seed=11)
    np.testing.assert_array_equal(A.indptr + 1, rp)
    np.testing.assert_array_equal(A.indices + 1, ci)
    np.testing.assert_allclose(cv, A.data, rtol=2e-15, atol=0)


def test_generator_properties():
    from hybridsmoothers import problems as pb
    for dims, seed in (((9, 7, 1), 3), ((5, 4, 3), 4)):
        rp, ci, cv = pb.lognormal_grid(*dims, seed=seed)
        A = pb.csr_to_scipy(rp, ci, cv)
        n = A.shape[0]
        nx, ny, nz = dims
        assert A.nnz == n + 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1))
        assert abs(A - A.T).max() == 0.0
        assert np.all(np.linalg.eigvalsh(A.toarray()) > 0)
        assert np.all(A.sum(axis=1) >= -1e-12)       # Dirichlet faces: diagonally dominant
        # boundary rows are strictly dominant, interior rows sum to ~0
        assert np.all(np.asarray(A.diagonal()) > 0)
    # the size query reports nnz without writing arrays
    from hybridsmoothers import _lib as L
    nnz = C.c_int64(-1)
    L.check(L.lib.hs_problem_lognormal(10, 10, 1, 1.0, 0, None, None, None, C.byref(nnz)))
    assert nnz.value == 5 * 100 - 2 * 20
    with pytest.raises(L.HSError):
        L.check(L.lib.hs_problem_lognormal(0, 10, 1, 1.0, 0, None, None, None, C.byref(nnz)))


def test_generator_lognormal_statistics():
    """ln k over many faces is N(0, sigma^2): mean ~ 0, std ~ sigma."""
    from hybridsmoothers import problems as pb
    k = pb.lognormal_faces(200_000, seed=7, sigma=0.8)
    z = np.log(k)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 0.8) < 0.01
    # pairs from one Box-Muller draw are uncorrelated
    assert abs(np.corrcoef(z[0::2], z[1::2])[0, 1]) < 0.01


def test_stream_specs_cover_the_range():
    from hybridsmoothers import problems as pb
    specs = pb.serving_stream(4096, seed=5)
    n = np.array([pb.stream_grid(s)[0] * pb.stream_grid(s)[1] * pb.stream_grid(s)[2]
                  for s in specs])
    assert n.min() >= 700 and n.max() <= 3.0e5
    # log-uniform in [1e3, 2.6e5]: mean (b - a) / ln(b / a) ~ 46.6e3
    assert 40e3 < n.mean() < 53e3
    kinds = {s[0] for s in specs}
    assert kinds == {"2d", "3d"}


def test_poisson_arrivals():
    from hybridsmoothers import serve
    t = serve.poisson_arrivals(20000, rate=250.0, seed=1)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps >= 0)
    assert abs(gaps.mean() - 1 / 250.0) < 0.05 / 250.0
    assert abs(gaps.std() - 1 / 250.0) < 0.05 / 250.0   # exponential: std = mean
    assert np.all(serve.poisson_arrivals(5, rate=float("inf")) == 0)


def test_producer_order_and_lookahead():
    from hybridsmoothers import problems as pb, serve
    specs = [("2d", 8 + k, k) for k in range(12)]
    prod = serve.Producer(specs, workers=3, ahead=4)
    try:
        for i in range(len(specs)):
            (rp, ci, cv), b = prod.get(i)
            g = 8 + i
            assert rp.size - 1 == g * g == b.size
            np.testing.assert_array_equal(b, pb.uniform_rhs(g * g, 1000 + i))
            with prod._cv:
                assert prod._next <= prod._taken + prod.ahead
    finally:
        prod.close()


def test_batch_create_argument_errors():
    from hybridsmoothers import _lib as L
    base = dict(arena_bytes=1 << 20, max_systems=4, staging_bytes=1 << 20, partsize=32,
                precision=L.PREC_FP32, passes=4, grid=0, eta=1.5, triangles=0)
    h = C.c_void_p()
    for change, code in ((dict(partsize=20), L.HS_EUNSUPPORTED), (dict(passes=0), L.HS_EINVAL),
                         (dict(arena_bytes=0), L.HS_EINVAL), (dict(max_systems=0), L.HS_EINVAL),
                         (dict(precision=7), L.HS_EINVAL), (dict(grid=-1), L.HS_EINVAL),
                         (dict(triangles=3), L.HS_EINVAL), (dict(smoother=3), L.HS_EINVAL),
                         (dict(smoother=L.PCG_CHEB, degree=1), L.HS_EUNSUPPORTED),
                         (dict(smoother=L.PCG_CHEB, degree=17), L.HS_EUNSUPPORTED),
                         (dict(smoother=L.PCG_CHEB, degree=2, precision=L.PREC_BF16),
                          L.HS_EUNSUPPORTED),
                         (dict(smoother=L.PCG_CHEB, degree=2, power_iters=-1), L.HS_EINVAL)):
        cfg = L.BatchConfig(**{**base, **change})
        assert L.lib.hs_batch_create(C.byref(cfg), C.byref(h)) == code, change
    assert L.lib.hs_batch_create(None, C.byref(h)) == L.HS_EINVAL
    assert L.lib.hs_batch_step(None) == L.HS_EINVAL
    cnt = C.c_int64(0)
    assert L.lib.hs_batch_collect(None, None, 0, C.byref(cnt)) == L.HS_EINVAL
    assert C.sizeof(L.BatchConfig) == 72 and C.sizeof(L.BatchResult) == 80


def test_batch_solver_argument_errors():
    """The Python engine rejects unknown names before touching the device."""
    from hybridsmoothers import _lib as L
    from hybridsmoothers.batch import BatchSolver, iteration_bytes
    with pytest.raises(L.HSError):
        BatchSolver(precision="fp16")
    with pytest.raises(L.HSError):
        BatchSolver(smoother="jacobi")
    # the bytes model: Chebyshev degree 2 in FP32 on a 5-point grid is the
    # SpMV's 8 + c B/entry plus the smoother's 4 + c B/entry (c the column
    # bytes: 2 on this grid), 16 B per slice, 92 B/row
    from hybridsmoothers import problems as pb
    from hybridsmoothers.batch import column_bytes
    rp, ci, _ = pb.lognormal_grid(64, 64, 1, seed=1)
    n = rp.size - 1
    lens = np.diff(rp)
    w = lens.reshape(-1, 64).max(axis=1)
    nent = int(64 * w.sum())
    c = column_bytes(rp, ci)
    assert c == 2
    assert iteration_bytes(rp, 32, "fp32", 1, "cheb", 2, c) == \
        (12 + 2 * c) * nent + 16 * (w.size + 1) + 92 * n
    # a coupling 40,000 rows away needs 32-bit columns
    assert column_bytes(np.array([1, 2, 3]), np.array([40001, 1])) == 4
