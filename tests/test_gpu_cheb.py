"""GPU parity of the Chebyshev smoothers (through the C ABI) against the oracle,
plus the reference's own property tests (test/test_chebyshev.jl).

Tolerances: l1inv and the CSR operator are bit-identical; rho agrees to 1e-13
relative (the squared norm is summed in a different order); smoother outputs
agree to 1e-12 relative (FP64) -- their only difference is that rho.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _pois(n):
    from hybridsmoothers import problems as pb
    return pb.poisson_test_matrix(n)


def _cfg(hs, kind, degree, precision="fp64"):
    C = hs.ChebyshevFourth if kind == 4 else hs.ChebyshevFirst
    return C(degree=degree, backend=hs.ROCBackend(), precision=precision)


@pytest.mark.parametrize("kind", [4, 1])
def test_build_fields(gpu, kind):
    import oracle as O
    hs = gpu
    A = _pois(50)
    s = hs._build_smoother(_cfg(hs, kind, 4), A)
    ref = O.cheb_build(A, kind, 4)
    np.testing.assert_array_equal(s.l1inv, ref.l1inv)
    np.testing.assert_array_equal(s.l1inv, 1.0 / O.l1_rowsum(A))
    assert abs(s.rho - ref.rho) <= 1e-13 * ref.rho
    assert 0 < s.rho <= 1
    assert s.degree == 4 and len(s.r) == 50 and len(s.z) == 50 and len(s.tmp) == 50
    Acsr = s.A
    assert (Acsr != A.tocsr()).nnz == 0
    if kind == 1:
        assert s.a == ref.a and 0 < s.a < 1


@pytest.mark.parametrize("kind", [4, 1])
@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4, 8])
def test_smooth_parity(gpu, kind, degree):
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    for A in (_pois(300), pb.poisson3d(12), pb.lognormal2d(20, seed=3)):
        n = A.shape[0]
        rng = np.random.default_rng(degree + 10 * kind)
        b = rng.uniform(-1, 1, n)
        x0 = rng.uniform(-1, 1, n)
        s = hs._build_smoother(_cfg(hs, kind, degree), A)
        ref = O.cheb_build(A, kind, degree)
        ref.rho = s.rho  # isolate the recurrences from the norm's summation order
        xo, ro, zo = O.cheb_smooth(ref, x0, b)
        x = x0.copy()
        hs.smooth_(x, s, b)
        np.testing.assert_array_equal(x, xo)
        np.testing.assert_array_equal(s.z, zo)
        if kind == 4:
            np.testing.assert_array_equal(s.r, ro)
        # as a left preconditioner: z = 0; smooth!(z, S, r)
        z = np.empty(n)
        hs.ldiv_(z, s, b)
        zo2, _, _ = O.cheb_smooth(ref, np.zeros(n), b)
        np.testing.assert_array_equal(z, zo2)


@pytest.mark.parametrize("kind", [4, 1])
def test_full_rho_parity_and_device_vectors(gpu, kind):
    import torch
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(32)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 4)
    s = hs._build_smoother(_cfg(hs, kind, 3), A)
    ref = O.cheb_build(A, kind, 3)
    assert abs(s.rho - ref.rho) <= 1e-13 * ref.rho
    zo, _, _ = O.cheb_smooth(ref, np.zeros(n), b)
    bd = torch.tensor(b, device="cuda")
    zd = torch.empty_like(bd)
    hs.ldiv_(zd, s, bd)
    assert relerr(zd.cpu().numpy(), zo) < 1e-12


@pytest.mark.parametrize("kind", [4, 1])
def test_fp32_smoother(gpu, kind):
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal3d(16, seed=2)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 5)
    s32 = hs._build_smoother(_cfg(hs, kind, 3, "fp32"), A)
    ref = O.cheb_build(A, kind, 3)
    zo, _, _ = O.cheb_smooth(ref, np.zeros(n), b)
    z = np.empty(n)
    hs.ldiv_(z, s32, b)
    assert relerr(z, zo) < 2e-6


# ---- properties from test/test_chebyshev.jl ---------------------------------
@pytest.mark.parametrize("kind", [4, 1])
def test_residual_reduction(gpu, kind):
    """test_chebyshev.jl:50-63, 115-128 and :82-95, 146-158."""
    hs = gpu
    N = 300
    A = _pois(N)
    b = np.ones(N)
    res = []
    for degree in (1, 2, 4, 8):
        s = hs._build_smoother(_cfg(hs, kind, degree), A)
        x = np.zeros(N)
        hs.smooth_(x, s, b)
        r = np.linalg.norm(b - A @ x)
        assert r < np.linalg.norm(b)
        res.append(r)
    assert res == sorted(res, reverse=True)


@pytest.mark.parametrize("kind", [4, 1])
def test_monotone_over_applications(gpu, kind):
    """test_chebyshev.jl:65-80, 130-144 (own seeded rhs instead of MersenneTwister(42))."""
    hs = gpu
    N = 200
    A = _pois(N)
    b = np.random.default_rng(42).random(N)
    s = hs._build_smoother(_cfg(hs, kind, 4), A)
    x = np.zeros(N)
    prev = np.linalg.norm(b)
    for _ in range(10):
        hs.smooth_(x, s, b)
        cur = np.linalg.norm(b - A @ x)
        assert cur < prev
        prev = cur
    assert prev / np.linalg.norm(b) < 0.9


def test_astar_and_first_vs_fourth(gpu):
    """test_chebyshev.jl:160-190."""
    hs = gpu
    astars = [hs._build_smoother(_cfg(hs, 1, k), _pois(10)).a for k in range(1, 9)]
    assert astars == sorted(astars, reverse=True)
    N = 300
    A = _pois(N)
    b = np.random.default_rng(7).random(N)
    for degree in range(1, 6):
        s4 = hs._build_smoother(_cfg(hs, 4, degree), A)
        s1 = hs._build_smoother(_cfg(hs, 1, degree), A)
        x4 = np.zeros(N)
        x1 = np.zeros(N)
        hs.smooth_(x4, s4, b)
        hs.smooth_(x1, s1, b)
        r4 = np.linalg.norm(b - A @ x4) / np.linalg.norm(b)
        r1 = np.linalg.norm(b - A @ x1) / np.linalg.norm(b)
        assert r4 < 1 and r1 < 1
        assert abs(r1 - r4) / r4 < 0.02


@pytest.mark.parametrize("kind", [4, 1])
def test_repeatable(gpu, kind):
    """test_chebyshev.jl:192-218: identical results on repeated runs."""
    hs = gpu
    N = 500
    A = _pois(N)
    b = np.ones(N)
    s = hs._build_smoother(_cfg(hs, kind, 4), A)
    x1 = np.zeros(N)
    hs.smooth_(x1, s, b)
    x2 = np.zeros(N)
    hs.smooth_(x2, s, b)
    np.testing.assert_array_equal(x1, x2)
