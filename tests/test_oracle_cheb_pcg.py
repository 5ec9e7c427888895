"""Pins of the CPU oracle's Chebyshev smoothers and CG (CPU only).

* l1 row sums and a* against the reference's known answers / table;
* the smoothers against closed-form residual polynomials from the literature:
  from x0 = 0, degree k of the 4th kind (Lottes) leaves the residual
  W_k(I - 2S) b / (2k+1), and the 1st kind with left end a* leaves
  T_k(((1+a)I - 2S)/(1-a)) b / T_k((1+a)/(1-a)), where S = A M^-1 / rho and
  W_k, T_k are the Chebyshev polynomials of the 4th and 1st kind (built here
  from their own three-term recurrences, not from the smoother's);
* the power iteration against dense eigenvalues;
* CG with M = I against scipy's CG (same stopping rule when M = I) and
  against a direct solve.
The iteration counts of Krylov.jl itself are PARITY UNPINNED (the package is
not importable here); the rule restated is its documented one.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from hybridsmoothers import problems as pb

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def test_l1_rowsum_known_answer():
    np.testing.assert_array_equal(O.l1_rowsum(pb.poisson_test_matrix(5)), KA["l1_rowsum"]["l1"])


def test_astar():
    assert abs(O.cheb1_astar(1) - 1.0 / 3.0) < 1e-15  # 1 - 3x^2 = 0 for k = 1
    tab = KA["cheb1_astar_table"]
    vals = [O.cheb1_astar(k) for k in tab["k"]]
    np.testing.assert_allclose(vals, tab["astar_3digits"], atol=5e-4)
    assert vals == sorted(vals, reverse=True)
    for k in range(1, 9):  # a* is a root of the defining equation
        x = np.sqrt(O.cheb1_astar(k))
        phi = 8 * k * (1 - x * x) ** (2 * k) + x * ((1 - x) ** (4 * k) - (1 + x) ** (4 * k))
        assert abs(phi) < 1e-9 * 8 * k


def _W(S, k):
    I = np.eye(len(S))
    T = I - 2 * S
    W0, W1 = I, 2 * T + I  # W_0 = 1, W_1(t) = 2t + 1
    if k == 0:
        return W0
    for _ in range(k - 1):
        W0, W1 = W1, 2 * T @ W1 - W0
    return W1


def _T(X, k):
    I = np.eye(len(X))
    T0, T1 = I, X
    if k == 0:
        return T0
    for _ in range(k - 1):
        T0, T1 = T1, 2 * X @ T1 - T0
    return T1


@pytest.mark.parametrize("k", range(1, 8))
def test_cheb4_residual_polynomial(k):
    A = pb.lognormal2d(7, seed=k)
    n = A.shape[0]
    Ad = A.toarray()
    b = np.random.default_rng(k).uniform(-1, 1, n)
    c = O.cheb_build(A, 4, k)
    x, r, _ = O.cheb_smooth(c, np.zeros(n), b)
    S = Ad @ np.diag(c.l1inv) / c.rho
    rk = _W(S, k) @ b / (2 * k + 1)
    assert np.linalg.norm((b - Ad @ x) - rk) <= 1e-13 * np.linalg.norm(rk)
    # the smoother's own r is the residual it leaves (:179)
    np.testing.assert_allclose(r, b - Ad @ x, rtol=0, atol=1e-12 * np.linalg.norm(b))


@pytest.mark.parametrize("k", range(1, 8))
def test_cheb1_residual_polynomial(k):
    A = pb.lognormal2d(7, seed=10 + k)
    n = A.shape[0]
    Ad = A.toarray()
    b = np.random.default_rng(k).uniform(-1, 1, n)
    c = O.cheb_build(A, 1, k)
    a = c.a
    x, _, _ = O.cheb_smooth(c, np.zeros(n), b)
    S = Ad @ np.diag(c.l1inv) / c.rho
    X = ((1 + a) * np.eye(n) - 2 * S) / (1 - a)
    t0 = np.polynomial.chebyshev.chebval((1 + a) / (1 - a), [0] * k + [1])
    rk = _T(X, k) @ b / t0
    assert np.linalg.norm((b - Ad @ x) - rk) <= 1e-13 * np.linalg.norm(rk)


def test_smooth_from_nonzero_x_is_affine():
    """smooth!(x0) = x0 + smooth!(0) applied to the residual b - A x0."""
    A = pb.lognormal2d(9, seed=4)
    n = A.shape[0]
    rng = np.random.default_rng(3)
    b, x0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for kind in (4, 1):
        c = O.cheb_build(A, kind, 3)
        x1, _, _ = O.cheb_smooth(c, x0, b)
        dx, _, _ = O.cheb_smooth(c, np.zeros(n), b - A @ x0)
        np.testing.assert_allclose(x1, x0 + dx, rtol=1e-12, atol=1e-13)


def test_spectral_radius_vs_eigenvalues():
    rng = np.random.default_rng(1)
    n = 40
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.concatenate([[10.0], rng.uniform(0.1, 2.0, n - 1)])
    A = sp.csc_matrix(Q @ np.diag(ev) @ Q.T)
    l1inv = 1.0 / O.l1_rowsum(A)
    rho = O.spectral_radius(O.csr1(A), l1inv, 50)
    lam = np.max(np.abs(np.linalg.eigvals(np.diag(l1inv) @ A.toarray())))
    assert abs(rho - lam) < 1e-8 * lam
    # reference property (test_chebyshev.jl:21-29): 0 < rho <= 1 for the 1D Poisson matrix
    P = pb.poisson_test_matrix(100)
    r = O.spectral_radius(O.csr1(P), 1.0 / O.l1_rowsum(P), 50)
    assert 0 < r <= 1


def test_csr_conversion_and_spmv():
    A = pb.lognormal3d(6, seed=3)
    R = O.csr1(A)
    C = A.tocsr()
    C.sort_indices()
    np.testing.assert_array_equal(R.ptr - 1, C.indptr)
    np.testing.assert_array_equal(R.ind - 1, C.indices)
    np.testing.assert_array_equal(R.val, C.data)
    x = np.random.default_rng(0).uniform(-1, 1, A.shape[0])
    np.testing.assert_allclose(O.spmv(R, x), A @ x, rtol=1e-14, atol=1e-14)


def test_cg_matches_scipy():
    A = pb.poisson3d(16)
    b = pb.uniform_rhs(A.shape[0], 0)
    x, st = O.pcg(A, b, None, rtol=1e-8)
    its = [0]
    xs, info = spla.cg(A, b, rtol=1e-8, atol=0.0, maxiter=10000,
                       callback=lambda _: its.__setitem__(0, its[0] + 1))
    assert info == 0 and st.status == 0
    assert st.iters == its[0]
    np.testing.assert_allclose(x, xs, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("prec", ["l1gs", "cheb4", "cheb1"])
def test_pcg_solves(prec):
    A = pb.lognormal3d(10, seed=5)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 1)
    if prec == "l1gs":
        M = O.l1gs_build(A, 16, is_sym=True, sweep="symmetric", gpu=(64, 8))
    else:
        M = O.cheb_build(A, 4 if prec == "cheb4" else 1, 3)
    x, st = O.pcg(A, b, M, rtol=1e-10)
    xd = spla.spsolve(A.tocsc(), b)
    assert st.status == 0
    np.testing.assert_allclose(x, xd, rtol=1e-7, atol=1e-9)
    _, st0 = O.pcg(A, b, None, rtol=1e-10)
    assert st.iters <= st0.iters  # the reference's assertion (test_preconditioners.jl:103)
