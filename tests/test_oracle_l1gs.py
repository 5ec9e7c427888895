"""Pins of the CPU oracle's L1GS (CPU only).

Each check is independent of the oracle's own formulas: the reference's
known-answer vectors (tests/golden/known_answers.json), textbook reductions
(partsize = N is classical Gauss-Seidel, solved with scipy triangular solves),
a dense brute-force block solve whose D_Dl1 is computed from the definition in
numpy, the literal O(N^2) vs O(nnz) non-symmetric diagonal, closed forms on
stencils, and the reference's GS sanity iteration counts.
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


def _n9():
    return pb.poisson_test_matrix(9)


@pytest.mark.parametrize("sweep", ["forward", "backward", "symmetric"])
@pytest.mark.parametrize("cache", ["packed", "view"])
@pytest.mark.parametrize("fmt", ["csc", "csr"])
def test_known_answers_n9(sweep, cache, fmt):
    f = KA["n9_poisson"]
    A = _n9() if fmt == "csc" else _n9().tocsr()
    x = np.array(f["x"], dtype=np.float64)
    for chunks in f["chunks"]:
        p = O.l1gs_build(A, f["partsize"], is_sym=True, sweep=sweep, cache=cache, chunks=chunks)
        np.testing.assert_array_equal(O.l1gs_ldiv(p, x), f["y_" + sweep])
        np.testing.assert_array_equal(p.D, f["D_DL1"])
        if cache == "packed":
            if p.SL is not None:
                np.testing.assert_array_equal(p.SL, f["SLbuffer"])
            if p.SU is not None:
                np.testing.assert_array_equal(p.SU, f["SUbuffer"])


def test_known_answers_eta():
    f = KA["eta_cases"]
    A = _n9()
    for case in f["cases"]:
        p = O.l1gs_build(A, f["partsize"], eta=case["eta"], sweep="forward", chunks=f["chunks"])
        np.testing.assert_array_equal(p.D, case["D_DL1"])
        if "SLbuffer" in case:
            np.testing.assert_array_equal(p.SL, case["SLbuffer"])
    p = O.l1gs_build(A, 2, eta=2.0, sweep="forward", chunks=2)
    np.testing.assert_array_equal(O.l1gs_ldiv(p, np.arange(9.0)), f["y_forward_eta2"])


@pytest.mark.parametrize("fast", [False, True])
def test_known_answers_nonsym(fast):
    f = KA["nonsym_csc"]
    A = _n9().tolil()
    for i, j, v in f["extra_entries_1based"]:
        A[i - 1, j - 1] = v
    A = A.tocsc()
    x = np.array(f["x"], dtype=np.float64)
    for sweep in ("forward", "backward"):
        for cache in ("packed", "view"):
            p = O.l1gs_build(A, f["partsize"], sweep=sweep, cache=cache, chunks=f["chunks"],
                             fast_nonsym=fast)
            np.testing.assert_array_equal(p.D, f["D_DL1"])
            np.testing.assert_allclose(O.l1gs_ldiv(p, x), f["y_" + sweep], rtol=1e-15)
            if cache == "packed":
                buf = p.SL if sweep == "forward" else p.SU
                np.testing.assert_array_equal(buf, f["SLbuffer"])


def test_known_answers_partsize3():
    f = KA["partsize3"]
    p = O.l1gs_build(_n9(), f["partsize"], sweep="forward", chunks=f["chunks"])
    np.testing.assert_array_equal(p.D, f["D_DL1"])
    np.testing.assert_array_equal(p.SL, f["SLbuffer"])


def test_partitioning_formulas():
    # _blockpartitioning CPU (:340-351) and GPU (:353-365)
    assert O.partition_cpu(9, 2, 3) == (2, 5, 3, 2)
    assert O.partition_cpu(2048, 16, 8) == (16, 128, 8, 16)
    assert O.partition_gpu(2048, 16, 256, 20) == (16, 128, 20, 7)
    assert O.partition_gpu(2048, 16, 4, 20) == (16, 128, 20, 4)
    assert O.partition_gpu(128 ** 3, 64, 256, 1024) == (64, 32768, 1024, 32)
    # packed buffer size with a short last partition (:698-701)
    assert O.packed_size(9, 2, 5) == 4
    assert O.packed_size(9, 3, 3) == 9
    assert O.packed_size(10, 4, 3) == 6 + 6 + 1


def _rand_sym(n, seed, density=0.05):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csc")
    B.data -= 0.5
    A = (B + B.T + sp.diags(rng.uniform(0.5, 3.0, n))).tocsc()
    A.sort_indices()
    return A


def _rand_nonsym(n, seed, density=0.05):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csc")
    B.data -= 0.5
    A = (B + sp.diags(rng.uniform(0.5, 3.0, n))).tocsc()
    A.sort_indices()
    return A


def _dense_reference(A, partsize, eta, sweep, sym):
    """Brute force from the definitions: D_Dl1 from |a_ij| sums outside the
    block, then M^-1 for M = D+L, D+U or (D+L) D^-1 (D+U) per diagonal block,
    inverted densely.  For isSymA the reference reads column i as row i, i.e.
    it works with A^T."""
    Ad = A.toarray()
    if sym:
        Ad = Ad.T
    n = Ad.shape[0]
    D = np.zeros(n)
    for i in range(n):
        k = i // partsize
        s, e = k * partsize, min(n, (k + 1) * partsize)
        off = np.abs(Ad[i, :s]).sum() + np.abs(Ad[i, e:]).sum()
        a = Ad[i, i]
        D[i] = a + (0.0 if a >= eta * off else off / 2)
    Minv = np.zeros((n, n))
    for s in range(0, n, partsize):
        e = min(n, s + partsize)
        B = Ad[s:e, s:e]
        Dk = np.diag(D[s:e])
        Lk = np.tril(B, -1)
        Uk = np.triu(B, 1)
        if sweep == "forward":
            M = Dk + Lk
        elif sweep == "backward":
            M = Dk + Uk
        else:
            M = (Dk + Lk) @ np.linalg.inv(Dk) @ (Dk + Uk)
        Minv[s:e, s:e] = np.linalg.inv(M)
    return D, Minv


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("partsize", [1, 3, 7, 16, 50])
def test_dense_bruteforce(seed, partsize):
    n = 53
    x = np.random.default_rng(100 + seed).uniform(-1, 1, n)
    for sym in (True, False):
        A = _rand_sym(n, seed) if sym else _rand_nonsym(n, seed)
        for sweep in ("forward", "backward", "symmetric"):
            D, Minv = _dense_reference(A, partsize, 1.5, sweep, sym)
            for cache in ("packed", "view"):
                p = O.l1gs_build(A, partsize, is_sym=sym, sweep=sweep, cache=cache, chunks=3)
                np.testing.assert_allclose(p.D, D, rtol=1e-14, atol=0)
                y = O.l1gs_ldiv(p, x)
                np.testing.assert_allclose(y, Minv @ x, rtol=1e-9, atol=1e-12)


def test_partsize_one_is_jacobi():
    A = _rand_sym(40, 9)
    x = np.random.default_rng(2).uniform(-1, 1, 40)
    for sweep in ("forward", "backward"):
        p = O.l1gs_build(A, 1, is_sym=True, sweep=sweep, chunks=2)
        np.testing.assert_array_equal(O.l1gs_ldiv(p, x), x / p.D)


@pytest.mark.parametrize("fmt", ["csc", "csr"])
def test_partsize_n_is_gauss_seidel(fmt):
    """Single partition: D_Dl1 = diag(A) and the sweeps are triangular solves."""
    A = pb.poisson2d_matrix(12, eps=0.3)
    if fmt == "csr":
        A = A.tocsr()
    n = A.shape[0]
    x = np.random.default_rng(5).uniform(-1, 1, n)
    Dm = sp.diags(A.diagonal())
    Lo = sp.tril(A, -1)
    Up = sp.triu(A, 1)
    yf = spla.spsolve_triangular((Dm + Lo).tocsr(), x, lower=True)
    yb = spla.spsolve_triangular((Dm + Up).tocsr(), x, lower=False)
    ys = spla.spsolve_triangular((Dm + Up).tocsr(), A.diagonal() * yf, lower=False)
    for cache in ("packed", "view"):
        for sweep, ye in (("forward", yf), ("backward", yb), ("symmetric", ys)):
            p = O.l1gs_build(A, n, is_sym=True, sweep=sweep, cache=cache, chunks=1)
            np.testing.assert_array_equal(p.D, A.diagonal())
            np.testing.assert_allclose(O.l1gs_ldiv(p, x), ye, rtol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_nonsym_diag_literal_equals_fast(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 300))
    A = _rand_nonsym(n, seed, density=float(rng.uniform(0.02, 0.3)))
    for partsize in (1, 2, 5, 17, n):
        lit = O.l1gs_build(A, partsize, sweep="forward", chunks=2, fast_nonsym=False)
        fast = O.l1gs_build(A, partsize, sweep="forward", chunks=2, fast_nonsym=True)
        np.testing.assert_array_equal(lit.D, fast.D)


def test_diag_indices_errors():
    A = _n9().tolil()
    A[3, 3] = 0.0  # LIL drops the entry: a missing diagonal
    with pytest.raises(O.SingularException) as e:
        O.l1gs_build(A.tocsc(), 2, sweep="forward")
    assert e.value.col == 4
    B = _n9().copy()
    B.data[B.indptr[6] + 1] = 0.0  # stored zero on the diagonal of column 7
    with pytest.raises(O.SingularException) as e:
        O.l1gs_build(B, 2, sweep="forward")
    assert e.value.col == 7
    # the symmetric path does not build DiagonalIndices (:620)
    O.l1gs_build(B, 2, is_sym=True, sweep="forward")


def test_stencil_closed_forms():
    p = O.l1gs_build(pb.laplace1d(2048), 16, is_sym=True, gpu=(256, 20))
    np.testing.assert_array_equal(p.D, 2.0)
    A = pb.poisson3d(24)
    p = O.l1gs_build(A, 12, is_sym=True, gpu=(64, 8))
    n = 24
    D = p.D.reshape(n, n, n)  # [z, y, x]
    vals, counts = np.unique(p.D, return_counts=True)
    assert set(vals) == {6.0, 8.5}
    # 8.5 where the x-neighbour across a partition boundary adds to 4 in-plane
    # neighbours: x in {11, 12}, interior y and z
    assert counts[list(vals).index(8.5)] == 2 * (n - 2) ** 2
    assert np.all(D[1:-1, 1:-1, [11, 12]] == 8.5)


def _gs_iterations(A, b, sweep_pass, tol=1e-10, maxiter=10000):
    """Stationary GS iteration written with single sweeps (:614-709)."""
    n = A.shape[0]
    Lo = sp.tril(A, -1).tocsr()
    Up = sp.triu(A, 1).tocsr()
    x = np.zeros(n)
    for it in range(1, maxiter + 1):
        x = sweep_pass(x, b, Lo, Up)
        if np.linalg.norm(A @ x - b) < tol:
            return it
    return maxiter


def test_gs_sanity_iteration_counts():
    """test/test_preconditioners.jl:614-709: single-partition L1GS as a GS solver."""
    N = 50
    A = pb.poisson_test_matrix(N)
    b = np.ones(N)
    pf = O.l1gs_build(A, N, is_sym=True, sweep="forward", cache="view", chunks=1)
    pb_ = O.l1gs_build(A, N, is_sym=True, sweep="backward", cache="view", chunks=1)

    def fwd(x, b, Lo, Up):
        y = b - Up @ x
        O.l1gs_sweep_inplace(pf, y, False)
        return y

    def bwd(x, b, Lo, Up):
        y = b - Lo @ x
        O.l1gs_sweep_inplace(pb_, y, True)
        return y

    def sym(x, b, Lo, Up):
        return bwd(fwd(x, b, Lo, Up), b, Lo, Up)

    it_f = _gs_iterations(A, b, fwd)
    it_b = _gs_iterations(A, b, bwd)
    it_s = _gs_iterations(A, b, sym)
    assert it_f == it_b
    assert abs(it_s - it_f // 2) <= it_f // 100


def test_symmetric_wrapper_matches_full():
    from hybridsmoothers.l1gs import Symmetric
    A = _rand_sym(30, 4)
    for uplo in ("U", "L"):
        half = sp.triu(A) if uplo == "U" else sp.tril(A)
        full = Symmetric(half.tocsc(), uplo).full()
        np.testing.assert_array_equal(full.toarray(), A.toarray())


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("partsize", [1, 4, 7, 16, 53])
def test_dense_bruteforce_nonsym_csr(seed, partsize):
    """SparseMatrixCSR input is read by rows whatever isSymA says
    (l1_gauss_seidel.jl:1231-1242, 1351-1374, 1496-1519, 1588-1597): on a
    non-symmetric matrix the result is the dense block inverse of A itself,
    not of A^T -- which a row/column swap on the CSR path would produce."""
    n = 53
    A = _rand_nonsym(n, 40 + seed)
    assert abs(A - A.T).max() > 0.1
    x = np.random.default_rng(200 + seed).uniform(-1, 1, n)
    R = A.tocsr()
    for sweep in ("forward", "backward", "symmetric"):
        D, Minv = _dense_reference(A, partsize, 1.5, sweep, sym=False)
        Dt, Minvt = _dense_reference(A, partsize, 1.5, sweep, sym=True)  # A^T
        for is_sym in (False, True):
            for cache in ("packed", "view"):
                p = O.l1gs_build(R, partsize, is_sym=is_sym, sweep=sweep, cache=cache, chunks=3)
                np.testing.assert_allclose(p.D, D, rtol=1e-14, atol=0)
                y = O.l1gs_ldiv(p, x)
                np.testing.assert_allclose(y, Minv @ x, rtol=1e-9, atol=1e-12)
                if partsize > 1:
                    assert np.abs(y - Minvt @ x).max() > 1e-6
        # the CSC of the same matrix (DiagonalIndices path) agrees with the CSR
        p = O.l1gs_build(A, partsize, sweep=sweep, chunks=2)
        q = O.l1gs_build(R, partsize, sweep=sweep, chunks=2)
        np.testing.assert_array_equal(p.D, q.D)
        np.testing.assert_allclose(O.l1gs_ldiv(p, x), O.l1gs_ldiv(q, x), rtol=1e-13, atol=1e-15)
