"""GPU parity of the device setup path (hs_setup.hip): the CSC -> CSR
transpose, the sliced-ELL operator built on the device, and the validation of
the caller's sparse arrays.

Bar: the operator read back through hs_csr_get is scipy's CSR of the same
matrix exactly (structure and values, columns ascending in every row); SpMV
equals the oracle's row-order products bit for bit; the non-symmetric l1
diagonal built from the device transpose equals the oracle's literal O(N^2)
restatement bit for bit; malformed arrays are refused with HS_EINVAL before any
kernel indexes through them.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _awkward(n, seed):
    """Random non-symmetric matrix with empty rows and columns, one dense row
    and one dense column, unsorted construction (scipy sorts on conversion)."""
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=min(1.0, 4.0 / max(n, 1)), random_state=rng, format="lil")
    if n > 3:
        A[1, :] = rng.uniform(-1, 1, n)          # dense row
        A[:, n - 2] = rng.uniform(-1, 1, (n, 1))  # dense column
        A[2, :] = 0                               # empty row
        A[:, 3 % n] = 0                           # empty column
    A = A.tocsc()
    A.eliminate_zeros()
    A.sort_indices()
    return A


@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 517, 5000])
def test_device_operator_equals_scipy(gpu, n):
    import oracle as O
    from hybridsmoothers.chebyshev import csr_download
    hs = gpu
    A = _awkward(n, n)
    R = A.tocsr()
    R.sort_indices()
    x = np.random.default_rng(n + 1).uniform(-1, 1, n)
    ref = O.spmv(O.csr1(A), x)
    for M in (A, R):  # CSC input (device transpose) and CSR input
        D = hs.DeviceMatrix(M)
        B = csr_download(D.handle)
        np.testing.assert_array_equal(B.indptr, R.indptr)
        np.testing.assert_array_equal(B.indices, R.indices)
        np.testing.assert_array_equal(B.data, R.data)
        y = np.zeros(n)
        D.mul(y, x)
        np.testing.assert_array_equal(y, ref)


def test_device_operator_large(gpu):
    """2048^2 upwind convection-diffusion (4.2 M rows): exact structure and
    values after the device transpose."""
    from hybridsmoothers import problems as pb
    from hybridsmoothers.chebyshev import csr_download
    hs = gpu
    A = pb.convdiff2d(2048)
    D = hs.DeviceMatrix(A)
    B = csr_download(D.handle)
    R = A.tocsr()
    R.sort_indices()
    np.testing.assert_array_equal(B.indptr, R.indptr)
    np.testing.assert_array_equal(B.indices, R.indices)
    np.testing.assert_array_equal(B.data, R.data)


@pytest.mark.parametrize("partsize", [3, 32, 64])
def test_nonsym_l1_diagonal_device_transpose(gpu, partsize):
    import oracle as O
    hs = gpu
    n = 700
    rng = np.random.default_rng(partsize)
    B = sp.random(n, n, density=0.01, random_state=rng, format="lil")
    B[5, :] = rng.uniform(-1, 1, n)  # dense row: many off-partition terms in one sum
    B = B.tocsc()
    A = (B + sp.diags(np.abs(B).sum(axis=1).A1 + 1.0)).tocsc()
    A.sort_indices()
    ref = O.l1gs_build(A, partsize, is_sym=False, sweep="forward", gpu=(64, 7),
                       fast_nonsym=False)  # the reference's literal O(N^2) scan
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=7)(
        A, partsize, isSymA=False, sweep=hs.ForwardSweep())
    np.testing.assert_array_equal(P.sweep.op.D_DL1, ref.D)
    np.testing.assert_array_equal(P.sweep.op.L.SLbuffer, ref.SL)
    x = rng.uniform(-1, 1, n)
    np.testing.assert_array_equal(hs.backslash(P, x), O.l1gs_ldiv(ref, x))


def _arrays(n):
    from hybridsmoothers import problems as pb
    A = pb.poisson_test_matrix(n)
    return (A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1,
            A.data.astype(np.float64))


@pytest.mark.parametrize("bad", ["ptr_decreases", "ptr_start", "ind_zero", "ind_past_n"])
def test_invalid_sparse_arrays_rejected(gpu, bad):
    from hybridsmoothers import _lib as L
    n = 9
    ptr, ind, val = _arrays(n)
    if bad == "ptr_decreases":
        ptr[4], ptr[5] = ptr[5], ptr[4]
    elif bad == "ptr_start":
        ptr[0] = 2
    elif bad == "ind_zero":
        ind[3] = 0
    else:
        ind[3] = n + 1
    h = C.c_void_p()
    st = L.lib.hs_csr_create(n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data, L.FMT_CSC,
                             None, C.byref(h))
    assert st == L.HS_EINVAL, bad
    assert "invalid sparse matrix" in L.lib.hs_last_error().decode()
    st = L.lib.hs_l1gs_build(n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data, L.FMT_CSC, 0,
                             3, 1.5, L.SWEEP_SYMMETRIC, L.CACHE_PACKED, L.PREC_FP64, 4, 4, None,
                             C.byref(h))
    assert st == L.HS_EINVAL, bad
    assert "invalid sparse matrix" in L.lib.hs_last_error().decode()


@pytest.mark.parametrize("far", [False, True])
def test_operator_column_formats(gpu, far):
    """Products are the oracle's bit for bit whether the operator stores 16-bit
    column offsets (every column within 32,767 of its row) or, with one
    coupling 39,999 apart, the 32-bit columns."""
    import oracle as O
    from hybridsmoothers.chebyshev import csr_download
    hs = gpu
    n = 40000
    rng = np.random.default_rng(7)
    A = sp.diags([rng.uniform(-1, 1, n - k) for k in (1, 300, 20000)], [1, 300, 20000],
                 shape=(n, n), format="lil")
    if far:
        A[0, n - 1] = 0.5
    A = (A + A.T + sp.identity(n) * 4.0).tocsc()
    A.sort_indices()
    x = rng.uniform(-1, 1, n)
    D = hs.DeviceMatrix(A)
    y = np.zeros(n)
    D.mul(y, x)
    np.testing.assert_array_equal(y, O.spmv(O.csr1(A), x))
    B = csr_download(D.handle)
    R = A.tocsr()
    R.sort_indices()
    np.testing.assert_array_equal(B.indices, R.indices)
    np.testing.assert_array_equal(B.data, R.data)


@pytest.mark.parametrize("far", [32767, 32768, 70000])
def test_column_format_boundary(gpu, far):
    """Column formats at their boundaries.  A coupling exactly 32,767 rows away
    still fits the 16-bit offsets from the row; one 32,768 away does not, and
    the device operator then stores 16-bit deltas from a base per (slice, k)
    -- the k-th entries of a 64-row slice span 32,768 columns -- while the
    batched engine's upload falls back to 32-bit columns; one 70,000 away fits
    neither 16-bit format.  The formats are decided on the device (the
    operator) and on the host (the batched upload, and the bytes model the
    benchmark uses), and the products are the same in every format."""
    import oracle as O
    from hybridsmoothers.batch import BatchSolver, column_bytes
    hs = gpu
    n = 80000
    rng = np.random.default_rng(far)
    A = sp.diags([rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - far)], [1, far],
                 shape=(n, n), format="csc")
    A = (A + A.T + 8.0 * sp.identity(n)).tocsc()
    A.sort_indices()
    want_row = 2 if far <= 32767 else 4        # 16-bit offsets from the row
    want_dev = 2 if far <= 65535 else 4        # ... or deltas from the (slice, k) base
    assert column_bytes(A.indptr + 1, A.indices + 1) == want_row
    M = hs.DeviceMatrix(A)
    assert M.column_bytes == want_dev
    x = rng.uniform(-1, 1, n)
    y = np.empty(n)
    M.mul(y, x)
    np.testing.assert_array_equal(y, O.spmv(O.csr1(A), x))
    b = rng.uniform(-1, 1, n)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(A, 64, isSymA=True)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 64))
    _, so = O.pcg(A, b, ref, rtol=1e-8)
    sol = hs.pcg_solve(M, b, Pl=P, reltol=1e-8)
    assert sol.iters == so.iters
    for kw in (dict(partsize=32), dict(smoother="cheb", degree=2)):
        with BatchSolver(precision="fp64", passes=4, arena_bytes=128 << 20, max_systems=2,
                         **kw) as eng:
            sid = eng.submit(A, b)
            eng.step()
            assert int(eng.peek(sid, "column_bytes")[0]) == want_row
            while eng.pending():
                eng.step()
                eng.collect()
