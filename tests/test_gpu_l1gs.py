"""GPU parity of the L1GS build and sweeps (through the C ABI) against the oracle.

Tolerance: FP64 results agree with the oracle within 1e-12 relative error;
forward sweeps and every D_Dl1 / packed buffer are expected bit-identical
(same summation order, no FMA contraction on either side), and are checked so.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12
SWEEPS = ["forward", "backward", "symmetric"]


def _hs_sweep(hs, name):
    return {"forward": hs.ForwardSweep(), "backward": hs.BackwardSweep(),
            "symmetric": hs.SymmetricSweep()}[name]


def _random_sym(n, density, seed):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csc")
    A = B + B.T
    A = A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0 + rng.random(n))
    A = A.tocsc()
    A.sort_indices()
    return A


def _random_nonsym(n, density, seed):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csc")
    B.data -= 0.5
    A = B + sp.diags(np.abs(B).sum(axis=1).A1 * rng.uniform(0.3, 2.0, n) + 0.5)
    A = A.tocsc()
    A.sort_indices()
    return A


def test_known_answers_n9(gpu):
    """test/test_preconditioners.jl:110-276 and test/gpu/gpu.jl:113-357."""
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson_test_matrix(9)
    x = np.arange(9, dtype=np.float64)
    exp = {"forward": [0, 0.5, 1.0, 2.0, 2.0, 3.5, 3.0, 5.0, 4.0],
           "backward": [0.25, 0.5, 1.75, 1.5, 3.25, 2.5, 4.75, 3.5, 4.0],
           "symmetric": [0.25, 0.5, 2.0, 2.0, 3.75, 3.5, 5.5, 5.0, 4.0]}
    for threads in (1, 3, 10):
        for blocks in (1, 4, 10):
            b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=threads, blocks=blocks)
            for M in (A, A.tocsr()):
                for sw in SWEEPS:
                    for cache in (hs.PackedBufferCache(), hs.MatrixViewCache()):
                        P = b(M, 2, isSymA=True, sweep=_hs_sweep(hs, sw), cache_strategy=cache)
                        y = hs.backslash(P, x)
                        np.testing.assert_array_equal(y, exp[sw])
                        if sw == "symmetric":
                            np.testing.assert_array_equal(P.sweep.lop.D_DL1, 2.0)
                            np.testing.assert_array_equal(P.sweep.uop.D_DL1, 2.0)
                            if isinstance(cache, hs.PackedBufferCache):
                                np.testing.assert_array_equal(P.sweep.lop.L.SLbuffer, -1.0)
                                np.testing.assert_array_equal(P.sweep.uop.U.SUbuffer, -1.0)
                        else:
                            np.testing.assert_array_equal(P.sweep.op.D_DL1, 2.0)


def test_eta_nonsym_partsize_known_answers(gpu):
    """test/test_preconditioners.jl:278-364."""
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson_test_matrix(9)
    x = np.arange(9, dtype=np.float64)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=2, blocks=2)
    for eta, D in ((2.0, [2.0] * 9), (3.0, [2.0] + [2.5] * 8), (1.0, [2.0] * 9)):
        P = b(A, 2, eta=eta, sweep=hs.ForwardSweep())
        np.testing.assert_array_equal(P.sweep.op.D_DL1, D)
        np.testing.assert_array_equal(P.sweep.op.L.SLbuffer, [-1.0] * 4)
    A2 = A.tolil()
    A2[0, 7] = -1.0
    A2[1, 7] = -1.0
    A2 = A2.tocsc()
    fwd = [0, 1 / 3, 1.0, 2.0, 2.0, 3.5, 3.0, 5.0, 4.0]
    bwd = [1 / 6, 1 / 3, 1.75, 1.5, 3.25, 2.5, 4.75, 3.5, 4.0]
    for cache in (hs.PackedBufferCache(), hs.MatrixViewCache()):
        P = b(A2, 2, sweep=hs.ForwardSweep(), cache_strategy=cache)
        np.testing.assert_array_equal(P.sweep.op.D_DL1, [2, 3, 2, 2, 2, 2, 2, 2, 2])
        np.testing.assert_allclose(hs.backslash(P, x), fwd, rtol=1e-15)
        P = b(A2, 2, sweep=hs.BackwardSweep(), cache_strategy=cache)
        np.testing.assert_allclose(hs.backslash(P, x), bwd, rtol=1e-15)
    P = b(A, 3, sweep=hs.ForwardSweep())
    np.testing.assert_array_equal(P.sweep.op.L.SLbuffer, [-1, 0, -1, -1, 0, -1, -1, 0, -1])


@pytest.mark.parametrize("partsize", [1, 2, 3, 5, 10, 16, 17, 31, 32, 33, 63, 64, 65, 100, 128,
                                      129, 200, 256, 300])
@pytest.mark.parametrize("kind", ["sym_csc", "csr", "nonsym_csc"])
def test_random_parity(gpu, partsize, kind):
    import oracle as O
    hs = gpu
    n = 517  # ragged last partition for most partsizes
    if kind == "nonsym_csc":
        A = _random_nonsym(n, 0.02, partsize)
        is_sym = False
    else:
        A = _random_sym(n, 0.02, partsize)
        is_sym = True
    M = A.tocsr() if kind == "csr" else A
    x = np.random.default_rng(7 + partsize).uniform(-1, 1, n)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=7)
    for sw in SWEEPS:
        if kind == "nonsym_csc" and sw == "symmetric":
            pass  # allowed by the reference too
        for cache in ("packed", "view"):
            ref = O.l1gs_build(M, partsize, is_sym=is_sym, sweep=sw, cache=cache,
                               gpu=(64, 7), fast_nonsym=False)
            yo = O.l1gs_ldiv(ref, x)
            P = b(M, partsize, isSymA=is_sym, sweep=_hs_sweep(hs, sw),
                  cache_strategy=hs.PackedBufferCache() if cache == "packed" else hs.MatrixViewCache())
            assert (P.partitioning.partsize, P.partitioning.nparts, P.partitioning.nchunks,
                    P.partitioning.chunksize) == ref.part
            y = hs.backslash(P, x)
            D = P.sweep.lop.D_DL1 if sw == "symmetric" else P.sweep.op.D_DL1
            np.testing.assert_array_equal(D, ref.D)
            if sw == "forward":
                np.testing.assert_array_equal(y, yo)
            else:
                assert relerr(y, yo) < TOL, (sw, cache, relerr(y, yo))
            if cache == "packed":
                if sw in ("forward", "symmetric"):
                    L = P.sweep.op.L.SLbuffer if sw == "forward" else P.sweep.lop.L.SLbuffer
                    np.testing.assert_array_equal(L, ref.SL)
                if sw in ("backward", "symmetric"):
                    U = P.sweep.op.U.SUbuffer if sw == "backward" else P.sweep.uop.U.SUbuffer
                    np.testing.assert_array_equal(U, ref.SU)


def test_device_vectors_and_inplace(gpu):
    import torch
    import oracle as O
    hs = gpu
    n = 1000
    A = _random_sym(n, 0.01, 3)
    x = np.random.default_rng(1).uniform(-1, 1, n)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 20))
    yo = O.l1gs_ldiv(ref, x)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)(A, 64, isSymA=True)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    hs.ldiv_(yd, P, xd)
    assert relerr(yd.cpu().numpy(), yo) < TOL
    np.testing.assert_array_equal(xd.cpu().numpy(), x)  # x unchanged
    hs.ldiv_(xd, P, xd)  # aliasing y = x
    np.testing.assert_array_equal(xd.cpu().numpy(), yd.cpu().numpy())


def test_errors(gpu):
    from hybridsmoothers import problems as pb
    hs = gpu
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=4, blocks=4)
    A = pb.poisson_test_matrix(9)
    with pytest.raises(hs.HSError, match="partsize must be greater than 0"):
        b(A, 0)
    A3 = A.copy()
    A3.data[A3.indptr[4] + 1] = 0.0  # explicitly stored zero on the diagonal of column 5
    assert A3.indices[A3.indptr[4] + 1] == 4
    with pytest.raises(hs.SingularException) as ei:
        b(A3, 2)  # non-symmetric path builds DiagonalIndices
    assert ei.value.col == 5
    A4 = A.tolil()
    A4[6, 6] = 0.0
    A4 = A4.tocsc()
    A4.eliminate_zeros()
    with pytest.raises(hs.SingularException) as ei:
        b(A4, 2)  # missing diagonal entry
    assert ei.value.col == 7


def test_c2_poisson128_parity(gpu):
    """3D 7-point Poisson 128^3, P=64, symmetric sweep, packed FP64 -- full size."""
    import torch
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(128)
    n = A.shape[0]
    x = pb.uniform_rhs(n, 0)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 1024))
    yo = O.l1gs_ldiv(ref, x)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)(A, 64, isSymA=True)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    hs.ldiv_(yd, P, xd)
    y = yd.cpu().numpy()
    # closed form of D_Dl1 on this stencil: 6, or 8.5 on the rows whose x-neighbour
    # across the partition boundary counts twice (x in {63, 64}, interior y and z)
    D = P.sweep.lop.D_DL1
    np.testing.assert_array_equal(D, ref.D)
    assert set(np.unique(D)) == {6.0, 8.5}
    assert int((D == 8.5).sum()) == 2 * 126 * 126
    np.testing.assert_array_equal(y, yo)  # one in-partition neighbour per side: exact


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_p64_symmetric_kernel_variants(gpu, precision):
    """Every kernel variant of the partsize-64 symmetric sweep gives the same bits."""
    import oracle as O
    hs = gpu
    n = 64 * 37 + 19  # ragged last partition takes the generic path
    A = _random_sym(n, 0.004, 11)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 64))
    yo = O.l1gs_ldiv(ref, x)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(A, 64, isSymA=True,
                                                                  precision=precision)
    ys = []
    for v in (0, 1, 2):
        P.set_variant(v)
        ys.append(hs.backslash(P, x))
    np.testing.assert_array_equal(ys[0], ys[1])
    np.testing.assert_array_equal(ys[0], ys[2])
    tol = {"fp64": TOL, "fp32": 2e-6, "bf16": 2e-2}[precision]
    assert relerr(ys[0], yo) < tol
    if precision == "fp64":
        np.testing.assert_array_equal(P.sweep.lop.L.SLbuffer, ref.SL)
        np.testing.assert_array_equal(P.sweep.uop.U.SUbuffer, ref.SU)


@pytest.mark.parametrize("case", ["sym64_fp64", "sym32_fp32", "fwd16_bf16", "bwd_view_nonsym"])
def test_save_load_roundtrip(gpu, tmp_path, case):
    """A saved preconditioner reloads without the matrix and applies bit for
    bit as the original; getters agree; a corrupted payload is rejected."""
    from hybridsmoothers import problems as pb
    hs = gpu
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)
    if case == "sym64_fp64":
        A, kw = pb.poisson3d(12), dict(partsize=64, isSymA=True, sweep=hs.SymmetricSweep())
    elif case == "sym32_fp32":
        A, kw = pb.lognormal2d(30, seed=4), dict(partsize=32, isSymA=True, precision="fp32")
    elif case == "fwd16_bf16":
        A, kw = pb.lognormal3d(9, seed=5), dict(partsize=16, isSymA=True,
                                                sweep=hs.ForwardSweep(), precision="bf16")
    else:
        A, kw = pb.convdiff2d(24), dict(partsize=8, isSymA=False, sweep=hs.BackwardSweep(),
                                        cache_strategy=hs.MatrixViewCache())
    P = builder(A, kw.pop("partsize"), **kw)
    path = tmp_path / f"{case}.hsl1gs"
    P.save(path)
    Q = hs.load_preconditioner(path)
    x = pb.uniform_rhs(A.shape[0], 3)
    np.testing.assert_array_equal(hs.backslash(Q, x), hs.backslash(P, x))
    assert Q.partitioning.nparts == P.partitioning.nparts
    assert Q.precision == P.precision and Q.sweep_kind == P.sweep_kind
    np.testing.assert_array_equal(Q._get(0), P._get(0))
    if not isinstance(kw.get("cache_strategy"), hs.MatrixViewCache):
        what = 2 if isinstance(kw.get("sweep"), hs.BackwardSweep) else 1
        np.testing.assert_array_equal(Q._get(what), P._get(what))
    raw = bytearray(path.read_bytes())
    raw[-3] ^= 0x40
    path.write_bytes(bytes(raw))
    with pytest.raises(hs.HSError, match="checksum"):
        hs.load_preconditioner(path)


@pytest.mark.parametrize("kind", ["sym_csc", "csr", "nonsym_csc"])
@pytest.mark.parametrize("partsize,density", [(5, 0.05), (32, 0.02), (64, 0.2), (64, 0.004)])
def test_view_wave_equals_packed(gpu, kind, partsize, density):
    """MatrixViewCache at partsize <= 64 (one wave per partition, reading the
    matrix by rows) gives the packed cache's results bit for bit: the same
    terms in the same order, the packed zeros skipped.  Density 0.2 puts about
    13 in-partition entries in a row, past the kernel's 4-entry register cache."""
    hs = gpu
    n = 64 * 9 + 23
    A = _random_nonsym(n, density, partsize) if kind == "nonsym_csc" else \
        _random_sym(n, density, partsize)
    M = A.tocsr() if kind == "csr" else A
    is_sym = kind != "nonsym_csc"
    x = np.random.default_rng(partsize).uniform(-1, 1, n)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=5)
    for sw in SWEEPS:
        Pp = b(M, partsize, isSymA=is_sym, sweep=_hs_sweep(hs, sw))
        Pv = b(M, partsize, isSymA=is_sym, sweep=_hs_sweep(hs, sw),
               cache_strategy=hs.MatrixViewCache())
        np.testing.assert_array_equal(hs.backslash(Pv, x), hs.backslash(Pp, x))


def _banded_sym(n, b, seed, keep=0.8):
    """Symmetric band matrix of half-bandwidth b, each off-diagonal entry kept
    with probability `keep`: row i has at most 2 b entries inside its
    partition, and the counts vary from row to row."""
    rng = np.random.default_rng(seed)
    R, C, V = [], [], []
    for k in range(1, b + 1):
        i = np.arange(n - k)
        m = rng.random(n - k) < keep
        R.append(i[m] + k)
        C.append(i[m])
        V.append(rng.uniform(-1, 0, m.sum()))
    L = sp.csc_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(n, n))
    A = L + L.T
    A = A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0 + rng.random(n))
    A = A.tocsc()
    A.sort_indices()
    return A


@pytest.mark.parametrize("band", [1, 2, 3, 5, 10])
def test_view_ring_entry_slots(gpu, band):
    """MatrixViewCache at partsize 64 holds each lane's in-partition entries
    in 2, 4 or (8, 16, 32 by slice width) slots, the fewest that fit the
    wave: half-bandwidths 1, 2, 3, 5, 10 put at most 2, 4, 6, 10, 20 entries
    in a row's partition (slices 3, 5, 7, 11, 21 wide), and rows kept at 80 %
    mix counts within a wave.  Bit-equal to the packed cache, and to the
    oracle's MatrixViewCache (forward bit-equal, the others within 1e-12)."""
    import oracle as O
    hs = gpu
    n = 64 * 10 + 17  # ragged last partition: the general form
    A = _banded_sym(n, band, band)
    x = np.random.default_rng(band).uniform(-1, 1, n)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=7)
    print(f"[view ring slots] band={band}", flush=True)
    for sw in SWEEPS:
        Pp = b(A, 64, isSymA=True, sweep=_hs_sweep(hs, sw))
        Pv = b(A, 64, isSymA=True, sweep=_hs_sweep(hs, sw), cache_strategy=hs.MatrixViewCache())
        yv = hs.backslash(Pv, x)
        np.testing.assert_array_equal(yv, hs.backslash(Pp, x), err_msg=sw)
        ref = O.l1gs_build(A, 64, is_sym=True, sweep=sw, cache="view", gpu=(64, 7),
                           fast_nonsym=False)
        yo = O.l1gs_ldiv(ref, x)
        if sw == "forward":
            np.testing.assert_array_equal(yv, yo)
        else:
            assert relerr(yv, yo) < TOL, (sw, relerr(yv, yo))


@pytest.mark.parametrize("partsize", [64, 128])
def test_view_nonfinite_input(gpu, partsize):
    """Infinite and NaN inputs through the MatrixViewCache rings: a wave whose
    result is not finite is redone by rows with true divisions, so the output
    follows the reference's MatrixViewCache arithmetic -- only stored entries
    are added (no 0 * inf from the ring's empty slots) and an infinite
    numerator divides to +-inf -- and matches the oracle's view sweep: the same
    NaN and infinite entries, the finite ones within 1e-12."""
    import oracle as O
    hs = gpu
    n = 64 * 10 + 17
    A = _banded_sym(n, 2, 1)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    x[70], x[300], x[500] = np.inf, -np.inf, np.nan
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=7)
    print(f"[view nonfinite] P={partsize}", flush=True)
    for sw in SWEEPS:
        Pv = b(A, partsize, isSymA=True, sweep=_hs_sweep(hs, sw),
               cache_strategy=hs.MatrixViewCache())
        yv = hs.backslash(Pv, x)
        ref = O.l1gs_build(A, partsize, is_sym=True, sweep=sw, cache="view", gpu=(64, 7),
                           fast_nonsym=False)
        yo = O.l1gs_ldiv(ref, x)
        assert np.isinf(yo).any() and np.isnan(yo).any()
        np.testing.assert_array_equal(np.isnan(yv), np.isnan(yo), err_msg=sw)
        inf = np.isinf(yo)
        np.testing.assert_array_equal(yv[inf], yo[inf], err_msg=sw)
        fin = np.isfinite(yo)
        assert np.isfinite(yv[fin]).all(), sw
        assert relerr(yv[fin], yo[fin]) < TOL, (sw, relerr(yv[fin], yo[fin]))


@pytest.mark.parametrize("kind", ["sym_csc", "csr", "nonsym_csc"])
@pytest.mark.parametrize("partsize,density", [(65, 0.004), (100, 0.006), (128, 0.004),
                                              (128, 0.009), (128, 0.03), (160, 0.004),
                                              (192, 0.006), (200, 0.004), (256, 0.004),
                                              (256, 0.012)])
def test_view_wide_equals_packed(gpu, kind, partsize, density):
    """MatrixViewCache at partsize 65..256 (one wave per partition, R rows per
    lane): rows of at most 8 stored entries (16 at R = 2) take the LDS step
    ring fed from the sliced-ELL rows, wider rows the register-queue form; both
    give the packed cache's results bit for bit.  Partitions do not start at
    64-row slices for P = 65, 100, 160, 200; n = 599 leaves a ragged last
    partition; densities 0.009 and 0.012 put 9..16 entries in some rows, 0.03
    more than 16."""
    hs = gpu
    n = 64 * 9 + 23
    A = _random_nonsym(n, density, partsize) if kind == "nonsym_csc" else \
        _random_sym(n, density, partsize)
    M = A.tocsr() if kind == "csr" else A
    is_sym = kind != "nonsym_csc"
    x = np.random.default_rng(partsize).uniform(-1, 1, n)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=5)
    print(f"[view wide] {kind} P={partsize} density={density}", flush=True)
    for sw in SWEEPS:
        Pp = b(M, partsize, isSymA=is_sym, sweep=_hs_sweep(hs, sw))
        Pv = b(M, partsize, isSymA=is_sym, sweep=_hs_sweep(hs, sw),
               cache_strategy=hs.MatrixViewCache())
        np.testing.assert_array_equal(hs.backslash(Pv, x), hs.backslash(Pp, x), err_msg=sw)


@pytest.mark.parametrize("partsize", [96, 128, 192, 256])
def test_view_wide_stencil_equals_packed(gpu, partsize):
    """A 7-point lognormal grid (28 x 25 x 23: partitions cross grid lines and
    planes at every partsize) through the wide MatrixViewCache ring, on device
    vectors, against the packed cache bit for bit, in all three sweeps."""
    import torch
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.csr_to_scipy(*pb.lognormal_grid(28, 25, 23, seed=4)).tocsc()
    n = A.shape[0]
    xd = torch.tensor(np.random.default_rng(3).uniform(-1, 1, n), device="cuda")
    yp, yv = torch.empty_like(xd), torch.empty_like(xd)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)
    print(f"[view wide stencil] n={n} P={partsize}", flush=True)
    for sw in SWEEPS:
        Pp = b(A, partsize, isSymA=True, sweep=_hs_sweep(hs, sw))
        Pv = b(A, partsize, isSymA=True, sweep=_hs_sweep(hs, sw),
               cache_strategy=hs.MatrixViewCache())
        hs.ldiv_(yp, Pp, xd)
        hs.ldiv_(yv, Pv, xd)
        assert torch.equal(yp, yv), sw


def test_c3_convdiff2048_parity(gpu):
    """2048^2 upwind convection-diffusion (4.19 M rows, non-symmetric CSC),
    DiagonalIndices path, P=32 forward and backward -- full size: D_Dl1 and
    the forward sweep bit-exact with the oracle, the backward within 1e-12."""
    import torch
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.convdiff2d(2048)
    n = A.shape[0]
    x = pb.uniform_rhs(n, 1)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)
    for sw in ("forward", "backward"):
        ref = O.l1gs_build(A, 32, is_sym=False, sweep=sw, gpu=(256, 1024))
        P = b(A, 32, isSymA=False, sweep=_hs_sweep(hs, sw))
        op = P.sweep.op
        np.testing.assert_array_equal(op.D_DL1, ref.D)
        hs.ldiv_(yd, P, xd)
        y = yd.cpu().numpy()
        yo = O.l1gs_ldiv(ref, x)
        if sw == "forward":
            np.testing.assert_array_equal(y, yo)
        else:
            assert relerr(y, yo) < TOL


@pytest.mark.parametrize("partsize", [16, 32, 64, 128])
def test_unsafe_diagonal_true_division(gpu, partsize):
    """Rows whose l1 diagonal lies outside 2^+-400 (here ~1e130, the matrix
    scaled symmetrically by 1e65 on a few rows and columns) make their wave
    divide instead of multiplying by a reciprocal.  Every kernel that path
    reaches -- packed forward / backward / symmetric (all P = 64 variants), the
    one-triangle layout, the matrix view -- against the oracle element by
    element (tolerance relative to each entry, so the ~1e-130 entries count)."""
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.csr_to_scipy(*pb.lognormal_grid(20, 20, 20, seed=3)).tocsc()
    n = A.shape[0]
    s = np.ones(n)
    s[[5, 700, 4001, n - 1]] = 1e65
    S = sp.diags(s)
    A = (S @ A @ S).tocsc()
    A.sort_indices()
    x = np.random.default_rng(partsize).uniform(-1, 1, n)
    print(f"[unsafe diagonal] n={n} partsize={partsize}", flush=True)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)
    cases = [("symmetric", "packed", True), ("forward", "packed", False),
             ("backward", "packed", False), ("symmetric", "view", True)]
    if partsize <= 64:
        cases.append(("symmetric", "packed_sym", True))
    for sw, cache, sym in cases:
        ref = O.l1gs_build(A, partsize, is_sym=sym, sweep=sw,
                           cache="view" if cache == "view" else "packed", gpu=(256, 64))
        D = ref.D
        assert np.abs(D).max() > 2.0 ** 400
        yo = O.l1gs_ldiv(ref, x)
        cs = {"packed": hs.PackedBufferCache(), "view": hs.MatrixViewCache(),
              "packed_sym": hs.PackedSymmetricCache()}[cache]
        P = b(A, partsize, isSymA=sym, sweep=_hs_sweep(hs, sw), cache_strategy=cs)
        variants = (0, 1, 2) if (partsize == 64 and sw == "symmetric" and cache == "packed") else (None,)
        for v in variants:
            if v is not None:
                P.set_variant(v)
            y = hs.backslash(P, x)
            if sw == "forward":
                np.testing.assert_array_equal(y, yo)
            else:
                np.testing.assert_allclose(y, yo, rtol=1e-11, atol=0, err_msg=f"{sw} {cache} {v}")


@pytest.mark.parametrize("partsize", [1, 2, 5, 16, 31, 32, 33, 64, 65, 128, 129, 256, 300])
def test_nonsym_csr_parity(gpu, partsize):
    """SparseMatrixCSR input of a non-symmetric matrix (read by rows,
    l1_gauss_seidel.jl:1231-1242, 1351-1374, 1496-1519, 1588-1597): every
    sweep, packed and view caches, against the oracle on the same CSR, with
    isSymA either way (the CSR path ignores it).  The matrix is far from
    symmetric, so reading columns where rows are meant fails these checks
    (the oracle itself is pinned to the dense block inverse of A, not of A^T)."""
    import oracle as O
    hs = gpu
    n = 517
    A = _random_nonsym(n, 0.02, 1000 + partsize)
    assert abs(A - A.T).max() > 0.1
    R = A.tocsr()
    x = np.random.default_rng(11 + partsize).uniform(-1, 1, n)
    b = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=7)
    for sw in SWEEPS:
        for cache in ("packed", "view"):
            ref = O.l1gs_build(R, partsize, is_sym=False, sweep=sw, cache=cache, gpu=(64, 7))
            yo = O.l1gs_ldiv(ref, x)
            for is_sym in (False, True):
                P = b(R, partsize, isSymA=is_sym, sweep=_hs_sweep(hs, sw),
                      cache_strategy=hs.PackedBufferCache() if cache == "packed"
                      else hs.MatrixViewCache())
                y = hs.backslash(P, x)
                D = P.sweep.lop.D_DL1 if sw == "symmetric" else P.sweep.op.D_DL1
                np.testing.assert_array_equal(D, ref.D)
                if sw == "forward":
                    np.testing.assert_array_equal(y, yo)
                else:
                    assert relerr(y, yo) < TOL, (sw, cache, is_sym, relerr(y, yo))
                if cache == "packed":
                    if sw in ("forward", "symmetric"):
                        L = P.sweep.op.L.SLbuffer if sw == "forward" else P.sweep.lop.L.SLbuffer
                        np.testing.assert_array_equal(L, ref.SL)
                    if sw in ("backward", "symmetric"):
                        U = P.sweep.op.U.SUbuffer if sw == "backward" else P.sweep.uop.U.SUbuffer
                        np.testing.assert_array_equal(U, ref.SU)
