"""CPU-side checks of the C ABI and the host API (no GPU compute calls).

* libhsmooth.so loads and exports every function include/hsmooth.h declares;
* argument validation that happens before any device work returns the
  reference's errors (partsize, threads, blocks; unknown enums);
* the Python mirror rejects what the reference rejects, and reports a
  non-functional backend when no GPU is visible;
* the problem generators produce the documented matrices.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hsmooth.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hs_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    from hybridsmoothers import _lib as L
    names = _declared()
    assert len(names) >= 20
    missing = [s for s in names if not hasattr(L.lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\b(hs_[a-z0-9_]+)\b", out))
    assert set(names) <= exported
    assert set(L.EXPORTED) <= set(names)


def test_library_is_gfx950_only():
    """Every code object bundled in libhsmooth.so targets gfx950."""
    from hybridsmoothers import _lib as L
    data = open(L.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}


def _build_args(n=9, partsize=2, threads=4, blocks=4, sweep=2, cache=0, prec=0, fmt=0, sym=1):
    from hybridsmoothers import problems as pb
    A = pb.poisson_test_matrix(n)
    ptr = A.indptr.astype(np.int64) + 1
    ind = A.indices.astype(np.int64) + 1
    val = A.data.astype(np.float64)
    return (n, ptr, ind, val, fmt, sym, partsize, 1.5, sweep, cache, prec, threads, blocks)


def _call_build(args):
    from hybridsmoothers import _lib as L
    n, ptr, ind, val, fmt, sym, P, eta, sw, cache, prec, th, bl = args
    h = C.c_void_p()
    st = L.lib.hs_l1gs_build(n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data, fmt, sym, P,
                             eta, sw, cache, prec, th, bl, None, C.byref(h))
    return st, L.lib.hs_last_error().decode()


@pytest.mark.parametrize("kw,msg", [
    (dict(partsize=0), "partsize must be greater than 0"),
    (dict(threads=0), "`threads` must be greater than 0"),
    (dict(blocks=-1), "`blocks` must be greater than 0"),
    (dict(sweep=7), "unknown sweep"),
    (dict(cache=3), "unknown cache strategy"),
    (dict(prec=9), "unknown precision"),
    (dict(fmt=5), "unknown matrix format"),
])
def test_build_argument_errors(kw, msg):
    from hybridsmoothers import _lib as L
    st, err = _call_build(_build_args(**kw))
    assert st == L.HS_EINVAL
    assert msg in err


def test_matrix_view_precision_unsupported():
    from hybridsmoothers import _lib as L
    st, err = _call_build(_build_args(cache=1, prec=1))
    assert st == L.HS_EUNSUPPORTED


@pytest.mark.parametrize("kw", [dict(sweep=0), dict(sweep=1), dict(sym=0), dict(partsize=65),
                                dict(sym=0, fmt=1)])
def test_one_triangle_cache_limits(kw):
    """The one-triangle packed cache (U = L^T) needs isSymA, SymmetricSweep and
    partsize <= 64; anything else is refused before any device work."""
    from hybridsmoothers import _lib as L
    st, err = _call_build(_build_args(cache=L.CACHE_PACKED_SYM, n=130, **kw))
    assert st == L.HS_EUNSUPPORTED
    assert "one-triangle" in err


def test_null_handles():
    from hybridsmoothers import _lib as L
    assert L.lib.hs_l1gs_apply(None, None, None, 0, None, 0) == L.HS_EINVAL
    assert L.lib.hs_l1gs_info(None, None) == L.HS_EINVAL
    assert L.lib.hs_cheb_smooth(None, None, None, 0, None, 0) == L.HS_EINVAL
    assert L.lib.hs_csr_spmv(None, None, None, 0, None, 0) == L.HS_EINVAL
    assert L.lib.hs_l1gs_destroy(None) == L.HS_OK
    h = C.c_void_p()
    assert L.lib.hs_cheb_build(5, None, None, None, 3, 2, 50, 0, None, C.byref(h)) == L.HS_EINVAL
    assert "kind must be 1" in L.lib.hs_last_error().decode()


def _has_gpu():
    from hybridsmoothers import l1gs
    return l1gs.functional(l1gs.ROCBackend())


def test_configs_and_backends():
    import hybridsmoothers as hs
    with pytest.raises(hs.HSError, match="`threads` must be greater than 0"):
        hs.GPUConfig(0, 4)
    with pytest.raises(hs.HSError, match="`blocks` must be greater than 0"):
        hs.GPUConfig(4, 0)
    with pytest.raises(hs.HSError, match="`chunks` must be greater than 0"):
        hs.CPUConfig(0)
    with pytest.raises(hs.HSError, match="KA.CPU"):
        hs.L1GSPrecBuilder(hs.CPU(), chunks=4)
    with pytest.raises(TypeError):
        hs.L1GSPrecBuilder(hs.ROCBackend(), threads=4)
    if not _has_gpu():
        with pytest.raises(hs.HSError, match="not functional"):
            hs.L1GSPrecBuilder(hs.ROCBackend(), threads=4, blocks=4)
        with pytest.raises(hs.HSError, match="not functional"):
            hs._build_smoother(hs.ChebyshevFourth(degree=2), sp.identity(4, format="csc"))


def test_vector_checks():
    from hybridsmoothers import _lib as L
    with pytest.raises(L.HSError):
        L.vec_ptr(np.zeros(5, np.float32), 5)
    with pytest.raises(L.HSError):
        L.vec_ptr(np.zeros(4), 5)
    p, loc = L.vec_ptr(np.zeros(5), 5)
    assert loc == L.LOC_HOST and p


def test_sparse_arrays_one_based():
    from hybridsmoothers.l1gs import sparse_arrays
    from hybridsmoothers import problems as pb
    A = pb.poisson_test_matrix(6)
    n, ptr, ind, val, fmt = sparse_arrays(A)
    assert n == 6 and fmt == 0 and ptr.dtype == np.int64 and ptr[0] == 1 and ind.min() == 1
    n, ptr, ind, val, fmt = sparse_arrays(A.tocsr())
    assert fmt == 1
    with pytest.raises(Exception):
        sparse_arrays(sp.csc_matrix(np.ones((3, 4))))


def test_problem_generators():
    from hybridsmoothers import problems as pb
    A = pb.poisson3d(8)
    assert A.shape == (512, 512) and (A != A.T).nnz == 0
    K = sp.kron
    T = pb.poisson_test_matrix(8)
    I = sp.identity(8)
    ref = (K(I, K(I, T)) + K(I, K(T, I)) + K(T, K(I, I))).tocsc()
    assert abs(A - ref).max() == 0
    # 128^3: 7 n^3 - 6 n^2 stored entries
    n = 128
    assert 7 * n ** 3 - 6 * n ** 2 == 14_581_760
    C3 = pb.convdiff2d(16)
    D = C3.diagonal()
    assert np.all(D == 19.0)
    row = C3.tocsr()[16 * 5 + 5].toarray().ravel()
    assert row[16 * 5 + 4] == -11 and row[16 * 5 + 6] == -1
    assert row[16 * 4 + 5] == -6 and row[16 * 6 + 5] == -1
    # 2048^2 nnz: 5 n^2 - 4 n
    assert 5 * 2048 ** 2 - 4 * 2048 == 20_963_328
    L4 = pb.lognormal3d(6, seed=2)
    assert (abs(L4 - L4.T) > 0).nnz == 0
    assert np.all(np.linalg.eigvalsh(L4.toarray()) > 0)
    L2 = pb.lognormal2d(9, seed=1)
    assert (abs(L2 - L2.T) > 0).nnz == 0
    specs = pb.serving_stream(64, seed=5)
    sizes = [g * g if k == "2d" else g ** 3 for k, g, _ in specs]
    assert min(sizes) >= 700 and max(sizes) <= 3.0e5


def _fnv1a(data, h=0xcbf29ce484222325):
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


_HDR = "<8s6q8i7q"  # the saved preconditioner's header before its checksum


def _saved_file(path, n=100, P=10, nparts=10, nchunks=1, chunksize=10, S=None, sweep=2, cache=0,
                prec=0, fmt=0, sym=1, rowwise=1, variant=2, nbytes=None, payload=None,
                checksum=None):
    """A file in the layout hs_l1gs_save writes (header, FNV-1a checksum of the
    header and payload), with every field settable, for the loader's checks."""
    import struct
    if S is None:
        last = n - (nparts - 1) * P
        S = P * (P - 1) * (nparts - 1) // 2 + last * (last - 1) // 2
    if nbytes is None:
        tri = 2 * S if sweep == 2 and cache == 0 else S
        nbytes = [n * 8, 0, tri * 8, 0, 0, 0, 0]
    if payload is None:
        payload = bytes(sum(nbytes))
    head = struct.pack(_HDR, b"HSL1GS02", n, P, nparts, nchunks, chunksize, S, sweep, cache,
                       prec, fmt, sym, rowwise, variant, 0, *nbytes)
    ck = _fnv1a(payload, _fnv1a(head)) if checksum is None else checksum
    path.write_bytes(head + struct.pack("<Q", ck) + payload)
    return path


def test_l1gs_load_rejects_bad_files(tmp_path):
    """hs_l1gs_load checks the file before any device work: missing file,
    foreign magic, short header, a header that disagrees with the partitioning
    hs_l1gs_build derives from (n, partsize) -- forged packed size S, row-wise
    flag, partition count, chunking, array sizes -- and a checksum that does
    not cover the header are all HS_EINVAL."""
    import hybridsmoothers._lib as L
    h = C.c_void_p()

    def load(path):
        h.value = None
        st = L.lib.hs_l1gs_load(str(path).encode(), None, C.byref(h))
        return st, L.lib.hs_last_error().decode()

    assert load(tmp_path / "missing.bin")[0] == L.HS_EINVAL
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTL1GS0" + bytes(200))
    st, err = load(bad)
    assert st == L.HS_EINVAL and "magic" in err
    old = tmp_path / "old.bin"  # the previous format, whose checksum skipped the header
    old.write_bytes(b"HSL1GS01" + bytes(200))
    st, err = load(old)
    assert st == L.HS_EINVAL and "magic" in err
    short = tmp_path / "short.bin"
    short.write_bytes(b"HSL1GS02" + bytes(10))
    assert load(short)[0] == L.HS_EINVAL
    # well-formed apart from one field (checksums recomputed, so only the
    # header check can refuse them)
    forged = {
        # S too small for (n, P), with the triangle bytes to match: the packed
        # sweeps would index past the buffer
        "packed size": dict(S=300, nbytes=[800, 0, 2 * 300 * 8, 0, 0, 0, 0]),
        # a CSC without isSymA read as rows: skips the diagonal-position checks
        "row-wise flag": dict(sym=0, rowwise=1),
        "row-wise flag ": dict(fmt=1, sym=0, rowwise=0),
        "partition count": dict(nparts=9),
        "chunking": dict(nchunks=2, chunksize=6),
        "array sizes": dict(nbytes=[800, 0, 2 * 450 * 8, 8, 0, 0, 0]),
        "rows or partsize out of range": dict(n=1 << 41, nparts=(1 << 41) // 10,
                                              nbytes=[0] * 7),
        "unknown matrix format": dict(fmt=3),
        "one-triangle cache": dict(cache=2, sym=0, rowwise=0),
        "matrix sizes": dict(cache=1, nbytes=[800, 0, 0, 101 * 8, 24, 0, 16]),
    }
    for what, kw in forged.items():
        st, err = load(_saved_file(tmp_path / "forged.bin", **kw))
        assert st == L.HS_EINVAL, what
        assert what.strip() in err, (what, err)
        assert not h.value
    # a consistent header with a wrong checksum
    g = _saved_file(tmp_path / "ck2.bin", checksum=12345)
    st, err = load(g)
    assert st == L.HS_EINVAL and "checksum" in err
    # a header field changed after saving (variant 2 -> 1, both valid): only
    # the header's share of the checksum catches it
    raw = bytearray(_saved_file(tmp_path / "ck3.bin").read_bytes())
    raw[8 + 6 * 8 + 6 * 4] = 1
    (tmp_path / "ck3.bin").write_bytes(bytes(raw))
    st, err = load(tmp_path / "ck3.bin")
    assert st == L.HS_EINVAL and "checksum" in err
