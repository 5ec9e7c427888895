"""Batched persistent PCG engine and the serving loop (hs_batch_* through the C ABI).

Parity bar:
* every system's solution, iteration count and true residual are identical
  (tolerance 0) to hs_pcg_solve with the same L1GS preconditioner -- the
  engine runs the same per-system arithmetic and reductions;
* with an FP64 smoother, identical iteration counts to the oracle's CG with
  the oracle's L1GS, solutions within 1e-9 relative;
* a system's result is the same (tolerance 0) alone and in any mix, in the
  engine and in the serving loop.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

# grids of uneven size: n = 2, 5, 153, 256 (one full tile), 1023, 210, 1728, 1600, 8000
SIZES = [(2, 1, 1), (5, 1, 1), (17, 9, 1), (16, 16, 1), (33, 31, 1), (7, 6, 5), (12, 12, 12),
         (40, 40, 1), (20, 20, 20)]


def _systems(seed0=3):
    from hybridsmoothers import problems as pb
    out = []
    for k, d in enumerate(SIZES):
        arrs = pb.lognormal_grid(*d, seed=seed0 + k)
        A = pb.csr_to_scipy(*arrs)
        b = pb.uniform_rhs(A.shape[0], 100 + seed0 + k)
        out.append((arrs, A, b))
    return out


@pytest.mark.parametrize("k", [3, 4, 6])
def test_first_pass_equals_single_apply(gpu, k):
    """One pass admits a system: r = b, x = 0, z = M^-1 b (bitwise the
    single-system FP32 apply), D_Dl1 (the FP64 rule rounded to FP32) and
    per-tile r'z partials."""
    from hybridsmoothers.batch import BatchSolver
    hs = gpu
    arrs, A, b = _systems()[k]
    n = A.shape[0]
    print(f"[batch first pass] n={n}", flush=True)
    Pl = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)(
        A, 32, isSymA=True, sweep=hs.SymmetricSweep(), precision="fp32")
    z_ref = hs.backslash(Pl, b)
    D_ref = Pl.sweep.lop.D_DL1.astype(np.float32).astype(np.float64)
    with BatchSolver(partsize=32, precision="fp32", passes=1, arena_bytes=64 << 20,
                     max_systems=4) as eng:
        sid = eng.submit(arrs, b)
        eng.step()
        got = {w: eng.peek(sid, w) for w in ("b", "D", "z", "r", "x", "partials")}
    np.testing.assert_array_equal(got["b"], b)
    bad_d = np.flatnonzero(got["D"] != D_ref)
    assert bad_d.size == 0, (bad_d[:10], got["D"][bad_d[:4]], D_ref[bad_d[:4]])
    bad_z = np.flatnonzero(got["z"] != z_ref)
    assert bad_z.size == 0, (bad_z.size, bad_z[:10], got["z"][bad_z[:4]], z_ref[bad_z[:4]])
    np.testing.assert_array_equal(got["r"], b)
    assert np.all(got["x"] == 0.0)
    nt = (n + 255) // 256
    tiles = np.array([b[256 * t:256 * (t + 1)] @ z_ref[256 * t:256 * (t + 1)] for t in range(nt)])
    np.testing.assert_allclose(got["partials"][0::2], tiles, rtol=1e-13)


@pytest.mark.parametrize("prec,P", [("fp32", 32), ("fp64", 16), ("bf16", 64), ("fp64", 64)])
def test_batch_equals_single_solve(gpu, prec, P):
    from hybridsmoothers.batch import BatchSolver
    hs = gpu
    systems = _systems()
    print(f"[batch] prec={prec} P={P} n={[A.shape[0] for _, A, _ in systems]}", flush=True)
    with BatchSolver(partsize=P, precision=prec, passes=5, arena_bytes=256 << 20,
                     max_systems=16) as eng:
        sols = eng.solve_all([(arrs, b) for arrs, _, b in systems], rtol=1e-8)
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)
    for (arrs, A, b), s in zip(systems, sols):
        Pl = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep(), precision=prec)
        if P == 64:
            Pl.set_variant(0)  # the engine's sweep: four partitions per 256-row workgroup
        ref = hs.pcg_solve(A, b, Pl=Pl, reltol=1e-8)
        assert s.status == "solved" and ref.status == "solved"
        assert s.iters == ref.iters, (A.shape[0], s.iters, ref.iters)
        np.testing.assert_array_equal(s.x, ref.u)
        assert s.true_relres == ref.true_relres
        assert s.rnorm0 == ref.resid0 and s.rnorm == ref.resid


def test_batch_fp64_vs_oracle(gpu):
    import oracle as O
    from hybridsmoothers.batch import BatchSolver
    systems = _systems(seed0=11)[2:]
    with BatchSolver(partsize=32, precision="fp64", passes=8, arena_bytes=256 << 20) as eng:
        sols = eng.solve_all([(arrs, b) for arrs, _, b in systems], rtol=1e-8)
    for (arrs, A, b), s in zip(systems, sols):
        ref = O.l1gs_build(A, 32, is_sym=True, sweep="symmetric")
        xo, so = O.pcg(A, b, ref, rtol=1e-8)
        assert s.iters == so.iters, (A.shape[0], s.iters, so.iters)
        assert relerr(s.x, xo) < 1e-9
        tr = np.linalg.norm(b - A @ s.x) / np.linalg.norm(b)
        assert abs(s.true_relres - tr) < 1e-12


def test_alone_vs_mix(gpu):
    """Small arena and few slots: systems are admitted and retired in waves,
    in a different company each time; results equal the solves alone."""
    from hybridsmoothers import autotune, problems as pb
    from hybridsmoothers.batch import BatchSolver
    specs = pb.serving_stream(24, seed=9, nmin=200, nmax=3e4)
    systems = autotune.sample_systems(specs)
    with BatchSolver(partsize=32, precision="fp32", passes=3, arena_bytes=48 << 20,
                     max_systems=6) as eng:
        mixed = eng.solve_all(systems, rtol=1e-8)
        assert eng.info()["steps"] > 10
        alone = [eng.solve_all([s], rtol=1e-8)[0] for s in systems]
    for m, a in zip(mixed, alone):
        assert m.status == "solved"
        assert m.iters == a.iters
        np.testing.assert_array_equal(m.x, a.x)
        assert m.true_relres == a.true_relres


def test_edge_cases(gpu):
    import hybridsmoothers._lib as L
    from hybridsmoothers import problems as pb
    from hybridsmoothers.batch import BatchSolver
    arrs = pb.lognormal_grid(30, 30, 1, seed=1)
    n = arrs[0].size - 1
    b = pb.uniform_rhs(n, 5)
    with BatchSolver(partsize=32, precision="fp32", passes=4, arena_bytes=64 << 20,
                     max_systems=1) as eng:
        # zero right-hand side: converged before the first iteration
        z = eng.solve_all([(arrs, np.zeros(n))])[0]
        assert z.iters == 0 and z.status == "solved" and z.true_relres == 0.0
        assert np.all(z.x == 0.0)
        # iteration limit
        s = eng.solve_all([(arrs, b)], itmax=3)[0]
        assert s.iters == 3 and s.status == "maxiters"
        # one slot: a second system waits until the first is collected
        i1 = eng.submit(arrs, b)
        assert i1 is not None
        assert eng.submit(arrs, b) is None
        got = []
        while eng.pending():
            eng.step()
            got += eng.collect()
        assert [g.id for g in got] == [i1] and got[0].status == "solved"
        assert eng.submit(arrs, b) is not None
        while eng.pending():
            eng.step()
            eng.collect()
        # bad input leaves the engine usable
        bad = (arrs[0], arrs[1].copy(), arrs[2])
        bad[1][3] = n + 5
        with pytest.raises(L.HSError):
            eng.submit(bad, b)
        assert eng.info()["free_slots"] == 1
        assert eng.solve_all([(arrs, b)])[0].status == "solved"
    # a system larger than the arena is refused outright
    with BatchSolver(partsize=32, precision="fp32", arena_bytes=1 << 20, max_systems=2) as eng:
        big = pb.lognormal_grid(128, 128, 1, seed=2)  # about 4 MB resident
        assert eng.footprint(big) > (1 << 20)
        with pytest.raises(L.HSError):
            eng.submit(big, np.ones(128 * 128))


def test_serve_stream_alone_vs_mix(gpu):
    from hybridsmoothers import autotune, problems as pb, serve
    from hybridsmoothers.batch import BatchSolver
    N = 32
    specs = pb.serving_stream(N, seed=21, nmin=1e3, nmax=4e4)
    arrivals = serve.poisson_arrivals(N, rate=400.0, seed=3)
    with BatchSolver(partsize=32, precision="fp32", passes=4, arena_bytes=128 << 20,
                     max_systems=8) as eng:
        rep, rec = serve.serve(eng, specs, arrivals, rtol=1e-8, keep_x=True)
        print(f"[serve] {rep.as_dict()}", flush=True)
        assert rep.count == N and rep.all_solved
        assert np.all(np.isfinite(rec["latency_s"])) and np.all(rec["latency_s"] >= 0)
        assert rep.latency_p50_ms <= rep.latency_p99_ms <= rep.latency_max_ms
        # latency splits into admission wait and service, both non-negative
        assert np.all(rec["wait_s"] >= 0) and np.all(rec["service_s"] > 0)
        np.testing.assert_allclose(rec["wait_s"] + rec["service_s"], rec["latency_s"],
                                   rtol=0, atol=1e-9)
        assert rep.wait_p99_ms <= rep.latency_max_ms and rep.service_p99_ms <= rep.latency_max_ms
        assert rep.max_true_relres < 1e-6
        systems = autotune.sample_systems(specs)  # same right-hand sides as the stream
        for i in range(0, N, 3):
            alone = eng.solve_all([systems[i]], rtol=1e-8)[0]
            assert alone.iters == rec["iters"][i]
            np.testing.assert_array_equal(alone.x, rec["x"][i])


def test_autotune(gpu):
    import torch
    from hybridsmoothers import autotune, problems as pb
    hs = gpu
    systems = autotune.sample_systems(pb.serving_stream(8, seed=4, nmin=500, nmax=5e3))
    res = autotune.tune_batch(systems, partsizes=(16, 32), passes=(2, 4),
                              arena_bytes=64 << 20)
    assert res["partsize"] in (16, 32) and res["passes"] in (2, 4)
    assert len(res["table"]) == 4 and res["solves_per_s"] > 0
    A = pb.poisson3d(16)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(A, 64, isSymA=True)
    x = torch.tensor(pb.uniform_rhs(A.shape[0], 1), device="cuda")
    times = autotune.tune_l1gs_variant(P, x, reps=5)
    assert set(times) == {0, 1, 2}


@pytest.mark.parametrize("smoother", ["l1gs", "cheb"])
def test_batch_column_formats(gpu, smoother):
    """A system whose columns all lie within 32,767 of their rows (uploaded as
    16-bit offsets) and one with a coupling 39,999 apart (32-bit columns),
    solved together: each equals the single-system solve bit for bit."""
    import scipy.sparse as sp
    from hybridsmoothers.batch import BatchSolver
    hs = gpu
    n = 40000
    rng = np.random.default_rng(8)
    systems = []
    for far in (False, True):
        A = sp.diags([rng.uniform(-1, 1, n - k) for k in (1, 300, 20000)], [1, 300, 20000],
                     shape=(n, n), format="lil")
        if far:
            A[0, n - 1] = 0.5
        A = (A + A.T + sp.identity(n) * 8.0).tocsc()
        A.sort_indices()
        systems.append((A, rng.uniform(-1, 1, n)))
    kw = (dict(smoother="cheb", degree=2) if smoother == "cheb" else dict(partsize=32))
    with BatchSolver(precision="fp64", passes=8, arena_bytes=256 << 20, max_systems=4,
                     **kw) as eng:
        sols = eng.solve_all(systems, rtol=1e-8)
    for (A, b), s in zip(systems, sols):
        if smoother == "cheb":
            Pl = hs._build_smoother(hs.ChebyshevFourth(degree=2, backend=hs.ROCBackend()), A)
        else:
            Pl = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)(
                A, 32, isSymA=True, sweep=hs.SymmetricSweep(), precision="fp64")
        ref = hs.pcg_solve(A, b, Pl=Pl, reltol=1e-8)
        assert s.status == "solved" and s.iters == ref.iters
        np.testing.assert_array_equal(s.x, ref.u)
        assert s.true_relres == ref.true_relres
