"""Batched persistent PCG with the Chebyshev (4th kind) smoother as the
preconditioner (hs_batch_* with smoother = HS_PCG_CHEB, through the C ABI).

Parity bar:
* a system's l1inv and rho (power iteration run inside the engine's passes)
  equal the single-system hs_cheb_build's bit for bit;
* every system's solution, iteration count, true residual and residual norms
  are identical (tolerance 0) to hs_pcg_solve with that smoother -- the engine
  runs the same per-row arithmetic and the same 256-row reductions;
* with an FP64 smoother, identical iteration counts to the oracle's CG with
  the oracle's Chebyshev recurrence (at the device's rho), solutions within
  1e-9 relative;
* a system's result is the same (tolerance 0) alone and in any mix.
"""
import numpy as np
import pytest

from conftest import relerr
from test_gpu_batch import _systems

pytestmark = pytest.mark.gpu


def _smoother(hs, A, degree, prec, power_iters=50):
    cfg = hs.ChebyshevFourth(degree=degree, backend=hs.ROCBackend(), precision=prec)
    return hs._build_smoother(cfg, A, power_iters=power_iters)


@pytest.mark.parametrize("power_iters", [50, 3, 0])
def test_power_iteration_equals_single_build(gpu, power_iters):
    """After exactly power_iters passes the system's rho is final: it equals
    hs_cheb_build's, as does l1inv."""
    from hybridsmoothers.batch import BatchSolver
    hs = gpu
    arrs, A, b = _systems()[6]  # 12^3, n = 1728
    print(f"[batch cheb power] n={A.shape[0]} power_iters={power_iters}", flush=True)
    S = _smoother(hs, A, 2, "fp64", power_iters)
    with BatchSolver(smoother="cheb", degree=2, precision="fp64", power_iters=power_iters,
                     passes=max(power_iters, 1), arena_bytes=64 << 20, max_systems=2) as eng:
        sid = eng.submit(arrs, b)
        eng.step()
        rho = eng.peek(sid, "rho")[0]
        l1 = eng.peek(sid, "l1inv")
        assert eng.info()["triangles"] == 0 and eng.info()["grid_sweep"] == 0
    assert rho == S.rho, (rho, S.rho)
    np.testing.assert_array_equal(l1, S.l1inv)


@pytest.mark.parametrize("prec,degree", [("fp64", 2), ("fp32", 2), ("fp64", 3), ("fp32", 4)])
def test_batch_cheb_equals_single_solve(gpu, prec, degree):
    from hybridsmoothers.batch import BatchSolver
    hs = gpu
    systems = _systems()
    print(f"[batch cheb] prec={prec} degree={degree} n={[A.shape[0] for _, A, _ in systems]}",
          flush=True)
    with BatchSolver(smoother="cheb", degree=degree, precision=prec, passes=7,
                     arena_bytes=256 << 20, max_systems=16) as eng:
        sols = eng.solve_all([(arrs, b) for arrs, _, b in systems], rtol=1e-8)
    for (arrs, A, b), s in zip(systems, sols):
        ref = hs.pcg_solve(A, b, Pl=_smoother(hs, A, degree, prec), reltol=1e-8)
        assert s.status == "solved" and ref.status == "solved"
        assert s.iters == ref.iters, (A.shape[0], s.iters, ref.iters)
        np.testing.assert_array_equal(s.x, ref.u)
        assert s.true_relres == ref.true_relres
        assert s.rnorm0 == ref.resid0 and s.rnorm == ref.resid


def test_batch_cheb_fp64_vs_oracle(gpu):
    import oracle as O
    from hybridsmoothers.batch import BatchSolver
    hs = gpu
    systems = _systems(seed0=11)[2:]
    with BatchSolver(smoother="cheb", degree=2, precision="fp64", passes=8,
                     arena_bytes=256 << 20) as eng:
        sols = eng.solve_all([(arrs, b) for arrs, _, b in systems], rtol=1e-8)
    for (arrs, A, b), s in zip(systems, sols):
        ref = O.cheb_build(A, 4, 2)
        ref.rho = _smoother(hs, A, 2, "fp64").rho  # the device's power iteration
        xo, so = O.pcg(A, b, ref, rtol=1e-8)
        assert s.iters == so.iters, (A.shape[0], s.iters, so.iters)
        assert relerr(s.x, xo) < 1e-9
        tr = np.linalg.norm(b - A @ s.x) / np.linalg.norm(b)
        assert abs(s.true_relres - tr) < 1e-12


def test_batch_cheb_alone_vs_mix(gpu):
    """Few slots and a small arena: systems enter and leave in waves (some
    still in their power iteration while others iterate); each result equals
    the system solved alone."""
    from hybridsmoothers import autotune, problems as pb
    from hybridsmoothers.batch import BatchSolver
    specs = pb.serving_stream(20, seed=13, nmin=200, nmax=3e4)
    systems = autotune.sample_systems(specs)
    with BatchSolver(smoother="cheb", degree=2, precision="fp32", passes=16,
                     arena_bytes=48 << 20, max_systems=5) as eng:
        mixed = eng.solve_all(systems, rtol=1e-8)
        assert eng.info()["steps"] > 10
        alone = [eng.solve_all([s], rtol=1e-8)[0] for s in systems]
    for m, a in zip(mixed, alone):
        assert m.status == "solved"
        assert m.iters == a.iters
        np.testing.assert_array_equal(m.x, a.x)
        assert m.true_relres == a.true_relres


def test_batch_cheb_edge_cases(gpu):
    import hybridsmoothers._lib as L
    from hybridsmoothers import problems as pb
    from hybridsmoothers.batch import BatchSolver
    arrs = pb.lognormal_grid(30, 30, 1, seed=1)
    n = arrs[0].size - 1
    b = pb.uniform_rhs(n, 5)
    with BatchSolver(smoother="cheb", degree=3, precision="fp32", passes=4,
                     arena_bytes=64 << 20, max_systems=2) as eng:
        z = eng.solve_all([(arrs, np.zeros(n))])[0]
        assert z.iters == 0 and z.status == "solved" and z.true_relres == 0.0
        assert np.all(z.x == 0.0)
        s = eng.solve_all([(arrs, b)], itmax=3)[0]
        assert s.iters == 3 and s.status == "maxiters"
        with pytest.raises(L.HSError):  # an L1GS quantity
            sid = eng.submit(arrs, b)
            eng.peek(sid, "D")
        while eng.pending():
            eng.step()
            eng.collect()
