"""The reference's "Solution with LinearSolve" cases (test/test_preconditioners.jl:366-531,
test/gpu/gpu.jl:360-400) with the GPU preconditioner.

For every case (convection_diffusion_matrix(60; eps=0.1), poisson2d_matrix(100; eps=0.1),
Symmetric(poisson2d_matrix(120; eps=0.1)) through the builder's Symmetric method;
partsize 10; both caches; each sweep): D_DL1 equals the oracle's, one application
equals the oracle's (forward bit for bit, backward and symmetric within 1e-12), and
the solve the reference runs -- GMRES (full, left-preconditioned) or CG -- meets the
reference's assertions (A x ~ b, no more iterations than unpreconditioned) with
the oracle-preconditioned solve's iteration count (GMRES within one iteration: the
backward and symmetric applications may differ from the oracle's in the last bit;
CG exactly, the device PCG against the oracle's).
"""
import numpy as np
import pytest

from conftest import relerr
from linearsolve_cases import PARTSIZE, cases, check_reference_properties, gmres

pytestmark = pytest.mark.gpu

_UNPREC = {}


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_linearsolve_case_gpu(gpu, case):
    import oracle as O
    hs = gpu
    name, A, is_sym, sw, cache, solver, wrapped = case
    print(f"[linearsolve] {name}", flush=True)
    n = A.shape[0]
    b = np.ones(n)
    sweep = {"forward": hs.ForwardSweep(), "backward": hs.BackwardSweep(),
             "symmetric": hs.SymmetricSweep()}[sw]
    cs = hs.PackedBufferCache() if cache == "packed" else hs.MatrixViewCache()
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=64, blocks=20)
    if wrapped:
        P = builder(hs.Symmetric(A), PARTSIZE, sweep=sweep, cache_strategy=cs)
    else:
        P = builder(A, PARTSIZE, isSymA=is_sym, sweep=sweep, cache_strategy=cs)
    ref = O.l1gs_build(A, PARTSIZE, is_sym=is_sym, sweep=sw, cache=cache, gpu=(64, 20))
    D = P.sweep.lop.D_DL1 if sw == "symmetric" else P.sweep.op.D_DL1
    np.testing.assert_array_equal(D, ref.D)
    v = np.random.default_rng(5).uniform(-1, 1, n)
    y, yo = hs.backslash(P, v), O.l1gs_ldiv(ref, v)
    if sw == "forward":
        np.testing.assert_array_equal(y, yo)
    else:
        assert relerr(y, yo) < 1e-12
    if solver == "cg":
        M = hs.DeviceMatrix(A)
        sol = hs.solve(hs.LinearProblem(M, b), hs.KrylovJL_CG(), Pl=P, reltol=1e-8)
        sol0 = hs.solve(hs.LinearProblem(M, b), hs.KrylovJL_CG(), reltol=1e-8)
        pp = O.l1gs_build(A, PARTSIZE, is_sym=is_sym, sweep=sw, cache="packed", gpu=(64, 20))
        _, so = O.pcg(A, b, pp, rtol=1e-8)
        assert sol.status == "solved" and sol.iters == so.iters, (sol.iters, so.iters)
        check_reference_properties(A, b, np.asarray(sol.u), sol.iters, sol0.iters)
        assert sol.true_relres < 1e-7
    else:
        x, it = gmres(A, b, lambda r: hs.backslash(P, r))
        _, ito = gmres(A, b, lambda r: O.l1gs_ldiv(ref, r))
        key = name.split()[0]
        if key not in _UNPREC:
            _UNPREC[key] = gmres(A, b)[1]
        assert abs(it - ito) <= 1, (it, ito)
        check_reference_properties(A, b, x, it, _UNPREC[key])
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-7
