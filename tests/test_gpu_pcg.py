"""GPU PCG (through the C ABI) against the oracle's CG.

Parity bar: identical iteration counts to the FP64 oracle at rtol 1e-8, and
solutions within 1e-9 relative (the dot products are summed in a different
order, so iterates drift by rounding); the FP64 true relative residual is
reported by the device and checked against a host recomputation.
Mixed precision (FP32/BF16 smoothers inside FP64 CG) is held to the same
final true residual, with its iteration-count delta bounded.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _true_relres(A, b, x):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def test_spmv(gpu):
    from hybridsmoothers import problems as pb
    import oracle as O
    hs = gpu
    A = pb.lognormal3d(20, seed=1)
    M = hs.DeviceMatrix(A)
    x = np.random.default_rng(0).uniform(-1, 1, A.shape[0])
    y = np.empty_like(x)
    M.mul(y, x)
    np.testing.assert_array_equal(y, O.spmv(O.csr1(A), x))
    Mr = hs.DeviceMatrix(A.tocsr())
    Mr.mul(y, x)
    np.testing.assert_array_equal(y, O.spmv(O.csr1(A), x))


def test_unpreconditioned(gpu):
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(24)
    b = pb.uniform_rhs(A.shape[0], 0)
    xo, so = O.pcg(A, b, None, rtol=1e-8)
    sol = hs.solve(hs.LinearProblem(A, b), hs.KrylovJL_CG(), reltol=1e-8)
    assert sol.status == "solved"
    assert sol.iters == so.iters
    assert relerr(sol.u, xo) < 1e-9
    assert abs(sol.true_relres - _true_relres(A, b, sol.u)) < 1e-12


def test_c1_readme_quickstart(gpu):
    """README.md:17-30: 1D Laplacian N=2048, b = ones, L1GS P=16 symmetric, CG."""
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.laplace1d(2048)
    b = np.ones(2048)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=20)(
        A, 16, isSymA=True, sweep=hs.SymmetricSweep())
    ref = O.l1gs_build(A, 16, is_sym=True, sweep="symmetric", gpu=(256, 20))
    xo, so = O.pcg(A, b, ref, rtol=1e-8)
    sol = hs.solve(hs.LinearProblem(A, b), hs.KrylovJL_CG(), Pl=P, reltol=1e-8)
    assert sol.status == "solved"
    assert sol.iters == so.iters
    assert relerr(sol.u, xo) < 1e-9
    assert sol.true_relres < 1e-6


@pytest.mark.parametrize("prec", ["l1gs64", "l1gs32_fwd", "cheb4_3", "cheb4_2", "cheb4_5",
                                  "cheb1_2", "cheb4_1"])
def test_preconditioned_parity(gpu, prec):
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal3d(24, seed=7)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 3)
    if prec.startswith("l1gs"):
        psz = 64 if prec == "l1gs64" else 32
        sw = "symmetric"
        P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(A, psz, isSymA=True)
        ref = O.l1gs_build(A, psz, is_sym=True, sweep=sw, gpu=(256, 64))
    else:
        kind, deg = (4 if prec.startswith("cheb4") else 1), int(prec[-1])
        C = hs.ChebyshevFourth if kind == 4 else hs.ChebyshevFirst
        P = hs._build_smoother(C(degree=deg, backend=hs.ROCBackend()), A)
        ref = O.cheb_build(A, kind, deg)
        ref.rho = P.rho
    xo, so = O.pcg(A, b, ref, rtol=1e-8)
    for rep in range(2):  # second solve replays the cached graph
        sol = hs.solve(hs.LinearProblem(A, b), hs.KrylovJL_CG(), Pl=P, reltol=1e-8)
        assert sol.status == "solved"
        assert sol.iters == so.iters, (sol.iters, so.iters)
        assert relerr(sol.u, xo) < 1e-9
        assert abs(sol.true_relres - _true_relres(A, b, sol.u)) < 1e-12


def test_device_vectors_and_reuse(gpu):
    import torch
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(20)
    n = A.shape[0]
    M = hs.DeviceMatrix(A)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(A, 32, isSymA=True)
    sols = []
    for seed in (1, 2, 1):
        b = pb.uniform_rhs(n, seed)
        bd = torch.tensor(b, device="cuda")
        s = hs.pcg_solve(M, bd, Pl=P, reltol=1e-8)
        sh = hs.pcg_solve(M, b, Pl=P, reltol=1e-8)
        np.testing.assert_array_equal(s.u.cpu().numpy(), sh.u)
        sols.append(sh)
    np.testing.assert_array_equal(sols[0].u, sols[2].u)  # deterministic
    assert sols[0].iters == sols[2].iters


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_mixed_precision_smoother(gpu, precision):
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal3d(32, seed=2)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 2)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 64))
    _, so = O.pcg(A, b, ref, rtol=1e-8)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)(
        A, 64, isSymA=True, precision=precision)
    sol = hs.solve(hs.LinearProblem(A, b), hs.KrylovJL_CG(), Pl=P, reltol=1e-8)
    assert sol.status == "solved"
    # FP64 outer CG: the final true residual matches the all-FP64 solve's scale
    assert sol.true_relres < 1e-7
    assert abs(sol.iters - so.iters) <= max(3, so.iters // 10), (sol.iters, so.iters)
    C32 = hs._build_smoother(hs.ChebyshevFourth(degree=3, backend=hs.ROCBackend(),
                                                precision="fp32"), A)
    ref4 = O.cheb_build(A, 4, 3)
    _, s4 = O.pcg(A, b, ref4, rtol=1e-8)
    sol4 = hs.solve(hs.LinearProblem(A, b), hs.KrylovJL_CG(), Pl=C32, reltol=1e-8)
    assert sol4.status == "solved" and sol4.true_relres < 1e-7
    assert abs(sol4.iters - s4.iters) <= max(3, s4.iters // 10)


def test_c2_poisson128_l1gs_pcg(gpu):
    """128^3 Poisson, L1GS P=64 symmetric as Pl, CG to rtol 1e-8: iteration parity."""
    import torch
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(128)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 0)
    ref = O.l1gs_build(A, 64, is_sym=True, sweep="symmetric", gpu=(256, 1024))
    xo, so = O.pcg(A, b, ref, rtol=1e-8)
    P = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)(A, 64, isSymA=True)
    sol = hs.pcg_solve(A, torch.tensor(b, device="cuda"), Pl=P, reltol=1e-8)
    assert sol.status == "solved"
    assert sol.iters == so.iters
    assert relerr(sol.u.cpu().numpy(), xo) < 1e-8


def test_c4_lognormal192_mixed_precision(gpu):
    """192^3 lognormal-coefficient Poisson (7.08 M rows), FP64 CG to the config's
    true relative residual 1e-8 (Krylov.jl's rule at rtol 1e-8, continued on the
    device while ||b - A x|| / ||b|| > 1e-8) with an L1GS P=64 smoother in FP64,
    FP32 and BF16, both triangle layouts -- full size, properties only (the
    oracle's CG would take minutes): every solve meets 1e-8, FP64 solves with
    one and two triangles are bit-identical, and the reduced-precision solves
    need at most 1 % more iterations."""
    import torch
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal3d(192, seed=2)
    n = A.shape[0]
    M = hs.DeviceMatrix(A)
    b = torch.tensor(pb.uniform_rhs(n, 2), device="cuda")
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)
    sols = {}
    for cache in ("two", "one"):
        strat = hs.PackedBufferCache() if cache == "two" else hs.PackedSymmetricCache()
        for prec in ("fp64", "fp32", "bf16"):
            P = builder(A, 64, isSymA=True, precision=prec, cache_strategy=strat)
            sols[cache, prec] = hs.pcg_solve(M, b, Pl=P, reltol=1e-8, true_reltol=1e-8)
            print(f"[c4] {cache} {prec}: {sols[cache, prec].iters} iterations, true relres "
                  f"{sols[cache, prec].true_relres:.3e}", flush=True)
            del P
    base = sols["two", "fp64"]
    assert base.status == "solved" and base.true_relres <= 1e-8
    np.testing.assert_array_equal(sols["one", "fp64"].u.cpu().numpy(), base.u.cpu().numpy())
    for key, s in sols.items():
        assert s.status == "solved", key
        assert s.true_relres <= 1e-8, (key, s.true_relres)
        assert s.iters <= base.iters * 1.01 + 1, (key, s.iters, base.iters)


@pytest.mark.parametrize("prec", ["l1gs_fp32", "l1gs128", "cheb1_3"])
def test_true_residual_continuation(gpu, prec):
    """hs_pcg_solve_true: the continued solve meets its true-residual target,
    never stops earlier than Krylov.jl's rule, and with a target the rule's stop
    already meets it is the plain solve bit for bit.  Covers the fused iteration
    (L1GS P <= 64) and the unfused one (P = 128, Chebyshev 1st kind), whose
    p-update the continuation redoes."""
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.lognormal3d(32, seed=5)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 6)
    M = hs.DeviceMatrix(A)
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=64)
    if prec == "l1gs_fp32":
        P = builder(A, 64, isSymA=True, precision="fp32")
    elif prec == "l1gs128":
        P = builder(A, 128, isSymA=True)
    else:
        P = hs._build_smoother(hs.ChebyshevFirst(degree=3, backend=hs.ROCBackend()), A)
    plain = hs.pcg_solve(M, b, Pl=P, reltol=1e-6)
    loose = hs.pcg_solve(M, b, Pl=P, reltol=1e-6, true_reltol=1.0)
    assert loose.iters == plain.iters
    np.testing.assert_array_equal(loose.u, plain.u)
    target = plain.true_relres / 20
    cont = hs.pcg_solve(M, b, Pl=P, reltol=1e-6, true_reltol=target)
    print(f"[continuation {prec}] plain {plain.iters} it {plain.true_relres:.2e}; continued "
          f"{cont.iters} it {cont.true_relres:.2e} (target {target:.2e})", flush=True)
    assert cont.status == "solved"
    assert cont.true_relres <= target
    assert abs(cont.true_relres - _true_relres(A, b, cont.u)) < 1e-12
    assert cont.iters > plain.iters
    # the continuation resumes the recurrence, it does not restart it: the
    # result is bit for bit the iterate an uninterrupted solve (no stopping
    # test, rtol 0) reaches after the same number of iterations
    ref = hs.pcg_solve(M, b, Pl=P, reltol=0.0, maxiters=cont.iters)
    assert ref.iters == cont.iters and ref.status != "solved"
    np.testing.assert_array_equal(cont.u, ref.u)


def test_c2_poisson128_cheb4_pcg(gpu):
    """128^3 Poisson, Chebyshev 4th kind degree 3 as Pl (the 50-step power
    iteration on the device), CG to rtol 1e-8: the iteration count equals the
    oracle's PCG run with the device's rho (rho itself agrees with the oracle's
    power iteration to 1e-13), solutions within 1e-8."""
    import torch
    import oracle as O
    from hybridsmoothers import problems as pb
    hs = gpu
    A = pb.poisson3d(128)
    n = A.shape[0]
    b = pb.uniform_rhs(n, 0)
    C = hs._build_smoother(hs.ChebyshevFourth(degree=3, backend=hs.ROCBackend()), A)
    ref = O.cheb_build(A, 4, 3)
    assert abs(C.rho - ref.rho) <= 1e-13 * ref.rho
    ref.rho = C.rho
    xo, so = O.pcg(A, b, ref, rtol=1e-8)
    sol = hs.pcg_solve(A, torch.tensor(b, device="cuda"), Pl=C, reltol=1e-8)
    print(f"[c2 cheb4 d3] iterations {sol.iters} (oracle {so.iters})", flush=True)
    assert sol.status == "solved"
    assert sol.iters == so.iters
    assert relerr(sol.u.cpu().numpy(), xo) < 1e-8
