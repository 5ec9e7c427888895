"""The reference's "Solution with LinearSolve" cases (test/test_preconditioners.jl:366-531),
shared by the oracle tests and the GPU tests.

Each case is (name, matrix, isSymA, sweep, cache, solver, wrapped): the matrices of
test/test_matrices.jl, b = ones, partsize 10.  `wrapped` marks the
`Symmetric(poisson2d_matrix(120; eps=0.1))` cases, built through the
builder's Symmetric method (:305-313).  The reference's assertions: the
preconditioned solve reaches A x ~ b (rtol 1e-1, atol 1e-1) and takes no more
iterations than the unpreconditioned one (:102-103).  Krylov.jl's GMRES is
restated by full left-preconditioned GMRES (`gmres` below) with
M^-1 = ldiv!(., P, .); CG by the device PCG and the oracle's CG.
"""
import numpy as np

PARTSIZE = 10


def cases():
    from hybridsmoothers import problems as pb
    cd = pb.convection_diffusion_matrix(60, eps=0.1)
    p2 = pb.poisson2d_matrix(100, eps=0.1)
    p3 = pb.poisson2d_matrix(120, eps=0.1)
    out = []
    for sw in ("forward", "backward"):
        for cache in ("packed", "view"):
            out.append((f"convdiff60 {sw} {cache}", cd, False, sw, cache, "gmres", False))
    for sw in ("forward", "backward", "symmetric"):
        for cache in ("packed", "view"):
            solver = "cg" if sw == "symmetric" else "gmres"
            out.append((f"poisson2d100 {sw} {cache}", p2, True, sw, cache, solver, False))
    for sw in ("forward", "backward", "symmetric"):
        for cache in ("packed", "view"):
            out.append((f"Symmetric(poisson2d120) {sw} {cache}", p3, True, sw, cache, "gmres",
                        True))
    return out


def gmres(A, b, apply=None, rtol=1e-8, maxiter=2000):
    """Full (unrestarted) GMRES from x0 = 0 with left preconditioner
    M^-1 = apply, stopping when ||M^-1 (b - A x)|| <= rtol ||M^-1 b|| -- the
    textbook method (Saad, Iterative Methods, alg. 6.9: Arnoldi with modified
    Gram-Schmidt, Givens rotations) as Krylov.jl's `gmres` runs it by default
    (no restart, left preconditioner through ldiv!).  Returns (x, iterations)."""
    n = A.shape[0]
    Mi = (lambda v: v) if apply is None else apply
    r0 = np.asarray(Mi(np.asarray(b, dtype=np.float64)), dtype=np.float64)
    beta = np.linalg.norm(r0)
    if beta == 0.0:
        return np.zeros(n), 0
    V = np.zeros((maxiter + 1, n))
    H = np.zeros((maxiter + 1, maxiter))
    cs = np.zeros(maxiter)
    sn = np.zeros(maxiter)
    g = np.zeros(maxiter + 1)
    V[0] = r0 / beta
    g[0] = beta
    k = 0
    for k in range(maxiter):
        w = np.asarray(Mi(A @ V[k]), dtype=np.float64)
        for i in range(k + 1):
            H[i, k] = w @ V[i]
            w = w - H[i, k] * V[i]
        H[k + 1, k] = np.linalg.norm(w)
        if H[k + 1, k] > 0:
            V[k + 1] = w / H[k + 1, k]
        for i in range(k):
            t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
            H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
            H[i, k] = t
        d = np.hypot(H[k, k], H[k + 1, k])
        cs[k], sn[k] = H[k, k] / d, H[k + 1, k] / d
        H[k, k] = d
        H[k + 1, k] = 0.0
        g[k + 1] = -sn[k] * g[k]
        g[k] = cs[k] * g[k]
        if abs(g[k + 1]) <= rtol * beta:
            break
    m = k + 1
    y = np.linalg.solve(np.triu(H[:m, :m]), g[:m])
    return V[:m].T @ y, m


def check_reference_properties(A, b, x_prec, it_prec, it_unprec):
    """test/test_preconditioners.jl:102-103."""
    r = A @ x_prec
    assert np.allclose(r, b, rtol=1e-1, atol=1e-1)
    assert it_prec <= it_unprec
