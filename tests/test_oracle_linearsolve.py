"""The reference's LinearSolve cases (test/test_preconditioners.jl:366-531) on the
oracle: its own assertions (A x ~ b, no more iterations than unpreconditioned)
and, for the Symmetric-wrapper cases, that the wrapper equals the full matrix
with isSymA = true.  The GPU tests run the same cases against these results."""
import numpy as np
import pytest

import oracle as O
from linearsolve_cases import PARTSIZE, cases, check_reference_properties, gmres

_UNPREC = {}


def _unprec(key, A, b):
    if key not in _UNPREC:
        _UNPREC[key] = gmres(A, b)
    return _UNPREC[key]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_linearsolve_case_oracle(case):
    name, A, is_sym, sw, cache, solver, wrapped = case
    n = A.shape[0]
    b = np.ones(n)
    M = A
    if wrapped:
        from hybridsmoothers.l1gs import Symmetric
        M = Symmetric(A).full()
        assert abs(M - A).max() == 0
    p = O.l1gs_build(M, PARTSIZE, is_sym=is_sym, sweep=sw, cache=cache, chunks=20)
    if solver == "cg":
        if cache == "view":
            # the oracle's CG takes the packed plan; the view plan applies the
            # same terms in the same order, bit for bit
            pp = O.l1gs_build(M, PARTSIZE, is_sym=is_sym, sweep=sw, cache="packed", chunks=20)
            v = np.random.default_rng(1).uniform(-1, 1, n)
            np.testing.assert_array_equal(O.l1gs_ldiv(p, v), O.l1gs_ldiv(pp, v))
            p = pp
        x, st = O.pcg(A, b, p, rtol=1e-8)
        _, st0 = O.pcg(A, b, None, rtol=1e-8)
        it, it0 = st.iters, st0.iters
    else:
        x, it = gmres(A, b, lambda v: O.l1gs_ldiv(p, v))
        _, it0 = _unprec(name.split()[0], A, b)
    check_reference_properties(A, b, x, it, it0)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-7


def test_gmres_restatement():
    """The GMRES restatement against a direct solve and against its definition:
    iteration m minimises ||M^-1 (b - A x)|| over the m-th Krylov space, so
    without a preconditioner on an n x n system it is exact by step n, and
    with M = A^-1 it converges in one step."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from linearsolve_cases import gmres
    rng = np.random.default_rng(3)
    n = 40
    A = sp.csc_matrix(rng.standard_normal((n, n)) + n * np.eye(n))
    b = rng.standard_normal(n)
    x, it = gmres(A, b, rtol=1e-12)
    np.testing.assert_allclose(A @ x, b, rtol=1e-10, atol=1e-10)
    assert it <= n
    lu = spla.splu(A)
    x1, it1 = gmres(A, b, apply=lu.solve, rtol=1e-12)
    assert it1 == 1
    np.testing.assert_allclose(x1, lu.solve(b), rtol=1e-10)
