"""Sparse test and benchmark problems (scipy CSC, float64, sorted indices).

The small generators restate test/test_matrices.jl:5-19; the large ones are the
synthetic workloads the benchmark and the parity tests run:

* 1D Laplacian tridiag(-1, 2, -1), N = 2048 (README quick start),
* 3D 7-point Poisson on an n^3 grid, Dirichlet, x fastest (128^3: 2.10M rows),
* 2D 5-point upwind convection-diffusion, wind (1, 0.5), cell Peclet 10 (2048^2),
* 3D 7-point variable-coefficient Poisson with lognormal face coefficients (192^3),
* random 2D/3D lognormal-coefficient SPD systems for the batched / serving stream.

Large matrices are assembled column by column straight into CSC arrays with
numpy (no kron, no sort), which keeps 128^3 to about a second.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def poisson_test_matrix(n: int) -> sp.csc_matrix:
    """spdiagm(0 => 2, -1 => -1, 1 => -1) (test/test_matrices.jl:5)."""
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csc")
    A.sort_indices()
    return A


def poisson2d_matrix(n: int, eps: float = 1.0) -> sp.csc_matrix:
    """kron(I, T) + eps*kron(T, I) (test/test_matrices.jl:9-12)."""
    T = poisson_test_matrix(n)
    I = sp.identity(n, format="csc")
    A = (sp.kron(I, T) + eps * sp.kron(T, I)).tocsc()
    A.sort_indices()
    return A


def convection_diffusion_matrix(n: int, eps: float = 1.0, c: float = 1.0) -> sp.csc_matrix:
    """poisson2d_matrix + c*kron(I, upwind) (test/test_matrices.jl:16-19)."""
    up = sp.diags([np.ones(n), -np.ones(n - 1)], [0, -1], format="csc")
    A = (poisson2d_matrix(n, eps) + c * sp.kron(sp.identity(n, format="csc"), up)).tocsc()
    A.sort_indices()
    return A


def _assemble_csc(n_rows: int, offsets, coefs, valid) -> sp.csc_matrix:
    """CSC from per-column candidate entries.

    offsets: ascending list of row offsets relative to the column index;
    coefs:   list of arrays (length n_rows) of values for column c at row c+off;
    valid:   list of boolean arrays (length n_rows).
    Because offsets are ascending, flattening (column, offset) row-major keeps
    row indices sorted inside every column.
    """
    cols = np.arange(n_rows, dtype=np.int64)
    k = len(offsets)
    rows = np.empty((n_rows, k), np.int64)
    vals = np.empty((n_rows, k), np.float64)
    mask = np.empty((n_rows, k), bool)
    for t, (off, cf, vd) in enumerate(zip(offsets, coefs, valid)):
        rows[:, t] = cols + off
        vals[:, t] = cf
        mask[:, t] = vd
    counts = mask.sum(axis=1)
    indptr = np.zeros(n_rows + 1, np.int64)
    np.cumsum(counts, out=indptr[1:])
    m = mask.ravel()
    indices = rows.ravel()[m]
    data = vals.ravel()[m]
    idx_dtype = np.int32 if indptr[-1] < 2**31 - 1 else np.int64
    A = sp.csc_matrix((data, indices.astype(idx_dtype), indptr.astype(idx_dtype)),
                      shape=(n_rows, n_rows))
    A.has_sorted_indices = True
    return A


def laplace1d(n: int = 2048) -> sp.csc_matrix:
    return poisson_test_matrix(n)


def poisson3d(n: int) -> sp.csc_matrix:
    """7-point Poisson on n^3: kron(I,I,T)+kron(I,T,I)+kron(T,I,I), diagonal 6, x fastest."""
    N = n ** 3
    c = np.arange(N, dtype=np.int64)
    x = c % n
    y = (c // n) % n
    z = c // (n * n)
    one = np.ones(N)
    offs = [-n * n, -n, -1, 0, 1, n, n * n]
    coefs = [-one, -one, -one, 6.0 * one, -one, -one, -one]
    valid = [z > 0, y > 0, x > 0, np.ones(N, bool), x < n - 1, y < n - 1, z < n - 1]
    return _assemble_csc(N, offs, coefs, valid)


def convdiff2d(n: int, wind=(1.0, 0.5), peclet: float = 10.0) -> sp.csc_matrix:
    """2D 5-point diffusion plus first-order upwind convection (non-symmetric).

    Row i: A_ii = 4 + Pe*wx + Pe*wy, west = -(1 + Pe*wx), east = -1,
    south = -(1 + Pe*wy), north = -1 (x fastest).  For wind (1, 0.5), Pe = 10:
    diagonal 19, west -11, east -1, south -6, north -1.  Built column-wise: the
    entry in column c at row c+off is the coupling of row c+off to column c.
    """
    N = n * n
    cx, cy = peclet * wind[0], peclet * wind[1]
    c = np.arange(N, dtype=np.int64)
    x = c % n
    y = c // n
    one = np.ones(N)
    # column c, row r = c + off: A[r, c]. r = c - n -> r's north neighbour is c -> -1.
    # r = c - 1 -> r's east neighbour is c -> -1. r = c + 1 -> r's west is c -> -(1+cx).
    # r = c + n -> r's south is c -> -(1+cy).
    offs = [-n, -1, 0, 1, n]
    coefs = [-one, -one, (4.0 + cx + cy) * one, -(1.0 + cx) * one, -(1.0 + cy) * one]
    valid = [y > 0, x > 0, np.ones(N, bool), x < n - 1, y < n - 1]
    return _assemble_csc(N, offs, coefs, valid)


def lognormal3d(n: int, seed: int = 2, sigma: float = 1.0) -> sp.csc_matrix:
    """7-point variable-coefficient Poisson with face coefficients exp(sigma*N(0,1)).

    Faces include the Dirichlet boundary faces; A_ii = sum of the 6 face
    coefficients of the cell, A_ij = -k(face between i and j).  SPD.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    kx = np.exp(sigma * rng.standard_normal((n, n, n + 1)))  # [z, y, xface]
    ky = np.exp(sigma * rng.standard_normal((n, n + 1, n)))  # [z, yface, x]
    kz = np.exp(sigma * rng.standard_normal((n + 1, n, n)))  # [zface, y, x]
    N = n ** 3
    diag = (kx[:, :, :-1] + kx[:, :, 1:] + ky[:, :-1, :] + ky[:, 1:, :]
            + kz[:-1, :, :] + kz[1:, :, :]).ravel()
    c = np.arange(N, dtype=np.int64)
    x = c % n
    y = (c // n) % n
    z = c // (n * n)
    # coupling of column c with row c+off equals the face between them
    k_xm = kx[:, :, :-1].ravel()  # face on the -x side of cell c
    k_xp = kx[:, :, 1:].ravel()
    k_ym = ky[:, :-1, :].ravel()
    k_yp = ky[:, 1:, :].ravel()
    k_zm = kz[:-1, :, :].ravel()
    k_zp = kz[1:, :, :].ravel()
    offs = [-n * n, -n, -1, 0, 1, n, n * n]
    coefs = [-k_zm, -k_ym, -k_xm, diag, -k_xp, -k_yp, -k_zp]
    valid = [z > 0, y > 0, x > 0, np.ones(N, bool), x < n - 1, y < n - 1, z < n - 1]
    return _assemble_csc(N, offs, coefs, valid)


def lognormal2d(n: int, seed: int = 0, sigma: float = 1.0) -> sp.csc_matrix:
    """5-point variable-coefficient Poisson on n^2 with lognormal face coefficients."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kx = np.exp(sigma * rng.standard_normal((n, n + 1)))
    ky = np.exp(sigma * rng.standard_normal((n + 1, n)))
    N = n * n
    diag = (kx[:, :-1] + kx[:, 1:] + ky[:-1, :] + ky[1:, :]).ravel()
    c = np.arange(N, dtype=np.int64)
    x = c % n
    y = c // n
    offs = [-n, -1, 0, 1, n]
    coefs = [-ky[:-1, :].ravel(), -kx[:, :-1].ravel(), diag, -kx[:, 1:].ravel(),
             -ky[1:, :].ravel()]
    valid = [y > 0, x > 0, np.ones(N, bool), x < n - 1, y < n - 1]
    return _assemble_csc(N, offs, coefs, valid)


def uniform_rhs(n: int, seed: int) -> np.ndarray:
    """b ~ U[-1, 1] from a numpy PCG64 stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-1.0, 1.0, n)


def serving_stream(count: int, seed: int = 5, nmin: float = 1e3, nmax: float = 2.6e5):
    """Specs of a stream of SPD systems with N log-uniform in [nmin, nmax].

    Systems alternate 2D 5-point and 3D 7-point lognormal (sigma = 1) grids of the
    nearest size.  Returns a list of (kind, grid_n, seed) tuples.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    logn = rng.uniform(np.log(nmin), np.log(nmax), count)
    specs = []
    for i, ln in enumerate(logn):
        N = float(np.exp(ln))
        if i % 2 == 0:
            specs.append(("2d", max(8, int(round(N ** 0.5))), seed * 100003 + i))
        else:
            specs.append(("3d", max(4, int(round(N ** (1.0 / 3.0)))), seed * 100003 + i))
    return specs


def make_system(spec):
    kind, g, s = spec
    return lognormal2d(g, seed=s) if kind == "2d" else lognormal3d(g, seed=s)


# ---------------------------------------------------------------------------
# counter-based lognormal grids (the serving stream's systems)
# ---------------------------------------------------------------------------
_M64 = 0xFFFFFFFFFFFFFFFF


def _mix64(seed: int, ctr: np.ndarray) -> np.ndarray:
    """splitmix64 finalizer of seed*phi + ctr*c2 + phi (wrapping uint64)."""
    with np.errstate(over="ignore"):
        z = (ctr.astype(np.uint64) * np.uint64(0xD1B54A32D192ED03)
             + np.uint64(((seed * 0x9E3779B97F4A7C15) + 0x9E3779B97F4A7C15) & _M64))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def lognormal_faces(count: int, seed: int, sigma: float) -> np.ndarray:
    """k_f = exp(sigma * g_f), g_f standard normal by Box-Muller on counter pairs."""
    f = np.arange(count, dtype=np.uint64)
    m = f >> np.uint64(1)
    h1 = _mix64(seed, np.uint64(2) * m)
    h2 = _mix64(seed, np.uint64(2) * m + np.uint64(1))
    two53 = 1.0 / 9007199254740992.0
    u1 = ((h1 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * two53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * two53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    g = np.where((f & np.uint64(1)) == 1, r * np.sin(th), r * np.cos(th))
    return np.exp(sigma * g)


def lognormal_grid_numpy(nx: int, ny: int, nz: int = 1, sigma: float = 1.0,
                         seed: int = 0) -> sp.csc_matrix:
    """numpy statement of hs_problem_lognormal (symmetric, so CSC = CSR).

    5-point (nz = 1) or 7-point Poisson with face coefficients from
    lognormal_faces (x faces, then y, then z, each family x fastest); the
    diagonal is k_xm + k_xp + k_ym + k_yp + k_zm + k_zp with Dirichlet faces,
    and axes of length 1 have no faces.
    """
    hx, hy, hz = nx > 1, ny > 1, nz > 1
    Fx = (nx + 1) * ny * nz if hx else 0
    Fy = nx * (ny + 1) * nz if hy else 0
    Fz = nx * ny * (nz + 1) if hz else 0
    k = lognormal_faces(Fx + Fy + Fz, seed, sigma)
    N = nx * ny * nz
    zero = np.zeros((nz, ny, nx))
    kx = k[:Fx].reshape(nz, ny, nx + 1) if hx else None
    ky = k[Fx:Fx + Fy].reshape(nz, ny + 1, nx) if hy else None
    kz = k[Fx + Fy:].reshape(nz + 1, ny, nx) if hz else None
    xm = kx[:, :, :-1] if hx else zero
    xp = kx[:, :, 1:] if hx else zero
    ym = ky[:, :-1, :] if hy else zero
    yp = ky[:, 1:, :] if hy else zero
    zm = kz[:-1, :, :] if hz else zero
    zp = kz[1:, :, :] if hz else zero
    diag = (xm + xp + ym + yp + zm + zp).ravel()
    c = np.arange(N, dtype=np.int64)
    x = c % nx
    y = (c // nx) % ny
    z = c // (nx * ny)
    offs = [-nx * ny, -nx, -1, 0, 1, nx, nx * ny]
    coefs = [-zm.ravel(), -ym.ravel(), -xm.ravel(), diag, -xp.ravel(), -yp.ravel(),
             -zp.ravel()]
    valid = [z > 0, y > 0, x > 0, np.ones(N, bool), x < nx - 1, y < ny - 1, z < nz - 1]
    return _assemble_csc(N, offs, coefs, valid)


def lognormal_grid(nx: int, ny: int, nz: int = 1, sigma: float = 1.0, seed: int = 0):
    """The same matrix from libhsmooth's threaded host generator, as 1-based
    Int64 CSR arrays (rowptr, colval, nzval) -- the arrays a client submits."""
    import ctypes as C
    from . import _lib as L
    nnz = C.c_int64(0)
    L.check(L.lib.hs_problem_lognormal(nx, ny, nz, sigma, seed, None, None, None, C.byref(nnz)))
    n = nx * ny * nz
    rp = np.empty(n + 1, np.int64)
    ci = np.empty(nnz.value, np.int64)
    cv = np.empty(nnz.value, np.float64)
    L.check(L.lib.hs_problem_lognormal(nx, ny, nz, sigma, seed, rp.ctypes.data, ci.ctypes.data,
                                       cv.ctypes.data, C.byref(nnz)))
    return rp, ci, cv


def csr_to_scipy(rp, ci, cv) -> sp.csr_matrix:
    n = rp.size - 1
    return sp.csr_matrix((cv, ci - 1, rp - 1), shape=(n, n))


def stream_grid(spec):
    """(nx, ny, nz) of a serving_stream spec."""
    kind, g, _ = spec
    return (g, g, 1) if kind == "2d" else (g, g, g)


def stream_system(spec, sigma: float = 1.0):
    """CSR arrays of a serving_stream spec from the native generator."""
    nx, ny, nz = stream_grid(spec)
    return lognormal_grid(nx, ny, nz, sigma=sigma, seed=spec[2])
