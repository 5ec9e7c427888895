"""Batched persistent PCG over many SPD systems (host side of hs_batch_*).

No counterpart in HybridSmoothers.jl, whose users solve one system per
`solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P)` call (README.md:27-29).
Each system submitted here is solved exactly as

    P = L1GSPrecBuilder(ROCBackend(), ...)(A, partsize, isSymA=True, sweep=SymmetricSweep(),
                                           precision=precision)
    # or, smoother="cheb":
    P = _build_smoother(ChebyshevFourth(degree=degree, precision=precision), A)
    pcg_solve(DeviceMatrix(A), b, Pl=P, reltol=rtol, abstol=atol)

would solve it (FP64 CG, Krylov.jl's stopping rule, x0 = 0), but many systems
share every kernel launch.  A Chebyshev system's rho comes from `power_iters`
power-iteration steps that the engine runs for it, inside its passes, before
its first CG iteration.  The engine keeps matrices, preconditioners and CG
state resident in one device arena; `submit` returns None when the arena or
the slot table is full (admission by device-memory footprint).

    eng = BatchSolver(partsize=32, precision="fp32")
    ids = [eng.submit(A, b) for A, b in systems]     # None: retry after collect()
    while eng.pending():
        eng.step()                                    # one graph launch, async
        for sol in eng.collect():                     # waits, retires finished
            ...
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import scipy.sparse as sp

from . import _lib as L

_PREC = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32, "bf16": L.PREC_BF16}
_SMOOTHER = {"l1gs": L.PCG_L1GS, "cheb": L.PCG_CHEB}
_STATUS = {0: "solved", 1: "maxiters", 2: "breakdown"}


@dataclass
class BatchSolution:
    id: int
    x: np.ndarray
    iters: int
    status: str
    rnorm0: float
    rnorm: float
    true_relres: float
    t_submit: float
    t_start: float
    t_done: float
    tag: Any = None


def csr_arrays(A):
    """1-based Int64 (rowptr, colval) and Float64 values of a symmetric matrix.

    A scipy CSR or CSC matrix (a symmetric CSC's columns are its rows) or a
    tuple of 1-based arrays (ptr, ind, val) as produced by problems.lognormal_grid.
    """
    if isinstance(A, tuple):
        ptr, ind, val = A
        return (np.ascontiguousarray(ptr, dtype=np.int64), np.ascontiguousarray(ind, dtype=np.int64),
                np.ascontiguousarray(val, dtype=np.float64))
    if not (sp.isspmatrix_csr(A) or sp.isspmatrix_csc(A)):
        A = sp.csr_matrix(A)
    if not A.has_sorted_indices:
        A = A.sorted_indices()
    if A.shape[0] != A.shape[1]:
        raise L.HSError(L.HS_EINVAL, "matrix must be square")
    return (A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1,
            np.ascontiguousarray(A.data, dtype=np.float64))


def column_bytes(ptr, ind) -> int:
    """Bytes per stored column in the engine: 2 when every column of the
    1-based arrays lies within 32,767 of its row (uploaded as 16-bit
    offsets), else 4."""
    ptr = np.asarray(ptr, dtype=np.int64)
    ind = np.asarray(ind, dtype=np.int64)
    rows = np.repeat(np.arange(ptr.size - 1, dtype=np.int64), np.diff(ptr))
    d = ind - 1 - rows
    return 2 if d.size == 0 or int(np.abs(d).max()) <= 32767 else 4


def iteration_bytes(ptr, partsize: int, precision: str, triangles: int = 2,
                    smoother: str = "l1gs", degree: int = 2, cbytes: int = 4) -> int:
    """Algorithmic HBM bytes of one fused CG iteration of one system in the
    engine: the sliced-ELL matrix (8-byte value + a `cbytes`-byte column per
    stored entry (column_bytes), slice offsets), z and p_{k-1} gathered once and p_k, q written
    (32 B/row) in the SpMV; with L1GS, the packed triangle(s) as stored
    (`triangles` = 1: L only, U = L^T), D_Dl1, r and x read and written, q
    and p read, z written (56 B/row) in the sweep; with the Chebyshev smoother
    of degree d (element size e), d-1 steps reading the matrix in the
    smoother's type: the first reads r, q, l1inv, x and p and writes r, x and
    z_cg (56 + e B/row), each later one reads z_k, r_{k-1}, l1inv and z_cg and
    writes z_cg (16 + 3e B/row), and every step but the last writes r_k and
    z_{k+1} (2e B/row; the last reads r instead, 8 B/row, after the first)."""
    ptr = np.asarray(ptr, dtype=np.int64)
    n = ptr.size - 1
    lens = np.diff(ptr)
    pad = (-n) % 64
    w = np.concatenate([lens, np.zeros(pad, np.int64)]).reshape(-1, 64).max(axis=1)
    nent = int(64 * w.sum())
    spmv = nent * (8 + cbytes) + (w.size + 1) * 8 + 32 * n
    if smoother == "cheb":
        e = 8 if precision == "fp64" else 4
        tot = spmv
        for k in range(1, degree):
            last = k == degree - 1
            row = 56 + e if k == 1 else 16 + 3 * e
            if not last:
                row += 2 * e
            elif k > 1:
                row += 8
            tot += nent * (e + cbytes) + (w.size + 1) * 8 + n * row
        return tot
    tsz = {"fp64": 8, "fp32": 4, "bf16": 2}[precision]
    dsz = 8 if precision == "fp64" else 4
    P = partsize
    nparts = (n + P - 1) // P
    tri = nparts * P * (P - 1) * triangles // 2  # the last block is stored full size
    return spmv + tri * tsz + n * (dsz + 56)


def default_arena_bytes(fraction: float = 0.5, cap: int = 96 << 30) -> int:
    try:
        import torch
        free, _ = torch.cuda.mem_get_info()
        return int(min(free * fraction, cap))
    except Exception:  # pragma: no cover
        return 8 << 30


class BatchSolver:
    def __init__(self, partsize: int = 32, precision: str = "fp32", arena_bytes: int | None = None,
                 max_systems: int = 1024, staging_bytes: int = 256 << 20, passes: int = 8,
                 grid: int = 0, eta: float = 1.5, triangles: int = 0, device: int = 0,
                 smoother: str = "l1gs", degree: int = 2, power_iters: int = 50):
        self._h = None
        if precision not in _PREC:
            raise L.HSError(L.HS_EINVAL, f"unknown precision {precision!r}")
        if smoother not in _SMOOTHER:
            raise L.HSError(L.HS_EINVAL, f"unknown smoother {smoother!r}")
        L.init(device)
        cfg = L.BatchConfig(arena_bytes=int(arena_bytes or default_arena_bytes()),
                            max_systems=int(max_systems), staging_bytes=int(staging_bytes),
                            partsize=int(partsize), precision=_PREC[precision],
                            passes=int(passes), grid=int(grid), eta=float(eta),
                            triangles=int(triangles), smoother=_SMOOTHER[smoother],
                            degree=int(degree), power_iters=int(power_iters))
        self.config = cfg
        self.partsize, self.precision, self.passes = int(partsize), precision, int(passes)
        self.smoother, self.degree = smoother, int(degree)
        self._h = C.c_void_p()
        L.check(L.lib.hs_batch_create(C.byref(cfg), C.byref(self._h)))
        self.triangles = self.info()["triangles"]
        self._live: dict[int, tuple[np.ndarray, Any]] = {}  # id -> (x buffer, tag)
        self._res = (L.BatchResult * 256)()

    # -- admission ---------------------------------------------------------
    def footprint(self, A) -> int:
        ptr, _, _ = csr_arrays(A)
        out = C.c_int64(0)
        L.check(L.lib.hs_batch_footprint(self._h, ptr.size - 1, ptr.ctypes.data, C.byref(out)))
        return out.value

    def submit(self, A, b, rtol: float = 1e-8, atol: float = 0.0, itmax: int = -1,
               tag: Any = None) -> int | None:
        """Uploads one system; its id, or None when the engine is full right now."""
        ptr, ind, val = csr_arrays(A)
        n = ptr.size - 1
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (n,):
            raise L.HSError(L.HS_EINVAL, f"b must have length {n}")
        x = np.empty(n, np.float64)
        sid = C.c_int64(-1)
        st = L.lib.hs_batch_submit(self._h, n, ptr.ctypes.data, ind.ctypes.data, val.ctypes.data,
                                   b.ctypes.data, x.ctypes.data, float(rtol), float(atol),
                                   int(itmax), C.byref(sid))
        if st == L.HS_EAGAIN:
            return None
        L.check(st)
        self._live[sid.value] = (x, tag)
        return sid.value

    # -- progress ------------------------------------------------------------
    def step(self) -> None:
        L.check(L.lib.hs_batch_step(self._h))

    def collect(self) -> list[BatchSolution]:
        out = []
        cnt = C.c_int64(0)
        while True:
            L.check(L.lib.hs_batch_collect(self._h, self._res, len(self._res), C.byref(cnt)))
            for k in range(cnt.value):
                r = self._res[k]
                x, tag = self._live.pop(r.id)
                out.append(BatchSolution(r.id, x, int(r.iters), _STATUS.get(int(r.status), "?"),
                                         r.rnorm0, r.rnorm, r.true_relres, r.t_submit, r.t_start,
                                         r.t_done, tag))
            if cnt.value < len(self._res):
                return out

    def busy(self) -> bool:
        """True while the last step is still running on the device."""
        r = L.lib.hs_batch_busy(self._h)
        if r < 0:
            L.check(r)
        return r == 1

    def pending(self) -> int:
        """Systems submitted and not yet collected."""
        return len(self._live)

    def info(self) -> dict:
        v = np.zeros(12, np.int64)
        L.check(L.lib.hs_batch_info(self._h, v.ctypes.data))
        keys = ("active", "uploading", "free_slots", "free_bytes", "arena_bytes",
                "largest_free_block", "steps", "retired", "results_waiting", "grid_spmv",
                "grid_sweep", "triangles")
        return dict(zip(keys, (int(t) for t in v)))

    _PEEK = {"x": 0, "r": 1, "z": 2, "b": 3, "q": 4, "D": 5, "partials": 6, "l1inv": 7,
             "rho": 8, "column_bytes": 9}

    def peek(self, sid: int, what: str) -> np.ndarray:
        """A device vector of a resident system after the last step (diagnostics)."""
        x, _ = self._live[sid]
        n = x.size
        size = {"partials": 2 * ((n + 255) // 256), "rho": 1, "column_bytes": 1}.get(what, n)
        out = np.empty(size, np.float64)
        L.check(L.lib.hs_batch_peek(self._h, sid, self._PEEK[what], out.ctypes.data))
        return out

    def clock(self) -> float:
        return L.lib.hs_batch_clock(self._h)

    # -- convenience -----------------------------------------------------------
    def solve_all(self, systems, rtol: float = 1e-8, atol: float = 0.0, itmax: int = -1):
        """Solves [(A, b), ...] with continuous admission; solutions in input order."""
        todo = list(enumerate(systems))
        done: dict[int, BatchSolution] = {}
        while todo or self.pending():
            while todo:
                k, (A, b) = todo[0]
                if self.submit(A, b, rtol=rtol, atol=atol, itmax=itmax, tag=k) is None:
                    if not self.pending():
                        raise L.HSError(L.HS_ENOMEM, "system does not fit an empty engine")
                    break
                todo.pop(0)
            self.step()
            for s in self.collect():
                done[s.tag] = s
        return [done[k] for k in range(len(systems))]

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.hs_batch_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
