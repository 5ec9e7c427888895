"""Autotuner for the launch parameters of the smoother paths, by measurement.

No counterpart in HybridSmoothers.jl, where partsize, threads and blocks are
the caller's choice (L1GSPrecBuilder(backend; threads, blocks)(A, partsize),
src/l1_gauss_seidel.jl:258-296).  Two tuners:

* tune_batch: the batched engine's partsize (16/32/64 -- it changes the
  preconditioner, hence the iteration counts, so it is tuned on whole solves),
  CG iterations per step (graph length: launch amortisation against how soon
  finished systems are retired and new ones start) and the tile-kernel grid.
  The score is solves/s on a sample of the workload, every solve checked to
  reach its tolerance.
* tune_l1gs_variant: the kernel variant of the partsize-64 symmetric sweep
  (results are identical across variants; only time differs).

Results can be cached in a JSON file keyed by device and candidate set.
"""
from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np

from . import problems as pb
from .batch import BatchSolver


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # pragma: no cover
        return "unknown"


def _load(cache_path, key):
    if not cache_path or not os.path.exists(cache_path):
        return None
    try:
        with open(cache_path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def _store(cache_path, key, value):
    if not cache_path:
        return
    data = {}
    if os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
    data[key] = value
    tmp = cache_path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, cache_path)


def sample_systems(specs, rhs_seed: int = 1000):
    """(CSR arrays, b) of a list of serving_stream specs."""
    out = []
    for i, s in enumerate(specs):
        A = pb.stream_system(s)
        out.append((A, pb.uniform_rhs(A[0].size - 1, rhs_seed + i)))
    return out


def time_batch(systems, partsize, passes, grid=0, precision="fp32", rtol=1e-8,
               arena_bytes=None, max_systems=1024, reps=1, **smoother):
    """Median wall time (s) of solving `systems` with one engine configuration,
    after one warm-up solve; raises if any solve misses its tolerance.
    `smoother`: BatchSolver's smoother, degree and power_iters."""
    with BatchSolver(partsize=partsize, precision=precision, passes=passes, grid=grid,
                     arena_bytes=arena_bytes, max_systems=max_systems, **smoother) as eng:
        eng.solve_all(systems[: max(1, len(systems) // 8)], rtol=rtol)  # warm-up
        times, sols = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            sols = eng.solve_all(systems, rtol=rtol)
            times.append(time.perf_counter() - t0)
    bad = [s for s in sols if s.status != "solved"]
    if bad:
        raise RuntimeError(f"{len(bad)} solves did not converge (partsize={partsize})")
    iters = float(np.mean([s.iters for s in sols]))
    return float(np.median(times)), iters


def tune_batch(systems, partsizes=(16, 32, 64), passes=(4, 8, 16), grids=(0,),
               precision="fp32", rtol=1e-8, arena_bytes=None, cache_path=None, log=None,
               **smoother):
    """Best (partsize, passes, grid) for `systems` by solves/s; returns
    {"partsize", "passes", "grid", "solves_per_s", "table": [...]}.  With
    smoother="cheb" partsize and grid do not apply: pass one of each."""
    sig = f"{len(systems)}:{sum(A[0].size - 1 for A, _ in systems)}"
    key = (f"batch|{_device_name()}|{precision}|{partsizes}|{passes}|{grids}|"
           f"{sorted(smoother.items())}|{sig}")
    hit = _load(cache_path, key)
    if hit:
        return hit
    table = []
    for p, k, g in itertools.product(partsizes, passes, grids):
        t, it = time_batch(systems, p, k, g, precision, rtol, arena_bytes, **smoother)
        row = {"partsize": p, "passes": k, "grid": g, "seconds": round(t, 5),
               "solves_per_s": round(len(systems) / t, 2), "iters_mean": round(it, 1)}
        table.append(row)
        if log:
            log(f"[autotune] {row}")
    best = max(table, key=lambda r: r["solves_per_s"])
    res = {"partsize": best["partsize"], "passes": best["passes"], "grid": best["grid"],
           "solves_per_s": best["solves_per_s"], "table": table}
    _store(cache_path, key, res)
    return res


def tune_l1gs_variant(prec, x, y=None, reps: int = 20, warmup: int = 3):
    """Fastest kernel variant (0-2) of a partsize-64 symmetric preconditioner
    on device vectors x (and y); sets it on `prec` and returns {variant: ms}."""
    import torch
    from .l1gs import ldiv_
    if prec.partitioning.partsize != 64:
        return {}
    y = torch.empty_like(x) if y is None else y
    times = {}
    for v in (0, 1, 2):
        prec.set_variant(v)
        for _ in range(warmup):
            ldiv_(y, prec, x, sync=False)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ldiv_(y, prec, x, sync=False)
        e1.record()
        torch.cuda.synchronize()
        times[v] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    prec.set_variant(best)
    return times
