"""Request loop over the batched PCG engine: continuous batching of a stream of
SPD solves with admission by device-memory footprint.

No counterpart in HybridSmoothers.jl (a library of smoothers, no service).
The workload: a stream of systems (problems.serving_stream specs, generated
by the native lognormal-grid generator) arriving as a Poisson process.  The
loop repeatedly

  1. moves requests whose arrival time has passed into the pending queue;
  2. launches one engine step (a graph of `passes` CG iterations of every
     resident system) without waiting for it;
  3. while the device works on that step (and at least once per round),
     admits pending requests in arrival order (with a small look-ahead window
     so one large system does not block smaller ones behind it) until the
     engine refuses one for lack of slots or memory -- each admission uploads
     the system and builds its preconditioner on a second stream; admissions
     stop when the step finishes, so the device never waits for a long
     admission burst;
  4. collects the systems that finished (their memory is freed at once, which
     is what lets the next requests in).

Latency of a request = completion time - scheduled arrival time (so time spent
waiting for admission, or for its matrix to be generated, counts).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import problems as pb
from .batch import BatchSolver, column_bytes, iteration_bytes


def poisson_arrivals(count: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of a Poisson process of `rate` requests/s; rate <= 0 or
    inf puts every arrival at t = 0 (a saturation run)."""
    if not np.isfinite(rate) or rate <= 0:
        return np.zeros(count)
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, count))


class Producer:
    """Generates the stream's systems ahead of the loop on worker threads (the
    native generator releases the GIL).  get(i) blocks until system i exists;
    at most `ahead` systems are held beyond the last one taken."""

    def __init__(self, specs, rhs_seed: int = 1000, workers: int = 4, ahead: int = 256):
        self.specs = specs
        self.rhs_seed = rhs_seed
        self.ahead = ahead
        self._items: dict[int, tuple] = {}
        self._next = 0        # next index to hand to a worker
        self._taken = 0       # lowest index not yet taken by get()
        self._cv = threading.Condition()
        self._stop = False
        self._threads = [threading.Thread(target=self._work, daemon=True) for _ in range(workers)]
        for t in self._threads:
            t.start()

    def _work(self):
        while True:
            with self._cv:
                while not self._stop and (self._next >= len(self.specs)
                                          or self._next >= self._taken + self.ahead):
                    self._cv.wait()
                if self._stop:
                    return
                i = self._next
                self._next += 1
            A = pb.stream_system(self.specs[i])
            b = pb.uniform_rhs(A[0].size - 1, self.rhs_seed + i)
            with self._cv:
                self._items[i] = (A, b)
                self._cv.notify_all()

    def get(self, i: int):
        with self._cv:
            while i not in self._items:
                self._cv.wait()
            item = self._items.pop(i)
            self._taken = max(self._taken, i + 1)
            self._cv.notify_all()
            return item

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()


@dataclass
class ServeReport:
    count: int
    wall_s: float
    solves_per_s: float
    latency_p50_ms: float
    latency_p99_ms: float
    latency_mean_ms: float
    latency_max_ms: float
    # latency = wait (arrival to admission: generation, queueing behind
    # resident systems) + service (admission to completion)
    wait_p50_ms: float
    wait_p99_ms: float
    service_p50_ms: float
    service_p99_ms: float
    rate: float
    iters_mean: float
    iters_max: int
    max_true_relres: float
    all_solved: bool
    rows_total: int
    steps: int
    algorithmic_bytes: int     # sum over systems of iteration bytes x (iterations + 1)
    achieved_GBs: float        # algorithmic_bytes / wall time

    def as_dict(self):
        return {k: (float(f"{v:.5g}") if isinstance(v, float) else v)
                for k, v in self.__dict__.items()}


class Prepared:
    """Requests whose matrices already exist (the clients built them before
    sending): same interface as Producer.get."""

    def __init__(self, systems):
        self.systems = systems

    def get(self, i: int):
        return self.systems[i]

    def close(self):
        pass


def generate(specs, rhs_seed: int = 1000, workers: int = 8):
    """All systems of a stream, generated in parallel: [(CSR arrays, b), ...]."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        A = pb.stream_system(specs[i])
        return A, pb.uniform_rhs(A[0].size - 1, rhs_seed + i)

    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(one, range(len(specs))))


def serve(engine: BatchSolver, specs, arrivals: np.ndarray, rtol: float = 1e-8,
          producer=None, window: int = 16, keep_x: bool = False, on_result=None,
          admit_idle_s: float = 0.005):
    """Runs the stream through `engine`; returns (ServeReport, per-request records).

    `producer` supplies request i's (matrix, b) through get(i): a Producer
    (generation on worker threads, overlapping the loop; its time counts in
    the latencies) or Prepared (pre-built requests).  Default: a Producer."""
    N = len(specs)
    own = producer is None
    prod = producer or Producer(specs)
    lat = np.full(N, np.nan)
    t_in = np.full(N, np.nan)  # admission time, on the loop's clock
    iters = np.zeros(N, np.int64)
    relres = np.full(N, np.nan)
    status = [""] * N
    xs = [None] * N if keep_x else None
    rows = 0
    ibytes = np.zeros(N, np.int64)
    pending: list[int] = []
    staged: dict[int, tuple] = {}
    nxt = 0
    completed = 0
    t0 = time.perf_counter()
    steps0 = engine.info()["steps"]
    try:
        while completed < N:
            now = time.perf_counter() - t0
            while nxt < N and arrivals[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            busy = engine.pending() > 0
            if busy:
                engine.step()
            # admission while the step runs; with nothing on the device, for a
            # few milliseconds before the first step starts (at least one
            # attempt per round either way)
            k = 0
            tried = 0
            t_adm = time.perf_counter()
            while k < min(window, len(pending)):
                if tried and (not engine.busy() if busy
                              else time.perf_counter() - t_adm > admit_idle_s):
                    break
                tried += 1
                i = pending[k]
                if i not in staged:
                    staged[i] = prod.get(i)
                A, b = staged[i]
                if engine.submit(A, b, rtol=rtol, tag=i) is None:
                    k += 1
                    continue
                t_in[i] = time.perf_counter() - t0
                del staged[i]
                pending.pop(k)
                rows += b.size
                ptr, ind = (A[0], A[1]) if isinstance(A, tuple) else (A.indptr + 1, A.indices + 1)
                ibytes[i] = iteration_bytes(ptr, engine.partsize, engine.precision,
                                            engine.triangles, engine.smoother, engine.degree,
                                            column_bytes(ptr, ind))
            if busy:
                for s in engine.collect():
                    i = s.tag
                    lat[i] = time.perf_counter() - t0 - arrivals[i]
                    iters[i] = s.iters
                    relres[i] = s.true_relres
                    status[i] = s.status
                    if keep_x:
                        xs[i] = s.x
                    if on_result is not None:
                        on_result(i, s)
                    completed += 1
            elif not pending and engine.pending() == 0 and nxt < N:
                time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.05)))
        wall = time.perf_counter() - t0
    finally:
        if own:
            prod.close()
    span = wall - float(arrivals[0])
    lat_ms = lat * 1e3
    wait_ms = (t_in - arrivals[:N]) * 1e3
    service_ms = lat_ms - wait_ms
    rate = float(N / arrivals[-1]) if arrivals[-1] > 0 else float("inf")
    total_bytes = int(np.sum(ibytes * (iters + 1)))  # + 1: the admission pass
    rep = ServeReport(count=N, wall_s=wall, solves_per_s=N / span if span > 0 else float("inf"),
                      latency_p50_ms=float(np.percentile(lat_ms, 50)),
                      latency_p99_ms=float(np.percentile(lat_ms, 99)),
                      latency_mean_ms=float(np.mean(lat_ms)), latency_max_ms=float(np.max(lat_ms)),
                      wait_p50_ms=float(np.percentile(wait_ms, 50)),
                      wait_p99_ms=float(np.percentile(wait_ms, 99)),
                      service_p50_ms=float(np.percentile(service_ms, 50)),
                      service_p99_ms=float(np.percentile(service_ms, 99)),
                      rate=rate, iters_mean=float(np.mean(iters)), iters_max=int(np.max(iters)),
                      max_true_relres=float(np.nanmax(relres)),
                      all_solved=all(s == "solved" for s in status), rows_total=int(rows),
                      steps=engine.info()["steps"] - steps0, algorithmic_bytes=total_bytes,
                      achieved_GBs=total_bytes / wall / 1e9)
    records = {"latency_s": lat, "wait_s": wait_ms * 1e-3, "service_s": service_ms * 1e-3,
               "iters": iters, "true_relres": relres, "status": status, "x": xs}
    return rep, records
