"""Benchmark of the L1GS / Chebyshev smoother hot path on one MI355X.

Default workload: 3D 7-point Poisson on a 128^3 grid (2,097,152 rows,
14,581,760 nnz, FP64 symmetric CSC), L1GS partsize 64, SymmetricSweep,
PackedBufferCache.  One "step" is one application y = M^-1 x of the
preconditioner (one fused kernel launch).  `value` is the achieved bandwidth
of the sweep against its algorithmic bytes

    2*S*8 (both packed triangles) + N*(8 + 8 + 8) (x, D_Dl1 read; y written)
    = 1,107,296,256 B per application, S = 66,060,288,

with inputs resident in HBM.  Further keys report the PCG solves (FP64
smoother and FP32 smoother inside FP64 CG, rtol 1e-8) and the CPU oracle
baseline timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--no-serve] [--no-cpu]
                    [--serve-only] [--no-serve-cheb] [--serve-count 4096]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hybridsmoothers.jl_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sweep_bytes(n, S, tri_bytes=8, vec_bytes=8, d_bytes=8, tris=2):
    return tris * S * tri_bytes + n * (vec_bytes + d_bytes + vec_bytes)


def measured_traffic(variant):
    """HBM bytes per launch of the headline sweep kernel, from the committed
    FETCH_SIZE / WRITE_SIZE counter passes (scaled by a plain copy's readings),
    or None when they do not cover the kernel variant being timed."""
    path = os.path.join(ROOT, "profiles", "l1gs_sym_sweep_fp64_poisson128_p64_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if variant != 2:  # the counters were taken on the ring-16 / run-24 kernel
        return None
    return int(rec["hbm_bytes_per_dispatch"])


def time_kernel(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms per launch


def time_launches(fn, steps, warmup, stream, flush=None):
    """Per-launch times (ms): p10, p50, p90 over `steps` launches, each between
    its own pair of events.  With `flush` (a device tensor, >= 512 MB), the
    tensor is read (summed) before every launch, outside the events, so the
    launch finds the 256 MB Infinity Cache holding other, clean data and its
    working set starts from HBM."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for a, b in ev:
        if flush is not None:
            flush.sum()
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    t = np.array([a.elapsed_time(b) for a, b in ev])
    return {"p10": round(float(np.percentile(t, 10)), 5), "p50": round(float(np.median(t)), 5),
            "p90": round(float(np.percentile(t, 90)), 5)}


def pcg_bench_interleaved(hs, M, b, configs, rounds=3, true_reltol=None):
    """Times PCG solves (x0 = 0, rtol 1e-8) of several preconditioners on one
    system: one warm-up solve each, then `rounds` rounds that solve once with
    every configuration in turn (so clock and temperature drift affect them
    alike); per configuration the median.  configs: [(label, P)]."""
    times = {label: [] for label, _ in configs}
    last = {}
    for label, P in configs:
        last[label] = hs.pcg_solve(M, b, Pl=P, reltol=1e-8, true_reltol=true_reltol)
    for _ in range(rounds):
        for label, P in configs:
            sol = hs.pcg_solve(M, b, Pl=P, reltol=1e-8, true_reltol=true_reltol)
            times[label].append(sol.seconds)
            last[label] = sol
    rows = []
    for label, _ in configs:
        sol, t = last[label], float(np.median(times[label]))
        rows.append({"name": label, "iters": sol.iters, "status": sol.status,
                     "true_relres": sol.true_relres, "ms_per_solve": round(t * 1e3, 3),
                     "solves_per_s": round(1.0 / t, 3),
                     "ms_per_iter": round(t * 1e3 / max(sol.iters, 1), 4),
                     "ms_all_rounds": [round(v * 1e3, 3) for v in times[label]]})
    return rows


# The CPU baselines use one OpenMP thread per host core up to 16: the GPU
# hosts give each GPU a share of 16 cores (whatever os.cpu_count() reports
# there), and more threads than cores would slow the reference's static
# chunk mapping down, not speed it up.  Both counts go into the record.
CPU_THREAD_CAP = 16


def cpu_threads():
    return min(os.cpu_count() or 1, CPU_THREAD_CAP)


def cpu_baseline(A, partsize, budget_s=12.0, A3=None, c4_rhs=None):
    """The oracle's packed symmetric sweep (KA.CPU mapping, chunks = host cores),
    its MatrixViewCache sweep and L1GS PCG at 128^3; the 2048^2 convection-
    diffusion forward and backward sweeps (P = 32); the Chebyshev 4th kind
    degree 3 PCG at 128^3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from hybridsmoothers import problems as pb
    n = A.shape[0]
    threads = cpu_threads()
    ref = O.l1gs_build(A, partsize, is_sym=True, sweep="symmetric", chunks=threads,
                       nthreads=threads)
    x = pb.uniform_rhs(n, 0)
    O.l1gs_ldiv(ref, x)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        O.l1gs_ldiv(ref, x)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 200:
            break
    dt = (time.perf_counter() - t0) / reps
    S = ref.SL.size
    gbs = sweep_bytes(n, S) / dt / 1e9
    out = {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
           "sample": f"{reps} symmetric L1GS applications of the 128^3 problem (P={partsize}, "
                     f"FP64 packed), oracle restatement with OpenMP, {dt * 1e3:.1f} ms each",
           "host_cpu_count": os.cpu_count(),
           "threads_note": f"OpenMP threads = min(os.cpu_count(), {CPU_THREAD_CAP}): the host's "
                           f"CPU share per GPU is {CPU_THREAD_CAP} cores"}
    # the CPU default layout (MatrixViewCache, :613), a few applications
    view = O.l1gs_build(A, partsize, is_sym=True, sweep="symmetric", cache="view", chunks=threads,
                        nthreads=threads)
    O.l1gs_ldiv(view, x)
    t0 = time.perf_counter()
    for _ in range(10):
        O.l1gs_ldiv(view, x)
    out["view_ms_per_apply"] = round((time.perf_counter() - t0) / 10 * 1e3, 2)
    # one PCG solve to rtol 1e-8 with the packed preconditioner (same restated CG)
    t0 = time.perf_counter()
    _, st = O.pcg(A, x, ref, rtol=1e-8, nthreads=threads)
    out["pcg_ms_per_solve"] = round((time.perf_counter() - t0) * 1e3, 1)
    out["pcg_iters"] = st.iters
    if c4_rhs is not None:
        # Chebyshev 4th kind degree 3 as Pl at 128^3 (its own power iteration)
        t0 = time.perf_counter()
        c4 = O.cheb_build(A, 4, 3, nthreads=threads)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, st4 = O.pcg(A, c4_rhs, c4, rtol=1e-8, nthreads=threads)
        out["cheb4_d3_pcg"] = {"ms_per_solve": round((time.perf_counter() - t0) * 1e3, 1),
                               "iters": st4.iters, "build_ms": round(tb * 1e3, 1)}
    if A3 is not None:
        # 2048^2 upwind convection-diffusion, P = 32, non-symmetric CSC (DiagonalIndices
        # path, O(nnz) l1 diagonal), forward and backward, FP64 packed
        x3 = pb.uniform_rhs(A3.shape[0], 1)
        c3 = {}
        for sw in ("forward", "backward"):
            t0 = time.perf_counter()
            r3 = O.l1gs_build(A3, 32, is_sym=False, sweep=sw, nthreads=threads)
            tb = time.perf_counter() - t0
            O.l1gs_ldiv(r3, x3)
            reps3, t0 = 0, time.perf_counter()
            while reps3 < 5 or (time.perf_counter() - t0 < 2.0 and reps3 < 50):
                O.l1gs_ldiv(r3, x3)
                reps3 += 1
            dt3 = (time.perf_counter() - t0) / reps3
            b3 = r3.SL.size * 8 + 3 * A3.shape[0] * 8 if sw == "forward" else \
                r3.SU.size * 8 + 3 * A3.shape[0] * 8
            c3[sw] = {"ms_per_apply": round(dt3 * 1e3, 3), "GB/s": round(b3 / dt3 / 1e9, 2),
                      "build_ms": round(tb * 1e3, 1), "applications": reps3}
        out["convdiff2048_p32"] = c3
    return out


def _serve_engine(args, specs, reqs, log_fn, label, engine_kw, tune_kw):
    """Autotune (on a separate 48-system sample), then a saturation run (all
    arrivals at t = 0: capacity) and a Poisson stream at 0.8 x capacity
    (latency percentiles) through one engine configuration."""
    from hybridsmoothers import autotune, problems as pb, serve
    from hybridsmoothers.batch import BatchSolver
    N = len(specs)
    log_fn(f"[bench] serving {label}: autotune on a 48-system sample")
    sample = autotune.sample_systems(pb.serving_stream(48, seed=77))
    tune = autotune.tune_batch(sample, passes=(8, 16, 32), precision="fp32",
                               arena_bytes=32 << 30, log=log_fn, **tune_kw, **engine_kw)
    out = {"autotune": tune}
    with BatchSolver(partsize=tune["partsize"], precision="fp32", passes=tune["passes"],
                     grid=tune["grid"], max_systems=args.serve_slots, **engine_kw) as eng:
        log_fn(f"[bench] serving {label}: saturation run ({N} systems at t=0)")
        rep_sat, _ = serve.serve(eng, specs, serve.poisson_arrivals(N, float("inf")),
                                 producer=reqs)
        log_fn(f"[bench] serving {label}: saturation {rep_sat.as_dict()}")
        lam = 0.8 * rep_sat.solves_per_s
        log_fn(f"[bench] serving {label}: Poisson arrivals at {lam:.1f}/s")
        rep, _ = serve.serve(eng, specs, serve.poisson_arrivals(N, lam, seed=7), producer=reqs)
        log_fn(f"[bench] serving {label}: poisson {rep.as_dict()}")
        out["arena_bytes"] = eng.info()["arena_bytes"]
    out["saturation"] = rep_sat.as_dict()
    out["poisson"] = rep.as_dict()
    out["solves_per_s"] = round(rep.solves_per_s, 2)
    out["latency_p50_ms"] = round(rep.latency_p50_ms, 3)
    out["latency_p99_ms"] = round(rep.latency_p99_ms, 3)
    out["capacity_solves_per_s"] = round(rep_sat.solves_per_s, 2)
    return out


def serving_bench(args, log_fn=log):
    """Batched persistent PCG serving the stream of SPD systems: N log-uniform in
    [1e3, 2.6e5] rows, alternating 2D 5-point / 3D 7-point lognormal (sigma 1)
    grids, FP64 CG to rtol 1e-8 preconditioned by FP32 L1GS partsize 32 (the
    top-level figures) and, as the stream's alternative smoother, by an FP32
    Chebyshev 4th-kind smoother of degree 2 (`chebyshev_d2`, rho from 50
    power-iteration steps in the engine).  Same requests for both."""
    from hybridsmoothers import problems as pb, serve
    N = args.serve_count
    specs = pb.serving_stream(N, seed=5)
    log_fn(f"[bench] serving: generating the {N} requests (clients' matrices, before the clock)")
    t0 = time.perf_counter()
    reqs = serve.Prepared(serve.generate(specs, workers=16))
    gen_s = round(time.perf_counter() - t0, 3)
    out = {"workload": f"{N} SPD systems, N log-uniform in [1e3, 2.6e5] (2D 5-point / 3D "
                       f"7-point lognormal sigma=1, alternating), b = U[-1,1]; FP64 CG, FP32 "
                       f"L1GS partsize 32 symmetric, rtol 1e-8"}
    out.update(_serve_engine(args, specs, reqs, log_fn, "L1GS P=32", {},
                             {"partsizes": (32,), "grids": (0, 2048)}))
    out["generation_s"] = gen_s
    if not args.no_serve_cheb:
        ch = _serve_engine(args, specs, reqs, log_fn, "Chebyshev d=2",
                           {"smoother": "cheb", "degree": 2, "power_iters": 50},
                           {"partsizes": (32,), "grids": (0,)})
        ch["workload"] = ("same stream and requests; FP64 CG, FP32 Chebyshev 4th kind degree 2 "
                          "(l1-Jacobi, rho from 50 power-iteration steps), rtol 1e-8")
        out["chebyshev_d2"] = ch
    return out


def mixed_precision_bench(hs, pb, builder, n3=192, log_fn=log):
    """3D 7-point variable-coefficient Poisson on n3^3 with lognormal face
    coefficients (seed 2), b = U[-1,1] (seed 2): FP64 CG to a true relative
    residual of 1e-8 with an L1GS partsize-64 symmetric smoother in FP64, FP32
    and BF16 (off-diagonals), both triangle layouts, the solves interleaved;
    each reported with its final true residual and its iteration delta
    against the all-FP64 solve."""
    import torch
    log_fn(f"[bench] mixed precision: building lognormal {n3}^3")
    A = pb.lognormal3d(n3, seed=2)
    n = A.shape[0]
    M = hs.DeviceMatrix(A)
    b = torch.tensor(pb.uniform_rhs(n, 2), device="cuda")
    configs = []
    for cache, tag in ((hs.PackedBufferCache(), ""), (hs.PackedSymmetricCache(), " one-triangle")):
        for prec in ("fp64", "fp32", "bf16"):
            log_fn(f"[bench] mixed precision: building the {prec}{tag} smoother")
            P = builder(A, 64, isSymA=True, sweep=hs.SymmetricSweep(), precision=prec,
                        cache_strategy=cache)
            configs.append((f"L1GS P=64 symmetric{tag}, {prec.upper()} smoother", P))
    log_fn("[bench] mixed precision: solves, interleaved")
    rows = pcg_bench_interleaved(hs, M, b, configs, rounds=2, true_reltol=1e-8)
    del configs
    base = rows[0]
    for k, r in enumerate(rows):
        same = rows[3 * (k // 3)]  # the FP64 solve with the same triangle storage
        r["iteration_delta"] = r["iters"] - base["iters"]
        r["speedup_vs_fp64"] = round(r["solves_per_s"] / base["solves_per_s"], 3)
        r["speedup_vs_fp64_same_storage"] = round(r["solves_per_s"] / same["solves_per_s"], 3)
    return {"workload": f"3D 7-point lognormal-coefficient Poisson {n3}^3 (N={n}, "
                        f"nnz={A.nnz}), seed 2; FP64 CG to a true relative residual of 1e-8 "
                        f"(Krylov.jl's rule at rtol 1e-8, continued on the device while "
                        f"||b - A x|| / ||b|| > 1e-8)", "solves": rows}


def cheb_apply_bench(hs, A, c4, x, y, steps, warmup, stream):
    """One Chebyshev 4th-kind degree-3 application as a left preconditioner
    (z = 0; smooth!(z, S, r); x0 = 0, the dead first and last SpMVs skipped) on
    the 128^3 problem, FP64: three launches (init and two SpMV-fused steps).
    Algorithmic bytes 2 B_A + 16 N 8 (two products, 16 vector streams), B_A the
    operator as the products read it: nnz (8-byte value + 2-byte column offset)
    + 8 (slices + 1); with the Int64 CSR layout the same formula gives
    768,606,224 B."""
    n = A.shape[0]
    M = hs.DeviceMatrix(A)
    cb = M.column_bytes
    nslices = (n + 63) // 64
    b_a = int(A.nnz) * (8 + cb) + (nslices + 1) * 8
    del M
    nbytes = 2 * b_a + 16 * n * 8
    ms = time_kernel(lambda: hs.ldiv_(y, c4, x, sync=False), steps, warmup, stream)
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"workload": "3D 7-point Poisson 128^3, ChebyshevFourth degree 3 FP64 as Pl "
                        "(x0 = 0), one application per step",
            "ms": round(ms, 5), "algorithmic_bytes": nbytes, "column_bytes": cb,
            "bytes_int64_csr": 2 * ((n + 1) * 8 + int(A.nnz) * 16) + 16 * n * 8,
            "GB/s": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
            "launches": 3,
            "launch_ms": time_launches(lambda: hs.ldiv_(y, c4, x, sync=False), min(steps, 100),
                                       warmup, stream)}


def amg_smoother_bench(hs, pb, A, builder, steps, warmup):
    """AMG.smooth! on the 128^3 problem (2 rounds of x -= M^-1 (A x - b), the
    extension's L1GS setup: isSymA = false, SymmetricSweep, P = 64): the fused
    launch per round against the same round composed of separate calls
    (SpMV, subtraction, sweep, subtraction)."""
    import torch
    from hybridsmoothers import amg
    n = A.shape[0]
    s = amg.setup_smoother(builder, A, None, partsize=64)
    b = torch.tensor(pb.uniform_rhs(n, 5), device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    fused = time_kernel(lambda: amg.smooth_(x, s, b, sync=False), steps, warmup, stream)
    r = torch.empty_like(x)
    z = torch.empty_like(x)

    def composed():
        for _ in range(s.iters):
            s.A.mul(r, x, sync=False)
            r.sub_(b)
            hs.ldiv_(z, s.s, r, sync=False)
            x.sub_(z)
    unfused = time_kernel(composed, steps, warmup, stream)
    S = s.s.packed_size
    nent = int(s.A.nnz)  # entries the rows read (the sliced-ELL padding is not counted)
    per_round = 2 * S * 8 + nent * 12 + n * (8 + 8 + 8 + 8)  # triangles, A, x, b, D, x_new
    return {"workload": f"3D 7-point Poisson 128^3, AMG.smooth! with L1GS P=64 (isSymA=false, "
                        f"SymmetricSweep), {s.iters} rounds per call",
            "fused_ms": round(fused, 4), "composed_ms": round(unfused, 4),
            "speedup": round(unfused / fused, 3),
            "bytes_per_round": per_round,
            "fused_GBs": round(per_round * s.iters / (fused * 1e-3) / 1e9, 1)}


def batch_steady(kind, nsys=256, iters=60):
    """Steady state of the batched engine for a kernel profile: `nsys` equal
    systems (2D 181^2 or 3D 32^3 lognormal, 8.4 M rows in all) resident
    together, FP32 L1GS partsize 32, a fixed `iters` CG iterations."""
    from hybridsmoothers import problems as pb
    from hybridsmoothers.batch import BatchSolver, column_bytes, iteration_bytes
    dims = (181, 181, 1) if kind == "2d" else (32, 32, 32)
    systems = []
    for i in range(nsys):
        A = pb.lognormal_grid(*dims, seed=100 + i)
        systems.append((A, pb.uniform_rhs(A[0].size - 1, 7 + i)))
    with BatchSolver(partsize=32, precision="fp32", passes=8, max_systems=1024) as eng:
        for A, b in systems:
            if eng.submit(A, b, itmax=iters) is None:
                raise RuntimeError("engine full")
        t0 = time.perf_counter()
        sols = []
        while eng.pending():
            eng.step()
            sols += eng.collect()
        dt = time.perf_counter() - t0
        A0 = systems[0][0]
        ib = iteration_bytes(A0[0], 32, "fp32", eng.triangles, cbytes=column_bytes(A0[0], A0[1]))
    rows = nsys * systems[0][1].size
    return {"workload": f"{nsys} x lognormal {'x'.join(map(str, dims))}", "rows": rows,
            "iterations": sols[0].iters, "bytes_per_iteration": ib * nsys, "wall_s": round(dt, 4),
            "GB/s_whole_run": round(ib * nsys * (sols[0].iters + 1) / dt / 1e9, 1)}


def serving_cpu_baseline(budget_s=10.0, smoother="l1gs"):
    """The oracle's PCG with its L1GS (FP64, partsize 32) or its Chebyshev 4th
    kind of degree 2 (FP64, 50 power-iteration steps), with OpenMP, on systems
    of the serving stream, in stream order, until the time budget is spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from hybridsmoothers import problems as pb
    specs = pb.serving_stream(4096, seed=5)
    threads = cpu_threads()
    done, rows, t0 = 0, 0, time.perf_counter()
    for i, spec in enumerate(specs):
        A = pb.csr_to_scipy(*pb.stream_system(spec))
        b = pb.uniform_rhs(A.shape[0], 1000 + i)
        if smoother == "cheb":
            ref = O.cheb_build(A, 4, 2, nthreads=threads)
        else:
            ref = O.l1gs_build(A, 32, is_sym=True, sweep="symmetric", nthreads=threads)
        O.pcg(A, b, ref, rtol=1e-8, nthreads=threads)
        done += 1
        rows += A.shape[0]
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    what = "Chebyshev 4th kind d=2" if smoother == "cheb" else "L1GS P=32"
    return {"value": round(done / dt, 3), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} systems of the stream ({rows} rows), oracle {what} FP64 "
                      f"+ CG with OpenMP, build included, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--partsize", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline sweep only")
    ap.add_argument("--sweeps-only", action="store_true",
                    help="the headline and the other sweep variants, no solves (for counter "
                         "passes)")
    ap.add_argument("--grid2d", type=int, default=2048)
    ap.add_argument("--serve-count", type=int, default=4096)
    ap.add_argument("--serve-slots", type=int, default=1024)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true", help="only the serving benchmark")
    ap.add_argument("--no-serve-cheb", action="store_true",
                    help="serve the stream with the L1GS engine only")
    ap.add_argument("--batch-steady", choices=("2d", "3d"), default=None,
                    help="only the batched engine's steady state (for kernel profiles)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 192^3 mixed-precision PCG")
    ap.add_argument("--mixed-only", action="store_true",
                    help="only the 192^3 mixed-precision PCG (for kernel profiles)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only time a plain device copy and a sum over MIB MiB per step "
                         "(counter calibration, read-only stream rate)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the smoother path has no multi-GPU code "
              "(partitions shard with no communication; see the design notes). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        ms = time_kernel(lambda: dst.copy_(src), args.steps, args.warmup,
                         torch.cuda.current_stream())
        # a read-only stream of the same size: a reduction over the buffer
        ms_r = time_kernel(lambda: src.sum(), args.steps, args.warmup,
                           torch.cuda.current_stream())
        print(json.dumps({"copy_bytes_each_way": nb, "ms": ms,
                          "GB/s_read+write": 2 * nb / (ms * 1e-3) / 1e9,
                          "sum_ms": ms_r, "GB/s_read_only": nb / (ms_r * 1e-3) / 1e9}),
              flush=True)
        return
    import hybridsmoothers as hs
    from hybridsmoothers import problems as pb

    hs.init(0)
    if args.batch_steady:
        print(json.dumps({"batch_steady": batch_steady(args.batch_steady)}), flush=True)
        return
    if args.mixed_only:
        builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)
        print(json.dumps({"mixed_precision_192": mixed_precision_bench(hs, pb, builder)}),
              flush=True)
        return
    if args.serve_only:
        res = serving_bench(args)
        if not args.no_cpu:
            res["cpu_baseline"] = serving_cpu_baseline()
            if "chebyshev_d2" in res:
                res["chebyshev_d2"]["cpu_baseline"] = serving_cpu_baseline(smoother="cheb")
        print(json.dumps({"serving": res}), flush=True)
        return
    g, P = args.grid, args.partsize
    log(f"[bench] building 3D Poisson {g}^3")
    A = pb.poisson3d(g)
    n = A.shape[0]
    builder = hs.L1GSPrecBuilder(hs.ROCBackend(), threads=256, blocks=1024)
    log("[bench] L1GS build (P=%d, symmetric, packed FP64)" % P)
    torch.cuda.synchronize()
    t_cold = time.perf_counter()
    prec = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep())
    t_cold = time.perf_counter() - t_cold
    del prec  # the second build is timed: the first also pays the process's first allocations
    torch.cuda.synchronize()
    t_build = time.perf_counter()
    prec = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep())
    t_build = time.perf_counter() - t_build
    S = prec.packed_size
    x = torch.tensor(pb.uniform_rhs(n, 0), device="cuda")
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    def step():
        hs.ldiv_(y, prec, x, sync=False)

    torch.cuda.synchronize()
    t_op = time.perf_counter()
    M_op = hs.DeviceMatrix(A)  # sliced-ELL operator built on the device
    t_op = time.perf_counter() - t_op
    del M_op
    log("[bench] timing symmetric sweep")
    nbytes = sweep_bytes(n, S)
    kernel_variants = {}
    for v in (0, 1, 2):
        prec.set_variant(v)
        mv = time_kernel(step, args.steps, args.warmup, stream)
        kernel_variants[str(v)] = round(mv, 5)
    best = min(kernel_variants, key=kernel_variants.get)
    prec.set_variant(int(best))
    ms = time_kernel(step, args.steps, args.warmup, stream)
    gbs = nbytes / (ms * 1e-3) / 1e9
    launch_ms = time_launches(step, min(args.steps, 100), args.warmup, stream)
    variants = []
    # working sets under 512 MB are also timed launch by launch with the
    # Infinity Cache flushed in between ("flushed_p50_ms", GB/s from it)
    flush = torch.zeros(512 << 17, dtype=torch.float64, device="cuda") if not args.quick else None

    def add_variant(name, fn, b, **extra):
        m = time_kernel(fn, args.steps, args.warmup, stream)
        row = {"name": name, "ms": round(m, 5), "bytes": b, "GB/s": round(b / (m * 1e-3) / 1e9, 1)}
        if b < (512 << 20):
            f = time_launches(fn, min(args.steps, 100), args.warmup, stream, flush)["p50"]
            row["flushed_p50_ms"] = f
            row["flushed_GB/s"] = round(b / (f * 1e-3) / 1e9, 1)
        row.update(extra)
        variants.append(row)

    if not args.quick:
        # reduced-precision smoothers on the same matrix (FP64 x in, FP64 y out)
        for pname, tb, db in (("fp32", 4, 4), ("bf16", 2, 4)):
            pv = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep(), precision=pname)
            pv.set_variant(int(best))
            add_variant(f"128^3 P={P} symmetric {pname}", lambda: hs.ldiv_(y, pv, x, sync=False),
                        sweep_bytes(n, S, tri_bytes=tb, d_bytes=db))
            del pv
        # MatrixViewCache: the matrix read by rows, no packed triangles; bytes
        # against the CSC-direct layout, (N+1) ix + nnz (ix + 8) + 3 N 8 (Int64)
        pv = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep(),
                     cache_strategy=hs.MatrixViewCache())
        add_variant(f"128^3 P={P} symmetric fp64 MatrixViewCache",
                    lambda: hs.ldiv_(y, pv, x, sync=False),
                    (n + 1) * 8 + int(A.nnz) * 16 + 3 * n * 8)
        del pv
        # partsizes 128, 192, 256: one wave per partition, two to four rows per lane
        for pw_size in (128, 192, 256):
            pw = builder(A, pw_size, isSymA=True, sweep=hs.SymmetricSweep())
            add_variant(f"128^3 P={pw_size} symmetric fp64", lambda: hs.ldiv_(y, pw, x, sync=False),
                        sweep_bytes(n, pw.packed_size))
            del pw
        # MatrixViewCache at partsize 128 and 256 (R = 2 and 4 rows per lane)
        for pv_size in (128, 256):
            pv = builder(A, pv_size, isSymA=True, sweep=hs.SymmetricSweep(),
                         cache_strategy=hs.MatrixViewCache())
            add_variant(f"128^3 P={pv_size} symmetric fp64 MatrixViewCache",
                        lambda: hs.ldiv_(y, pv, x, sync=False),
                        (n + 1) * 8 + int(A.nnz) * 16 + 3 * n * 8)
            del pv
        # one stored triangle (U = L^T): same results, half the packed bytes
        for pname, tb, db in (("fp64", 8, 8), ("fp32", 4, 4), ("bf16", 2, 4)):
            pv = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep(), precision=pname,
                         cache_strategy=hs.PackedSymmetricCache())
            add_variant(f"128^3 P={P} symmetric {pname} one-triangle",
                        lambda: hs.ldiv_(y, pv, x, sync=False),
                        sweep_bytes(n, S, tri_bytes=tb, d_bytes=db, tris=1))
            del pv
        # non-symmetric 2D upwind convection-diffusion 2048^2, P=32, DiagonalIndices path
        log("[bench] building 2048^2 convection-diffusion")
        A3 = pb.convdiff2d(args.grid2d)
        n3 = A3.shape[0]
        x3 = torch.tensor(pb.uniform_rhs(n3, 1), device="cuda")
        y3 = torch.empty_like(x3)
        for sw in (hs.ForwardSweep(), hs.BackwardSweep()):
            tb = time.perf_counter()
            p3 = builder(A3, 32, isSymA=False, sweep=sw)
            tb = time.perf_counter() - tb
            add_variant(f"{args.grid2d}^2 convdiff P=32 {type(sw).__name__}",
                        lambda: hs.ldiv_(y3, p3, x3, sync=False), p3.packed_size * 8 + 3 * n3 * 8,
                        build_ms=round(tb * 1e3, 1))
            del p3
    out = {
        "metric": "L1GS sweep GB/s (% HBM peak) at 128³; batched PCG solves/s + p99 at rtol 1e-8",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D 7-point Poisson {g}^3 (N={n}, nnz={A.nnz}), L1GS "
                               f"partsize={P} SymmetricSweep PackedBufferCache FP64, "
                               f"one application per step",
                   "rows": n, "nnz": int(A.nnz), "partsize": P, "packed_entries": S},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(int(best)) if g == 128 and P == 64 else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, scaled "
                                       "by a plain copy's readings, read from the committed "
                                       "profiles/l1gs_sym_sweep_fp64_poisson128_p64_traffic.json "
                                       "(not measured in this run)",
                     "algorithmic_bytes": nbytes},
        "sweeps": variants,
        "kernel_variant_ms": kernel_variants,
        "launch_ms": launch_ms,
        "build_ms": round(t_build * 1e3, 1),
        "build_ms_first": round(t_cold * 1e3, 1),
        "operator_build_ms": round(t_op * 1e3, 1),
        "kernel_variant": int(best),
    }
    if not args.quick and not args.sweeps_only:
        log("[bench] PCG on 128^3 (rtol 1e-8)")
        M = hs.DeviceMatrix(A)
        bvec = torch.tensor(pb.uniform_rhs(n, 0), device="cuda")
        # every preconditioner built first, then the solves interleaved round by round
        configs = [("L1GS P=64 symmetric, FP64 smoother", prec)]
        for pname in ("fp32", "bf16"):
            pv = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep(), precision=pname)
            pv.set_variant(int(best))
            configs.append((f"L1GS P=64 symmetric, {pname.upper()} smoother", pv))
        for pname in ("fp64", "fp32", "bf16"):
            pv = builder(A, P, isSymA=True, sweep=hs.SymmetricSweep(), precision=pname,
                         cache_strategy=hs.PackedSymmetricCache())
            configs.append((f"L1GS P=64 symmetric one-triangle, {pname.upper()} smoother", pv))
        c4 = hs._build_smoother(hs.ChebyshevFourth(degree=3, backend=hs.ROCBackend()), A)
        configs.append(("Chebyshev 4th kind degree 3, FP64", c4))
        c4f = hs._build_smoother(hs.ChebyshevFourth(degree=3, backend=hs.ROCBackend(),
                                                    precision="fp32"), A)
        configs.append(("Chebyshev 4th kind degree 3, FP32", c4f))
        pcg = pcg_bench_interleaved(hs, M, bvec, configs, rounds=3)
        out["chebyshev_apply_128"] = cheb_apply_bench(hs, A, c4, x, y, args.steps, args.warmup,
                                                      stream)
        del configs, c4f
        base = pcg[0]["solves_per_s"]
        base1 = next(r["solves_per_s"] for r in pcg if "one-triangle, FP64" in r["name"])
        for r in pcg:
            r["speedup_vs_fp64_l1gs"] = round(r["solves_per_s"] / base, 3)
            if "one-triangle" in r["name"]:
                r["speedup_vs_fp64_one_triangle"] = round(r["solves_per_s"] / base1, 3)
        out["pcg_128"] = pcg
        del M
        log("[bench] AMG smoother on 128^3")
        out["amg_smoother_128"] = amg_smoother_bench(hs, pb, A, builder, 50, 5)
        if not args.no_mixed:
            out["mixed_precision_192"] = mixed_precision_bench(hs, pb, builder)
    if not args.quick and not args.no_serve and not args.sweeps_only:
        out["serving"] = serving_bench(args)
        if not args.no_cpu:
            log("[bench] serving CPU baseline (oracle)")
            out["serving"]["cpu_baseline"] = serving_cpu_baseline()
            if "chebyshev_d2" in out["serving"]:
                out["serving"]["chebyshev_d2"]["cpu_baseline"] = serving_cpu_baseline(
                    smoother="cheb")
    if not args.no_cpu and not args.sweeps_only:
        log("[bench] CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(
            A, P, A3=None if args.quick else pb.convdiff2d(args.grid2d),
            c4_rhs=None if args.quick else pb.uniform_rhs(n, 0))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
