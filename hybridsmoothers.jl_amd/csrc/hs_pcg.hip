// hs_pcg.hip -- FP64 preconditioned CG on the device, the solver the
// reference's users reach through solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P)
// (README.md:27-29).  The iteration is the one of Krylov.jl `cg` with
// ldiv = true and x0 = 0:
//
//   r = b; z = M^-1 r; p = z; gamma = r'z; eps = atol + rtol*sqrt(gamma)
//   loop: q = A p; alpha = gamma / p'q; x += alpha p; r -= alpha q;
//         z = M^-1 r; gamma' = r'z; stop if sqrt(gamma') <= eps;
//         beta = gamma'/gamma; p = z + beta p
//
// All scalars live on the device; every kernel first reads the `done` flag,
// so blocks of iterations are launched (and captured in a HIP graph) without
// a host round trip per iteration, and launches past convergence do nothing.
// Dot products are deterministic (fixed block tree, fixed final order).
#include "hs_plans.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <mutex>

namespace hs {

__global__ void k_cg_start(int64_t n, const double *__restrict__ b, double *__restrict__ x,
                           double *__restrict__ r) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] = 0.0;
        r[i] = b[i];
    }
}

// p = z; gamma = r'z; rnorm0; eps
__global__ __launch_bounds__(256) void k_cg_dot0(int64_t n, const double *__restrict__ r,
                                                 const double *__restrict__ z,
                                                 double *__restrict__ p, double *partials,
                                                 CGDev *st) {
    __shared__ double sm[4];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double zi = z[i];
        p[i] = zi;
        v = r[i] * zi;
    }
    double bs = block_sum(v, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// gamma = r'z, rnorm0, eps and the trivial-convergence test (one block)
__global__ __launch_bounds__(FIN_BLOCK) void k_cg_fin0(const double *__restrict__ partials,
                                                       unsigned count, CGDev *st) {
    __shared__ double sm[16];
    double tot = finalize_sum(partials, count, 1, sm);
    if (threadIdx.x == 0) st->done = cg_scalars_start(*st, tot) ? 1 : 0;
}

// q = A p; delta = p'q
__global__ __launch_bounds__(256) void k_cg_spmv(int64_t n, SellView<double> A,
                                                 const double *__restrict__ p,
                                                 double *__restrict__ q, double *partials,
                                                 CGDev *st) {
    __shared__ double sm[4];
    if (st->done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n) {
        double qi = row_dot(A, p, i);
        q[i] = qi;
        d = p[i] * qi;
    }
    double bs = block_sum(d, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// delta = p'q (one block)
__global__ __launch_bounds__(FIN_BLOCK) void k_cg_fin_delta(const double *__restrict__ partials,
                                                            unsigned count, CGDev *st) {
    __shared__ double sm[16];
    if (st->done) return;
    double tot = finalize_sum(partials, count, 1, sm);
    // breakdown (p'q <= 0): A or M is not positive definite
    if (threadIdx.x == 0 && !cg_scalars_delta(*st, tot)) st->done = 1;
}

// Fused CG form with an L1GS preconditioner, iteration k (st->iter == k):
// p_k = z + beta p_{k-1} (p_0 = z) formed on the fly, q = A p_k, p_k'q
// partials.  p_k goes to pbuf[k & 1] while p_{k-1} is gathered from
// pbuf[(k + 1) & 1], so no thread reads a value another thread rewrites.
template <class TZ>
__global__ __launch_bounds__(256) void k_cg_spmv_fused(int64_t n, SellView<double> A,
                                                       const TZ *__restrict__ z,
                                                       double *__restrict__ pa,
                                                       double *__restrict__ pb,
                                                       double *__restrict__ q, double *partials,
                                                       const CGDev *st) {
    __shared__ double sm[4];
    if (st->done) return;
    const long long k = st->iter;
    const bool first = (k == 0);
    const double beta = first ? 0.0 : st->beta;
    const double *pold = (k & 1) ? pa : pb;
    double *pnew = (k & 1) ? pb : pa;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    int64_t i = (int64_t)tile * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n) d = cg_spmv_row(A, z, pold, pnew, q, first, beta, i);
    double bs = block_sum(d, sm);
    if (threadIdx.x == 0) partials[tile] = bs;
}

// x += alpha p; r -= alpha q
__global__ void k_cg_axpy(int64_t n, const double *__restrict__ p, const double *__restrict__ q,
                          double *__restrict__ x, double *__restrict__ r, const CGDev *st) {
    if (st->done) return;
    const double alpha = st->gamma / st->delta;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] = x[i] + alpha * p[i];
        r[i] = r[i] - alpha * q[i];
    }
}

// gamma' = r'z, stopping test, beta
__global__ __launch_bounds__(256) void k_cg_dot(int64_t n, const double *__restrict__ r,
                                                const double *__restrict__ z, double *partials,
                                                CGDev *st) {
    __shared__ double sm[4];
    if (st->done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = (i < n) ? r[i] * z[i] : 0.0;
    double bs = block_sum(v, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// gamma' = r'z, stopping test, beta (one block)
__global__ __launch_bounds__(FIN_BLOCK) void k_cg_fin_gamma(const double *__restrict__ partials,
                                                            unsigned count, CGDev *st) {
    __shared__ double sm[16];
    if (st->done) return;
    double tot = finalize_sum(partials, count, 1, sm);
    if (threadIdx.x == 0 && cg_scalars_next(*st, tot)) st->done = 1;
}

// p = z + beta p
__global__ void k_cg_pupd(int64_t n, const double *__restrict__ z, double *__restrict__ p,
                          const CGDev *st) {
    if (st->done) return;
    const double beta = st->beta;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_copy(int64_t n, const double *__restrict__ a, double *__restrict__ b,
                       const CGDev *st) {
    if (st && st->done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ||b - A x||^2 and ||b||^2
__global__ __launch_bounds__(256) void k_cg_resid(int64_t n, SellView<double> A,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ b,
                                                  double *partials, CGDev *st) {
    __shared__ double sm[4];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double t2 = 0.0, b2 = 0.0;
    if (i < n) {
        double bi = b[i];
        double t = bi - row_dot(A, x, i);
        t2 = t * t;
        b2 = bi * bi;
    }
    double s1 = block_sum(t2, sm);
    double s2 = block_sum(b2, sm);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s1;
        partials[2 * blockIdx.x + 1] = s2;
    }
}

__global__ __launch_bounds__(FIN_BLOCK) void k_cg_fin_resid(const double *__restrict__ partials,
                                                            unsigned count, CGDev *st) {
    __shared__ double sm[16];
    double a = finalize_sum(partials, count, 2, sm);
    double c = finalize_sum(partials + 1, count, 2, sm);
    if (threadIdx.x == 0) {
        st->res2 = a;
        st->b2 = c;
    }
}

// ---------------------------------------------------------------------------
// workspace: vectors, scalars, a private stream for graph capture
// ---------------------------------------------------------------------------
struct CGWork {
    int64_t n = 0;
    DBuf<double> x, r, z, p, p2, q, b, partials;
    DBuf<float> z32;  // z of an FP32 or BF16 L1GS smoother in the fused iteration
    DBuf<CGDev> st;
    CGDev *host = nullptr;  // pinned mirror of the scalars
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_key[3] = {0, 0, 0};  // matrix serial, preconditioner serial, kind
    int graph_block = 0;

    void ensure(int64_t nn) {
        if (!stream) {
            HS_CHECK_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            HS_CHECK_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
            HS_CHECK_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
            HS_CHECK_HIP(hipHostMalloc((void **)&host, sizeof(CGDev), hipHostMallocDefault));
            st.alloc(1);
        }
        if (nn != n) {
            n = nn;
            x.alloc(n); r.alloc(n); z.alloc(n); p.alloc(n); p2.alloc(n); q.alloc(n); b.alloc(n);
            z32.release();
            // block partials of the 256-row kernels (two values each for the
            // residual) or one per partition of the fused sweep (>= 1 row each)
            partials.alloc(std::max<size_t>(2 * (size_t)grid_for(n, 256), (size_t)n));
            drop_graph();
        }
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_key[0] = graph_key[1] = graph_key[2] = 0;
    }
    ~CGWork() {
        drop_graph();
        if (host) (void)hipHostFree(host);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct PrecRef {
    int kind;
    void *plan;
    uint64_t serial() const {
        if (kind == HS_PCG_L1GS) return l1gs_serial(*reinterpret_cast<L1GSPlan *>(plan));
        if (kind == HS_PCG_CHEB) return cheb_serial(*reinterpret_cast<ChebPlan *>(plan));
        return 0;
    }
};

static void apply_prec(const PrecRef &M, const double *r, double *z, int64_t n, hipStream_t st,
                       const int *done, const CGDev *dst) {
    switch (M.kind) {
    case HS_PCG_L1GS:
        l1gs_apply_device(*reinterpret_cast<L1GSPlan *>(M.plan), r, z, st, done);
        break;
    case HS_PCG_CHEB:
        cheb_apply_device(*reinterpret_cast<ChebPlan *>(M.plan), z, r, true, st, done);
        break;
    default:
        hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256)), dim3(256), 0, st, n, r, z, dst);
        check_launch();
    }
}

// One iteration of the fused form: SpMV, finalize delta, the preconditioner
// with the CG updates folded in (an L1GS sweep: one launch; a 4th-kind
// Chebyshev smoother of degree d: d launches), finalize gamma.  See
// k_cg_spmv_fused and CGFuse.
// Whether the fused iteration keeps z in FP32 (an FP32 or BF16 L1GS smoother).
static bool z_in_fp32(const PrecRef &M) {
    return M.kind == HS_PCG_L1GS &&
           l1gs_precision(*reinterpret_cast<const L1GSPlan *>(M.plan)) != HS_PREC_FP64;
}

static void enqueue_iteration_fused(const CSR &A, const PrecRef &M, CGWork &w, hipStream_t st) {
    const int64_t n = w.n;
    const unsigned g = grid_for(n, 256);
    CGDev *d = w.st.get();
    const bool z32 = z_in_fp32(M);
    if (z32)
        hipLaunchKernelGGL(k_cg_spmv_fused<float>, dim3(g), dim3(256), 0, st, n, A.view(),
                           w.z32.get(), w.p.get(), w.p2.get(), w.q.get(), w.partials.get(), d);
    else
        hipLaunchKernelGGL(k_cg_spmv_fused<double>, dim3(g), dim3(256), 0, st, n, A.view(),
                           w.z.get(), w.p.get(), w.p2.get(), w.q.get(), w.partials.get(), d);
    hipLaunchKernelGGL(k_cg_fin_delta, dim3(1), dim3(FIN_BLOCK), 0, st, w.partials.get(), g, d);
    check_launch();
    // the sweep reads p_k from the buffer the SpMV above wrote; a graph
    // replays fixed pointers, so both buffers are passed and the parity of
    // st->iter picks one on the device
    CGFuse f{w.r.get(), w.q.get(), w.p.get(), w.p2.get(), w.x.get(), w.partials.get(), d};
    if (z32) f.z32 = w.z32.get();
    const unsigned cnt =
        M.kind == HS_PCG_L1GS
            ? l1gs_apply_fused(*reinterpret_cast<const L1GSPlan *>(M.plan), w.z.get(), st, f)
            : cheb_apply_fused(*reinterpret_cast<ChebPlan *>(M.plan), w.z.get(), st, f);
    hipLaunchKernelGGL(k_cg_fin_gamma, dim3(1), dim3(FIN_BLOCK), 0, st, w.partials.get(), cnt, d);
    check_launch();
}

static void enqueue_iteration(const CSR &A, const PrecRef &M, CGWork &w, hipStream_t st) {
    const int64_t n = w.n;
    const unsigned g = grid_for(n, 256);
    CGDev *d = w.st.get();
    const int *done = &d->done;
    hipLaunchKernelGGL(k_cg_spmv, dim3(g), dim3(256), 0, st, n, A.view(), w.p.get(), w.q.get(),
                       w.partials.get(), d);
    hipLaunchKernelGGL(k_cg_fin_delta, dim3(1), dim3(FIN_BLOCK), 0, st, w.partials.get(), g, d);
    hipLaunchKernelGGL(k_cg_axpy, dim3(g), dim3(256), 0, st, n, w.p.get(), w.q.get(), w.x.get(),
                       w.r.get(), d);
    check_launch();
    apply_prec(M, w.r.get(), w.z.get(), n, st, done, d);
    hipLaunchKernelGGL(k_cg_dot, dim3(g), dim3(256), 0, st, n, w.r.get(), w.z.get(),
                       w.partials.get(), d);
    hipLaunchKernelGGL(k_cg_fin_gamma, dim3(1), dim3(FIN_BLOCK), 0, st, w.partials.get(), g, d);
    hipLaunchKernelGGL(k_cg_pupd, dim3(g), dim3(256), 0, st, n, w.z.get(), w.p.get(), d);
    check_launch();
}

// Solves with x0 = 0; bvec/xvec are device pointers.  Iterations run in
// blocks of `block` launches between host checks of the done flag; with
// use_graph the block is captured once into a HIP graph and replayed.
// true_rtol > 0: at each stop of Krylov.jl's rule the FP64 true relative
// residual ||b - A x|| / ||b|| is computed on the device, and while it exceeds
// true_rtol (and iterations remain) the same CG recurrence continues with the
// stopping threshold eps scaled down by the ratio true_rtol / true residual
// (times 0.95): no restart, the iterates are those Krylov.jl's rule would
// produce with a smaller rtol.
void pcg_device(const CSR &A, const PrecRef &M, CGWork &w, const double *bvec, double *xvec,
                double rtol, double atol, int64_t itmax, int block, bool use_graph,
                hipStream_t user, hs_pcg_stats *stats, double true_rtol) {
    const int64_t n = A.n;
    w.ensure(n);
    hipStream_t st = w.stream;
    const unsigned g = grid_for(n, 256);
    auto t0 = std::chrono::steady_clock::now();
    HS_CHECK_HIP(hipEventRecord(w.ev_in, user));
    HS_CHECK_HIP(hipStreamWaitEvent(st, w.ev_in, 0));

    CGDev init{};
    init.atol = atol;
    init.rtol = rtol;
    init.itmax = itmax;
    *w.host = init;
    HS_CHECK_HIP(hipMemcpyAsync(w.st.get(), w.host, sizeof(CGDev), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cg_start, dim3(g), dim3(256), 0, st, n, bvec, w.x.get(), w.r.get());
    check_launch();
    apply_prec(M, w.r.get(), w.z.get(), n, st, nullptr, nullptr);
    hipLaunchKernelGGL(k_cg_dot0, dim3(g), dim3(256), 0, st, n, w.r.get(), w.z.get(), w.p.get(),
                       w.partials.get(), w.st.get());
    hipLaunchKernelGGL(k_cg_fin0, dim3(1), dim3(FIN_BLOCK), 0, st, w.partials.get(), g, w.st.get());
    check_launch();

    if (block < 1) block = 1;
    const bool fused =
        (M.kind == HS_PCG_L1GS && l1gs_can_fuse(*reinterpret_cast<const L1GSPlan *>(M.plan))) ||
        (M.kind == HS_PCG_CHEB && cheb_can_fuse(*reinterpret_cast<const ChebPlan *>(M.plan)));
    if (fused && z_in_fp32(M)) {
        // the fused iterations read z in FP32: z0 (FP32 numbers) narrowed exactly
        if (w.z32.n < (size_t)n) w.z32.alloc(n);
        hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(256), 0, st, n, w.z.get(), w.z32.get());
        check_launch();
    }
    auto iteration = [&](hipStream_t s) {
        if (fused) enqueue_iteration_fused(A, M, w, s);
        else enqueue_iteration(A, M, w, s);
    };
    if (use_graph) {
        const uint64_t key[3] = {A.serial, M.serial(), (uint64_t)M.kind};
        if (!w.graph || w.graph_block != block || key[0] != w.graph_key[0] ||
            key[1] != w.graph_key[1] || key[2] != w.graph_key[2]) {
            w.drop_graph();
            hipGraph_t graph;
            HS_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < block; ++k) iteration(st);
            HS_CHECK_HIP(hipStreamEndCapture(st, &graph));
            HS_CHECK_HIP(hipGraphInstantiate(&w.graph, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            w.graph_block = block;
            for (int k = 0; k < 3; ++k) w.graph_key[k] = key[k];
        }
    }
    // iterations until the device's done flag is set
    auto run = [&] {
        for (;;) {
            HS_CHECK_HIP(
                hipMemcpyAsync(w.host, w.st.get(), sizeof(CGDev), hipMemcpyDeviceToHost, st));
            HS_CHECK_HIP(hipStreamSynchronize(st));
            if (w.host->done) break;
            if (use_graph) {
                HS_CHECK_HIP(hipGraphLaunch(w.graph, st));
            } else {
                for (int k = 0; k < block; ++k) iteration(st);
            }
        }
    };
    // true residual ||b - A x||^2, ||b||^2 in FP64 into the device scalars
    // (and their host mirror)
    auto true_residual = [&] {
        hipLaunchKernelGGL(k_cg_resid, dim3(g), dim3(256), 0, st, n, A.view(), w.x.get(), bvec,
                           w.partials.get(), w.st.get());
        hipLaunchKernelGGL(k_cg_fin_resid, dim3(1), dim3(FIN_BLOCK), 0, st, w.partials.get(), g,
                           w.st.get());
        check_launch();
        HS_CHECK_HIP(hipMemcpyAsync(w.host, w.st.get(), sizeof(CGDev), hipMemcpyDeviceToHost, st));
        HS_CHECK_HIP(hipStreamSynchronize(st));
        const CGDev &h = *w.host;
        return h.b2 > 0 ? std::sqrt(h.res2 / h.b2) : std::sqrt(h.res2);
    };
    run();
    double relres = true_residual();
    for (int check = 0; check < 64 && true_rtol > 0.0 && relres > true_rtol; ++check) {
        CGDev h = *w.host;
        // only a stop of the rule itself continues (not the iteration limit or a
        // breakdown, nor an exactly zero residual)
        if (h.status != 0 || h.iter >= h.itmax || !(h.gnext > 0.0) || !(h.gamma > 0.0)) break;
        h.eps = h.rnorm * (true_rtol / relres) * 0.95;
        h.beta = h.gnext / h.gamma;  // the update cg_scalars_next skips when it stops
        h.gamma = h.gnext;
        h.done = 0;
        *w.host = h;
        HS_CHECK_HIP(hipMemcpyAsync(w.st.get(), w.host, sizeof(CGDev), hipMemcpyHostToDevice, st));
        if (!fused) {
            // the unfused form updates p = z + beta p after its stopping test
            hipLaunchKernelGGL(k_cg_pupd, dim3(g), dim3(256), 0, st, n, w.z.get(), w.p.get(),
                               w.st.get());
            check_launch();
        }
        run();
        relres = true_residual();
    }
    HS_CHECK_HIP(hipMemcpyAsync(xvec, w.x.get(), n * 8, hipMemcpyDeviceToDevice, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    HS_CHECK_HIP(hipEventRecord(w.ev_out, st));
    HS_CHECK_HIP(hipStreamWaitEvent(user, w.ev_out, 0));
    auto t1 = std::chrono::steady_clock::now();
    const CGDev &h = *w.host;
    stats->iters = h.iter;
    stats->status = h.status;
    stats->rnorm0 = h.rnorm0;
    stats->rnorm = h.rnorm;
    stats->true_relres = relres;
    stats->seconds = std::chrono::duration<double>(t1 - t0).count();
}

void cgwork_free(CGWork *w) { delete w; }

// ---------------------------------------------------------------------------
// the AMG smoother form: `iters` times x -= M^-1 (A x - b)
// (ext/HybridSmoothersAMGExt.jl:29-35: mul!(residual, A, x); residual .-= b;
// ldiv!(correction, P, residual); x .-= correction)
// ---------------------------------------------------------------------------
// r_i = (A x)_i - b_i: the row sum in column order first (mul! with a zero
// start), then the subtraction, as the reference rounds them.
__global__ void k_resid_minus(SellView<double> A, const double *__restrict__ x,
                              const double *__restrict__ b, double *__restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n) r[i] = row_dot(A, x, i) - b[i];
}

__global__ void k_sub_inplace(int64_t n, double *__restrict__ x, const double *__restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] - z[i];
}

void l1gs_smooth_device(const CSR &A, const L1GSPlan &P, double *x, const double *b, int iters,
                        double *r, double *z, hipStream_t st) {
    const int64_t n = A.n;
    const unsigned g = grid_for(n, 256);
    if (l1gs_can_smooth_fused(P)) {
        // one launch per round: residual, sweep and update fused; the rounds
        // alternate between x and r, and an odd count ends with a copy to x
        double *cur = x, *next = r;
        for (int it = 0; it < iters; ++it) {
            l1gs_smooth_fused(P, SmoothFuse{A.view(), cur, b, next}, st);
            std::swap(cur, next);
        }
        if (cur != x) {
            hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, st, n, cur, x, (const CGDev *)nullptr);
            check_launch();
        }
        return;
    }
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_resid_minus, dim3(g), dim3(256), 0, st, A.view(), x, b, r);
        check_launch();
        l1gs_apply_device(P, r, z, st, nullptr);
        hipLaunchKernelGGL(k_sub_inplace, dim3(g), dim3(256), 0, st, n, x, z);
        check_launch();
    }
}

// ---------------------------------------------------------------------------
// scratch of the allocation-free smoother round (hs_l1gs_smooth_noalloc): the
// residual and the correction vector, NOALLOC_ROWS values each, in a
// file-scope device array.  Each GPU holds its own copy for the life of the
// process; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (noalloc_scratch), never
// through the array's name, which in host code is not a device address.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_ROWS = int64_t(1) << 21;  // 2,097,152 rows: 32 MiB in all
static __device__ __align__(256) double scratch[2 * NOALLOC_ROWS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

// Resolves the calling thread's current device's copy of `scratch`, once per
// device (the device must be current).
void noalloc_scratch_init() {
    int dev = -1;
    HS_CHECK_HIP(hipGetDevice(&dev));
    HS_REQUIRE(dev >= 0 && dev < kMaxDevices, HS_EUNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return;
    void *p = nullptr;
    HS_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HS_REQUIRE(p != nullptr, HS_EDEVICE, "hipGetSymbolAddress returned no address");
    g_scratch[dev] = static_cast<double *>(p);
}

// This device's scratch (resolved here on the device's first use if hs_init
// was not called for it).
static double *noalloc_scratch() {
    int dev = -1;
    HS_CHECK_HIP(hipGetDevice(&dev));
    HS_REQUIRE(dev >= 0 && dev < kMaxDevices, HS_EUNSUPPORTED, "device index beyond 63");
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (g_scratch[dev]) return g_scratch[dev];
    }
    noalloc_scratch_init();
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[dev];
}

// The CG workspace of a matrix handle: created on first use, freed with the
// handle.  Concurrent solves with one matrix handle are not supported.
static CGWork &work_for(CSR &A) {
    if (!A.cg) A.cg = new CGWork();
    return *A.cg;
}

}  // namespace hs

using namespace hs;

extern "C" {

int hs_l1gs_smooth(hs_l1gs_t P, hs_csr_t A, double *x, const double *b, int iters, int loc,
                   void *stream, int flags) {
    return guard([&] {
        HS_REQUIRE(P && A, HS_EINVAL, "NULL handle");
        HS_REQUIRE(iters >= 0, HS_EINVAL, "iterations must be non-negative");
        const int64_t n = A->n;
        // hs_l1gs is the plan type of hs_l1gs.hip (a complete type only there)
        const L1GSPlan &plan = *reinterpret_cast<const L1GSPlan *>(P);
        HS_REQUIRE(l1gs_rows(plan) == n, HS_EINVAL, "preconditioner size does not match the matrix");
        HS_REQUIRE(n == 0 || (x && b), HS_EINVAL, "NULL vector");
        if (n == 0 || iters == 0) return;
        hipStream_t st = as_stream(stream);
        // the residual and correction vectors of the smoother (HybridSmoother's
        // residual/correction fields), kept with the matrix between calls
        CGWork &w = work_for(*A);
        w.ensure(n);
        if (loc == HS_LOC_DEVICE) {
            l1gs_smooth_device(*A, plan, x, b, iters, w.r.get(), w.z.get(), st);
            finish(st, flags);
            return;
        }
        DBuf<double> dx(n);
        HS_CHECK_HIP(hipMemcpyAsync(dx.get(), x, n * 8, hipMemcpyHostToDevice, st));
        HS_CHECK_HIP(hipMemcpyAsync(w.b.get(), b, n * 8, hipMemcpyHostToDevice, st));
        l1gs_smooth_device(*A, plan, dx.get(), w.b.get(), iters, w.r.get(), w.z.get(), st);
        HS_CHECK_HIP(hipMemcpyAsync(x, dx.get(), n * 8, hipMemcpyDeviceToHost, st));
        HS_CHECK_HIP(hipStreamSynchronize(st));
    });
}

int hs_l1gs_smooth_noalloc(hs_l1gs_t P, hs_csr_t A, double *x, const double *b, int iters,
                           void *stream, int flags) {
    return guard([&] {
        HS_REQUIRE(P && A, HS_EINVAL, "NULL handle");
        HS_REQUIRE(iters >= 0, HS_EINVAL, "iterations must be non-negative");
        const int64_t n = A->n;
        const L1GSPlan &plan = *reinterpret_cast<const L1GSPlan *>(P);
        HS_REQUIRE(l1gs_rows(plan) == n, HS_EINVAL, "preconditioner size does not match the matrix");
        HS_REQUIRE(n <= NOALLOC_ROWS, HS_EINVAL,
                   "system larger than the allocation-free scratch (2,097,152 rows); use "
                   "hs_l1gs_smooth");
        HS_REQUIRE(n == 0 || (x && b), HS_EINVAL, "NULL vector");
        if (n == 0 || iters == 0) return;
        double *r = noalloc_scratch();
        // the residual / next iterate and the correction, as hs_l1gs_smooth's
        // matrix-owned workspace; every value is written before it is read
        l1gs_smooth_device(*A, plan, x, b, iters, r, r + NOALLOC_ROWS, as_stream(stream));
        finish(as_stream(stream), flags);
    });
}

int hs_pcg_solve(hs_csr_t A, int prec_kind, void *prec, const double *b, double *x, double rtol,
                 double atol, int64_t itmax, int loc, void *stream, hs_pcg_stats *stats) {
    return hs_pcg_solve_true(A, prec_kind, prec, b, x, rtol, atol, itmax, 0.0, loc, stream, stats);
}

int hs_pcg_solve_true(hs_csr_t A, int prec_kind, void *prec, const double *b, double *x,
                      double rtol, double atol, int64_t itmax, double true_rtol, int loc,
                      void *stream, hs_pcg_stats *stats) {
    return guard([&] {
        HS_REQUIRE(true_rtol >= 0.0, HS_EINVAL, "true_rtol must be non-negative");
        HS_REQUIRE(A && b && x && stats, HS_EINVAL, "NULL argument");
        HS_REQUIRE(prec_kind == HS_PCG_NONE || prec, HS_EINVAL, "NULL preconditioner");
        HS_REQUIRE(prec_kind >= HS_PCG_NONE && prec_kind <= HS_PCG_CHEB, HS_EINVAL,
                   "unknown preconditioner kind");
        const int64_t n = A->n;
        HS_REQUIRE(n > 0, HS_EINVAL, "empty system");
        if (prec_kind == HS_PCG_L1GS)
            HS_REQUIRE(l1gs_rows(*reinterpret_cast<L1GSPlan *>(prec)) == n, HS_EINVAL,
                       "preconditioner size does not match the matrix");
        if (prec_kind == HS_PCG_CHEB)
            HS_REQUIRE(cheb_rows(*reinterpret_cast<ChebPlan *>(prec)) == n, HS_EINVAL,
                       "preconditioner size does not match the matrix");
        hipStream_t user = as_stream(stream);
        CGWork &w = work_for(*A);
        PrecRef M{prec_kind, prec};
        if (itmax < 0) itmax = 2 * n;
        // iterations per graph replay between host checks of the done flag:
        // launches past convergence exit at the flag (about 15 us an iteration
        // at 128^3), a host check costs a copy and a synchronisation
        const int block = 16;
        if (loc == HS_LOC_DEVICE) {
            pcg_device(*A, M, w, b, x, rtol, atol, itmax, block, true, user, stats, true_rtol);
        } else {
            w.ensure(n);
            HS_CHECK_HIP(hipMemcpyAsync(w.b.get(), b, n * 8, hipMemcpyHostToDevice, user));
            DBuf<double> xd(n);
            pcg_device(*A, M, w, w.b.get(), xd.get(), rtol, atol, itmax, block, true, user, stats,
                       true_rtol);
            HS_CHECK_HIP(hipMemcpyAsync(x, xd.get(), n * 8, hipMemcpyDeviceToHost, user));
            HS_CHECK_HIP(hipStreamSynchronize(user));
        }
    });
}

}  // extern "C"
