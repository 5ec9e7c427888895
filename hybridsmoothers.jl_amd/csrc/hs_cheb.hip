// hs_cheb.hip -- l1-Jacobi Chebyshev smoothers of the 4th kind (Lottes) and
// the 1st kind with optimal a* (D'Ambra et al.), src/chebyshev.jl.
//
// Each SpMV of the reference's recurrences is fused with the pointwise
// kernels around it: one kernel per degree reads A once, gathers the current
// z, and updates r, x and the next z (z ping-pongs between two buffers so the
// gather never sees a half-written vector).  Row sums run in row order and no
// FMA is formed, as in the reference's CSR mul! and KA kernels.
//
//   4th kind, step k (reference :158-183):
//     r_k = r_{k-1} - A z_k;  x += z_k;  z_{k+1} = a_{k+1} z_k + b_{k+1} l1inv r_k
//   1st kind, step k (reference :278-322):
//     x += z_{k-1};  r_k = r_{k-1} - l1inv (A z_{k-1}) / rho;  z_k = c1 z_{k-1} + c2 r_k
//
// As a left preconditioner (x0 = 0) the first SpMV (A*0) and the last one of
// the 4th kind (whose r is never read) are skipped: degree d costs d-1 SpMVs.
#include "hs_plans.hpp"

#include <cmath>
#include <cstddef>
#include <memory>

namespace hs {

enum : int {
    F_X_ZERO = 1,    // x starts at zero: do not read it
    F_HAS_NEXT = 2,  // compute and store the next z
    F_ADD_NEXT = 4,  // also add the next z to x (it is the last one)
    F_WRITE_R = 8,   // store r
};

// ---- setup ----------------------------------------------------------------
// l1inv_i = 1 / sum_j |a_ij| in row order (_l1_rowsum, :27-36, then 1 ./ l1)
__global__ void k_l1inv(SellView<double> A, double *__restrict__ l1inv,
                        float *__restrict__ l1inv32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    double s = row_abs_sum(A, i);
    double r = 1.0 / s;
    l1inv[i] = r;
    if (l1inv32) l1inv32[i] = (float)r;
}

struct PowState {
    double rho;
    double sumsq;
    int stopped;
};

// w = (A v) .* l1inv and its squared norm (power iteration, :40-54)
__global__ __launch_bounds__(256) void k_pow_spmv(int64_t n, SellView<double> A,
                                                  const double *__restrict__ l1inv,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ w, double *partials,
                                                  const PowState *st) {
    __shared__ double sm[4];
    if (st->stopped) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0;
    if (i < n) {
        double wi = row_dot(A, x, i);
        wi = wi * l1inv[i];
        w[i] = wi;
        sq = wi * wi;
    }
    double bs = block_sum(sq, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// rho = norm(w) from the block partials (one block)
__global__ __launch_bounds__(256) void k_pow_fin(const double *__restrict__ partials,
                                                 unsigned count, PowState *st) {
    __shared__ double sm[4];
    if (st->stopped) return;
    double tot = finalize_sum(partials, count, 1, sm);
    if (threadIdx.x == 0) {
        double rho = sqrt(tot);
        st->sumsq = tot;
        st->rho = rho;
        if (rho == 0.0) st->stopped = 1;  // iszero(rho) && break
    }
}

__global__ void k_pow_scale(int64_t n, const double *__restrict__ w, double *__restrict__ v,
                            const PowState *st) {
    if (st->stopped) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = w[i] / st->rho;
}

__global__ void k_fill(int64_t n, double *__restrict__ v, double a) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a;
}

// ---- 4th kind -------------------------------------------------------------
// init: r0 = b - A x (or b when x0 = 0) and z1 = a1*0 + b1*l1inv*r0
template <class T>
__global__ void k_c4_init(int64_t n, SellView<T> A, const double *__restrict__ x,
                          const double *__restrict__ b, const T *__restrict__ l1inv,
                          T *__restrict__ r, T *__restrict__ z1, double *__restrict__ xout, T a1,
                          T b1, int flags, const int *__restrict__ done) {
    if (done && *done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T ri = (T)b[i];
    if (!(flags & F_X_ZERO)) ri = ri - row_dot(A, x, i);
    if (flags & F_WRITE_R) r[i] = ri;
    T zi = a1 * (T)0 + b1 * l1inv[i] * ri;
    if (flags & F_HAS_NEXT) z1[i] = zi;
    if (flags & F_ADD_NEXT) xout[i] = 0.0 + (double)zi;  // degree 1 as a preconditioner
}

// step k: r_k = r_{k-1} - A z_k; x += z_k; z_{k+1} = an*z_k + bn*l1inv*r_k
template <class T, class TR>
__global__ void k_c4_step(int64_t n, SellView<T> A, const T *__restrict__ zk,
                          const TR *__restrict__ rin, T *__restrict__ rout,
                          const T *__restrict__ l1inv, double *__restrict__ x,
                          T *__restrict__ znext, T an, T bn, int flags, const int *__restrict__ done) {
    if (done && *done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T ri = row_sub(A, zk, i, (T)rin[i]);
    if (flags & F_WRITE_R) rout[i] = ri;
    T zi = zk[i];
    double xi = (flags & F_X_ZERO) ? 0.0 : x[i];
    xi = xi + (double)zi;
    if (flags & F_HAS_NEXT) {
        T zn = an * zi + bn * l1inv[i] * ri;
        znext[i] = zn;
        if (flags & F_ADD_NEXT) xi = xi + (double)zn;
    }
    x[i] = xi;
}

// ---- 4th kind inside the fused CG iteration --------------------------------
// init, preconditioner mode: the CG updates r -= alpha q, x += alpha p
// (cg_prologue, the k_cg_axpy arithmetic) and then z1 = b1 l1inv r
template <class T>
__global__ void k_c4_init_cg(int64_t n, const T *__restrict__ l1inv, T *__restrict__ z1, T a1,
                             T b1, CGFuse f) {
    if (f.st->done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T ri = (T)cg_prologue(f, i);
    z1[i] = a1 * (T)0 + b1 * l1inv[i] * ri;
}

// the last step of the preconditioner (degree d: step d-1, which also adds
// z_d) with the r'z partial of its 256-row block, as k_cg_dot forms it
template <class T, class TR>
__global__ __launch_bounds__(256) void k_c4_step_dot(int64_t n, SellView<T> A,
                                                     const T *__restrict__ zk,
                                                     const TR *__restrict__ rin,
                                                     const T *__restrict__ l1inv,
                                                     double *__restrict__ x, T an, T bn, int flags,
                                                     const double *__restrict__ rcg,
                                                     double *__restrict__ partials,
                                                     const int *__restrict__ done) {
    __shared__ double sm[4];
    if (*done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const T ri = row_sub(A, zk, i, (T)rin[i]);
        const T zi = zk[i];
        double xi = (flags & F_X_ZERO) ? 0.0 : x[i];
        xi = xi + (double)zi;
        const T zn = an * zi + bn * l1inv[i] * ri;
        xi = xi + (double)zn;
        x[i] = xi;
        v = rcg[i] * xi;
    }
    v = block_sum(v, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// ---- 1st kind -------------------------------------------------------------
// init: r0 = l1inv*(b - A x)*inv_rho; z0 = c*r0
template <class T>
__global__ void k_c1_init(int64_t n, SellView<T> A, const double *__restrict__ x,
                          const double *__restrict__ b, const T *__restrict__ l1inv,
                          T *__restrict__ r, T *__restrict__ z0, double *__restrict__ xout,
                          T inv_rho, T c, int flags, const int *__restrict__ done) {
    if (done && *done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T tmp = (flags & F_X_ZERO) ? (T)0 : row_dot(A, x, i);
    T ri = l1inv[i] * ((T)b[i] - tmp) * inv_rho;
    r[i] = ri;
    T zi = c * ri;
    z0[i] = zi;
    if (flags & F_ADD_NEXT) xout[i] = 0.0 + (double)zi;  // degree 1 as a preconditioner
}

// step k < d: x += z_{k-1}; r -= l1inv*(A z_{k-1})*inv_rho; z_k = c1 z_{k-1} + c2 r
template <class T>
__global__ void k_c1_step(int64_t n, SellView<T> A, const T *__restrict__ zp, T *__restrict__ r,
                          const T *__restrict__ l1inv, double *__restrict__ x,
                          T *__restrict__ znext, T inv_rho, T c1, T c2, int flags,
                          const int *__restrict__ done) {
    if (done && *done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T zi = zp[i];
    double xi = (flags & F_X_ZERO) ? 0.0 : x[i];
    xi = xi + (double)zi;
    T tmp = row_dot(A, zp, i);
    T ri = r[i] - l1inv[i] * tmp * inv_rho;
    r[i] = ri;
    T zn = c1 * zi + c2 * ri;
    znext[i] = zn;
    if (flags & F_ADD_NEXT) xi = xi + (double)zn;
    x[i] = xi;
}

// x += z (the last 1st-kind update when it cannot be fused)
template <class T>
__global__ void k_zero(int64_t n, T *__restrict__ v, const int *__restrict__ done) {
    if (done && *done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (T)0;
}

template <class T>
__global__ void k_add(int64_t n, double *__restrict__ x, const T *__restrict__ z, int xzero,
                      const int *__restrict__ done) {
    if (done && *done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (xzero ? 0.0 : x[i]) + (double)z[i];
}

// ---- plan -----------------------------------------------------------------
double cheb1_astar(int k) {
    // _cheb1_astar (:81-89): bisection on x = sqrt(a*), 60 halvings
    double lo = 1e-10, hi = 1.0 - 1e-10;
    for (int it = 0; it < 60; ++it) {
        double mid = (lo + hi) / 2.0;
        double x = mid;
        double phi = 8.0 * k * std::pow(1.0 - x * x, 2.0 * k) +
                     x * (std::pow(1.0 - x, 4.0 * k) - std::pow(1.0 + x, 4.0 * k));
        if (phi > 0) lo = mid; else hi = mid;
    }
    double m = (lo + hi) / 2.0;
    return m * m;
}

struct ChebPlan {
    uint64_t serial = next_serial();
    int kind = 4, degree = 4, prec = HS_PREC_FP64;
    int64_t n = 0;
    CSR A;
    DBuf<double> l1inv;
    DBuf<float> l1inv32;
    double rho = 1.0, a = 0.0;
    DBuf<uint8_t> r, za, zb;  // workspaces, element type by prec
    int zcur = 0;             // which of za/zb holds the current z
    Staging stage;

    size_t es() const { return prec == HS_PREC_FP64 ? 8 : 4; }
    template <class T> T *ptr(DBuf<uint8_t> &b) { return reinterpret_cast<T *>(b.get()); }
};

template <class T>
static void cheb_run(ChebPlan &s, double *x, const double *b, bool as_prec, hipStream_t st,
                     const int *done) {
    const int64_t n = s.n;
    const unsigned g = grid_for(n, 256);
    const SellView<T> Av = s.A.template view_as<T>();
    const T *l1;
    if constexpr (sizeof(T) == 8) l1 = s.l1inv.get();
    else l1 = s.l1inv32.get();
    T *r = s.ptr<T>(s.r);
    T *zb[2] = {s.ptr<T>(s.za), s.ptr<T>(s.zb)};
    const int d = s.degree;
    const int xz = as_prec ? F_X_ZERO : 0;
    if (s.kind == 4) {
        auto ab = [&](int k, T &ak, T &bk) {
            ak = (T)((double)(2 * k - 3) / (double)(2 * k + 1));
            bk = (T)((double)(8 * k - 4) / ((double)(2 * k + 1) * s.rho));
        };
        if (d == 0) {
            // r = b - A x; z = 0; x unchanged
            hipLaunchKernelGGL((k_c4_init<T>), dim3(g), dim3(256), 0, st, n, Av, x, b, l1, r,
                               zb[0], x, (T)0, (T)0, xz | F_WRITE_R, done);
            hipLaunchKernelGGL((k_zero<T>), dim3(g), dim3(256), 0, st, n, zb[0], done);
            if (as_prec) hipLaunchKernelGGL((k_zero<double>), dim3(g), dim3(256), 0, st, n, x, done);
            s.zcur = 0;
            check_launch();
            return;
        }
        T a1, b1;
        ab(1, a1, b1);
        // In preconditioner mode r0 = b is read straight from b by step 1.
        int fl = xz | F_HAS_NEXT | (as_prec ? 0 : F_WRITE_R);
        if (as_prec && d == 1) fl |= F_ADD_NEXT;
        hipLaunchKernelGGL((k_c4_init<T>), dim3(g), dim3(256), 0, st, n, Av, x, b, l1, r,
                           zb[0], x, a1, b1, fl, done);
        check_launch();
        int cur = 0;
        if (as_prec && d == 1) { s.zcur = 0; return; }
        for (int k = 1; k <= d; ++k) {
            const bool last = (k == d);
            if (as_prec && last) break;  // final r = r - A z_d is never read
            T an = 0, bn = 0;
            // r_k is read by step k+1; in preconditioner mode the last step's r is dead
            int f = (as_prec && k + 1 >= d) ? 0 : F_WRITE_R;
            if (k == 1) f |= xz;
            if (!last) {
                ab(k + 1, an, bn);
                f |= F_HAS_NEXT;
                if (as_prec && k + 1 == d) f |= F_ADD_NEXT;
            }
            if (as_prec && k == 1) {
                hipLaunchKernelGGL((k_c4_step<T, double>), dim3(g), dim3(256), 0, st, n, Av,
                                   zb[cur], b, r, l1, x, zb[cur ^ 1], an, bn, f, done);
            } else {
                hipLaunchKernelGGL((k_c4_step<T, T>), dim3(g), dim3(256), 0, st, n, Av,
                                   zb[cur], r, r, l1, x, zb[cur ^ 1], an, bn, f, done);
            }
            check_launch();
            if (!last) cur ^= 1;
        }
        s.zcur = cur;
        return;
    }
    // 1st kind
    const T inv_rho = (T)(1.0 / s.rho);
    const T c = (T)(2.0 / (1.0 + s.a));
    int fl = xz;
    if (as_prec && d == 1) fl |= F_ADD_NEXT;
    hipLaunchKernelGGL((k_c1_init<T>), dim3(g), dim3(256), 0, st, n, Av, x, b, l1, r,
                       zb[0], x, inv_rho, c, fl, done);
    check_launch();
    int cur = 0;
    if (d == 0) {
        if (as_prec) hipLaunchKernelGGL((k_zero<double>), dim3(g), dim3(256), 0, st, n, x, done);
        check_launch();
        s.zcur = 0;
        return;
    }
    if (as_prec && d == 1) { s.zcur = 0; return; }
    if (d == 1) {
        hipLaunchKernelGGL((k_add<T>), dim3(g), dim3(256), 0, st, n, x, zb[0], 0, done);
        check_launch();
        s.zcur = 0;
        return;
    }
    // scalar recurrence in FP64 (the reference's eltype(x)), :296-316
    const double ad = s.a;
    double rs = (1.0 - ad) / (1.0 + ad);
    for (int k = 1; k < d; ++k) {
        double rsn = 1.0 / (2.0 * (1.0 + ad) / (1.0 - ad) - rs);
        T c1 = (T)(rsn * rs);
        T c2 = (T)(4.0 * rsn / (1.0 - ad));
        int f = (k == 1) ? xz : 0;
        if (k + 1 == d) f |= F_ADD_NEXT;  // the final x += z needs no SpMV
        hipLaunchKernelGGL((k_c1_step<T>), dim3(g), dim3(256), 0, st, n, Av, zb[cur], r, l1,
                           x, zb[cur ^ 1], inv_rho, c1, c2, f, done);
        check_launch();
        cur ^= 1;
        rs = rsn;
    }
    s.zcur = cur;
}

// The 4th-kind preconditioner (degree >= 2) inside the fused CG iteration:
// same launches and arithmetic as cheb_run in preconditioner mode, with the
// CG's r and x updates in the init kernel and r'z in the last step.
template <class T>
static unsigned cheb_run_cg(ChebPlan &s, double *z, hipStream_t st, const CGFuse &f) {
    const int64_t n = s.n;
    const unsigned g = grid_for(n, 256);
    const SellView<T> Av = s.A.template view_as<T>();
    const T *l1;
    if constexpr (sizeof(T) == 8) l1 = s.l1inv.get();
    else l1 = s.l1inv32.get();
    T *r = s.ptr<T>(s.r);
    T *zb[2] = {s.ptr<T>(s.za), s.ptr<T>(s.zb)};
    const int d = s.degree;
    const int *done = reinterpret_cast<const int *>(reinterpret_cast<const char *>(f.st) +
                                                    offsetof(CGDev, done));
    auto ab = [&](int k, T &ak, T &bk) {
        ak = (T)((double)(2 * k - 3) / (double)(2 * k + 1));
        bk = (T)((double)(8 * k - 4) / ((double)(2 * k + 1) * s.rho));
    };
    T a1, b1;
    ab(1, a1, b1);
    hipLaunchKernelGGL((k_c4_init_cg<T>), dim3(g), dim3(256), 0, st, n, l1, zb[0], a1, b1, f);
    check_launch();
    int cur = 0;
    for (int k = 1; k <= d - 1; ++k) {
        const bool last = (k == d - 1);
        T an, bn;
        ab(k + 1, an, bn);
        int fl = F_HAS_NEXT | (k == 1 ? F_X_ZERO : 0) | (last ? F_ADD_NEXT : F_WRITE_R);
        if (!last) {
            if (k == 1)
                hipLaunchKernelGGL((k_c4_step<T, double>), dim3(g), dim3(256), 0, st, n, Av,
                                   zb[cur], (const double *)f.r, r, l1, z, zb[cur ^ 1], an, bn, fl,
                                   done);
            else
                hipLaunchKernelGGL((k_c4_step<T, T>), dim3(g), dim3(256), 0, st, n, Av, zb[cur],
                                   r, r, l1, z, zb[cur ^ 1], an, bn, fl, done);
            cur ^= 1;
        } else if (k == 1) {
            hipLaunchKernelGGL((k_c4_step_dot<T, double>), dim3(g), dim3(256), 0, st, n, Av,
                               zb[cur], (const double *)f.r, l1, z, an, bn, fl, f.r, f.partials,
                               done);
        } else {
            hipLaunchKernelGGL((k_c4_step_dot<T, T>), dim3(g), dim3(256), 0, st, n, Av, zb[cur],
                               r, l1, z, an, bn, fl, f.r, f.partials, done);
        }
        check_launch();
    }
    s.zcur = cur;
    return g;
}

bool cheb_can_fuse(const ChebPlan &s) { return s.kind == 4 && s.degree >= 2 && s.n > 0; }

unsigned cheb_apply_fused(ChebPlan &s, double *z, hipStream_t st, const CGFuse &f) {
    if (s.prec == HS_PREC_FP64) return cheb_run_cg<double>(s, z, st, f);
    return cheb_run_cg<float>(s, z, st, f);
}

void cheb_apply_device(ChebPlan &s, double *x, const double *b, bool as_prec, hipStream_t st,
                       const int *done) {
    if (s.n == 0) return;
    if (s.prec == HS_PREC_FP64) cheb_run<double>(s, x, b, as_prec, st, done);
    else cheb_run<float>(s, x, b, as_prec, st, done);
}

int64_t cheb_rows(const ChebPlan &s) { return s.n; }
uint64_t cheb_serial(const ChebPlan &s) { return s.serial; }

}  // namespace hs

struct hs_cheb : hs::ChebPlan {};

using namespace hs;

extern "C" {

int hs_cheb_build(int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                  int kind, int degree, int power_iters, int precision, void *stream,
                  hs_cheb_t *out) {
    return guard([&] {
        HS_REQUIRE(out, HS_EINVAL, "out handle pointer is NULL");
        *out = nullptr;
        HS_REQUIRE(kind == 1 || kind == 4, HS_EINVAL, "kind must be 1 (ChebyshevFirst) or 4 (ChebyshevFourth)");
        HS_REQUIRE(degree >= 0, HS_EINVAL, "degree must be non-negative");
        HS_REQUIRE(power_iters >= 0, HS_EINVAL, "power_iters must be non-negative");
        HS_REQUIRE(precision == HS_PREC_FP64 || precision == HS_PREC_FP32, HS_EUNSUPPORTED,
                   "Chebyshev smoothers run in FP64 or FP32");
        HS_REQUIRE(n > 0 && colptr && rowval && nzval, HS_EINVAL, "empty or NULL matrix");
        hipStream_t st = as_stream(stream);
        std::unique_ptr<hs_cheb> s(new hs_cheb());
        s->kind = kind;
        s->degree = degree;
        s->prec = precision;
        s->n = n;
        csr_create_device(s->A, n, colptr, rowval, nzval, HS_FMT_CSC, st);  // SparseMatrixCSR(A) (:144)
        s->l1inv.alloc(n);
        if (precision == HS_PREC_FP32) {
            s->l1inv32.alloc(n);
            s->A.ensure_fp32(st);
        }
        const unsigned g = grid_for(n, 256);
        hipLaunchKernelGGL(k_l1inv, dim3(g), dim3(256), 0, st, s->A.view(), s->l1inv.get(),
                           s->l1inv32.get());
        check_launch();
        // power iteration for rho(M^-1 A) in FP64 (:40-54)
        {
            DBuf<double> v(n), w(n), partials(g);
            DBuf<PowState> ps(1);
            PowState init{1.0, 0.0, 0};
            HS_CHECK_HIP(hipMemcpyAsync(ps.get(), &init, sizeof init, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, st, n, v.get(), 1.0 / std::sqrt((double)n));
            check_launch();
            for (int it = 0; it < power_iters; ++it) {
                hipLaunchKernelGGL(k_pow_spmv, dim3(g), dim3(256), 0, st, n, s->A.view(),
                                   s->l1inv.get(), v.get(), w.get(), partials.get(), ps.get());
                hipLaunchKernelGGL(k_pow_fin, dim3(1), dim3(256), 0, st, partials.get(), g, ps.get());
                hipLaunchKernelGGL(k_pow_scale, dim3(g), dim3(256), 0, st, n, w.get(), v.get(),
                                   ps.get());
            }
            check_launch();
            PowState res;
            HS_CHECK_HIP(hipMemcpyAsync(&res, ps.get(), sizeof res, hipMemcpyDeviceToHost, st));
            HS_CHECK_HIP(hipStreamSynchronize(st));
            s->rho = res.rho;
        }
        s->a = (kind == 1) ? cheb1_astar(degree) : 0.0;
        const size_t es = s->es();
        s->r.alloc(n * es);
        s->za.alloc(n * es);
        s->zb.alloc(n * es);
        HS_CHECK_HIP(hipMemsetAsync(s->r.get(), 0, n * es, st));
        HS_CHECK_HIP(hipMemsetAsync(s->za.get(), 0, n * es, st));
        HS_CHECK_HIP(hipMemsetAsync(s->zb.get(), 0, n * es, st));
        HS_CHECK_HIP(hipStreamSynchronize(st));
        *out = s.release();
    });
}

static int cheb_call(hs_cheb_t s, double *x, const double *b, int loc, void *stream, int flags,
                     bool as_prec) {
    return guard([&] {
        HS_REQUIRE(s, HS_EINVAL, "NULL smoother handle");
        HS_REQUIRE(x && b, HS_EINVAL, "NULL vector");
        HS_REQUIRE((const double *)x != b, HS_EINVAL, "x and b must not alias");
        hipStream_t st = as_stream(stream);
        if (loc == HS_LOC_DEVICE) {
            cheb_apply_device(*s, x, b, as_prec, st, nullptr);
            finish(st, flags);
            return;
        }
        const size_t bytes = s->n * 8;
        s->stage.ensure(s->n);
        if (!as_prec)
            HS_CHECK_HIP(hipMemcpyAsync(s->stage.a.get(), x, bytes, hipMemcpyHostToDevice, st));
        HS_CHECK_HIP(hipMemcpyAsync(s->stage.b.get(), b, bytes, hipMemcpyHostToDevice, st));
        cheb_apply_device(*s, s->stage.a.get(), s->stage.b.get(), as_prec, st, nullptr);
        HS_CHECK_HIP(hipMemcpyAsync(x, s->stage.a.get(), bytes, hipMemcpyDeviceToHost, st));
        HS_CHECK_HIP(hipStreamSynchronize(st));
    });
}

int hs_cheb_smooth(hs_cheb_t s, double *x, const double *b, int loc, void *stream, int flags) {
    return cheb_call(s, x, b, loc, stream, flags, false);
}

int hs_cheb_apply(hs_cheb_t s, double *z, const double *r, int loc, void *stream, int flags) {
    return cheb_call(s, z, r, loc, stream, flags, true);
}

int hs_cheb_info(hs_cheb_t s, double *out) {
    return guard([&] {
        HS_REQUIRE(s && out, HS_EINVAL, "NULL argument");
        out[0] = s->rho;
        out[1] = s->a;
        out[2] = s->degree;
        out[3] = s->kind;
        out[4] = (double)s->n;
    });
}

int hs_cheb_get(hs_cheb_t s, int what, double *out) {
    return guard([&] {
        HS_REQUIRE(s && out, HS_EINVAL, "NULL argument");
        HS_REQUIRE(what >= 0 && what <= 3, HS_EINVAL, "unknown field");
        const int64_t n = s->n;
        if (what == 3) {  // tmp: the fused kernels never materialise A*z
            for (int64_t i = 0; i < n; ++i) out[i] = 0.0;
            return;
        }
        const void *src = nullptr;
        if (what == 0) {
            HS_CHECK_HIP(hipMemcpy(out, s->l1inv.get(), n * 8, hipMemcpyDeviceToHost));
            return;
        }
        src = (what == 1) ? s->r.get() : (s->zcur == 0 ? s->za.get() : s->zb.get());
        if (s->prec == HS_PREC_FP64) {
            HS_CHECK_HIP(hipMemcpy(out, src, n * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp(n);
            HS_CHECK_HIP(hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) out[i] = tmp[i];
        }
    });
}

hs_csr_t hs_cheb_matrix(hs_cheb_t s) {
    return s ? reinterpret_cast<hs_csr_t>(static_cast<CSR *>(&s->A)) : nullptr;
}

int hs_cheb_destroy(hs_cheb_t s) {
    return guard([&] { delete s; });
}

}  // extern "C"
