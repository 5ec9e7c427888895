// hs_plans.hpp -- device entry points of the preconditioner plans, for the
// translation units that drive them (CG, batched and serving paths).
#pragma once

#include "hs_sparse.hpp"

namespace hs {

// Device-resident scalars of one CG solve.
struct CGDev {
    double gamma, delta, gnext, beta;
    double rnorm, rnorm0, eps, atol, rtol;
    double res2, b2;
    long long iter, itmax;
    int done, status;
};

// Scalar steps of the CG recurrence (one thread), shared by the single-system
// and the batched solvers so that both run the same arithmetic.
// After gamma0 = r0'z0: sets rnorm0, eps and the trivial-convergence test.
// Returns true when the solve is finished (st.status says why).
__device__ __forceinline__ bool cg_scalars_start(CGDev &st, double gamma0) {
    st.gamma = gamma0;
    const double rn = sqrt(gamma0);
    st.rnorm0 = rn;
    st.rnorm = rn;
    st.eps = st.atol + st.rtol * rn;
    st.iter = 0;
    st.status = 0;
    bool done = (gamma0 == 0.0 || rn <= st.eps);
    if (!done && st.itmax <= 0) {
        done = true;
        st.status = 1;
    }
    return done;
}

// After gamma' = r'z of an iteration: Krylov.jl's stopping rule
// sqrt(r'z) <= eps, then beta = gamma'/gamma and the iteration limit.
__device__ __forceinline__ bool cg_scalars_next(CGDev &st, double gnext) {
    st.gnext = gnext;
    const double rn = sqrt(gnext);
    st.rnorm = rn;
    st.iter += 1;
    if ((rn <= st.eps) || (rn + 1.0 <= 1.0)) {
        st.status = 0;
        return true;
    }
    st.beta = gnext / st.gamma;
    st.gamma = gnext;
    if (st.iter >= st.itmax) {
        st.status = 1;
        return true;
    }
    return false;
}

// delta = p'q; false (breakdown, status 2) when it is not positive.
__device__ __forceinline__ bool cg_scalars_delta(CGDev &st, double delta) {
    st.delta = delta;
    if (!(delta > 0.0)) {
        st.status = 2;
        return false;
    }
    return true;
}

// Row i of the fused SpMV of CG iteration k: p_k = z + beta p_{k-1} (p_0 = z)
// formed on the fly for row i and its columns, q_i = (A p_k)_i in row order.
// Stores p_k[i] and q[i]; returns p_k[i] * q_i (the p'q partial of the row).
// The row's entries are taken in batches of 8: all column indices and values
// of a batch are loaded before any gather (entries past the slice width load
// the slice's last entry again and are not added), so a wave keeps its loads
// in flight instead of one dependent index -> value chain per entry; the
// products are still added one by one in row order.
// CF: 0 32-bit columns, 1 16-bit offsets from the row, 2 16-bit deltas from the
// (slice, k) bases
template <int CF, class TZ>
__device__ __forceinline__ double cg_spmv_row_t(const SellView<double> &A,
                                                const TZ *__restrict__ z,
                                                const double *__restrict__ pold,
                                                double *__restrict__ pnew, double *__restrict__ q,
                                                bool first, double beta, int64_t i) {
    constexpr int B = 8;
    const int64_t s = i >> 6;
    const int64_t p0 = A.sptr[s] + (i & 63);
    const int w = (int)((A.sptr[s + 1] - A.sptr[s]) >> 6);  // slice width, same in the wave
    double acc = 0.0;
    for (int k0 = 0; k0 < w; k0 += B) {
        int32_t c[B];
        double v[B], zc[B], pc[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            // every load is issued, with no branch around it: past the slice
            // width (the same in the whole wave) the slice's last entry is
            // loaded again (its lines are in cache) and never added.  Loads
            // under a per-entry branch were each waited for inside the branch,
            // so the batch's entries reached memory one after another.
            const int ke = (k0 + k < w) ? k0 + k : w - 1;
            const int64_t p = p0 + 64 * (int64_t)ke;
            // the matrix is streamed once per product: nontemporal, so it
            // does not displace the gathered z and p from L2
            if constexpr (CF == 1)
                c[k] = (int32_t)(i + __builtin_nontemporal_load(A.col16 + p));
            else if constexpr (CF == 2)
                c[k] = A.kbase[p >> 6] + (int32_t)__builtin_nontemporal_load(A.cdelta + p);
            else
                c[k] = __builtin_nontemporal_load(A.col + p);
            v[k] = __builtin_nontemporal_load(A.val + p);
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            zc[k] = (double)z[c[k]];
            pc[k] = first ? 0.0 : pold[c[k]];
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            if (k0 + k < w) {
                const double pv = first ? zc[k] : zc[k] + beta * pc[k];
                acc = acc + v[k] * pv;
            }
        }
    }
    const double zi = (double)z[i];
    const double pi = first ? zi : zi + beta * pold[i];
    pnew[i] = pi;
    q[i] = acc;
    return pi * acc;
}
// z in FP64, or in FP32 when the preconditioner computes in FP32 (its values
// are then FP32 numbers, and widening them gives the same doubles)
template <class TZ>
__device__ __forceinline__ double cg_spmv_row(const SellView<double> &A,
                                              const TZ *__restrict__ z,
                                              const double *__restrict__ pold,
                                              double *__restrict__ pnew, double *__restrict__ q,
                                              bool first, double beta, int64_t i) {
    if (A.col16) return cg_spmv_row_t<1, TZ>(A, z, pold, pnew, q, first, beta, i);
    if (A.kbase) return cg_spmv_row_t<2, TZ>(A, z, pold, pnew, q, first, beta, i);
    return cg_spmv_row_t<0, TZ>(A, z, pold, pnew, q, first, beta, i);
}

// CG vector updates fused into a symmetric L1GS sweep: before the sweep each
// row does r -= alpha q and x += alpha p (alpha = gamma/delta), the sweep
// reads the updated r, and afterwards each block stores its partial of r'z
// (partials[block]; the count is returned by l1gs_apply_fused).  p of
// iteration k is pb when k is odd and pa otherwise (k = st->iter).
struct CGFuse {
    double *r;
    const double *q;
    const double *pa;
    const double *pb;
    double *x;
    double *partials;
    const CGDev *st;
    // when set (FP32 and BF16 smoothers, whose results are FP32 numbers) z is
    // stored here in FP32 instead of the FP64 vector: the same values, 4 bytes
    // a row less written here and read by the next product
    float *z32 = nullptr;
};

// Row prologue of the fused CG form: r -= alpha q; x += alpha p (the k_cg_axpy
// arithmetic); returns the updated r, the preconditioner's input.
__device__ __forceinline__ double cg_prologue(const CGFuse &f, int64_t i) {
    const double alpha = f.st->gamma / f.st->delta;
    const double *p = (f.st->iter & 1) ? f.pb : f.pa;
    const double ri = f.r[i] - alpha * f.q[i];
    f.r[i] = ri;
    f.x[i] = f.x[i] + alpha * p[i];
    return ri;
}

// The fused smoother form, one round of AMG.smooth! (ext/HybridSmoothersAMGExt.jl:
// 29-35): r_i = (A x)_i - b_i computed in the sweep's prologue (row sum in
// column order, then the subtraction, as mul! followed by .-= rounds them),
// z = M^-1 r, and x_new_i = x_i - z_i stored to xout (not x: other partitions
// still read x for their residuals).
struct SmoothFuse {
    SellView<double> A;
    const double *x;
    const double *b;
    double *xout;
};

struct L1GSPlan;
struct ChebPlan;

// y = M^-1 x with an L1GS plan (x, y device; may alias).  When `done` is not
// NULL and *done != 0 on the device the launch does nothing.
void l1gs_apply_device(const L1GSPlan &h, const double *x, double *y, hipStream_t st,
                       const int *done);
// The fused CG form (symmetric packed sweeps with partsize <= 64 only; check
// l1gs_can_fuse first): z = M^-1 (r - alpha q) with the updates above.
// Returns the number of r'z partials written.
bool l1gs_can_fuse(const L1GSPlan &h);
unsigned l1gs_apply_fused(const L1GSPlan &h, double *z, hipStream_t st, const CGFuse &f);
int64_t l1gs_rows(const L1GSPlan &h);
// smoother precision of the plan (HS_PREC_*)
int l1gs_precision(const L1GSPlan &h);
uint64_t l1gs_serial(const L1GSPlan &h);

// The 4th-kind smoother (degree >= 2) as the fused CG's preconditioner: z =
// M^-1 (r - alpha q) with the CGFuse updates; returns the r'z partial count.
bool cheb_can_fuse(const ChebPlan &s);
unsigned cheb_apply_fused(ChebPlan &s, double *z, hipStream_t st, const CGFuse &f);

// Chebyshev smoother: smooth!(x, s, b), or z = 0; smooth!(z, s, b) when as_prec.
void cheb_apply_device(ChebPlan &s, double *x, const double *b, bool as_prec, hipStream_t st,
                       const int *done);
int64_t cheb_rows(const ChebPlan &s);
uint64_t cheb_serial(const ChebPlan &s);

void spmv_device(const CSR &A, const double *x, double *y, hipStream_t st);

// One fused round xout = x - M^-1 (A x - b) (packed layouts, partsize <= 64:
// check l1gs_can_smooth_fused first); x and xout must not alias.
bool l1gs_can_smooth_fused(const L1GSPlan &h);
void l1gs_smooth_fused(const L1GSPlan &h, const SmoothFuse &sf, hipStream_t st);

// `iters` times x -= M^-1 (A x - b) on device vectors (r, z: n-vectors of scratch).
void l1gs_smooth_device(const CSR &A, const L1GSPlan &P, double *x, const double *b, int iters,
                        double *r, double *z, hipStream_t st);

}  // namespace hs
