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
uint64_t l1gs_serial(const L1GSPlan &h);

// Chebyshev smoother: smooth!(x, s, b), or z = 0; smooth!(z, s, b) when as_prec.
void cheb_apply_device(ChebPlan &s, double *x, const double *b, bool as_prec, hipStream_t st,
                       const int *done);
int64_t cheb_rows(const ChebPlan &s);
uint64_t cheb_serial(const ChebPlan &s);

void spmv_device(const CSR &A, const double *x, double *y, hipStream_t st);

}  // namespace hs
