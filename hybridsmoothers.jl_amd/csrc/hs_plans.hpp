// hs_plans.hpp -- device entry points of the preconditioner plans, for the
// translation units that drive them (CG, batched and serving paths).
#pragma once

#include "hs_sparse.hpp"

namespace hs {

struct L1GSPlan;
struct ChebPlan;

// y = M^-1 x with an L1GS plan (x, y device; may alias).  When `done` is not
// NULL and *done != 0 on the device the launch does nothing.
void l1gs_apply_device(const L1GSPlan &h, const double *x, double *y, hipStream_t st,
                       const int *done);
int64_t l1gs_rows(const L1GSPlan &h);
uint64_t l1gs_serial(const L1GSPlan &h);

// Chebyshev smoother: smooth!(x, s, b), or z = 0; smooth!(z, s, b) when as_prec.
void cheb_apply_device(ChebPlan &s, double *x, const double *b, bool as_prec, hipStream_t st,
                       const int *done);
int64_t cheb_rows(const ChebPlan &s);
uint64_t cheb_serial(const ChebPlan &s);

void spmv_device(const CSR &A, const double *x, double *y, hipStream_t st);

}  // namespace hs
