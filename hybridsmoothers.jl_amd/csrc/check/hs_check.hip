// hs_check.hip -- host builds of device arithmetic of the sweeps, for the CPU
// tests (libhs_check.so, a library of its own: not part of libhsmooth.so).
// The functions below run the very inline code the kernels compile
// (hs_sweep.hpp) on the host, element by element.
#include "../hs_sweep.hpp"

extern "C" {

// q[i] = z[i] / d[i] as the FP64 sweeps form it: the reciprocal r = 1/d
// (correctly rounded, once per row) and quot_fast; safe[i] = whether the
// sweep's per-partition guard lets d[i] take that path.
void hs_check_quot_f64(int64_t n, const double *z, const double *d, double *q, int32_t *safe) {
    for (int64_t i = 0; i < n; ++i) {
        const double r = 1.0 / d[i];
        q[i] = hs::quot_fast(z[i], d[i], r);
        safe[i] = hs::safe_divisor(d[i]) ? 1 : 0;
    }
}

void hs_check_quot_f32(int64_t n, const float *z, const float *d, float *q, int32_t *safe) {
    for (int64_t i = 0; i < n; ++i) {
        const float r = 1.0f / d[i];
        q[i] = hs::quot_fast(z[i], d[i], r);
        safe[i] = hs::safe_divisor(d[i]) ? 1 : 0;
    }
}

}  // extern "C"
