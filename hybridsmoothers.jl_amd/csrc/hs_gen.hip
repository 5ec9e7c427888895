// hs_gen.hip -- host generator of the serving workload's systems: 5-point
// (nz = 1) or 7-point variable-coefficient Poisson on an nx x ny x nz grid
// with lognormal face coefficients, Dirichlet boundary faces included in the
// diagonal.  Host code only (no kernels): it produces the 1-based CSR arrays a
// client submits, fast enough to stream thousands of systems.
//
// Coefficients: face f (x faces first, then y, then z; each family numbered
// x fastest) has k_f = exp(sigma * g_f) with g_f standard normal from a
// counter-based generator: the pair (2m, 2m+1) comes from Box-Muller on two
// splitmix64 words of counters 2m and 2m+1 under the seed, cos for the even
// face and sin for the odd one.  Row i (x fastest) holds, in ascending
// column order, -k of its -z, -y, -x faces, the diagonal
// k_xm + k_xp + k_ym + k_yp + k_zm + k_zp (boundary faces included, axes of
// length 1 have no faces), then -k of its +x, +y, +z faces.  The matrix is
// symmetric, so the same arrays are its CSC.
#include "hs_common.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace hs {
namespace gen {

static inline uint64_t mix64(uint64_t seed, uint64_t ctr) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + ctr * 0xD1B54A32D192ED03ULL + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// k of faces [f0, f1)
static void faces(uint64_t seed, double sigma, int64_t f0, int64_t f1, double *k) {
    const double two53 = 1.0 / 9007199254740992.0;
    const double twopi = 6.283185307179586;
    for (int64_t f = f0; f < f1; ++f) {
        const uint64_t m = (uint64_t)f >> 1;
        const double u1 = (double)((mix64(seed, 2 * m) >> 11) + 1) * two53;  // (0, 1]
        const double u2 = (double)(mix64(seed, 2 * m + 1) >> 11) * two53;    // [0, 1)
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double th = twopi * u2;
        const double g = (f & 1) ? r * std::sin(th) : r * std::cos(th);
        k[f - f0] = std::exp(sigma * g);
    }
}

template <class F>
static void parallel_for(int64_t n, int64_t grain, F &&body) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / grain));
    if (nt <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] { body(n * t / nt, n * (t + 1) / nt); });
    for (auto &x : th) x.join();
}

}  // namespace gen
}  // namespace hs

using namespace hs;

extern "C" {

int hs_problem_lognormal(int64_t nx, int64_t ny, int64_t nz, double sigma, uint64_t seed,
                         int64_t *rowptr, int64_t *colval, double *nzval, int64_t *nnz_out) {
    return guard([&] {
        HS_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, HS_EINVAL, "grid sizes must be >= 1");
        HS_REQUIRE(nnz_out, HS_EINVAL, "NULL nnz_out");
        const int64_t n = nx * ny * nz;
        // couplings: (nx-1) per x line, etc.; each appears twice (both rows)
        const int64_t off = 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1));
        *nnz_out = n + off;
        if (!colval) return;  // size query
        HS_REQUIRE(rowptr && nzval, HS_EINVAL, "NULL output array");
        const bool hx = nx > 1, hy = ny > 1, hz = nz > 1;
        const int64_t Fx = hx ? (nx + 1) * ny * nz : 0;
        const int64_t Fy = hy ? nx * (ny + 1) * nz : 0;
        const int64_t Fz = hz ? nx * ny * (nz + 1) : 0;
        std::vector<double> k(Fx + Fy + Fz);
        gen::parallel_for(Fx + Fy + Fz, 1 << 16, [&](int64_t a, int64_t b) {
            gen::faces(seed, sigma, a, b, k.data() + a);
        });
        const double *kx = k.data(), *ky = kx + Fx, *kz = ky + Fy;
        // row lengths are 1 + the number of interior neighbours
        rowptr[0] = 1;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            const int64_t len = 1 + (x > 0) + (x < nx - 1) + (y > 0) + (y < ny - 1) + (z > 0) +
                                (z < nz - 1);
            rowptr[i + 1] = rowptr[i] + len;
        }
        gen::parallel_for(n, 1 << 15, [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) {
                const int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
                const double xm = hx ? kx[(z * ny + y) * (nx + 1) + x] : 0.0;
                const double xp = hx ? kx[(z * ny + y) * (nx + 1) + x + 1] : 0.0;
                const double ym = hy ? ky[(z * (ny + 1) + y) * nx + x] : 0.0;
                const double yp = hy ? ky[(z * (ny + 1) + y + 1) * nx + x] : 0.0;
                const double zm = hz ? kz[(z * ny + y) * nx + x] : 0.0;
                const double zp = hz ? kz[((z + 1) * ny + y) * nx + x] : 0.0;
                int64_t p = rowptr[i] - 1;
                auto put = [&](int64_t c, double v) {
                    colval[p] = c + 1;
                    nzval[p] = v;
                    ++p;
                };
                if (z > 0) put(i - nx * ny, -zm);
                if (y > 0) put(i - nx, -ym);
                if (x > 0) put(i - 1, -xm);
                put(i, xm + xp + ym + yp + zm + zp);
                if (x < nx - 1) put(i + 1, -xp);
                if (y < ny - 1) put(i + nx, -yp);
                if (z < nz - 1) put(i + nx * ny, -zp);
            }
        });
    });
}

}  // extern "C"
