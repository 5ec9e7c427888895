// hs_batch.hip -- batched persistent PCG: many SPD systems of uneven size
// solved together, each with an L1GS symmetric preconditioner or a 4th-kind
// l1-Jacobi Chebyshev smoother, as a serving engine (submit / step / collect).
//
// No counterpart in the reference, whose users solve one system per call
// (solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P), README.md:27-29); every
// system here runs exactly the arithmetic of hs_pcg_solve with that
// preconditioner (FP64 CG, Krylov.jl's stopping rule, x0 = 0).
//
// Device state:
//  * one arena of HBM, carved by a host first-fit allocator into one block
//    per resident system: its sliced-ELL matrix, row lengths, right-hand side,
//    the smoother's arrays (packed triangles and D_Dl1, or l1inv, FP32 values
//    and the Chebyshev z and r vectors), CG vectors and reduction partials;
//  * a slot table (descriptor + CG scalars per resident system);
//  * a tile table: one entry per (system, 256-row tile) of every active system.
//
// One CG iteration of every active system is one "pass" of four launches:
//   k_b_spmv      tiles: p = z + beta p, q = A p, p'q partial   (or the true
//                 residual partials of a converged system)
//   k_b_fin_delta systems: delta = p'q (or ||b - Ax||^2, ||b||^2)
//   k_b_sweep     tiles: r -= alpha q, x += alpha p, z = M^-1 r, r'z partial
//                 (a newly admitted system: r = b, x = 0, z = M^-1 b)
//   k_b_fin_gamma systems: the CG scalars and the stopping test
// With the Chebyshev smoother of degree d the sweep is replaced by d - 1
// launches (chebyshev.jl:158-183 as a left preconditioner, x0 = 0: the A*0 of
// the first step and the last step's SpMV are dead):
//   k_bc_first    tiles: r -= alpha q, x += alpha p, z_1 = b_1 l1inv r formed
//                 at every column it is gathered at, r_1 = r - A z_1, z_2,
//                 z_cg = z_1 (+ z_2 and the r'z partial when d = 2)
//   k_bc_step     tiles, k = 2 .. d-1: r_k = r_{k-1} - A z_k, the next z,
//                 z_cg += z_k (+ z_d and the r'z partial at k = d-1)
// and a system's rho comes from power-iteration steps (chebyshev.jl:40-54) run
// by the same launches before its first CG step (phase PH_POWER: w = (A v) .*
// l1inv and ||w||^2 in k_b_spmv, rho = ||w|| in k_b_fin_delta, v = w / rho in
// k_bc_first, the step count in k_b_fin_gamma).
// The sweep is persistent: a fixed grid, each workgroup walking a contiguous
// run of the tile table (neighbouring tiles of a system share the workgroup's
// XCD and its L2); the SpMV takes one tile per workgroup of a large grid.  A
// fixed number of passes is captured once per SpMV grid size (SPMV_GRIDS) in
// HIP graphs; admission and retirement only rewrite the device tables between
// graph launches, so no graph is ever re-captured.
//
// Every per-system reduction sums that system's own tile partials in a fixed
// order, so a system's result does not depend on which others share the
// batch, and it equals hs_pcg_solve's bit for bit (same kernels' arithmetic,
// same 256-row partials, same finalize tree).
#include "hs_sweep.hpp"

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

namespace hs {
namespace batch {

constexpr int TILE = 256;         // rows per tile (= one 256-thread workgroup)
// Workgroups of the SpMV launch: one graph is captured per grid size and a
// step replays the smallest that covers the tile table (past 16.8 M rows the
// largest strides).  A workgroup past the tile count exits at once, but
// dispatching 65,536 of them still costs about 14 us per launch, which a
// nearly empty engine would pay on every pass.
constexpr int SPMV_GRIDS[3] = {1024, 8192, 65536};

enum { PH_EMPTY = 0, PH_INIT = 1, PH_ITER = 2, PH_RESID = 3, PH_DONE = 4, PH_POWER = 5 };

// Descriptor of a resident system (device array, one per slot): byte offsets
// of its arrays in the arena.  The kernels add them to the arena base, a
// kernel argument, so the compiler knows every access is to global memory
// (pointers loaded from memory would be generic: flat loads, which are also
// counted against the LDS counter).
struct Sys {
    int64_t sptr, col, val, T, D, x, r, z, pa, pb, q, b, part;
    int64_t l1, l1s, v32, rb, za, zb, rc;  // Chebyshev
    int64_t n, nparts;
    int32_t ntiles, c16;  // c16: columns stored as 16-bit offsets from the row
};

// The same, resolved against the arena base.
struct SysP {
    const int64_t *sptr;
    const int32_t *col;
    const int16_t *col16;  // the column section as 16-bit offsets, or null
    const double *val;
    const uint8_t *T;  // packed triangles (storage type of the engine's precision)
    const uint8_t *D;  // D_Dl1 (FP64, or FP32 for the reduced-precision smoothers)
    double *x, *r, *z, *pa, *pb, *q;
    const double *b;
    double *part;  // two partials per tile
    // Chebyshev: l1inv in FP64 (power iteration) and in the smoother's type,
    // the FP32 matrix values (FP32 smoother), the CG residual's second buffer
    // (see cheb_rnew), the z ping-pong and r of the recurrence (smoother's type)
    const double *l1;
    const uint8_t *l1s, *v32;
    double *rb;
    uint8_t *za, *zb, *rc;
    int64_t n, nparts;
    int32_t ntiles;
};

__device__ __forceinline__ SysP load_sys(const Sys *__restrict__ sys, int s, uint8_t *arena) {
    const Sys d = sys[s];
    SysP S;
    S.sptr = reinterpret_cast<const int64_t *>(arena + d.sptr);
    S.col = reinterpret_cast<const int32_t *>(arena + d.col);
    S.col16 = d.c16 ? reinterpret_cast<const int16_t *>(arena + d.col) : nullptr;
    S.val = reinterpret_cast<const double *>(arena + d.val);
    S.T = arena + d.T;
    S.D = arena + d.D;
    S.x = reinterpret_cast<double *>(arena + d.x);
    S.r = reinterpret_cast<double *>(arena + d.r);
    S.z = reinterpret_cast<double *>(arena + d.z);
    S.pa = reinterpret_cast<double *>(arena + d.pa);
    S.pb = reinterpret_cast<double *>(arena + d.pb);
    S.q = reinterpret_cast<double *>(arena + d.q);
    S.b = reinterpret_cast<const double *>(arena + d.b);
    S.part = reinterpret_cast<double *>(arena + d.part);
    S.l1 = reinterpret_cast<const double *>(arena + d.l1);
    S.l1s = arena + d.l1s;
    S.v32 = arena + d.v32;
    S.rb = reinterpret_cast<double *>(arena + d.rb);
    S.za = arena + d.za;
    S.zb = arena + d.zb;
    S.rc = arena + d.rc;
    S.n = d.n;
    S.nparts = d.nparts;
    S.ntiles = d.ntiles;
    return S;
}

struct State {
    CGDev cg;
    int phase, pad;
    double rho;      // Chebyshev: the power iteration's current estimate (1 before it)
    int pw, pmax;    // power-iteration steps done and to do
    int pstop, pad2; // rho == 0 ended the power iteration
};

struct Tile {
    int32_t slot, tile;
};

struct Counts {
    int32_t ntiles, nactive;
};

// contiguous run of the tile table for this workgroup
__device__ __forceinline__ void tile_range(int nt, int &t0, int &t1) {
    const int per = (nt + (int)gridDim.x - 1) / (int)gridDim.x;
    t0 = (int)blockIdx.x * per;
    t1 = min(t0 + per, nt);
}

// A workgroup's run of tiles mostly belongs to one or a few systems, in tile
// order: the system's descriptor and scalars are loaded when the slot changes
// (not per tile), and the next tile entry is fetched one tile ahead, so a
// tile starts its vector loads without a chain of dependent scalar loads.

// The SpMV runs one tile per workgroup over a large fixed grid (workgroups
// past the tile count exit at once; a grid smaller than the tile count
// strides): the hardware dispatcher then balances the tiles, which measured
// faster than persistent runs for this short, load-bound kernel.
__global__ __launch_bounds__(256) void k_b_spmv(const Tile *__restrict__ tiles,
                                                const Counts *__restrict__ cnt,
                                                const Sys *__restrict__ sys,
                                                const State *__restrict__ st,
                                                uint8_t *__restrict__ arena) {
    __shared__ double sm[4];
    const int nt = cnt->ntiles;
    int cur = -1, ph = PH_EMPTY;
    SysP S{};
    bool first = false;
    double beta = 0.0;
    const double *pold = nullptr;
    double *pnew = nullptr;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.slot != cur) {
            cur = tl.slot;
            ph = st[cur].phase;
            S = load_sys(sys, cur, arena);
            const long long k = st[cur].cg.iter;
            first = (k == 0);
            beta = first ? 0.0 : st[cur].cg.beta;
            pold = (k & 1) ? S.pa : S.pb;
            pnew = (k & 1) ? S.pb : S.pa;
        }
        if (ph != PH_ITER && ph != PH_RESID && ph != PH_POWER) continue;  // uniform in the workgroup
        const SellView<double> A{S.sptr, S.col, S.val, S.n, S.col16};
        const int64_t i = (int64_t)tl.tile * TILE + threadIdx.x;
        if (ph == PH_ITER) {
            double d = 0.0;
            if (i < S.n) d = cg_spmv_row(A, S.z, pold, pnew, S.q, first, beta, i);
            d = block_sum(d, sm);
            if (threadIdx.x == 0) S.part[2 * tl.tile] = d;
        } else if (ph == PH_POWER) {
            // power-iteration step of a Chebyshev system: w = (A v) .* l1inv
            // (v in pa, w to pb) and the tile's share of ||w||^2, as k_pow_spmv
            double sq = 0.0;
            if (i < S.n) {
                double wi = row_dot(A, S.pa, i);
                wi = wi * S.l1[i];
                S.pb[i] = wi;
                sq = wi * wi;
            }
            sq = block_sum(sq, sm);
            if (threadIdx.x == 0) S.part[2 * tl.tile] = sq;
        } else {
            // true residual of a finished system, ||b - A x||^2 and ||b||^2
            double t2 = 0.0, b2 = 0.0;
            if (i < S.n) {
                const double bi = S.b[i];
                const double e = bi - row_dot(A, S.x, i);
                t2 = e * e;
                b2 = bi * bi;
            }
            t2 = block_sum(t2, sm);
            b2 = block_sum(b2, sm);
            if (threadIdx.x == 0) {
                S.part[2 * tl.tile] = t2;
                S.part[2 * tl.tile + 1] = b2;
            }
        }
    }
}

__global__ __launch_bounds__(FIN_BLOCK) void k_b_fin_delta(const int32_t *__restrict__ act,
                                                           const Counts *__restrict__ cnt,
                                                           const Sys *__restrict__ sys,
                                                           State *st, uint8_t *__restrict__ arena) {
    __shared__ double sm[16];
    const int na = cnt->nactive;
    for (int a = blockIdx.x; a < na; a += gridDim.x) {
        const int s = act[a];
        const int ph = st[s].phase;  // read by every thread before the barriers below
        const SysP S = load_sys(sys, s, arena);
        if (ph == PH_ITER) {
            const double tot = finalize_sum(S.part, (unsigned)S.ntiles, 2, sm);
            if (threadIdx.x == 0 && !cg_scalars_delta(st[s].cg, tot)) st[s].phase = PH_RESID;
        } else if (ph == PH_POWER) {  // rho = ||w|| (k_pow_fin)
            const double tot = finalize_sum(S.part, (unsigned)S.ntiles, 2, sm);
            if (threadIdx.x == 0) {
                const double rho = sqrt(tot);
                st[s].rho = rho;
                if (rho == 0.0) st[s].pstop = 1;
            }
        } else if (ph == PH_RESID) {
            const double e2 = finalize_sum(S.part, (unsigned)S.ntiles, 2, sm);
            const double b2 = finalize_sum(S.part + 1, (unsigned)S.ntiles, 2, sm);
            if (threadIdx.x == 0) {
                st[s].cg.res2 = e2;
                st[s].cg.b2 = b2;
                st[s].phase = PH_DONE;
            }
        }
    }
}

// The one-triangle layout (MODE_SYMT) is available when the segments' LDS
// copies of L take at most 36 KiB per workgroup (four workgroups per CU):
// partsize 16 in every precision, partsize 32 in FP32 and BF16, partsize 64
// in BF16.  It halves the triangle bytes; the two-triangle kernel needs no
// LDS and keeps more waves resident.
template <int PM, class TS>
constexpr bool symt_fits() {
    return (TILE / PM) * symt_lds(PM) * (int)sizeof(TS) <= 36 * 1024;
}

// PM = partsize (16, 32 or 64): 256/PM partitions per tile, one per PM-lane segment.
template <int PM, class TS, class TC, bool SYMT>
__global__ __launch_bounds__(256) void k_b_sweep(const Tile *__restrict__ tiles,
                                                 const Counts *__restrict__ cnt,
                                                 const Sys *__restrict__ sys,
                                                 const State *__restrict__ st,
                                                 uint8_t *__restrict__ arena) {
    __shared__ double sm[4];
    constexpr int SEGS = TILE / PM;
    constexpr int MODE = SYMT ? MODE_SYMT : MODE_SYM;
    __shared__ __align__(16) TS lds_all[SYMT ? SEGS * symt_lds(PM) : 1];  // staged word stores
    TS *lds = SYMT ? lds_all + (threadIdx.x / PM) * symt_lds(PM) : nullptr;
    int t0, t1;
    tile_range(cnt->ntiles, t0, t1);
    if (t0 >= t1) return;
    Tile nxt = tiles[t0];
    int cur = -1, ph = PH_EMPTY;
    SysP S{};
    double alpha = 0.0;
    const double *p = nullptr;
    for (int t = t0; t < t1; ++t) {
        const Tile tl = nxt;
        if (t + 1 < t1) nxt = tiles[t + 1];
        if (tl.slot != cur) {
            cur = tl.slot;
            ph = st[cur].phase;
            S = load_sys(sys, cur, arena);
            // the alpha and p of cg_prologue, once per system
            alpha = st[cur].cg.gamma / st[cur].cg.delta;
            p = (st[cur].cg.iter & 1) ? S.pb : S.pa;
        }
        if (ph != PH_INIT && ph != PH_ITER) continue;
        const TS *T = reinterpret_cast<const TS *>(S.T);
        const TC *D = reinterpret_cast<const TC *>(S.D);
        int lane = threadIdx.x % PM;
        // Opaque per tile: the sweep's per-lane offsets (dozens per step) are
        // otherwise hoisted out of the tile loop and held in registers.
        asm volatile("" : "+v"(lane));
        const int64_t k = (int64_t)tl.tile * SEGS + threadIdx.x / PM;
        double v = 0.0;
        if (!__all(k >= S.nparts)) {  // whole waves: see sweep_seg_run
            const bool init = (ph == PH_INIT);
            // admission: r = b, x = 0, z = M^-1 b (partial b'z = r'z);
            // iteration: r -= alpha q, x += alpha p, z = M^-1 r (cg_prologue)
            auto input = [&](int64_t i) -> double {
                if (init) {
                    const double bi = S.b[i];
                    S.r[i] = bi;
                    S.x[i] = 0.0;
                    return bi;
                }
                const double ri = S.r[i] - alpha * S.q[i];
                S.r[i] = ri;
                S.x[i] = S.x[i] + alpha * p[i];
                return ri;
            };
            v = sweep_seg_run<PM, MODE, TS, TC, true>(input, S.z, T, D, S.n, PM, S.nparts, k, lane,
                                                       lds);
        }
        v = block_sum(v, sm);
        if (threadIdx.x == 0) S.part[2 * tl.tile] = v;
    }
}

__global__ __launch_bounds__(FIN_BLOCK) void k_b_fin_gamma(const int32_t *__restrict__ act,
                                                           const Counts *__restrict__ cnt,
                                                           const Sys *__restrict__ sys,
                                                           State *st, uint8_t *__restrict__ arena) {
    __shared__ double sm[16];
    const int na = cnt->nactive;
    for (int a = blockIdx.x; a < na; a += gridDim.x) {
        const int s = act[a];
        const int ph = st[s].phase;
        if (ph == PH_POWER) {  // one more power-iteration step done
            __syncthreads();   // every thread has read the phase before it changes
            if (threadIdx.x == 0) {
                const int pw = st[s].pw + 1;
                st[s].pw = pw;
                if (st[s].pstop || pw >= st[s].pmax) st[s].phase = PH_INIT;
            }
            continue;
        }
        if (ph != PH_INIT && ph != PH_ITER) continue;
        const SysP S = load_sys(sys, s, arena);
        const double tot = finalize_sum(S.part, (unsigned)S.ntiles, 2, sm);
        if (threadIdx.x == 0) {
            const bool done = (ph == PH_INIT) ? cg_scalars_start(st[s].cg, tot)
                                              : cg_scalars_next(st[s].cg, tot);
            st[s].phase = done ? PH_RESID : PH_ITER;
        }
    }
}

// ---------------------------------------------------------------------------
// the Chebyshev smoother (4th kind, degree d >= 2) as the preconditioner
// ---------------------------------------------------------------------------
// Coefficients of step k, formed in FP64 from the system's rho and rounded to
// the smoother's type, as the single-system plan forms them on the host:
// a_k = (2k-3)/(2k+1), b_k = (8k-4)/((2k+1) rho)  (chebyshev.jl:158-183).
template <class T>
__device__ __forceinline__ void cheb_ab(int k, double rho, T &ak, T &bk) {
    ak = (T)((double)(2 * k - 3) / (double)(2 * k + 1));
    bk = (T)((double)(8 * k - 4) / ((double)(2 * k + 1) * rho));
}

template <class T>
__device__ __forceinline__ SellView<T> cheb_view(const SysP &S) {
    if constexpr (sizeof(T) == 8) return {S.sptr, S.col, S.val, S.n, S.col16};
    else return {S.sptr, S.col, reinterpret_cast<const float *>(S.v32), S.n, S.col16};
}

// The CG residual of a Chebyshev system alternates between two buffers,
// r_k in r[k & 1] (k = the CG iteration; the admission pass writes r_0 = b to
// r[0]), because k_bc_first gathers r_k at neighbouring rows while it writes
// r_{k+1}.  The buffer the current pass writes:
__device__ __forceinline__ double *cheb_rnew(const SysP &S, int ph, long long iter) {
    return (ph == PH_INIT || (iter & 1)) ? S.r : S.rb;
}

// Step 1 with the start of the preconditioner folded in, one tile per
// workgroup over the SpMV grid.  The start is pointwise -- r' = r - alpha q,
// x += alpha p (the CG's own update, cg_prologue; on admission r' = b,
// x = 0), z_1 = a_1 0 + b_1 l1inv r' (k_c4_init_cg) -- so z_1 is not stored:
// r'_j and z_1,j are formed from r, q and l1inv wherever column j is gathered,
// with the same operations, hence the same values, as the stored z_1 of the
// single-system kernels.  Then k_c4_step (or, for d = 2, k_c4_step_dot):
// r_1 = r' - A z_1, z_cg = 0 + z_1, z_2 = a_2 z_1 + b_2 l1inv r_1 (stored for
// the next step, or added to z_cg with the tile's r'z partial when d = 2).
// Power iteration: v = w / rho (k_pow_scale).
template <class T>
__global__ __launch_bounds__(256) void k_bc_first(const Tile *__restrict__ tiles,
                                                  const Counts *__restrict__ cnt,
                                                  const Sys *__restrict__ sys,
                                                  const State *__restrict__ st,
                                                  uint8_t *__restrict__ arena, int d) {
    __shared__ double sm[4];
    const bool last = (d == 2);
    const int nt = cnt->ntiles;
    int cur = -1, ph = PH_EMPTY;
    SysP S{};
    T a1 = (T)0, b1 = (T)0, a2 = (T)0, b2 = (T)0;
    double alpha = 0.0, rho = 1.0;
    const double *p = nullptr, *rsrc = nullptr;
    double *rdst = nullptr;
    bool stop = false;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.slot != cur) {
            cur = tl.slot;
            ph = st[cur].phase;
            S = load_sys(sys, cur, arena);
            rho = st[cur].rho;
            stop = st[cur].pstop != 0;
            cheb_ab<T>(1, rho, a1, b1);
            cheb_ab<T>(2, rho, a2, b2);
            const long long it = st[cur].cg.iter;
            alpha = st[cur].cg.gamma / st[cur].cg.delta;
            p = (it & 1) ? S.pb : S.pa;
            rsrc = (ph == PH_INIT) ? S.b : ((it & 1) ? S.rb : S.r);
            rdst = cheb_rnew(S, ph, it);
        }
        const int64_t i = (int64_t)tl.tile * TILE + threadIdx.x;
        if (ph == PH_POWER) {
            if (i < S.n && !stop) S.pa[i] = S.pb[i] / rho;
            continue;
        }
        if (ph != PH_INIT && ph != PH_ITER) continue;  // uniform in the workgroup
        const bool init = (ph == PH_INIT);
        const T *l1 = reinterpret_cast<const T *>(S.l1s);
        auto rstart = [&](int64_t j) -> double {
            return init ? rsrc[j] : rsrc[j] - alpha * S.q[j];
        };
        auto z1_at = [&](int64_t j) -> T { return a1 * (T)0 + b1 * l1[j] * (T)rstart(j); };
        double v = 0.0;
        if (i < S.n) {
            const double rn = rstart(i);
            rdst[i] = rn;
            S.x[i] = init ? 0.0 : S.x[i] + alpha * p[i];
            const SellView<T> A = cheb_view<T>(S);
            const T ri = row_fold(
                A, i, (T)rn, [&](int64_t c) { return z1_at(c); },
                [](T acc, T a, T z1) { return acc - a * z1; });
            const T zi = z1_at(i);
            double xi = 0.0 + (double)zi;
            const T zn = a2 * zi + b2 * l1[i] * ri;
            if (!last) {
                reinterpret_cast<T *>(S.rc)[i] = ri;
                reinterpret_cast<T *>(S.zb)[i] = zn;
            } else {
                xi = xi + (double)zn;
                v = rn * xi;
            }
            S.z[i] = xi;
        }
        if (last) {
            v = block_sum(v, sm);
            if (threadIdx.x == 0) S.part[2 * tl.tile] = v;
        }
    }
}

// Step k of 2 .. d-1 (k_c4_step; the last one is k_c4_step_dot): r_k =
// r_{k-1} - A z_k, z_cg += z_k, z_{k+1} = a_{k+1} z_k + b_{k+1} l1inv r_k; the
// last step adds z_d to z_cg instead of storing it and forms the tile's r'z
// partial.  z_k is read from za (k odd) or zb and z_{k+1} written to the other.
template <class T>
__global__ __launch_bounds__(256) void k_bc_step(const Tile *__restrict__ tiles,
                                                 const Counts *__restrict__ cnt,
                                                 const Sys *__restrict__ sys,
                                                 const State *__restrict__ st,
                                                 uint8_t *__restrict__ arena, int k, int d) {
    __shared__ double sm[4];
    const bool last = (k == d - 1);
    const int nt = cnt->ntiles;
    int cur = -1, ph = PH_EMPTY;
    SysP S{};
    T an = (T)0, bn = (T)0;
    const double *rnew = nullptr;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile tl = tiles[t];
        if (tl.slot != cur) {
            cur = tl.slot;
            ph = st[cur].phase;
            S = load_sys(sys, cur, arena);
            cheb_ab<T>(k + 1, st[cur].rho, an, bn);
            rnew = cheb_rnew(S, ph, st[cur].cg.iter);
        }
        if (ph != PH_INIT && ph != PH_ITER) continue;  // uniform in the workgroup
        const SellView<T> A = cheb_view<T>(S);
        const T *l1 = reinterpret_cast<const T *>(S.l1s);
        const T *zk = reinterpret_cast<const T *>((k & 1) ? S.za : S.zb);
        T *znext = reinterpret_cast<T *>((k & 1) ? S.zb : S.za);
        T *rc = reinterpret_cast<T *>(S.rc);
        const int64_t i = (int64_t)tl.tile * TILE + threadIdx.x;
        double v = 0.0;
        if (i < S.n) {
            const T ri = row_sub(A, zk, i, rc[i]);
            if (!last) rc[i] = ri;
            const T zi = zk[i];
            double xi = S.z[i];
            xi = xi + (double)zi;
            const T zn = an * zi + bn * l1[i] * ri;
            if (!last) {
                znext[i] = zn;
            } else {
                xi = xi + (double)zn;
                v = rnew[i] * xi;
            }
            S.z[i] = xi;
        }
        if (last) {
            v = block_sum(v, sm);
            if (threadIdx.x == 0) S.part[2 * tl.tile] = v;
        }
    }
}

// D_Dl1 and the packed triangles of one system from its sliced-ELL rows (the
// rows of a symmetric matrix, summed in ascending column order exactly as
// k_setup_rows does for a CSR or symmetric CSC).
template <class TS, class TC>
__global__ void k_b_setup(int64_t n, int64_t P, const int64_t *__restrict__ sptr,
                          const int32_t *__restrict__ col, const int16_t *__restrict__ col16,
                          const double *__restrict__ val,
                          const int32_t *__restrict__ rlen, double eta, TC *__restrict__ D,
                          TS *__restrict__ T, int mode) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t k = g / P;
    const int64_t s = k * P;
    const int64_t e = (s + P < n ? s + P : n) - 1;
    const int64_t Pa = P;  // every block full size (the last one padded, see sweep_seg_run)
    const int64_t i = g - s;
    int64_t lb, ub;
    block_bases(k, P, Pa, mode, lb, ub);
    const int64_t base = sptr[g >> 6] + (g & 63);
    double a = 0.0, d = 0.0;
    const int len = rlen[g];
    for (int kk = 0; kk < len; ++kk) {
        const int64_t p = base + 64 * (int64_t)kk;
        const int64_t j = col16 ? g + col16[p] : (int64_t)col[p];
        const double v = val[p];
        if (j == g) {
            a = v;
        } else if (j < s || j > e) {
            d += fabs(v);
        } else if (j < g) {
            const int64_t jl = j - s;
            T[lb + colL_off(jl, Pa) + (i - jl - 1)] = Scalar<TS>::from_f64(v);
        } else if (mode == MODE_SYM) {  // the one-triangle layout stores no U
            const int64_t jl = j - s;
            T[ub + colU_off(jl, Pa) + i] = Scalar<TS>::from_f64(v);
        }
    }
    D[g] = (TC)dl1_rule(a, d, eta);
}

// Chebyshev setup of one system: l1inv_i = 1 / sum_j |a_ij| in row order
// (k_l1inv: _l1_rowsum, chebyshev.jl:27-36) and its FP32 copy.
__global__ void k_bc_setup(SellView<double> A, double *__restrict__ l1,
                           float *__restrict__ l1s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double r = 1.0 / row_abs_sum(A, i);
    l1[i] = r;
    if (l1s) l1s[i] = (float)r;
}

// FP32 matrix values of the FP32 smoother (as CSR::ensure_fp32) and the
// power iteration's start vector v0 = 1/sqrt(n)
__global__ void k_bc_f32(int64_t m, const double *__restrict__ v, float *__restrict__ o) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) o[i] = (float)v[i];
}
__global__ void k_bc_fill(int64_t n, double *__restrict__ v, double a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

// Placement of one system inside its arena block.  The uploaded part (matrix,
// row lengths, right-hand side) comes first so that one copy moves it.
struct Layout {
    int64_t n = 0, nslices = 0, nent = 0, nparts = 0, ntiles = 0, tri = 0;
    int64_t o_sptr = 0, o_col = 0, o_val = 0, o_rlen = 0, o_b = 0, upload = 0;
    int64_t o_T = 0, o_D = 0, o_x = 0, o_r = 0, o_z = 0, o_pa = 0, o_pb = 0, o_q = 0, o_part = 0;
    int64_t o_l1 = 0, o_l1s = 0, o_v32 = 0, o_rb = 0, o_za = 0, o_zb = 0, o_rc = 0;
    int64_t total = 0;
};

// What a system's block holds besides the matrix and the CG vectors: the
// L1GS arrays (partsize P, triangle and D element sizes, one or two
// triangles) or the Chebyshev ones (element size tsz, degree).
struct Shape {
    int64_t P = 32;
    int tsz = 8, dsz = 8;
    bool one_tri = false, cheb = false;
    int degree = 2;
};

static Layout make_layout(int64_t n, const int64_t *ptr, const Shape &sh) {
    const int64_t P = sh.P;
    const int tsz = sh.tsz, dsz = sh.dsz;
    Layout L;
    L.n = n;
    L.nslices = (n + 63) / 64;
    for (int64_t s = 0; s < L.nslices; ++s) {
        int64_t w = 0;
        for (int64_t i = 64 * s; i < std::min<int64_t>(64 * s + 64, n); ++i)
            w = std::max<int64_t>(w, ptr[i + 1] - ptr[i]);
        L.nent += 64 * w;
    }
    L.nparts = sh.cheb ? 0 : (n + P - 1) / P;
    L.ntiles = (n + TILE - 1) / TILE;
    L.tri = L.nparts * P * (P - 1);  // full blocks, the last one padded
    if (sh.one_tri) L.tri /= 2;
    L.o_sptr = 0;
    L.o_col = align256(L.o_sptr + (L.nslices + 1) * 8);
    L.o_val = align256(L.o_col + L.nent * 4);
    L.o_rlen = align256(L.o_val + L.nent * 8);
    L.o_b = align256(L.o_rlen + n * 4);
    L.upload = L.o_b + n * 8;
    L.o_T = align256(L.upload);
    L.o_D = align256(L.o_T + L.tri * tsz);
    L.o_x = align256(L.o_D + (sh.cheb ? 0 : n * dsz));
    L.o_r = align256(L.o_x + n * 8);
    L.o_z = align256(L.o_r + n * 8);
    L.o_pa = align256(L.o_z + n * 8);
    L.o_pb = align256(L.o_pa + n * 8);
    L.o_q = align256(L.o_pb + n * 8);
    L.o_part = align256(L.o_q + n * 8);
    L.total = align256(L.o_part + 2 * L.ntiles * 8);
    if (sh.cheb) {
        const bool f32 = tsz == 4;
        L.o_l1 = L.total;
        L.o_l1s = f32 ? align256(L.o_l1 + n * 8) : L.o_l1;  // FP64: the same array
        L.o_v32 = align256(L.o_l1s + (f32 ? n * 4 : n * 8));
        L.o_rb = align256(L.o_v32 + (f32 ? L.nent * 4 : 0));
        // z_1 is never stored; z_2 (zb), z_3 (za), ... from degree 3 on
        L.o_za = align256(L.o_rb + n * 8);
        L.o_zb = align256(L.o_za + (sh.degree >= 4 ? n * tsz : 0));
        L.o_rc = align256(L.o_zb + (sh.degree >= 3 ? n * tsz : 0));
        L.total = align256(L.o_rc + (sh.degree >= 3 ? n * tsz : 0));
    }
    return L;
}

// Sliced-ELL of the rows (1-based CSR, or the columns of a symmetric CSC),
// row lengths and b, written into `dst` at the layout's offsets.  The slice
// format is the device operator's (padding: the row itself, value 0); the
// columns are stored as 16-bit offsets from the row when every one fits
// (returns true; the column section then holds int16 values), else as int32.
static bool fill_upload(const Layout &L, const int64_t *ptr, const int64_t *ind,
                        const double *val, const double *b, uint8_t *dst) {
    const int64_t n = L.n;
    bool c16 = true;
    for (int64_t i = 0; i < n && c16; ++i)
        for (int64_t p = ptr[i] - 1; p < ptr[i + 1] - 1; ++p) {
            const int64_t d = ind[p] - 1 - i;
            if (d < -32767 || d > 32767) {
                c16 = false;
                break;
            }
        }
    int64_t *sp = reinterpret_cast<int64_t *>(dst + L.o_sptr);
    int32_t *col = reinterpret_cast<int32_t *>(dst + L.o_col);
    int16_t *col16 = reinterpret_cast<int16_t *>(dst + L.o_col);
    double *v = reinterpret_cast<double *>(dst + L.o_val);
    int32_t *rl = reinterpret_cast<int32_t *>(dst + L.o_rlen);
    sp[0] = 0;
    for (int64_t s = 0; s < L.nslices; ++s) {
        int64_t w = 0;
        const int64_t i1 = std::min<int64_t>(64 * s + 64, n);
        for (int64_t i = 64 * s; i < i1; ++i) w = std::max<int64_t>(w, ptr[i + 1] - ptr[i]);
        sp[s + 1] = sp[s] + 64 * w;
        for (int l = 0; l < 64; ++l) {
            const int64_t i = 64 * s + l;
            const bool row = i < n;
            const int64_t len = row ? ptr[i + 1] - ptr[i] : 0;
            const int64_t p0 = row ? ptr[i] - 1 : 0;
            if (row) rl[i] = (int32_t)len;
            for (int64_t k = 0; k < w; ++k) {
                const int64_t q = sp[s] + 64 * k + l;
                int64_t c;
                if (k < len) {
                    c = ind[p0 + k] - 1;
                    HS_REQUIRE(c >= 0 && c < n, HS_EINVAL, "column index out of range");
                    v[q] = val[p0 + k];
                } else {
                    c = row ? i : n - 1;
                    v[q] = 0.0;
                }
                if (c16) col16[q] = (int16_t)(c - i);
                else col[q] = (int32_t)c;
            }
        }
    }
    std::memcpy(dst + L.o_b, b, n * 8);
    return c16;
}

// First-fit allocator of arena offsets with coalescing of free neighbours.
class Arena {
  public:
    explicit Arena(int64_t cap = 0) { reset(cap); }
    void reset(int64_t cap) {
        cap_ = cap;
        free_.clear();
        if (cap > 0) free_[0] = cap;
        used_ = 0;
    }
    bool alloc(int64_t size, int64_t &off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= size) {
                off = it->first;
                const int64_t rest = it->second - size;
                free_.erase(it);
                if (rest > 0) free_[off + size] = rest;
                used_ += size;
                return true;
            }
        }
        return false;
    }
    void release(int64_t off, int64_t size) {
        used_ -= size;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                size += prev->second;
                free_.erase(prev);
            }
        }
        next = free_.lower_bound(off);
        if (next != free_.end() && off + size == next->first) {
            size += next->second;
            free_.erase(next);
        }
        free_[off] = size;
    }
    int64_t used() const { return used_; }
    int64_t capacity() const { return cap_; }
    int64_t largest() const {
        int64_t m = 0;
        for (auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }

  private:
    std::map<int64_t, int64_t> free_;
    int64_t cap_ = 0, used_ = 0;
};

enum { SLOT_FREE = 0, SLOT_UPLOADING = 1, SLOT_ACTIVE = 2 };

struct Slot {
    int state = SLOT_FREE;
    int64_t id = -1, off = 0;
    Layout lay;
    double *x_out = nullptr;
    double rtol = 0, atol = 0;
    int64_t itmax = 0;
    double t_submit = 0, t_start = 0;
    hipEvent_t ready = nullptr;
    bool c16 = false;  // columns uploaded as 16-bit offsets
};

struct Engine {
    hs_batch_config cfg{};
    int tsz = 8, dsz = 8;
    int grid_spmv = 0, grid_sweep = 0;  // workgroups of the two tile kernels
    bool one_tri = false;               // packed lower triangle only (MODE_SYMT)
    bool cheb = false;                  // Chebyshev smoother instead of L1GS
    uint8_t *arena_dev = nullptr;
    Arena arena;
    uint8_t *staging = nullptr;
    int64_t stage_off = 0;
    hipStream_t cs = nullptr, us = nullptr;  // compute and upload streams
    DBuf<Sys> d_sys;
    DBuf<State> d_st;
    DBuf<Tile> d_tiles;
    DBuf<int32_t> d_act;
    DBuf<Counts> d_cnt;
    // pinned host mirrors; rewritten only after the previous step has finished
    Sys *h_sys = nullptr;
    State *h_st = nullptr;
    Tile *h_tiles = nullptr;
    int32_t *h_act = nullptr;
    Counts *h_cnt = nullptr;
    State *h_out = nullptr;  // scalars copied back after every step
    int64_t max_tiles = 0;
    std::vector<Slot> slots;
    std::vector<int> free_slots;
    std::vector<int> activated;  // slots whose descriptor must be uploaded
    bool tables_dirty = false;
    int32_t nactive = 0;
    hipGraphExec_t graphs[3] = {nullptr, nullptr, nullptr};  // by SpMV grid, SPMV_GRIDS order
    int32_t ntiles_host = 0;
    hipEvent_t ev_step = nullptr;
    bool step_pending = false;
    int64_t next_id = 0, steps = 0, retired = 0;
    std::deque<hs_batch_result> results;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    Shape shape() const {
        Shape sh;
        sh.P = cfg.partsize;
        sh.tsz = tsz;
        sh.dsz = dsz;
        sh.one_tri = one_tri;
        sh.cheb = cheb;
        sh.degree = cfg.degree;
        return sh;
    }
    ~Engine();
    void create(const hs_batch_config &c);
    void enqueue_pass(hipStream_t s, unsigned spmv_grid);
    void capture();
    int submit(int64_t n, const int64_t *ptr, const int64_t *ind, const double *val,
               const double *b, double *x_out, double rtol, double atol, int64_t itmax,
               int64_t *id);
    void wait_step();
    void step();
    void collect();
};

Engine::~Engine() {
    if (cs) (void)hipStreamSynchronize(cs);
    if (us) (void)hipStreamSynchronize(us);
    for (hipGraphExec_t g : graphs)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto &s : slots)
        if (s.ready) (void)hipEventDestroy(s.ready);
    if (ev_step) (void)hipEventDestroy(ev_step);
    if (arena_dev) {
        (void)hipFree(arena_dev);
        g_device_bytes -= arena.capacity();
    }
    for (void *p : {(void *)staging, (void *)h_sys, (void *)h_st, (void *)h_tiles, (void *)h_act,
                    (void *)h_cnt, (void *)h_out})
        if (p) (void)hipHostFree(p);
    if (cs) (void)hipStreamDestroy(cs);
    if (us) (void)hipStreamDestroy(us);
}

template <class T>
static void host_alloc(T *&p, size_t count) {
    HS_CHECK_HIP(hipHostMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T),
                               hipHostMallocDefault));
}

template <int PM>
static bool symt_fits_p(int prec) {
    if (prec == HS_PREC_FP64) return symt_fits<PM, double>();
    if (prec == HS_PREC_FP32) return symt_fits<PM, float>();
    return symt_fits<PM, __hip_bfloat16>();
}

static bool symt_fits_rt(int P, int prec) {
    return P == 16 ? symt_fits_p<16>(prec) : (P == 32 ? symt_fits_p<32>(prec)
                                                      : symt_fits_p<64>(prec));
}

static int resident_grid(const void *kernel, int cus) {
    int nb = 0;
    HS_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, TILE, 0));
    return std::max(1, nb) * cus;
}

template <int PM, class TS, class TC>
static const void *sweep_fn_t(bool symt) {
    if constexpr (symt_fits<PM, TS>())
        if (symt) return reinterpret_cast<const void *>(&k_b_sweep<PM, TS, TC, true>);
    return reinterpret_cast<const void *>(&k_b_sweep<PM, TS, TC, false>);
}

template <int PM>
static const void *sweep_fn_p(int prec, bool symt) {
    if (prec == HS_PREC_FP64) return sweep_fn_t<PM, double, double>(symt);
    if (prec == HS_PREC_FP32) return sweep_fn_t<PM, float, float>(symt);
    return sweep_fn_t<PM, __hip_bfloat16, float>(symt);
}

static const void *sweep_fn(int P, int prec, bool symt) {
    return P == 16 ? sweep_fn_p<16>(prec, symt)
                   : (P == 32 ? sweep_fn_p<32>(prec, symt) : sweep_fn_p<64>(prec, symt));
}

void Engine::create(const hs_batch_config &c) {
    cfg = c;
    cheb = c.smoother == HS_PCG_CHEB;
    tsz = c.precision == HS_PREC_FP64 ? 8 : (c.precision == HS_PREC_FP32 ? 4 : 2);
    dsz = c.precision == HS_PREC_FP64 ? 8 : 4;
    if (cheb) {
        one_tri = false;
    } else {
        const bool fits = symt_fits_rt((int)c.partsize, c.precision);
        HS_REQUIRE(c.triangles != 1 || fits, HS_EUNSUPPORTED,
                   "one-triangle storage needs partsize 16, or partsize 32 or 64 in FP32/BF16");
        one_tri = c.triangles == 1 || (c.triangles == 0 && fits);
    }
    int dev = 0;
    HS_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HS_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    // The persistent sweep: by default exactly as many workgroups as fit on
    // the chip at once (each walks a fixed run of the tile table; a second
    // round of workgroups would wait for the first to finish).
    const int cus = prop.multiProcessorCount;
    grid_spmv = SPMV_GRIDS[2];
    grid_sweep = cheb ? 0
                      : (c.grid > 0 ? c.grid
                                    : resident_grid(sweep_fn((int)c.partsize, c.precision, one_tri),
                                                    cus));
    HS_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    HS_CHECK_HIP(hipStreamCreateWithFlags(&us, hipStreamNonBlocking));
    HS_CHECK_HIP(hipEventCreateWithFlags(&ev_step, hipEventDisableTiming));
    HS_CHECK_HIP(hipMalloc((void **)&arena_dev, c.arena_bytes));
    g_device_bytes += c.arena_bytes;
    arena.reset(c.arena_bytes);
    host_alloc(staging, (size_t)c.staging_bytes);
    const int64_t ms = c.max_systems;
    // every resident row takes at least 64 arena bytes (seven FP64 vectors)
    max_tiles = c.arena_bytes / (64 * TILE) + ms;
    d_sys.alloc(ms);
    d_st.alloc(ms);
    d_tiles.alloc(max_tiles);
    d_act.alloc(ms);
    d_cnt.alloc(1);
    host_alloc(h_sys, ms);
    host_alloc(h_st, ms);
    host_alloc(h_tiles, max_tiles);
    host_alloc(h_act, ms);
    host_alloc(h_cnt, 1);
    host_alloc(h_out, ms);
    std::memset(h_st, 0, ms * sizeof(State));
    std::memset(h_out, 0, ms * sizeof(State));
    std::memset(h_sys, 0, ms * sizeof(Sys));
    h_cnt->ntiles = 0;
    h_cnt->nactive = 0;
    HS_CHECK_HIP(hipMemcpyAsync(d_st.get(), h_st, ms * sizeof(State), hipMemcpyHostToDevice, cs));
    HS_CHECK_HIP(hipMemcpyAsync(d_sys.get(), h_sys, ms * sizeof(Sys), hipMemcpyHostToDevice, cs));
    HS_CHECK_HIP(hipMemcpyAsync(d_cnt.get(), h_cnt, sizeof(Counts), hipMemcpyHostToDevice, cs));
    slots.resize(ms);
    for (int64_t s = ms - 1; s >= 0; --s) {
        HS_CHECK_HIP(hipEventCreateWithFlags(&slots[s].ready, hipEventDisableTiming));
        free_slots.push_back((int)s);
    }
    HS_CHECK_HIP(hipStreamSynchronize(cs));
    capture();
}

template <int PM, class TS, class TC>
static void launch_sweep_t(bool symt, unsigned grid, hipStream_t s, const Tile *tiles,
                           const Counts *cnt, const Sys *sys, const State *st, uint8_t *arena) {
    if constexpr (symt_fits<PM, TS>()) {
        if (symt) {
            hipLaunchKernelGGL((k_b_sweep<PM, TS, TC, true>), dim3(grid), dim3(TILE), 0, s, tiles,
                               cnt, sys, st, arena);
            return;
        }
    }
    hipLaunchKernelGGL((k_b_sweep<PM, TS, TC, false>), dim3(grid), dim3(TILE), 0, s, tiles, cnt,
                       sys, st, arena);
}

template <int PM>
static void launch_sweep(int prec, bool symt, unsigned grid, hipStream_t s, const Tile *tiles,
                         const Counts *cnt, const Sys *sys, const State *st, uint8_t *arena) {
    if (prec == HS_PREC_FP64)
        launch_sweep_t<PM, double, double>(symt, grid, s, tiles, cnt, sys, st, arena);
    else if (prec == HS_PREC_FP32)
        launch_sweep_t<PM, float, float>(symt, grid, s, tiles, cnt, sys, st, arena);
    else
        launch_sweep_t<PM, __hip_bfloat16, float>(symt, grid, s, tiles, cnt, sys, st, arena);
}

void Engine::enqueue_pass(hipStream_t s, unsigned spmv_grid) {
    // one finalize workgroup per active system up to 1,024 of them (each
    // system's sum is a dependent chain of loads and barriers; a workgroup
    // walking several systems pays it once per system)
    const unsigned fin_grid = (unsigned)std::min<int64_t>(cfg.max_systems, 1024);
    const Tile *tl = d_tiles.get();
    const Counts *cn = d_cnt.get();
    const Sys *sy = d_sys.get();
    State *stt = d_st.get();
    hipLaunchKernelGGL(k_b_spmv, dim3(spmv_grid), dim3(TILE), 0, s, tl, cn, sy, stt, arena_dev);
    hipLaunchKernelGGL(k_b_fin_delta, dim3(fin_grid), dim3(FIN_BLOCK), 0, s, d_act.get(), cn, sy,
                       stt, arena_dev);
    if (cheb) {
        const dim3 g(spmv_grid), t(TILE);
        if (cfg.precision == HS_PREC_FP64) {
            hipLaunchKernelGGL(k_bc_first<double>, g, t, 0, s, tl, cn, sy, stt, arena_dev,
                               cfg.degree);
            for (int k = 2; k < cfg.degree; ++k)
                hipLaunchKernelGGL(k_bc_step<double>, g, t, 0, s, tl, cn, sy, stt, arena_dev, k,
                                   cfg.degree);
        } else {
            hipLaunchKernelGGL(k_bc_first<float>, g, t, 0, s, tl, cn, sy, stt, arena_dev,
                               cfg.degree);
            for (int k = 2; k < cfg.degree; ++k)
                hipLaunchKernelGGL(k_bc_step<float>, g, t, 0, s, tl, cn, sy, stt, arena_dev, k,
                                   cfg.degree);
        }
    } else if (cfg.partsize == 16)
        launch_sweep<16>(cfg.precision, one_tri, grid_sweep, s, tl, cn, sy, stt, arena_dev);
    else if (cfg.partsize == 32)
        launch_sweep<32>(cfg.precision, one_tri, grid_sweep, s, tl, cn, sy, stt, arena_dev);
    else
        launch_sweep<64>(cfg.precision, one_tri, grid_sweep, s, tl, cn, sy, stt, arena_dev);
    hipLaunchKernelGGL(k_b_fin_gamma, dim3(fin_grid), dim3(FIN_BLOCK), 0, s, d_act.get(), cn, sy,
                       stt, arena_dev);
    check_launch();
}

void Engine::capture() {
    for (int v = 0; v < 3; ++v) {
        hipGraph_t g;
        HS_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < cfg.passes; ++k) enqueue_pass(cs, (unsigned)SPMV_GRIDS[v]);
        HS_CHECK_HIP(hipStreamEndCapture(cs, &g));
        HS_CHECK_HIP(hipGraphInstantiate(&graphs[v], g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
    }
}

int Engine::submit(int64_t n, const int64_t *ptr, const int64_t *ind, const double *val,
                   const double *b, double *x_out, double rtol, double atol, int64_t itmax,
                   int64_t *id) {
    HS_REQUIRE(n > 0 && n < (int64_t)INT32_MAX, HS_EINVAL, "system size must be in [1, 2^31-1)");
    HS_REQUIRE(ptr[0] == 1, HS_EINVAL, "pointer array must be 1-based (ptr[0] == 1)");
    for (int64_t i = 0; i < n; ++i)
        HS_REQUIRE(ptr[i + 1] >= ptr[i], HS_EINVAL, "pointer array must be non-decreasing");
    const Layout L = make_layout(n, ptr, shape());
    HS_REQUIRE(L.upload <= cfg.staging_bytes, HS_EINVAL,
               "system larger than the engine's upload staging buffer");
    HS_REQUIRE(L.total <= cfg.arena_bytes, HS_EINVAL, "system larger than the engine's arena");
    if (free_slots.empty()) return HS_EAGAIN;
    int64_t off = 0;
    if (!arena.alloc(L.total, off)) return HS_EAGAIN;
    const int s = free_slots.back();
    bool c16 = false;
    try {
        if (stage_off + L.upload > cfg.staging_bytes) {
            HS_CHECK_HIP(hipStreamSynchronize(us));  // earlier uploads have left the staging buffer
            stage_off = 0;
        }
        uint8_t *src = staging + stage_off;
        c16 = fill_upload(L, ptr, ind, val, b, src);
        stage_off = align256(stage_off + L.upload);
        uint8_t *blk = arena_dev + off;
        HS_CHECK_HIP(hipMemcpyAsync(blk, src, L.upload, hipMemcpyHostToDevice, us));
        if (L.tri) HS_CHECK_HIP(hipMemsetAsync(blk + L.o_T, 0, L.tri * tsz, us));
        const dim3 g(grid_for(n, 256)), t(256);
        const int mode = one_tri ? MODE_SYMT : MODE_SYM;
        const auto *sp = reinterpret_cast<const int64_t *>(blk + L.o_sptr);
        const auto *cl = reinterpret_cast<const int32_t *>(blk + L.o_col);
        const auto *c16p = c16 ? reinterpret_cast<const int16_t *>(blk + L.o_col) : nullptr;
        const auto *vl = reinterpret_cast<const double *>(blk + L.o_val);
        const auto *rl = reinterpret_cast<const int32_t *>(blk + L.o_rlen);
        if (cheb) {
            const bool f32 = cfg.precision == HS_PREC_FP32;
            // the view of the uploaded format: 16-bit offsets from the row, or
            // 32-bit columns (never the 16-bit section read as 32-bit columns)
            const SellView<double> av{sp, c16p ? nullptr : cl, vl, n, c16p};
            hipLaunchKernelGGL(k_bc_setup, g, t, 0, us, av,
                               reinterpret_cast<double *>(blk + L.o_l1),
                               f32 ? reinterpret_cast<float *>(blk + L.o_l1s) : nullptr);
            if (f32)
                hipLaunchKernelGGL(k_bc_f32, dim3(grid_for(L.nent, 256)), t, 0, us, L.nent, vl,
                                   reinterpret_cast<float *>(blk + L.o_v32));
            // the power iteration's v0 (in pa), as hs_cheb_build forms it
            if (cfg.power_iters > 0)
                hipLaunchKernelGGL(k_bc_fill, g, t, 0, us, n,
                                   reinterpret_cast<double *>(blk + L.o_pa),
                                   1.0 / std::sqrt((double)n));
        } else if (cfg.precision == HS_PREC_FP64)
            hipLaunchKernelGGL((k_b_setup<double, double>), g, t, 0, us, n, cfg.partsize, sp, cl,
                               c16p, vl, rl, cfg.eta, reinterpret_cast<double *>(blk + L.o_D),
                               reinterpret_cast<double *>(blk + L.o_T), mode);
        else if (cfg.precision == HS_PREC_FP32)
            hipLaunchKernelGGL((k_b_setup<float, float>), g, t, 0, us, n, cfg.partsize, sp, cl,
                               c16p, vl, rl, cfg.eta, reinterpret_cast<float *>(blk + L.o_D),
                               reinterpret_cast<float *>(blk + L.o_T), mode);
        else
            hipLaunchKernelGGL((k_b_setup<__hip_bfloat16, float>), g, t, 0, us, n, cfg.partsize,
                               sp, cl, c16p, vl, rl, cfg.eta, reinterpret_cast<float *>(blk + L.o_D),
                               reinterpret_cast<__hip_bfloat16 *>(blk + L.o_T), mode);
        check_launch();
        HS_CHECK_HIP(hipEventRecord(slots[s].ready, us));
    } catch (...) {
        arena.release(off, L.total);
        throw;
    }
    free_slots.pop_back();
    Slot &S = slots[s];
    S.state = SLOT_UPLOADING;
    S.id = next_id++;
    S.off = off;
    S.lay = L;
    S.c16 = c16;
    S.x_out = x_out;
    S.rtol = rtol;
    S.atol = atol;
    S.itmax = itmax;
    S.t_submit = now();
    *id = S.id;
    return HS_OK;
}

void Engine::wait_step() {
    if (!step_pending) return;
    HS_CHECK_HIP(hipEventSynchronize(ev_step));
    step_pending = false;
}

void Engine::step() {
    wait_step();
    bool any_active = false, any_uploading = false;
    for (const Slot &S : slots) {
        any_active |= S.state == SLOT_ACTIVE;
        any_uploading |= S.state == SLOT_UPLOADING;
    }
    // nothing to iterate yet: wait for the uploads instead of returning empty-handed
    if (!any_active && any_uploading) HS_CHECK_HIP(hipStreamSynchronize(us));
    // admit uploaded systems
    for (size_t s = 0; s < slots.size(); ++s) {
        Slot &S = slots[s];
        if (S.state != SLOT_UPLOADING) continue;
        const hipError_t q = hipEventQuery(S.ready);
        if (q == hipErrorNotReady) continue;
        HS_CHECK_HIP(q);
        S.state = SLOT_ACTIVE;
        S.t_start = now();
        activated.push_back((int)s);
        tables_dirty = true;
    }
    if (tables_dirty) {
        int64_t want = 0;
        for (const Slot &S : slots)
            if (S.state == SLOT_ACTIVE) want += S.lay.ntiles;
        // cannot trigger: every resident row holds >= 64 arena bytes
        HS_REQUIRE(want <= max_tiles, HS_EDEVICE, "tile table overflow");
        int32_t nt = 0, na = 0;
        for (size_t s = 0; s < slots.size(); ++s) {
            const Slot &S = slots[s];
            if (S.state != SLOT_ACTIVE) continue;
            h_act[na++] = (int32_t)s;
            for (int64_t t = 0; t < S.lay.ntiles; ++t) h_tiles[nt++] = Tile{(int32_t)s, (int32_t)t};
        }
        h_cnt->ntiles = nt;
        h_cnt->nactive = na;
        nactive = na;
        ntiles_host = nt;
        for (int s : activated) {
            const Slot &S = slots[s];
            const Layout &L = S.lay;
            const int64_t o = S.off;
            Sys &d = h_sys[s];
            d.sptr = o + L.o_sptr;
            d.col = o + L.o_col;
            d.val = o + L.o_val;
            d.T = o + L.o_T;
            d.D = o + L.o_D;
            d.x = o + L.o_x;
            d.r = o + L.o_r;
            d.z = o + L.o_z;
            d.pa = o + L.o_pa;
            d.pb = o + L.o_pb;
            d.q = o + L.o_q;
            d.b = o + L.o_b;
            d.part = o + L.o_part;
            d.l1 = o + L.o_l1;
            d.l1s = o + L.o_l1s;
            d.v32 = o + L.o_v32;
            d.rb = o + L.o_rb;
            d.za = o + L.o_za;
            d.zb = o + L.o_zb;
            d.rc = o + L.o_rc;
            d.n = L.n;
            d.nparts = L.nparts;
            d.ntiles = (int32_t)L.ntiles;
            d.c16 = S.c16 ? 1 : 0;
            State &h = h_st[s];
            std::memset(&h, 0, sizeof(State));
            h.cg.atol = S.atol;
            h.cg.rtol = S.rtol;
            h.cg.itmax = S.itmax;
            h.rho = 1.0;  // hs_cheb_build's estimate before (or without) the power iteration
            h.pmax = cheb ? cfg.power_iters : 0;
            h.phase = h.pmax > 0 ? PH_POWER : PH_INIT;
            HS_CHECK_HIP(hipStreamWaitEvent(cs, S.ready, 0));
            HS_CHECK_HIP(hipMemcpyAsync(d_sys.get() + s, &h_sys[s], sizeof(Sys),
                                        hipMemcpyHostToDevice, cs));
            HS_CHECK_HIP(hipMemcpyAsync(d_st.get() + s, &h_st[s], sizeof(State),
                                        hipMemcpyHostToDevice, cs));
        }
        activated.clear();
        if (nt) HS_CHECK_HIP(hipMemcpyAsync(d_tiles.get(), h_tiles, nt * sizeof(Tile),
                                            hipMemcpyHostToDevice, cs));
        if (na) HS_CHECK_HIP(hipMemcpyAsync(d_act.get(), h_act, na * sizeof(int32_t),
                                            hipMemcpyHostToDevice, cs));
        HS_CHECK_HIP(hipMemcpyAsync(d_cnt.get(), h_cnt, sizeof(Counts), hipMemcpyHostToDevice, cs));
        tables_dirty = false;
    }
    if (nactive == 0) return;
    int v = 0;
    while (v < 2 && SPMV_GRIDS[v] < ntiles_host) ++v;
    grid_spmv = SPMV_GRIDS[v];
    HS_CHECK_HIP(hipGraphLaunch(graphs[v], cs));
    HS_CHECK_HIP(hipMemcpyAsync(h_out, d_st.get(), slots.size() * sizeof(State),
                                hipMemcpyDeviceToHost, cs));
    HS_CHECK_HIP(hipEventRecord(ev_step, cs));
    step_pending = true;
    ++steps;
}

void Engine::collect() {
    if (!step_pending) return;
    wait_step();
    std::vector<int> fin;
    for (size_t s = 0; s < slots.size(); ++s)
        if (slots[s].state == SLOT_ACTIVE && h_out[s].phase == PH_DONE) fin.push_back((int)s);
    if (fin.empty()) return;
    for (int s : fin) {
        const Slot &S = slots[s];
        if (S.x_out)
            HS_CHECK_HIP(hipMemcpyAsync(S.x_out, arena_dev + S.off + S.lay.o_x, S.lay.n * 8,
                                        hipMemcpyDeviceToHost, cs));
    }
    HS_CHECK_HIP(hipStreamSynchronize(cs));
    const double t = now();
    for (int s : fin) {
        Slot &S = slots[s];
        const CGDev &c = h_out[s].cg;
        hs_batch_result r{};
        r.id = S.id;
        r.n = S.lay.n;
        r.iters = c.iter;
        r.status = c.status;
        r.rnorm0 = c.rnorm0;
        r.rnorm = c.rnorm;
        r.true_relres = c.b2 > 0 ? std::sqrt(c.res2 / c.b2) : std::sqrt(c.res2);
        r.t_submit = S.t_submit;
        r.t_start = S.t_start;
        r.t_done = t;
        results.push_back(r);
        arena.release(S.off, S.lay.total);
        S.state = SLOT_FREE;
        S.x_out = nullptr;
        free_slots.push_back(s);
        ++retired;
    }
    tables_dirty = true;
}

}  // namespace batch
}  // namespace hs

struct hs_batch : hs::batch::Engine {};

using namespace hs;

extern "C" {

int hs_batch_create(const hs_batch_config *cfg, hs_batch_t *out) {
    return guard([&] {
        HS_REQUIRE(cfg && out, HS_EINVAL, "NULL argument");
        *out = nullptr;
        HS_REQUIRE(cfg->arena_bytes > 0, HS_EINVAL, "arena_bytes must be positive");
        HS_REQUIRE(cfg->max_systems > 0 && cfg->max_systems <= (1 << 20), HS_EINVAL,
                   "max_systems must be in [1, 2^20]");
        HS_REQUIRE(cfg->staging_bytes > 0, HS_EINVAL, "staging_bytes must be positive");
        HS_REQUIRE(cfg->smoother == 0 || cfg->smoother == HS_PCG_L1GS ||
                       cfg->smoother == HS_PCG_CHEB,
                   HS_EINVAL, "smoother must be 0 or HS_PCG_L1GS (L1GS) or HS_PCG_CHEB");
        const bool cheb = cfg->smoother == HS_PCG_CHEB;
        HS_REQUIRE(cheb || cfg->partsize == 16 || cfg->partsize == 32 || cfg->partsize == 64,
                   HS_EUNSUPPORTED, "the batched engine takes partsize 16, 32 or 64");
        HS_REQUIRE(cfg->precision == HS_PREC_FP64 || cfg->precision == HS_PREC_FP32 ||
                       cfg->precision == HS_PREC_BF16,
                   HS_EINVAL, "unknown precision");
        HS_REQUIRE(!cheb || cfg->precision != HS_PREC_BF16, HS_EUNSUPPORTED,
                   "Chebyshev smoothers run in FP64 or FP32");
        HS_REQUIRE(!cheb || (cfg->degree >= 2 && cfg->degree <= 16), HS_EUNSUPPORTED,
                   "the batched Chebyshev smoother takes degree 2 to 16");
        HS_REQUIRE(!cheb || cfg->power_iters >= 0, HS_EINVAL,
                   "power_iters must be non-negative");
        HS_REQUIRE(cfg->passes >= 1 && cfg->passes <= 1024, HS_EINVAL,
                   "passes must be in [1, 1024]");
        HS_REQUIRE(cfg->grid >= 0, HS_EINVAL, "grid must be >= 0");
        HS_REQUIRE(cfg->triangles >= 0 && cfg->triangles <= 2, HS_EINVAL,
                   "triangles must be 0 (automatic), 1 or 2");
        std::unique_ptr<hs_batch> e(new hs_batch());
        e->create(*cfg);
        *out = e.release();
    });
}

int hs_batch_footprint(hs_batch_t e, int64_t n, const int64_t *ptr, int64_t *bytes) {
    return guard([&] {
        HS_REQUIRE(e && ptr && bytes, HS_EINVAL, "NULL argument");
        HS_REQUIRE(n > 0, HS_EINVAL, "system size must be positive");
        *bytes = batch::make_layout(n, ptr, e->shape()).total;
    });
}

int hs_batch_submit(hs_batch_t e, int64_t n, const int64_t *ptr, const int64_t *ind,
                    const double *val, const double *b, double *x_out, double rtol, double atol,
                    int64_t itmax, int64_t *id) {
    int status = HS_OK;
    int rc = guard([&] {
        HS_REQUIRE(e && ptr && ind && val && b && id, HS_EINVAL, "NULL argument");
        status = e->submit(n, ptr, ind, val, b, x_out, rtol, atol, itmax < 0 ? 2 * n : itmax, id);
        if (status == HS_EAGAIN) set_error(HS_EAGAIN, "no free slot or arena space; retry later");
    });
    return rc != HS_OK ? rc : status;
}

int hs_batch_step(hs_batch_t e) {
    return guard([&] {
        HS_REQUIRE(e, HS_EINVAL, "NULL engine");
        e->step();
    });
}

int hs_batch_collect(hs_batch_t e, hs_batch_result *out, int64_t max, int64_t *count) {
    return guard([&] {
        HS_REQUIRE(e && count && (max == 0 || out), HS_EINVAL, "NULL argument");
        e->collect();
        int64_t k = 0;
        while (k < max && !e->results.empty()) {
            out[k++] = e->results.front();
            e->results.pop_front();
        }
        *count = k;
    });
}

int hs_batch_info(hs_batch_t e, int64_t *out) {
    return guard([&] {
        HS_REQUIRE(e && out, HS_EINVAL, "NULL argument");
        int64_t active = 0, uploading = 0;
        for (auto &s : e->slots) {
            active += s.state == batch::SLOT_ACTIVE;
            uploading += s.state == batch::SLOT_UPLOADING;
        }
        out[0] = active;
        out[1] = uploading;
        out[2] = (int64_t)e->free_slots.size();
        out[3] = e->arena.capacity() - e->arena.used();
        out[4] = e->arena.capacity();
        out[5] = e->arena.largest();
        out[6] = e->steps;
        out[7] = e->retired;
        out[8] = (int64_t)e->results.size();
        out[9] = e->grid_spmv;
        out[10] = e->grid_sweep;
        out[11] = e->cheb ? 0 : (e->one_tri ? 1 : 2);  // packed triangles per partition
    });
}

int hs_batch_peek(hs_batch_t e, int64_t id, int what, double *out) {
    return guard([&] {
        HS_REQUIRE(e && out, HS_EINVAL, "NULL argument");
        HS_REQUIRE(what >= 0 && what <= 9, HS_EINVAL, "unknown quantity");
        HS_REQUIRE((what != 5 || !e->cheb) && (what < 7 || what > 8 || e->cheb), HS_EINVAL,
                   "quantity of the other smoother");
        e->wait_step();
        HS_CHECK_HIP(hipStreamSynchronize(e->us));
        const batch::Slot *S = nullptr;
        for (const auto &s : e->slots)
            if (s.state != batch::SLOT_FREE && s.id == id) S = &s;
        HS_REQUIRE(S, HS_EINVAL, "no resident system with this id");
        const batch::Layout &L = S->lay;
        const uint8_t *blk = e->arena_dev + S->off;
        int64_t off[5] = {L.o_x, L.o_r, L.o_z, L.o_b, L.o_q};
        const int64_t slot = S - e->slots.data();
        batch::State h;
        if (e->cheb && (what == 1 || what == 8))
            HS_CHECK_HIP(hipMemcpy(&h, e->d_st.get() + slot, sizeof h, hipMemcpyDeviceToHost));
        if (e->cheb && what == 1 && (h.cg.iter & 1)) off[1] = L.o_rb;  // r_k is in r[k & 1]
        if (what <= 4) {
            HS_CHECK_HIP(hipMemcpy(out, blk + off[what], L.n * 8, hipMemcpyDeviceToHost));
        } else if (what == 5) {  // D_Dl1 widened to FP64
            if (e->dsz == 8) {
                HS_CHECK_HIP(hipMemcpy(out, blk + L.o_D, L.n * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> f(L.n);
                HS_CHECK_HIP(hipMemcpy(f.data(), blk + L.o_D, L.n * 4, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < L.n; ++i) out[i] = f[i];
            }
        } else if (what == 6) {  // the two partials of every tile
            HS_CHECK_HIP(hipMemcpy(out, blk + L.o_part, 2 * L.ntiles * 8, hipMemcpyDeviceToHost));
        } else if (what == 7) {  // l1inv (FP64)
            HS_CHECK_HIP(hipMemcpy(out, blk + L.o_l1, L.n * 8, hipMemcpyDeviceToHost));
        } else if (what == 9) {  // bytes per stored column as uploaded
            out[0] = S->c16 ? 2.0 : 4.0;
        } else {  // rho
            out[0] = h.rho;
        }
    });
}

int hs_batch_busy(hs_batch_t e) {
    if (!e || !e->step_pending) return 0;
    const hipError_t q = hipEventQuery(e->ev_step);
    if (q == hipErrorNotReady) return 1;
    if (q != hipSuccess) {
        set_error(HS_EDEVICE, std::string("hipEventQuery: ") + hipGetErrorString(q));
        return HS_EDEVICE;
    }
    return 0;
}

double hs_batch_clock(hs_batch_t e) { return e ? e->now() : 0.0; }

int hs_batch_destroy(hs_batch_t e) {
    return guard([&] { delete e; });
}

}  // extern "C"
