// hs_l1gs.hip -- partitioned l1 Gauss-Seidel (L1GS) on gfx950.
//
// Build: D_Dl1 and the in-partition strict triangles on the device
// (reference: _precompute_and_pack_kernel!, src/l1_gauss_seidel.jl:1028-1066;
// _diag_offpart :1101-1242; packing :1248-1519; DiagonalIndices :422-436).
// Apply: forward, backward and symmetric sweeps, one partition per wave
// segment (reference: _apply_sweep_kernel! :1068-1095, _apply_sweep! :846-906).
//
// Device layout of the packed triangles: ONE buffer, one block per partition,
// holding the triangles the sweep reads in the order it reads them.  Block k
// starts at k*P*(P-1) for a symmetric sweep (lower then upper triangle) and at
// k*P*(P-1)/2 for a forward (lower only) or backward (upper only) sweep, like
// the reference's packed buffers (:698-701) laid end to end.  Inside a block
// the triangles are COLUMN-major, so that step j of the right-looking
// recurrence reads one contiguous run of values, one per lane, and the upper
// triangle's columns are stored in descending order, the order the backward
// recurrence visits them.  A symmetric sweep therefore streams each block once,
// front to back:
//   L(i,j), j < i : lb + j*(Pa-1) - j*(j-1)/2 + (i-j-1)
//   U(i,j), i < j : ub + (Pa*(Pa-1) - j*(j+1))/2 + i
// with lb = block start and ub = lb + Pa*(Pa-1)/2 (symmetric) or lb (backward);
// Pa = rows of this partition (only the last partition can be shorter).
// The reference row-major layout is produced on request by hs_l1gs_get.
// The one-triangle cache (HS_CACHE_PACKED_SYM, symmetric A) stores only the
// lower triangle, k*P*(P-1)/2 per block like a forward sweep, and reads
// U(i,j) = L(j,i) back from an LDS copy made during the forward pass.
//
// Recurrence.  Lane i of a segment owns row i.  Forward step j: every lane
// forms (x_i - acc_i)/d_i, the value of lane j is broadcast, and lanes i > j
// add L(i,j)*y_j.  Row i therefore sums its terms in ascending j -- the order
// of the reference's _accumulate_from_cache (:1547-1562) -- and the file is
// compiled with -ffp-contract=off, so forward sweeps are bit-identical to the
// reference arithmetic.  The backward recurrence adds in descending j.
#include "hs_sweep.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstddef>
#include <memory>
#include <type_traits>
#include <vector>

namespace hs {

// ---------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------
// One thread per row g.  Reads row g of a 1-based CSR (or column g of a
// symmetric CSC, which the reference reads as row g, :1218-1229), computes
// D_Dl1[g] and scatters the in-partition strict entries of the row into the
// column-major triangles.  want_L / want_U select the triangles to pack.
template <class TS>
__global__ void k_setup_rows(int64_t n, int64_t P, const int64_t *__restrict__ ptr,
                             const int64_t *__restrict__ ind, const double *__restrict__ val,
                             double eta, double *__restrict__ D, TS *__restrict__ T, int mode,
                             int want_L, int want_U) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int64_t k = g / P;
    int64_t s = k * P;
    int64_t e = (s + P < n ? s + P : n) - 1;
    int64_t Pa = e - s + 1;
    int64_t i = g - s;
    int64_t lb, ub;
    block_bases(k, P, Pa, mode, lb, ub);
    double a = 0.0, d = 0.0;
    int64_t p0 = ptr[g] - 1, p1 = ptr[g + 1] - 1;
    for (int64_t p = p0; p < p1; ++p) {
        int64_t j = ind[p] - 1;
        double v = val[p];
        if (j == g) {
            a = v;
        } else if (j < s || j > e) {
            d += fabs(v);
        } else if (j < g) {
            if (want_L) {
                int64_t jl = j - s;
                T[lb + colL_off(jl, Pa) + (i - jl - 1)] = Scalar<TS>::from_f64(v);
            }
        } else {
            if (want_U) {
                int64_t jl = j - s;
                T[ub + colU_off(jl, Pa) + i] = Scalar<TS>::from_f64(v);
            }
        }
    }
    D[g] = dl1_rule(a, d, eta);
}

// One thread per column c of a non-symmetric 1-based CSC: packs the column's
// in-partition entries (reference :1279-1324 lower, :1415-1469 upper).  A CSC
// column c holds column c of both triangles, so lower writes are contiguous.
template <class TS>
__global__ void k_pack_cols(int64_t n, int64_t P, const int64_t *__restrict__ colptr,
                            const int64_t *__restrict__ rowval, const double *__restrict__ val,
                            TS *__restrict__ T, int mode, int want_L, int want_U) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int64_t k = c / P;
    int64_t s = k * P;
    int64_t e = (s + P < n ? s + P : n) - 1;
    int64_t Pa = e - s + 1;
    int64_t jl = c - s;
    int64_t lb, ub;
    block_bases(k, P, Pa, mode, lb, ub);
    int64_t p0 = colptr[c] - 1, p1 = colptr[c + 1] - 1;
    for (int64_t p = p0; p < p1; ++p) {
        int64_t r = rowval[p] - 1;
        if (r < s) continue;
        if (r > e) break;
        if (r > c) {
            if (want_L) T[lb + colL_off(jl, Pa) + (r - c - 1)] = Scalar<TS>::from_f64(val[p]);
        } else if (r < c) {
            if (want_U) T[ub + colU_off(jl, Pa) + (r - s)] = Scalar<TS>::from_f64(val[p]);
        }
    }
}

// DiagonalIndices (:422-436): position of a_cc in column c by binary search;
// the smallest failing column (missing or zero diagonal) is kept in *bad.
__global__ void k_diag_indices(int64_t n, const int64_t *__restrict__ colptr,
                               const int64_t *__restrict__ rowval, const double *__restrict__ val,
                               int64_t *__restrict__ diag, unsigned long long *__restrict__ bad) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int64_t lo = colptr[c] - 1, hi = colptr[c + 1] - 1;  // [lo, hi)
    int64_t end = hi;
    int64_t want = c + 1;  // 1-based row index
    while (lo < hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if (rowval[mid] < want) lo = mid + 1; else hi = mid;
    }
    if (lo >= end || rowval[lo] != want || val[lo] == 0.0) {
        atomicMin(bad, (unsigned long long)(c + 1));
        diag[c] = -1;
    } else {
        diag[c] = lo;
    }
}

template <class TD>
__global__ void k_convert(int64_t n, const double *__restrict__ src, TD *__restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (TD)src[i];
}

// ---------------------------------------------------------------------------
// P = 64 symmetric sweep streamed through LDS.
//
// A full partition's block is 64*63 = 4032 values = 63 chunks of 64 values,
// the lower triangle's columns then the upper triangle's (descending).  The
// wave copies chunk c with one aligned full-wave load (lane i takes value
// 64c + i) into a ring of NR chunks in LDS, RUN steps before the first step
// that reads it, and every recurrence step reads its column back from LDS
// (one value per lane, consecutive addresses).  Compared with loading each
// column straight from global memory this halves the vector-memory
// instructions and makes every one of them a 512-byte aligned transfer; the
// whole schedule is fixed at compile time.
// ---------------------------------------------------------------------------
namespace ring64 {
constexpr int P = 64;
constexpr int NSTEP = 2 * (P - 1);          // 63 forward + 63 backward columns
constexpr int LTRI = P * (P - 1) / 2;       // 2016
constexpr int NCHUNK = P - 1;               // 4032 / 64
constexpr int colL(int j) { return j * (P - 1) - j * (j - 1) / 2; }
constexpr int colU(int j) { return (P * (P - 1) - j * (j + 1)) / 2; }
// column read at step t and the block offset of lane 0's value
constexpr int step_col(int t) { return t < P - 1 ? t : (P - 1) - (t - (P - 1)); }
constexpr int step_base(int t) {
    return t < P - 1 ? colL(t) - t - 1 : LTRI + colU(step_col(t));
}
// live values of step t: [lo, hi)
constexpr int step_lo(int t) { return t < P - 1 ? colL(t) : LTRI + colU(step_col(t)); }
constexpr int step_hi(int t) {
    return t < P - 1 ? colL(t + 1) : LTRI + colU(step_col(t)) + step_col(t);
}
constexpr int first_use(int c) {
    for (int t = 0; t < NSTEP; ++t)
        if (step_lo(t) < 64 * (c + 1) && step_hi(t) > 64 * c) return t;
    return NSTEP;
}
constexpr int last_use(int c) {
    int last = -1;
    for (int t = 0; t < NSTEP; ++t)
        if (step_lo(t) < 64 * (c + 1) && step_hi(t) > 64 * c) last = t;
    return last;
}
// Ring slots must not be rewritten while an older chunk is still live.  A
// chunk first read at step t is written to LDS while the group of G steps
// before t's group computes, after that earlier group's columns were read
// into registers: the older chunk of the slot must not be read at or after
// the start of the writing group's successor, i.e. last use < G*floor(t/G).
template <int NR, int G>
constexpr bool ring_ok() {
    for (int c = NR; c < NCHUNK; ++c)
        if (last_use(c - NR) >= (first_use(c) / G) * G) return false;
    return true;
}
}  // namespace ring64

namespace ring64 {
struct Sched {
    int first[NCHUNK];  // first step reading chunk c
    int issue[NCHUNK];  // step whose group issues the load of chunk c (0: prologue)
};
constexpr Sched make_sched(int run) {
    Sched sc{};
    for (int c = 0; c < NCHUNK; ++c) {
        sc.first[c] = first_use(c);
        int is = first_use(c) - run;
        sc.issue[c] = is < 0 ? 0 : is;
    }
    return sc;
}
// k-th chunk (ascending) whose first use, or whose issue step, is t; -1 if none
constexpr int kth_first(const Sched &sc, int t, int k) {
    for (int c = 0; c < NCHUNK; ++c)
        if (sc.first[c] == t && k-- == 0) return c;
    return -1;
}
constexpr int kth_issue(const Sched &sc, int t, int k) {
    for (int c = 0; c < NCHUNK; ++c)
        if (sc.issue[c] == t && k-- == 0) return c;
    return -1;
}
constexpr int max_per_step(const Sched &sc) {
    int m = 0;
    for (int t = 1; t < NSTEP; ++t) {
        int a = 0, b = 0;
        for (int c = 0; c < NCHUNK; ++c) {
            a += sc.first[c] == t;
            b += sc.issue[c] == t;
        }
        m = a > m ? a : m;
        m = b > m ? b : m;
    }
    return m;
}
constexpr int count_issue0(const Sched &sc) {
    int a = 0;
    for (int c = 0; c < NCHUNK; ++c) a += sc.issue[c] == 0;
    return a;
}
}  // namespace ring64

template <class TS, class TC, int NR, int RUN>
__device__ __forceinline__ TC sweep_ring64(TC xr, TC dr, TC rr, int lane,
                                           const TS *__restrict__ blk, TS *ring) {
    using namespace ring64;
    constexpr int G = 6;                       // steps per register group
    static_assert(ring_ok<NR, G>(), "LDS ring too small for the chunk schedule");
    static_assert((NR & (NR - 1)) == 0, "ring size must be a power of two");
    constexpr Sched SC = make_sched(RUN);
    static_assert(max_per_step(SC) <= 2, "more than two chunks scheduled in one step");
    constexpr int NG = (NSTEP + G - 1) / G;
    constexpr int RING = NR * 64;
    TS chunk[NCHUNK];                          // in-flight chunk registers
    TS col[2][G];                              // column values of two groups

    // global loads issued at step t >= 1 (first use RUN steps later)
    auto issue_at = [&](auto T) {
        constexpr int t = decltype(T)::value;
        constexpr int c0 = kth_issue(SC, t, 0), c1 = kth_issue(SC, t, 1);
        if constexpr (c0 >= 0) chunk[c0] = ld_stream(blk + 64 * c0 + lane);
        if constexpr (c1 >= 0) chunk[c1] = ld_stream(blk + 64 * c1 + lane);
    };
    // LDS writes of the chunks first read at step t
    auto write_at = [&](auto T) {
        constexpr int t = decltype(T)::value;
        constexpr int c0 = kth_first(SC, t, 0), c1 = kth_first(SC, t, 1);
        if constexpr (c0 >= 0) ring[(c0 % NR) * 64 + lane] = chunk[c0];
        if constexpr (c1 >= 0) ring[(c1 % NR) * 64 + lane] = chunk[c1];
    };
    auto read_at = [&](auto T) -> TS {
        constexpr int t = decltype(T)::value;
        constexpr int b = step_base(t);
        constexpr int m = ((b % RING) + RING) % RING;
        if constexpr (m + 63 < RING) return ring[m + lane];  // immediate offset
        else return ring[(lane + m) & (RING - 1)];
    };

    // Lanes exchange values through LDS: a lane reads what other lanes wrote.
    // From one thread's view ring[a + lane] and ring[b + lane] never alias, so
    // without a fence the compiler may reorder a write and a read; a
    // wavefront-scope fence orders them (the LDS itself executes a wave's
    // operations in order, so no wait instruction is needed).
    auto lds_fence = [] { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); };

    // prologue: the loads first used within RUN steps, group 0 into LDS and registers
    static_for<0, count_issue0(SC)>([&](auto K) {
        constexpr int c = kth_issue(SC, 0, decltype(K)::value);
        chunk[c] = ld_stream(blk + 64 * c + lane);
    });
    static_for<0, G>([&](auto T) { write_at(T); });
    lds_fence();
    static_for<0, G>([&](auto T) { col[0][decltype(T)::value] = read_at(T); });

    RowAcc<TC> row(xr);
    TC yfin = (TC)0, ysym = xr;
    static_for<0, NG>([&](auto Gi) {
        constexpr int g = decltype(Gi)::value;
        constexpr int t0 = g * G;
        // keep each group's loads, LDS traffic and arithmetic in place: left
        // alone the scheduler hoists every chunk load to the top
        __builtin_amdgcn_sched_barrier(0);
        static_for<t0 + 1, t0 + G + 1>([&](auto T) {
            if constexpr (decltype(T)::value < NSTEP) issue_at(T);
        });
        lds_fence();  // earlier reads of the slots rewritten below
        static_for<t0 + G, t0 + 2 * G>([&](auto T) {
            if constexpr (decltype(T)::value < NSTEP) write_at(T);
        });
        lds_fence();  // the writes above before the reads below
        static_for<t0 + G, t0 + 2 * G>([&](auto T) {
            constexpr int t = decltype(T)::value;
            if constexpr (t < NSTEP) col[(g + 1) & 1][t - t0 - G] = read_at(T);
        });
        static_for<t0, t0 + G>([&](auto T) {
            constexpr int t = decltype(T)::value;
            if constexpr (t < NSTEP) {
                if constexpr (t == P - 1) {
                    // last forward row, y .= D_Dl1 .* y (:897), then the backward pass
                    TC q = row.q(dr, rr);
                    TC yf = (lane == P - 1) ? q : yfin;
                    ysym = dr * yf;
                    row = RowAcc<TC>(ysym);
                    yfin = (TC)0;
                }
                constexpr int j = step_col(t);
                TC q = row.q(dr, rr);
                TC yj = bcast_keep<64>(q, j, yfin);
                row.add(load_as<TC>(col[g & 1][t - t0]), yj);  // dead for finished lanes
            }
        });
    });
    TC q = row.q(dr, rr);
    return (lane == 0) ? q : yfin;
}

// LDS elements allocated per workgroup: the ring uses NR * 64 of them; the
// rest holds a CU to 8 workgroups in FP64 (20 KB each), 16 in FP32.  Fewer
// block streams at once run faster: the 128^3 FP64 apply took 176-178 us with
// the 8 KB ring alone (19 workgroups per CU), 171 us at 8 per CU, 172 us at
// 9-11, 179 us at 6 and 217 us at 4; FP32 92.5 -> 90.3 us.  The fused
// smoother form, whose prologue gathers x for the residual, keeps the ring
// alone (its round took 0.40 ms so, 0.42 ms at 8 per CU).
constexpr int RING64_ALLOC = 2560;

template <class TS, class TC, int NR, int RUN, int FORM>
__global__ __launch_bounds__(64) void k_sweep_ring64(const double *x, double *y,
                                                     const TS *__restrict__ T,
                                                     const TC *__restrict__ D, int64_t n,
                                                     int64_t nparts, const int *__restrict__ done,
                                                     CGFuse f, SmoothFuse sf = SmoothFuse{}) {
    if (done && *done) return;
    static_assert(NR * 64 <= RING64_ALLOC, "the ring fits its allocation");
    __shared__ TS ring[FORM == FORM_SMOOTH ? NR * 64 : RING64_ALLOC];
    const int lane = threadIdx.x;
    const int64_t k = blockIdx.x;
    if (k >= nparts) return;
    const int64_t s = k * 64;
    const int64_t rem = n - s;
    const int Pa = (int)(rem < 64 ? rem : 64);
    int64_t lb, ub;
    block_bases(k, 64, Pa, MODE_SYM, lb, ub);
    const bool live = lane < Pa;
    double xin = 0.0;
    if (live) {
        if constexpr (FORM == FORM_CG) xin = cg_prologue(f, s + lane);
        else if constexpr (FORM == FORM_SMOOTH) xin = row_dot(sf.A, sf.x, s + lane) - sf.b[s + lane];
        else xin = x[s + lane];
    }
    const TC xr = (TC)xin;
    const TC dr = live ? D[s + lane] : (TC)1;
    const TC rr = (TC)1 / dr;
    TC yr;
    if (__any(!safe_divisor(dr))) {
        yr = sweep_generic<64, MODE_SYM, false, TS, TC>(xr, dr, rr, lane, Pa, live, T + lb, T + ub);
    } else if (Pa == 64) {
        yr = sweep_ring64<TS, TC, NR, RUN>(xr, dr, rr, lane, T + lb, ring);
    } else {
        yr = sweep_generic<64, MODE_SYM, true, TS, TC>(xr, dr, rr, lane, Pa, live, T + lb, T + ub);
    }
    if (live) {
        if constexpr (FORM == FORM_SMOOTH) {
            sf.xout[s + lane] = sf.x[s + lane] - (double)yr;
        } else if constexpr (FORM == FORM_CG && std::is_same_v<TC, float>) {
            if (f.z32) f.z32[s + lane] = yr;  // an FP32 smoother's z, kept in FP32
            else y[s + lane] = (double)yr;
        } else {
            y[s + lane] = (double)yr;
        }
    }
    if constexpr (FORM == FORM_CG) {
        double v = live ? xin * (double)yr : 0.0;  // r'z partial of this partition
        v = wave_sum(v);
        if (lane == 0) f.partials[k] = v;
    }
}

// One partition per PM-lane segment, 256 threads per workgroup.
// TS: storage type of the triangles, TC: arithmetic type (double or float).
// x and y may alias (each partition reads its rows before writing them).
template <int PM, int MODE, class TS, class TC, int FORM = FORM_PLAIN>
__global__ __launch_bounds__(256) void k_sweep_seg(const double *x, double *y,
                                                   const TS *__restrict__ T,
                                                   const TC *__restrict__ D, int64_t n, int P,
                                                   int64_t nparts, const int *__restrict__ done,
                                                   CGFuse f = CGFuse{},
                                                   SmoothFuse sf = SmoothFuse{}) {
    if (done && *done) return;  // converged CG: speculative launches do nothing
    constexpr int SEGS = 256 / PM;
    const int lane = threadIdx.x % PM;
    const int64_t k = (int64_t)blockIdx.x * SEGS + threadIdx.x / PM;
    const bool wave_live = !__all(k >= nparts);  // the same in every lane of a wave
    if constexpr (FORM == FORM_CG) {
        // whole blocks stay for the block-wide r'z partial
        __shared__ double sm[4];
        double v = 0.0;
        if (wave_live)
            v = sweep_seg_body<PM, MODE, TS, TC, FORM>(x, y, T, D, n, P, nparts, k, lane, f, sf);
        v = block_sum(v, sm);
        if (threadIdx.x == 0) f.partials[blockIdx.x] = v;
        return;
    }
    if (!wave_live) return;
    (void)sweep_seg_body<PM, MODE, TS, TC, FORM>(x, y, T, D, n, P, nparts, k, lane, f, sf);
}

// One-sided sweeps with P == PM (16 or 32): each wave first copies the packed
// blocks of its 64 / PM partitions -- consecutive in the buffer -- into LDS
// with full-wave loads (NCH aligned 64-element chunks), then runs the same
// recurrence reading its columns from LDS.  Loading columns straight from
// memory issues one partly overlapping wave load per step; this issues
// (64/PM) P(P-1)/2 / 64 aligned ones.  A wave holding the last (short)
// partition keeps the direct path.  WPB waves per workgroup, each with its own
// LDS area and no workgroup barrier.
template <int PM, int MODE, class TS, class TC, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_sweep_seg_lds(const double *x, double *y,
                                                            const TS *__restrict__ T,
                                                            const TC *__restrict__ D, int64_t n,
                                                            int P, int64_t nparts,
                                                            const int *__restrict__ done) {
    static_assert(MODE == MODE_FWD || MODE == MODE_BWD, "one triangle per block");
    if (done && *done) return;
    constexpr int SW = 64 / PM;             // partitions per wave
    constexpr int BLK = PM * (PM - 1) / 2;  // elements per block
    constexpr int WE = SW * BLK;            // elements per wave
    constexpr int NCH = (WE + 63) / 64;
    __shared__ TS lds_all[WPB * NCH * 64];
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int lane = l64 % PM, seg = l64 / PM;
    const int64_t kw = ((int64_t)blockIdx.x * WPB + wave) * SW;  // the wave's first partition
    const int64_t k = kw + seg;
    if (__all(k >= nparts)) return;
    const bool staged = (kw + SW) * (int64_t)PM <= n;  // the same in every lane of the wave
    if (!staged) {
        (void)sweep_seg_body<PM, MODE, TS, TC, FORM_PLAIN>(x, y, T, D, n, P, nparts, k, lane,
                                                           CGFuse{}, SmoothFuse{});
        return;
    }
    TS *const lw = lds_all + wave * NCH * 64;
    const TS *const src = T + kw * (int64_t)BLK;
    TS chunk[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int idx = c * 64 + l64;
        chunk[c] = ld_stream(src + (idx < WE ? idx : WE - 1));
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int idx = c * 64 + l64;
        if (idx < WE) lw[idx] = chunk[c];
    }
    // the stores above before other lanes' reads below
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int64_t g = k * PM + lane;
    const TC xr = (TC)x[g];
    const TC dr = D[g];
    const TC rr = (TC)1 / dr;
    const TS *const blk = lw + seg * BLK;
    TC yr;
    if (__any(!safe_divisor(dr)))
        yr = sweep_generic<PM, MODE, false, TS, TC>(xr, dr, rr, lane, PM, true, blk, blk);
    else
        yr = sweep_full<PM, MODE, 16, TS, TC>(xr, dr, rr, lane, blk, blk);
    y[g] = (double)yr;
}

// One-triangle symmetric sweep (HS_CACHE_PACKED_SYM): one wave per workgroup,
// 64 / PM partitions per wave, each with its LDS copy of L (symt_lds(PM)
// elements; at PM = 64 the staged block alone, 16,128 B per wave in FP64), so
// that LDS does not limit the waves per CU below what the recurrence's
// latency needs.
template <int PM, class TS, class TC, int FORM = FORM_PLAIN>
__global__ __launch_bounds__(64) void k_sweep_symt(const double *x, double *y,
                                                   const TS *__restrict__ T,
                                                   const TC *__restrict__ D, int64_t n, int P,
                                                   int64_t nparts, const int *__restrict__ done,
                                                   CGFuse f = CGFuse{},
                                                   SmoothFuse sf = SmoothFuse{}) {
    if (done && *done) return;
    constexpr int SEGS = 64 / PM;
    // PM = 64: one partition per wave, so a wave either stages its full block
    // or holds the short last partition (register form); LDS holds only the
    // block, 2,016 values (measured within run-to-run spread of the larger
    // 2,144-value area, which the staged form does not use)
    constexpr int LSTR = PM == 64 ? PM * (PM - 1) / 2 : symt_lds(PM);
    __shared__ __align__(16) TS lds[SEGS * LSTR];  // word stores of the staged form
    const int lane = threadIdx.x % PM;
    const int seg = threadIdx.x / PM;
    const int64_t k = (int64_t)blockIdx.x * SEGS + seg;
    if (__all(k >= nparts)) {
        if constexpr (FORM == FORM_CG)
            if (threadIdx.x == 0) f.partials[blockIdx.x] = 0.0;
        return;
    }
    double v = sweep_seg_body<PM, MODE_SYMT, TS, TC, FORM, LSTR>(x, y, T, D, n, P, nparts, k,
                                                                 lane, f, sf, lds + seg * LSTR);
    if constexpr (FORM == FORM_CG) {
        v = wave_sum(v);  // r'z partial of this workgroup
        if (threadIdx.x == 0) f.partials[blockIdx.x] = v;
    }
}

// Partitions of 65 to 64 R rows: one wave per partition, lane l owning rows
// l, l + 64, ..., l + 64 (R - 1).  Step j of the right-looking recurrence:
// the owner of row j (lane j mod 64, slot j / 64) forms its quotient, it is
// broadcast from that lane, and every lane adds L(i, j) y_j to its rows
// i > j -- one coalesced load per slot from the packed column j.  Each row
// therefore sums its terms in ascending j (descending in the backward pass),
// the order of the 64-row kernels and of the reference; with FAST = false
// (a wave holding an unsafe diagonal) quotients are true divisions.
template <int R, int MODE, bool FAST, class TS, class TC>
__device__ __forceinline__ void sweep_wide(TC (&xr)[R], const TC (&dr)[R], const TC (&rr)[R],
                                           int lane, int Pa, const TS *__restrict__ Lk,
                                           const TS *__restrict__ Uk) {
    // Steps run in groups of U; the column values of the next group are
    // loaded while the current group computes (U R values per lane in
    // flight).  Lanes without a term at a step (and steps past the last
    // column) load element 0 of the block.
    constexpr int U = 16 / R > 2 ? 16 / R : 2;
    auto pass = [&](auto Upper) {
        constexpr bool UP = decltype(Upper)::value;
        const TS *__restrict__ B = UP ? Uk : Lk;
        auto step_of = [&](int t) { return UP ? Pa - 1 - t : t; };  // t-th step's column
        auto load_group = [&](int t0, TS (&buf)[U][R]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = step_of(t0 + u);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = lane + 64 * r;
                    const bool has = (t0 + u < Pa) && (UP ? i < j : (i > j && i < Pa));
                    const int64_t off = UP ? colU_off(j, Pa) + i : colL_off(j, Pa) + (i - j - 1);
                    buf[u][r] = B[has ? off : 0];
                }
            }
        };
        RowAcc<TC> row[R];
        TC yfin[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            row[r] = RowAcc<TC>(xr[r]);
            yfin[r] = (TC)0;
        }
        TS cur[U][R], nxt[U][R];
        load_group(0, cur);
        for (int t0 = 0; t0 < Pa; t0 += U) {
            load_group(t0 + U, nxt);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u;
                if (t < Pa) {
                    const int j = step_of(t);
                    const int js = j >> 6, jl = j & 63;
                    TC q = (TC)0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (r == js) q = row[r].template q<FAST>(dr[r], rr[r]);
                    const TC yj = seg_bcast<64>(q, jl);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if (r == js && lane == jl) yfin[r] = q;
                        const int i = lane + 64 * r;
                        if (UP ? i < j : (i > j && i < Pa)) row[r].add(load_as<TC>(cur[u][r]), yj);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int r = 0; r < R; ++r) cur[u][r] = nxt[u][r];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = yfin[r];
    };
    if constexpr (MODE != MODE_BWD) pass(std::false_type{});
    if constexpr (MODE == MODE_SYM) {
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = dr[r] * xr[r];  // y .= D_Dl1 .* y (:897)
    }
    if constexpr (MODE != MODE_FWD) pass(std::true_type{});
}

// sweep_wide for a full partition (64 R rows): the steps run in R blocks of
// 64 whose owner row slot js is known at compile time, so the quotient is the
// owner slot's, slots already finished take no work, and slots past the
// owner (forward; before it, backward) have a term in every lane with no
// test.  In the owner slot the lanes whose rows are finished add a term of
// the packed block's element 0 (their quotients were kept when they formed:
// bcast_keep), as in the 64-row kernels.  Loads, products and their order
// per row are those of sweep_wide, so the results are the same bit for bit.
template <int R, int MODE, bool FAST, class TS, class TC>
__device__ __forceinline__ void sweep_wide_full(TC (&xr)[R], const TC (&dr)[R],
                                                const TC (&rr)[R], int lane,
                                                const TS *__restrict__ Lk,
                                                const TS *__restrict__ Uk) {
    constexpr int P = 64 * R;
    constexpr int U = 4;                    // steps per group, a divisor of 64
    constexpr int NGB = 64 / U;             // groups per block
    static_assert(NGB * U == 64, "the groups cover the block");
    auto pass = [&](auto Upper) {
        constexpr bool UP = decltype(Upper)::value;
        const TS *__restrict__ B = UP ? Uk : Lk;
        RowAcc<TC> row[R];
        TC yfin[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            row[r] = RowAcc<TC>(xr[r]);
            yfin[r] = (TC)0;
        }
        static_for<0, R>([&](auto Bk) {
            constexpr int b = decltype(Bk)::value;
            constexpr int js = UP ? R - 1 - b : b;  // row slot owning the block's steps
            constexpr int r0 = UP ? 0 : js, r1 = UP ? js + 1 : R;  // slots with terms
            auto load_group = [&](int g, TS (&buf)[U][R]) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jl = UP ? 63 - (g * U + u) : g * U + u;
                    const int j = 64 * js + jl;
#pragma unroll
                    for (int r = r0; r < r1; ++r) {
                        const int i = lane + 64 * r;
                        const bool has = (r != js) || (UP ? lane < jl : lane > jl);
                        const int64_t off = UP ? colU_off(j, P) + i : colL_off(j, P) + (i - j - 1);
                        buf[u][r] = B[has ? off : 0];
                    }
                }
            };
            TS cur[U][R], nxt[U][R];
            load_group(0, cur);
            // rolled: unrolled, the blocks' loads are hoisted and the kernel
            // runs out of registers
#pragma unroll 1
            for (int g = 0; g < NGB; ++g) {
                if (g + 1 < NGB) load_group(g + 1, nxt);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jl = UP ? 63 - (g * U + u) : g * U + u;
                    const TC q = row[js].template q<FAST>(dr[js], rr[js]);
                    const TC yj = bcast_keep<64>(q, jl, yfin[js]);
#pragma unroll
                    for (int r = r0; r < r1; ++r) row[r].add(load_as<TC>(cur[u][r]), yj);
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int r = r0; r < r1; ++r) cur[u][r] = nxt[u][r];
            }
        });
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = yfin[r];
    };
    if constexpr (MODE != MODE_BWD) pass(std::false_type{});
    if constexpr (MODE == MODE_SYM) {
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = dr[r] * xr[r];  // y .= D_Dl1 .* y (:897)
    }
    if constexpr (MODE != MODE_FWD) pass(std::true_type{});
}

template <int R, int MODE, class TS, class TC>
__global__ __launch_bounds__(64) void k_sweep_wide(const double *x, double *y,
                                                   const TS *__restrict__ T,
                                                   const TC *__restrict__ D, int64_t n, int64_t P,
                                                   int64_t nparts, const int *__restrict__ done) {
    if (done && *done) return;
    const int64_t k = blockIdx.x;
    if (k >= nparts) return;
    const int lane = threadIdx.x;
    const int64_t s = k * P;
    const int Pa = (int)(n - s < P ? n - s : P);
    int64_t lb, ub;
    block_bases(k, P, Pa, MODE, lb, ub);
    TC xr[R], dr[R], rr[R];
    bool unsafe = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        const bool live = i < Pa;
        xr[r] = live ? (TC)x[s + i] : (TC)0;
        dr[r] = live ? D[s + i] : (TC)1;
        rr[r] = (TC)1 / dr[r];
        unsafe = unsafe || !safe_divisor(dr[r]);
    }
    if (__any(unsafe))
        sweep_wide<R, MODE, false, TS, TC>(xr, dr, rr, lane, Pa, T + lb, T + ub);
    else if (Pa == 64 * R)
        sweep_wide_full<R, MODE, true, TS, TC>(xr, dr, rr, lane, T + lb, T + ub);
    else
        sweep_wide<R, MODE, true, TS, TC>(xr, dr, rr, lane, Pa, T + lb, T + ub);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < Pa) y[s + i] = (double)xr[r];
    }
}

// Partitions larger than a wave: one thread per partition, rows in sequence,
// summing in the reference's order (left-looking).  Correctness path for
// large partsize (e.g. partsize = N in the Gauss-Seidel sanity checks).
template <int MODE, class TS, class TC>
__global__ void k_sweep_thread(const double *x, double *y, const TS *__restrict__ T,
                               const TC *__restrict__ D, int64_t n, int64_t P, int64_t nparts,
                               const int *__restrict__ done) {
    if (done && *done) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nparts) return;
    int64_t s = k * P;
    int64_t Pa = (n - s < P) ? n - s : P;
    int64_t lb, ub;
    block_bases(k, P, Pa, MODE, lb, ub);
    if (x != y)
        for (int64_t i = 0; i < Pa; ++i) y[s + i] = x[s + i];
    if (MODE != MODE_BWD) {
        for (int64_t i = 0; i < Pa; ++i) {
            RowAcc<TC> row((TC)y[s + i]);
            for (int64_t j = 0; j < i; ++j)
                row.add(load_as<TC>(T[lb + colL_off(j, Pa) + (i - j - 1)]), (TC)y[s + j]);
            y[s + i] = (double)row.template q<false>(D[s + i], (TC)0);
        }
    }
    if (MODE == MODE_SYM)
        for (int64_t i = 0; i < Pa; ++i) y[s + i] = (double)(D[s + i] * (TC)y[s + i]);
    if (MODE != MODE_FWD) {
        for (int64_t i = Pa - 1; i >= 0; --i) {
            RowAcc<TC> row((TC)y[s + i]);
            for (int64_t j = i + 1; j < Pa; ++j)
                row.add(load_as<TC>(T[ub + colU_off(j, Pa) + i]), (TC)y[s + j]);
            y[s + i] = (double)row.template q<false>(D[s + i], (TC)0);
        }
    }
}

// MatrixViewCache: the in-partition entries are read straight from the matrix
// (reference :1588-1737), one thread per partition as in the reference's
// kernel.  rowwise=1: 1-based CSR or symmetric CSC read by rows;
// rowwise=0: non-symmetric CSC with DiagonalIndices binary searches.
__device__ __forceinline__ int64_t bsearch_row(const int64_t *rowval, int64_t lo, int64_t hi,
                                               int64_t want) {
    // 0-based positions, inclusive [lo, hi]; rowval is 1-based; want 1-based
    int64_t left = lo, right = hi;
    while (left <= right) {
        int64_t mid = (left + right) >> 1;
        int64_t r = rowval[mid];
        if (r < want) left = mid + 1;
        else if (r > want) right = mid - 1;
        else return mid;
    }
    return -1;
}

__device__ double view_acc(int rowwise, int upper, const int64_t *ptr, const int64_t *ind,
                           const double *val, const int64_t *diag, const double *y, int64_t g,
                           int64_t s, int64_t e) {
    double acc = 0.0;
    if (rowwise) {
        for (int64_t p = ptr[g] - 1; p < ptr[g + 1] - 1; ++p) {
            int64_t j = ind[p] - 1;
            if (!upper) {
                if (s <= j && j < g) acc += val[p] * y[j];
            } else {
                if (g < j && j <= e) acc += val[p] * y[j];
            }
        }
        return acc;
    }
    int64_t c0 = upper ? g + 1 : s;
    int64_t c1 = upper ? e : (g - 1 < e ? g - 1 : e);
    for (int64_t col = c0; col <= c1; ++col) {
        int64_t di = diag[col];
        int64_t lo = upper ? ptr[col] - 1 : di + 1;
        int64_t hi = upper ? di - 1 : ptr[col + 1] - 2;
        if (lo <= hi && ind[lo] <= g + 1 && ind[hi] >= g + 1) {
            int64_t m = bsearch_row(ind, lo, hi, g + 1);
            if (m >= 0) acc += val[m] * y[col];
        }
    }
    return acc;
}

template <int MODE>
__global__ void k_sweep_view(const double *x, double *y, const int64_t *__restrict__ ptr,
                             const int64_t *__restrict__ ind, const double *__restrict__ val,
                             const int64_t *__restrict__ diag, int rowwise,
                             const double *__restrict__ D, int64_t n, int64_t P, int64_t nparts,
                             const int *__restrict__ done) {
    if (done && *done) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nparts) return;
    int64_t s = k * P;
    int64_t e = (s + P < n ? s + P : n) - 1;
    if (x != y)
        for (int64_t g = s; g <= e; ++g) y[g] = x[g];
    if (MODE != MODE_BWD)
        for (int64_t g = s; g <= e; ++g)
            y[g] = (y[g] - view_acc(rowwise, 0, ptr, ind, val, diag, y, g, s, e)) / D[g];
    if (MODE == MODE_SYM)
        for (int64_t g = s; g <= e; ++g) y[g] = D[g] * y[g];
    if (MODE != MODE_FWD)
        for (int64_t g = e; g >= s; --g)
            y[g] = (y[g] - view_acc(rowwise, 1, ptr, ind, val, diag, y, g, s, e)) / D[g];
}

// MatrixViewCache with one wave per partition (partsize <= 64), the matrix
// read by rows (a CSR, a symmetric CSC read as rows, or the CSR transpose of
// a non-symmetric CSC).  Lane i owns row g = s + i and the right-looking
// recurrence runs as in the packed kernels, but a lane adds a term only at
// the steps where its row has an in-partition entry: it holds the next KV
// entries of that range (column, value) in registers, compares the next
// column with the step and, on a match, adds the product and shifts the
// next entry in (reloading KV more from memory when its cache runs dry).
// Terms are added in ascending column order going forward and descending
// going backward -- the packed kernels' order -- and the zeros a packed
// triangle would add are skipped, which changes nothing for finite values:
// the results equal PackedBufferCache's bit for bit.
constexpr int KV = 4;

struct ViewCache {
    int c[KV];
    double v[KV];
};

// loads up to KV entries starting at position p, walking by `dir` (+1 or -1)
// while inside [lo, hi); columns relative to the partition start s
__device__ __forceinline__ void view_fill(ViewCache &vc, int64_t p, int dir, int64_t lo, int64_t hi,
                                          const int64_t *__restrict__ ind,
                                          const double *__restrict__ val, int64_t s) {
#pragma unroll
    for (int k = 0; k < KV; ++k) {
        const int64_t q = p + dir * k;
        const bool in = q >= lo && q < hi;
        vc.c[k] = in ? (int)(ind[q] - 1 - s) : -1;
        vc.v[k] = in ? val[q] : 0.0;
    }
}

// first position in [lo, hi) whose 0-based column is >= c (columns ascend)
__device__ __forceinline__ int64_t row_lower_bound(const int64_t *__restrict__ ind, int64_t lo,
                                                   int64_t hi, int64_t c) {
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (ind[mid] - 1 < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int MODE, bool FAST>
__device__ __forceinline__ double view_wave(double xr, double dr, double rr, int lane, int Pa,
                                            int64_t s, int64_t plo, int64_t pmid, int64_t pmid1,
                                            int64_t phi, const int64_t *__restrict__ ind,
                                            const double *__restrict__ val) {
    double yr = xr;
    if constexpr (MODE != MODE_BWD) {
        // lower entries [plo, pmid): columns s .. g-1, consumed in ascending order
        RowAcc<double> row(xr);
        double yfin = 0.0;
        ViewCache vc;
        int64_t p = plo;
        view_fill(vc, p, 1, plo, pmid, ind, val, s);
        for (int j = 0; j < Pa; ++j) {
            const double q = row.template q<FAST>(dr, rr);
            const double yj = seg_bcast<64>(q, j);
            yfin = (lane == j) ? q : yfin;
            if (vc.c[0] == j) {
                row.add(vc.v[0], yj);
                ++p;
#pragma unroll
                for (int k = 0; k + 1 < KV; ++k) {
                    vc.c[k] = vc.c[k + 1];
                    vc.v[k] = vc.v[k + 1];
                }
                vc.c[KV - 1] = -1;
                if (vc.c[0] < 0 && p < pmid) view_fill(vc, p, 1, plo, pmid, ind, val, s);
            }
        }
        yr = yfin;
    }
    if constexpr (MODE == MODE_SYM) xr = dr * yr;  // y .= D_Dl1 .* y (:897)
    if constexpr (MODE != MODE_FWD) {
        // upper entries [pmid1, phi): columns g+1 .. e, consumed in descending order
        RowAcc<double> row(xr);
        double yfin = 0.0;
        ViewCache vc;
        int64_t p = phi - 1;
        view_fill(vc, p, -1, pmid1, phi, ind, val, s);
        for (int j = Pa - 1; j >= 0; --j) {
            const double q = row.template q<FAST>(dr, rr);
            const double yj = seg_bcast<64>(q, j);
            yfin = (lane == j) ? q : yfin;
            if (vc.c[0] == j) {
                row.add(vc.v[0], yj);
                --p;
#pragma unroll
                for (int k = 0; k + 1 < KV; ++k) {
                    vc.c[k] = vc.c[k + 1];
                    vc.v[k] = vc.v[k + 1];
                }
                vc.c[KV - 1] = -1;
                if (vc.c[0] < 0 && p >= pmid1) view_fill(vc, p, -1, pmid1, phi, ind, val, s);
            }
        }
        yr = yfin;
    }
    return yr;
}

template <int MODE>
__global__ __launch_bounds__(64) void k_sweep_view_wave(const double *x, double *y,
                                                        const int64_t *__restrict__ ptr,
                                                        const int64_t *__restrict__ ind,
                                                        const double *__restrict__ val,
                                                        const double *__restrict__ D, int64_t n,
                                                        int64_t P, int64_t nparts,
                                                        const int *__restrict__ done) {
    if (done && *done) return;
    const int64_t k = blockIdx.x;
    if (k >= nparts) return;
    const int lane = threadIdx.x;
    const int64_t s = k * P;
    const int Pa = (int)(n - s < P ? n - s : P);
    const bool live = lane < Pa;
    const int64_t g = s + lane;
    int64_t plo = 0, pmid = 0, pmid1 = 0, phi = 0;
    double xr = 0.0, dr = 1.0;
    if (live) {
        xr = x[g];
        dr = D[g];
        const int64_t p0 = ptr[g] - 1, p1 = ptr[g + 1] - 1;
        plo = row_lower_bound(ind, p0, p1, s);             // first column >= s
        pmid = row_lower_bound(ind, plo, p1, g);           // first column >= g
        pmid1 = (pmid < p1 && ind[pmid] - 1 == g) ? pmid + 1 : pmid;  // past the diagonal
        phi = row_lower_bound(ind, pmid1, p1, s + Pa);     // first column past the partition
    }
    const double rr = 1.0 / dr;
    double yr;
    if (__any(!safe_divisor(dr)))
        yr = view_wave<MODE, false>(xr, dr, rr, lane, Pa, s, plo, pmid, pmid1, phi, ind, val);
    else
        yr = view_wave<MODE, true>(xr, dr, rr, lane, Pa, s, plo, pmid, pmid1, phi, ind, val);
    if (live) y[g] = yr;
}

// MatrixViewCache at partsize 64 from the sliced-ELL copy of the rows (the
// device operator's layout, hs_sparse.hpp): partition k is slice k, so a
// full-wave load of the slice's entry k gives lane i the k-th stored entry of
// its own row (value, and its 32-bit column).  The wave loads its rows'
// W <= WMAX entries once, maps each in-partition entry to the recurrence step
// that uses it (forward step j for L(i, j), backward step 126 - j for U(i, j)
// in the symmetric sweep), keeps those that have a step in C <= WMAX slots
// (C = 2, 4 or WMAX, the fewest that hold every lane's), and runs the packed
// kernels' step sequence with each step's coefficient taken from a per-lane
// ring of NR = 8 step slots in LDS (4 KB): every G = 4 steps, in the group of
// slots that has been read, the slots that held an entry are zeroed and the
// entries of the group two ahead are written.
// A lane only ever reads and writes its own column of the ring
// (slot * 64 + lane), so no cross-lane ordering is involved.  Slots without an
// entry hold zero, so each step adds L(i, j) y_j or 0 * y_j exactly as the
// packed triangles do: the results equal PackedBufferCache's bit for bit.
// A lane's own quotient is not kept at its step: the row sum starts at +0 and
// a rounded sum is -0 only when both addends are, so the sum is never -0, and
// adding 0 * y_j = +-0 (y_j finite) leaves it unchanged bit for bit.  Each
// lane's quotient formed once after the last step is therefore the one of its
// own step.  A non-finite y_j makes 0 * y_j a NaN; it also leaves the output
// of lane j (or of a lane that took the NaN) non-finite, and the caller then
// redoes the wave by rows (view_wave), which adds stored entries only.
// Matrix bytes read once per application: nnz (8 + 4) + 8 (slices + 1).
namespace viewring {
constexpr int LG = 2;
constexpr int G = 1 << LG;  // steps per group
constexpr int NR = 2 * G;   // ring slots (two groups)
template <int MODE>
constexpr int nstep() { return MODE == MODE_SYM ? 126 : 63; }
// column of step t
template <int MODE>
__host__ __device__ constexpr int step_col(int t) {
    return MODE == MODE_FWD ? t : (MODE == MODE_BWD ? 63 - t : (t < 63 ? t : 126 - t));
}
}  // namespace viewring

// The ring run over a lane's C entries (steps tk, -1: none; values v).
template <int MODE, int C>
__device__ __forceinline__ double view_ring_run(double xr, double dr, double rr, int lane,
                                                const int (&tk)[C], const double (&v)[C],
                                                double *ring) {
    using namespace viewring;
    constexpr int NS = nstep<MODE>();
    constexpr int NG = (NS + G - 1) / G;
    static_assert(NR == 2 * G, "the ring holds two groups");
    double *const mine = ring + lane;  // slot u of this lane: mine[64 u]
    // each entry's group and slot, once
    int gk[C], sk[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
        gk[k] = tk[k] >= 0 ? tk[k] >> LG : -1;
        sk[k] = 64 * (tk[k] & (NR - 1));
    }
    // groups 0 and 1 into the cleared ring
#pragma unroll
    for (int u = 0; u < NR; ++u) mine[64 * u] = 0.0;
#pragma unroll
    for (int k = 0; k < C; ++k)
        if (gk[k] == 0 || gk[k] == 1) mine[sk[k]] = v[k];
    double col[2][G];
#pragma unroll
    for (int u = 0; u < G; ++u) col[0][u] = mine[64 * u];

    RowAcc<double> row(xr);
    static_for<0, NG>([&](auto Gi) {
        constexpr int g = decltype(Gi)::value;
        constexpr int t0 = g * G;
        if constexpr ((g + 2) * G < NS) {
            // group g's slots were read into registers: zero the ones that
            // held an entry, then write the entries of group g + 2 (the same
            // slots); slots that held none are still zero
#pragma unroll
            for (int k = 0; k < C; ++k)
                if (gk[k] == g) mine[sk[k]] = 0.0;
#pragma unroll
            for (int k = 0; k < C; ++k)
                if (gk[k] == g + 2) mine[sk[k]] = v[k];
        }
        if constexpr ((g + 1) * G < NS) {
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (t0 + G + u < NS) col[(g + 1) & 1][u] = mine[64 * ((t0 + G + u) % NR)];
        }
        static_for<t0, t0 + G>([&](auto T) {
            constexpr int t = decltype(T)::value;
            if constexpr (t < NS) {
                if constexpr (MODE == MODE_SYM && t == 63) {
                    // every row's forward quotient, y .= D_Dl1 .* y (:897),
                    // then the backward pass
                    row = RowAcc<double>(dr * row.q(dr, rr));
                }
                constexpr int j = step_col<MODE>(t);
                const double yj = seg_bcast<64>(row.q(dr, rr), j);
                row.add(col[g & 1][t - t0], yj);  // zero for lanes without a term
            }
        });
    });
    return row.q(dr, rr);
}

// The first C entries with a step, in stored order (at most C per lane).
template <int C, int W>
__device__ __forceinline__ void view_ring_compact(const int (&tk)[W], const double (&v)[W],
                                                  int (&ct)[C], double (&cv)[C]) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
        ct[q] = -1;
        cv[q] = 0.0;
    }
#pragma unroll
    for (int k = W - 1; k >= 0; --k) {
        if (tk[k] >= 0) {
#pragma unroll
            for (int q = C - 1; q > 0; --q) {
                ct[q] = ct[q - 1];
                cv[q] = cv[q - 1];
            }
            ct[0] = tk[k];
            cv[0] = v[k];
        }
    }
}

template <int MODE, int WMAX>
__device__ __forceinline__ double view_ring(double xr, double dr, double rr, int lane, int64_t s,
                                            const SellView<double> &A, int64_t e0, int w,
                                            double *ring) {
    // this lane's entries, their steps (-1: none) and values
    double v[WMAX];
    int tk[WMAX];
#pragma unroll
    for (int k = 0; k < WMAX; ++k) {
        v[k] = 0.0;
        tk[k] = -1;
    }
    int cnt = 0;
    if (w > 0) {  // wave-uniform
        // every load is issued with no branch around it (loads under a
        // per-entry branch were each waited for inside it): past the slice
        // width the slice's last entry is loaded again and not used
#pragma unroll
        for (int k = 0; k < WMAX; ++k) {
            const int64_t p = e0 + 64 * (k < w ? k : w - 1) + lane;
            const int j = (int)((int64_t)A.col[p] - s);  // column within the partition
            const double val = A.val[p];
            int t = -1;
            if (MODE != MODE_BWD && j >= 0 && j < lane) t = j;
            if (MODE == MODE_SYM && j > lane && j < 64) t = 126 - j;
            if (MODE == MODE_BWD && j > lane && j < 64) t = 63 - j;
            if (k < w) {
                tk[k] = t;
                v[k] = val;
                cnt += t >= 0;
            }
        }
    }
    // the ring upkeep costs a compare per held entry and group: hold only the
    // entries with a step (a row of the 3D Poisson stencil has two, its
    // neighbours on the same grid line), in the fewest slots that fit every
    // lane of the wave
    if (!__any(cnt > 2)) {
        int ct[2];
        double cv[2];
        view_ring_compact<2>(tk, v, ct, cv);
        return view_ring_run<MODE, 2>(xr, dr, rr, lane, ct, cv, ring);
    }
    if constexpr (WMAX > 4) {
        if (!__any(cnt > 4)) {
            int ct[4];
            double cv[4];
            view_ring_compact<4>(tk, v, ct, cv);
            return view_ring_run<MODE, 4>(xr, dr, rr, lane, ct, cv, ring);
        }
    }
    return view_ring_run<MODE, WMAX>(xr, dr, rr, lane, tk, v, ring);
}

__device__ __forceinline__ bool finite64(double v) {
    return (__double2hiint(v) & 0x7ff00000) != 0x7ff00000;
}

// partsize 64: full partitions whose slice is at most WMAX wide and whose
// diagonals are safe take view_ring; others (the short last partition, wider
// rows, unsafe diagonals) the general view_wave on the 1-based row arrays.
template <int MODE, int WMAX>
__global__ __launch_bounds__(64) void k_sweep_view_ring(const double *x, double *y,
                                                        SellView<double> A,
                                                        const int64_t *__restrict__ ptr,
                                                        const int64_t *__restrict__ ind,
                                                        const double *__restrict__ val,
                                                        const double *__restrict__ D, int64_t n,
                                                        int64_t nparts,
                                                        const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ double ring[viewring::NR * 64];
    const int64_t k = blockIdx.x;
    if (k >= nparts) return;
    const int lane = threadIdx.x;
    const int64_t s = k * 64;
    const int Pa = (int)(n - s < 64 ? n - s : 64);
    const bool live = lane < Pa;
    const int64_t g = s + lane;
    const double xr = live ? x[g] : 0.0;
    const double dr = live ? D[g] : 1.0;
    const double rr = 1.0 / dr;
    const int64_t e0 = A.sptr[k];
    const int w = (int)((A.sptr[k + 1] - e0) >> 6);
    double yr = 0.0;
    bool ring_done = false, nonfinite = false;
    if (Pa == 64 && w <= WMAX && !__any(!safe_divisor(dr))) {
        yr = view_ring<MODE, WMAX>(xr, dr, rr, lane, s, A, e0, w, ring);
        nonfinite = __any(!finite64(yr));
        ring_done = !nonfinite;
    }
    if (!ring_done) {
        int64_t plo = 0, pmid = 0, pmid1 = 0, phi = 0;
        if (live) {
            const int64_t p0 = ptr[g] - 1, p1 = ptr[g + 1] - 1;
            plo = row_lower_bound(ind, p0, p1, s);
            pmid = row_lower_bound(ind, plo, p1, g);
            pmid1 = (pmid < p1 && ind[pmid] - 1 == g) ? pmid + 1 : pmid;
            phi = row_lower_bound(ind, pmid1, p1, s + Pa);
        }
        // true divisions where a diagonal is unsafe, or where a quotient is
        // not finite (the fast quotient of an infinite x - sum is a NaN)
        if (nonfinite || __any(!safe_divisor(dr)))
            yr = view_wave<MODE, false>(xr, dr, rr, lane, Pa, s, plo, pmid, pmid1, phi, ind, val);
        else
            yr = view_wave<MODE, true>(xr, dr, rr, lane, Pa, s, plo, pmid, pmid1, phi, ind, val);
    }
    if (live) y[g] = yr;
}

// MatrixViewCache for partsize 65 to 256: one wave per partition in the
// packed k_sweep_wide shape, lane l owning rows l, l + 64, ..., l + 64 (R - 1)
// of the partition, and the right-looking recurrence of sweep_wide (the owner
// of row j forms its quotient, it is broadcast, every row i past j adds
// L(i, j) y_j, ascending j forward and descending backward).
//
// view_ring_wide: a row's entries inside the partition are contiguous in its
// stored order (columns ascend), so each lane counts, from its row's 32-bit
// columns in the sliced-ELL copy, the entries before the partition and before
// its end, and loads only that window (at most KS entries: value, and the step that
// uses it -- forward step j for L(i, j), backward step 64 R - 1 - j for
// U(i, j)).  The steps run in 64-step blocks, one per row slot, so the owner's
// slot is known at compile time and rows already finished in a block take no
// work; every partition runs all 64 R steps (rows past the end hold x = 0 and
// no entries, and come after the live rows going forward and before them,
// adding exact zeros, going back).  Within a block, groups of G steps read
// their coefficients from a per-lane, per-row-slot LDS ring of G slots: a
// group reads its slots, zeroes the ones that held an entry and writes the
// entries of the next group (every entry is written, to its slot or to a
// per-lane trash slot, so the fill has no branches; the scans stop at the
// widest window of the wave).
// A lane only reads and writes its own ring columns.  Slots without an entry
// hold zero and finished rows take dead additions, as in the partsize-64 ring:
// the results equal PackedBufferCache's bit for bit, and each row's quotient
// is formed once after its pass (the additions 0 * y_j leave a finished row's
// sum as it was while every y_j is finite).  Returns 0 when done; 1 for a wave
// with a row of more than KS in-partition entries and 2 for a wave with a
// non-finite result (the caller then takes the queue form on the original x,
// with true divisions for the latter).
namespace viewwide {
constexpr int G = 8;                // steps per group
constexpr int LG = 3;               // log2 G
template <int R>
constexpr int KS = R <= 2 ? 8 : 6;  // in-partition entries held per row
constexpr int UPPER = 1024;         // step tag offset of the backward pass
template <int R>
constexpr int ring_doubles() { return (R * G + 1) * 64; }  // + one trash slot per lane
}  // namespace viewwide

template <int R, int MODE>
__device__ __forceinline__ int view_ring_wide(double (&xr)[R], const double (&dr)[R],
                                               const double (&rr)[R], int lane, int Pa, int64_t s,
                                               const SellView<double> &A, double *ring) {
    using namespace viewwide;
    constexpr int K = KS<R>, NB = 64 / G;
    double v[R][K];
    int tk[R][K];
    int cnt = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        const int64_t g = s + i;
        int64_t base = 0;
        int klo = 0, khi = 0, len = 0;
        if (i < Pa) {
            const int64_t sl = g >> 6;
            base = A.sptr[sl] + (g & 63);
            len = A.rowlen[g];
        }
        // count the row's entries before the partition and before its end,
        // eight (clamped, unconditional) column loads at a time
#pragma unroll 1
        for (int k0 = 0; k0 < len; k0 += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + u < len ? k0 + u : len - 1;
                const int64_t c = A.col[base + 64 * (int64_t)k];
                klo += (k0 + u < len) & (c < s);
                khi += (k0 + u < len) & (c < s + Pa);
            }
        }
        cnt = max(cnt, khi - klo);
        // the window [klo, khi): entry klo + q, clamped loads
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const bool in = klo + q < khi;
            const int k = in ? klo + q : (len > 0 ? len - 1 : 0);
            const int64_t p = len > 0 ? base + 64 * (int64_t)k : 0;
            const int j = (int)((int64_t)A.col[p] - s);  // 0 <= j < Pa inside the window
            const double val = A.val[p];
            int t = -1;
            if (in && MODE != MODE_BWD && j < i) t = j;
            else if (in && MODE != MODE_FWD && j > i) t = UPPER + (64 * R - 1 - j);
            tk[r][q] = t;
            v[r][q] = t >= 0 ? val : 0.0;
        }
    }
    if (__any(cnt > K)) return 1;
    // the widest window in the wave bounds the ring-fill scans (wave-uniform)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt = max(cnt, __shfl_xor(cnt, off));
    const int kw = __builtin_amdgcn_readfirstlane(cnt);

    double *const mine = ring + lane;  // slot u of row slot r: mine[64 (r G + u)]
    double *const trash = ring + R * G * 64 + lane;
    // write each entry of the given group into its slot (its value, or zero to
    // clear it), the others into the trash slot
    auto fill = [&](int grp, int rlo, int rhi, bool zero) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            if (q >= kw) break;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r < rlo || r >= rhi) continue;
                const int t = tk[r][q];
                double *const dst = (t >> LG) == grp ? mine + 64 * (r * G + (t & (G - 1))) : trash;
                *dst = zero ? 0.0 : v[r][q];
            }
        }
    };
    RowAcc<double> row[R];
    auto pass = [&](auto Upper) {
        constexpr bool UP = decltype(Upper)::value;
        constexpr int TAG = UP ? UPPER : 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            row[r] = RowAcc<double>(xr[r]);
#pragma unroll
            for (int u = 0; u < G; ++u) mine[64 * (r * G + u)] = 0.0;
        }
        fill(TAG >> LG, 0, R, false);
        static_for<0, R>([&](auto B) {
            constexpr int b = decltype(B)::value;  // block of the pass
            constexpr int js = UP ? R - 1 - b : b;  // row slot owning its steps
            // row slots with terms left: past js going forward, up to js going back
            constexpr int r0 = UP ? 0 : js, r1 = UP ? js + 1 : R;
#pragma unroll 1
            for (int gb = 0; gb < NB; ++gb) {
                double cur[G][R];
#pragma unroll
                for (int r = r0; r < r1; ++r)
#pragma unroll
                    for (int u = 0; u < G; ++u) cur[u][r] = mine[64 * (r * G + u)];
                // the slots just read take the next group's entries: zero the
                // ones that held an entry (the others are still zero), then fill
                const int grp = (TAG >> LG) + b * NB + gb;
                fill(grp, r0, r1, true);
                fill(grp + 1, r0, r1, false);
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const int jl = UP ? 63 - (gb * G + u) : gb * G + u;
                    const double yj = seg_bcast<64>(row[js].q(dr[js], rr[js]), jl);
#pragma unroll
                    for (int r = r0; r < r1; ++r) row[r].add(cur[u][r], yj);
                }
            }
        });
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = row[r].q(dr[r], rr[r]);
    };
    if constexpr (MODE != MODE_BWD) pass(std::false_type{});
    if constexpr (MODE == MODE_SYM) {
        // y .= D_Dl1 .* y (:897); rows past the end restart from zero
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = (lane + 64 * r < Pa) ? dr[r] * xr[r] : 0.0;
    }
    if constexpr (MODE != MODE_FWD) pass(std::true_type{});
    bool fin = true;
#pragma unroll
    for (int r = 0; r < R; ++r) fin = fin && finite64(xr[r]);
    return __any(!fin) ? 2 : 0;
}

// view_wave_wide: the general form (any row length, unsafe diagonals) on the
// 1-based row arrays, view_wave's register queues with R rows per lane.
template <int R, int MODE, bool FAST>
__device__ __forceinline__ void view_wave_wide(double (&xr)[R], const double (&dr)[R],
                                               const double (&rr)[R], int lane, int Pa, int64_t s,
                                               const int64_t *__restrict__ ptr,
                                               const int64_t *__restrict__ ind,
                                               const double *__restrict__ val) {
    int64_t plo[R], pmid[R], pmid1[R], phi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        const int64_t g = s + i;
        plo[r] = pmid[r] = pmid1[r] = phi[r] = 0;
        if (i < Pa) {
            const int64_t p0 = ptr[g] - 1, p1 = ptr[g + 1] - 1;
            plo[r] = row_lower_bound(ind, p0, p1, s);
            pmid[r] = row_lower_bound(ind, plo[r], p1, g);
            pmid1[r] = (pmid[r] < p1 && ind[pmid[r]] - 1 == g) ? pmid[r] + 1 : pmid[r];
            phi[r] = row_lower_bound(ind, pmid1[r], p1, s + Pa);
        }
    }
    auto pass = [&](auto Upper) {
        constexpr bool UP = decltype(Upper)::value;
        RowAcc<double> row[R];
        double yfin[R];
        ViewCache vc[R];
        int64_t p[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            row[r] = RowAcc<double>(xr[r]);
            yfin[r] = 0.0;
            p[r] = UP ? phi[r] - 1 : plo[r];
            if (UP) view_fill(vc[r], p[r], -1, pmid1[r], phi[r], ind, val, s);
            else view_fill(vc[r], p[r], 1, plo[r], pmid[r], ind, val, s);
        }
        for (int t = 0; t < Pa; ++t) {
            const int j = UP ? Pa - 1 - t : t;
            const int js = j >> 6, jl = j & 63;
            double yj = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r == js)
                    yj = bcast_keep<64>(row[r].template q<FAST>(dr[r], rr[r]), jl, yfin[r]);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (vc[r].c[0] == j) {
                    row[r].add(vc[r].v[0], yj);
                    p[r] += UP ? -1 : 1;
#pragma unroll
                    for (int k = 0; k + 1 < KV; ++k) {
                        vc[r].c[k] = vc[r].c[k + 1];
                        vc[r].v[k] = vc[r].v[k + 1];
                    }
                    vc[r].c[KV - 1] = -1;
                    if (vc[r].c[0] < 0) {
                        if (UP && p[r] >= pmid1[r])
                            view_fill(vc[r], p[r], -1, pmid1[r], phi[r], ind, val, s);
                        if (!UP && p[r] < pmid[r])
                            view_fill(vc[r], p[r], 1, plo[r], pmid[r], ind, val, s);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = yfin[r];
    };
    if constexpr (MODE != MODE_BWD) pass(std::false_type{});
    if constexpr (MODE == MODE_SYM) {
#pragma unroll
        for (int r = 0; r < R; ++r) xr[r] = dr[r] * xr[r];  // y .= D_Dl1 .* y (:897)
    }
    if constexpr (MODE != MODE_FWD) pass(std::true_type{});
}

// RING: waves whose diagonals are safe take the ring; the others, waves with
// a row of too many in-partition entries and waves with a non-finite result,
// the queue form (true divisions where a diagonal is unsafe or a result is
// not finite; both round correctly, so the results are the same).
template <int R, int MODE, bool RING>
__global__ __launch_bounds__(64) void k_sweep_view_wide(const double *x, double *y,
                                                        SellView<double> A,
                                                        const int64_t *__restrict__ ptr,
                                                        const int64_t *__restrict__ ind,
                                                        const double *__restrict__ val,
                                                        const double *__restrict__ D, int64_t n,
                                                        int64_t P, int64_t nparts,
                                                        const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ double ring[RING ? viewwide::ring_doubles<R>() : 1];
    const int64_t k = blockIdx.x;
    if (k >= nparts) return;
    const int lane = threadIdx.x;
    const int64_t s = k * P;
    const int Pa = (int)(n - s < P ? n - s : P);
    double xr[R], dr[R], rr[R];
    bool unsafe = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        const bool live = i < Pa;
        xr[r] = live ? x[s + i] : 0.0;
        dr[r] = live ? D[s + i] : 1.0;
        rr[r] = 1.0 / dr[r];
        unsafe = unsafe || !safe_divisor(dr[r]);
    }
    int ring_status = 1;
    if constexpr (RING) {
        if (!__any(unsafe)) ring_status = view_ring_wide<R, MODE>(xr, dr, rr, lane, Pa, s, A, ring);
    }
    if (ring_status != 0) {
        if constexpr (RING) {
#pragma unroll
            for (int r = 0; r < R; ++r) xr[r] = (lane + 64 * r < Pa) ? x[s + lane + 64 * r] : 0.0;
        }
        // true divisions where a diagonal is unsafe or a quotient not finite
        if (ring_status == 2 || __any(unsafe))
            view_wave_wide<R, MODE, false>(xr, dr, rr, lane, Pa, s, ptr, ind, val);
        else
            view_wave_wide<R, MODE, true>(xr, dr, rr, lane, Pa, s, ptr, ind, val);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < Pa) y[s + i] = xr[r];
    }
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct L1GSPlan {
    uint64_t serial = next_serial();
    int64_t n = 0, P = 1, nparts = 0, nchunks = 1, chunksize = 1, S = 0;
    int sweep = HS_SWEEP_SYMMETRIC, cache = HS_CACHE_PACKED, prec = HS_PREC_FP64;
    int fmt = HS_FMT_CSC, sym = 0, rowwise = 1;
    // kernel variant of the P = 64 symmetric sweep: 0 column loads straight to
    // registers, 1-2 LDS-ring configurations (see hs_l1gs_set_variant)
    int variant = 2;
    DBuf<double> D;   // D_Dl1 in FP64 (always kept: getter and FP64 sweeps)
    DBuf<float> Df;   // FP32 copy for the reduced-precision sweeps
    DBuf<uint8_t> T;  // packed triangles (block layout above), element type by `prec`
    // MatrixViewCache: device copy of the matrix (1-based, as given), and for
    // a non-symmetric CSC its CSR transpose (the rows the wave kernel reads)
    DBuf<int64_t> vptr, vind, vdiag;
    DBuf<double> vval;
    DBuf<int64_t> wptr, wind;
    DBuf<double> wval;
    // MatrixViewCache at partsize 64 to 256: the rows as sliced ELL
    // (k_sweep_view_ring, k_sweep_view_wide)
    // and the widest slice's entry count
    CSR vsell;
    int vsell_w = 0;
    Staging stage;

    size_t elem() const { return prec == HS_PREC_FP64 ? 8 : (prec == HS_PREC_FP32 ? 4 : 2); }
    // block layout of the packed buffer (block_bases) and its element count
    int tri_mode() const { return cache == HS_CACHE_PACKED_SYM ? (int)MODE_SYMT : sweep; }
    int64_t tri_elems() const {
        if (cache == HS_CACHE_VIEW) return 0;
        return (cache == HS_CACHE_PACKED && sweep == HS_SWEEP_SYMMETRIC) ? 2 * S : S;
    }
};

template <int PM, class TS, class TC>
static void launch_seg(const L1GSPlan &h, const double *x, double *y, const TC *D, hipStream_t st,
                       const int *done) {
    constexpr int SEGS = 256 / PM;
    unsigned grid = grid_for(h.nparts, SEGS);
    const TS *T = reinterpret_cast<const TS *>(h.T.get());
    int P = (int)h.P;
    if constexpr (PM == 16 || PM == 32) {
        if (P == PM && h.sweep != HS_SWEEP_SYMMETRIC) {
            // one wave per workgroup: a wave stages 8 KB (FP64); workgroups
            // of four (32 KB) fit four to a CU, 16 waves, single waves
            // 16 to 19 (2048^2 forward / backward 118 -> 108 us)
            constexpr int WPB = 1;
            const unsigned g1 = grid_for(h.nparts, WPB * 64 / PM);
            if (h.sweep == HS_SWEEP_FORWARD)
                hipLaunchKernelGGL((k_sweep_seg_lds<PM, MODE_FWD, TS, TC, WPB>), dim3(g1),
                                   dim3(64 * WPB), 0, st, x, y, T, D, h.n, P, h.nparts, done);
            else
                hipLaunchKernelGGL((k_sweep_seg_lds<PM, MODE_BWD, TS, TC, WPB>), dim3(g1),
                                   dim3(64 * WPB), 0, st, x, y, T, D, h.n, P, h.nparts, done);
            return;
        }
    }
    switch (h.sweep) {
    case HS_SWEEP_FORWARD:
        hipLaunchKernelGGL((k_sweep_seg<PM, MODE_FWD, TS, TC>), dim3(grid), dim3(256), 0, st, x, y,
                           T, D, h.n, P, h.nparts, done);
        break;
    case HS_SWEEP_BACKWARD:
        hipLaunchKernelGGL((k_sweep_seg<PM, MODE_BWD, TS, TC>), dim3(grid), dim3(256), 0, st, x, y,
                           T, D, h.n, P, h.nparts, done);
        break;
    default:
        hipLaunchKernelGGL((k_sweep_seg<PM, MODE_SYM, TS, TC>), dim3(grid), dim3(256), 0, st, x, y,
                           T, D, h.n, P, h.nparts, done);
        break;
    }
}

template <int PM, class TS, class TC, int FORM>
static unsigned launch_symt_pm(const L1GSPlan &h, const double *x, double *y, const TC *D,
                               hipStream_t st, const int *done, const CGFuse &f,
                               const SmoothFuse &sf) {
    constexpr int SEGS = 64 / PM;
    const unsigned grid = grid_for(h.nparts, SEGS);
    hipLaunchKernelGGL((k_sweep_symt<PM, TS, TC, FORM>), dim3(grid), dim3(64), 0, st, x, y,
                       reinterpret_cast<const TS *>(h.T.get()), D, h.n, (int)h.P, h.nparts, done,
                       f, sf);
    return grid;
}

// HS_CACHE_PACKED_SYM (partsize <= 64, checked at build); returns the number
// of workgroups (= r'z partials in the CG form)
template <class TS, class TC, int FORM>
static unsigned launch_symt(const L1GSPlan &h, const double *x, double *y, const TC *D,
                            hipStream_t st, const int *done, const CGFuse &f,
                            const SmoothFuse &sf = SmoothFuse{}) {
    if (h.P <= 4) return launch_symt_pm<4, TS, TC, FORM>(h, x, y, D, st, done, f, sf);
    if (h.P <= 8) return launch_symt_pm<8, TS, TC, FORM>(h, x, y, D, st, done, f, sf);
    if (h.P <= 16) return launch_symt_pm<16, TS, TC, FORM>(h, x, y, D, st, done, f, sf);
    if (h.P <= 32) return launch_symt_pm<32, TS, TC, FORM>(h, x, y, D, st, done, f, sf);
    return launch_symt_pm<64, TS, TC, FORM>(h, x, y, D, st, done, f, sf);
}

// The fused smoother round with both triangles or a one-sided sweep.
template <int PM, class TS, class TC>
static void launch_seg_smooth(const L1GSPlan &h, const TC *D, hipStream_t st,
                              const SmoothFuse &sf) {
    constexpr int SEGS = 256 / PM;
    const unsigned grid = grid_for(h.nparts, SEGS);
    const TS *T = reinterpret_cast<const TS *>(h.T.get());
    const int P = (int)h.P;
    const double *x = nullptr;
    double *y = nullptr;
    const int *done = nullptr;
    if (h.sweep == HS_SWEEP_FORWARD)
        hipLaunchKernelGGL((k_sweep_seg<PM, MODE_FWD, TS, TC, FORM_SMOOTH>), dim3(grid), dim3(256),
                           0, st, x, y, T, D, h.n, P, h.nparts, done, CGFuse{}, sf);
    else if (h.sweep == HS_SWEEP_BACKWARD)
        hipLaunchKernelGGL((k_sweep_seg<PM, MODE_BWD, TS, TC, FORM_SMOOTH>), dim3(grid), dim3(256),
                           0, st, x, y, T, D, h.n, P, h.nparts, done, CGFuse{}, sf);
    else
        hipLaunchKernelGGL((k_sweep_seg<PM, MODE_SYM, TS, TC, FORM_SMOOTH>), dim3(grid), dim3(256),
                           0, st, x, y, T, D, h.n, P, h.nparts, done, CGFuse{}, sf);
}

template <class TS, class TC>
static void launch_smooth(const L1GSPlan &h, const TC *D, hipStream_t st, const SmoothFuse &sf) {
    if (h.cache == HS_CACHE_PACKED_SYM) {
        launch_symt<TS, TC, FORM_SMOOTH>(h, nullptr, nullptr, D, st, nullptr, CGFuse{}, sf);
        return;
    }
    if (h.P == 64 && h.sweep == HS_SWEEP_SYMMETRIC && h.variant > 0) {
        const TS *T = reinterpret_cast<const TS *>(h.T.get());
        const dim3 grid(grid_for(h.nparts, 1)), block(64);
        if (h.variant == 1)
            hipLaunchKernelGGL((k_sweep_ring64<TS, TC, 8, 12, FORM_SMOOTH>), grid, block, 0, st,
                               (const double *)nullptr, (double *)nullptr, T, D, h.n, h.nparts,
                               (const int *)nullptr, CGFuse{}, sf);
        else
            hipLaunchKernelGGL((k_sweep_ring64<TS, TC, 16, 24, FORM_SMOOTH>), grid, block, 0, st,
                               (const double *)nullptr, (double *)nullptr, T, D, h.n, h.nparts,
                               (const int *)nullptr, CGFuse{}, sf);
        return;
    }
    if (h.P <= 4) launch_seg_smooth<4, TS, TC>(h, D, st, sf);
    else if (h.P <= 8) launch_seg_smooth<8, TS, TC>(h, D, st, sf);
    else if (h.P <= 16) launch_seg_smooth<16, TS, TC>(h, D, st, sf);
    else if (h.P <= 32) launch_seg_smooth<32, TS, TC>(h, D, st, sf);
    else launch_seg_smooth<64, TS, TC>(h, D, st, sf);
}

template <int R, class TS, class TC>
static void launch_wide(const L1GSPlan &h, const double *x, double *y, const TC *D,
                        hipStream_t st, const int *done) {
    const TS *T = reinterpret_cast<const TS *>(h.T.get());
    const dim3 grid(grid_for(h.nparts, 1)), block(64);
    if (h.sweep == HS_SWEEP_FORWARD)
        hipLaunchKernelGGL((k_sweep_wide<R, MODE_FWD, TS, TC>), grid, block, 0, st, x, y, T, D,
                           h.n, h.P, h.nparts, done);
    else if (h.sweep == HS_SWEEP_BACKWARD)
        hipLaunchKernelGGL((k_sweep_wide<R, MODE_BWD, TS, TC>), grid, block, 0, st, x, y, T, D,
                           h.n, h.P, h.nparts, done);
    else
        hipLaunchKernelGGL((k_sweep_wide<R, MODE_SYM, TS, TC>), grid, block, 0, st, x, y, T, D,
                           h.n, h.P, h.nparts, done);
}

template <class TS, class TC>
static void launch_packed(const L1GSPlan &h, const double *x, double *y, const TC *D,
                          hipStream_t st, const int *done) {
    if (h.cache == HS_CACHE_PACKED_SYM) {
        launch_symt<TS, TC, FORM_PLAIN>(h, x, y, D, st, done, CGFuse{});
        return;
    }
    if (h.P == 64 && h.sweep == HS_SWEEP_SYMMETRIC && h.variant > 0) {
        const TS *T = reinterpret_cast<const TS *>(h.T.get());
        const dim3 grid(grid_for(h.nparts, 1)), block(64);
        if (h.variant == 1)
            hipLaunchKernelGGL((k_sweep_ring64<TS, TC, 8, 12, FORM_PLAIN>), grid, block, 0, st, x, y, T,
                               D, h.n, h.nparts, done, CGFuse{});
        else
            hipLaunchKernelGGL((k_sweep_ring64<TS, TC, 16, 24, FORM_PLAIN>), grid, block, 0, st, x, y, T,
                               D, h.n, h.nparts, done, CGFuse{});
        return;
    }
    if (h.P <= 4) launch_seg<4, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 8) launch_seg<8, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 16) launch_seg<16, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 32) launch_seg<32, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 64) launch_seg<64, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 128) launch_wide<2, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 192) launch_wide<3, TS, TC>(h, x, y, D, st, done);
    else if (h.P <= 256) launch_wide<4, TS, TC>(h, x, y, D, st, done);
    else {
        const TS *T = reinterpret_cast<const TS *>(h.T.get());
        unsigned grid = grid_for(h.nparts, 64);
        if (h.sweep == HS_SWEEP_FORWARD)
            hipLaunchKernelGGL((k_sweep_thread<MODE_FWD, TS, TC>), dim3(grid), dim3(64), 0, st, x,
                               y, T, D, h.n, h.P, h.nparts, done);
        else if (h.sweep == HS_SWEEP_BACKWARD)
            hipLaunchKernelGGL((k_sweep_thread<MODE_BWD, TS, TC>), dim3(grid), dim3(64), 0, st, x,
                               y, T, D, h.n, h.P, h.nparts, done);
        else
            hipLaunchKernelGGL((k_sweep_thread<MODE_SYM, TS, TC>), dim3(grid), dim3(64), 0, st, x,
                               y, T, D, h.n, h.P, h.nparts, done);
    }
}

template <int PM, class TS, class TC>
static unsigned launch_seg_fused(const L1GSPlan &h, double *z, const TC *D, hipStream_t st,
                                 const int *done, const CGFuse &f) {
    constexpr int SEGS = 256 / PM;
    const unsigned grid = grid_for(h.nparts, SEGS);
    hipLaunchKernelGGL((k_sweep_seg<PM, MODE_SYM, TS, TC, true>), dim3(grid), dim3(256), 0, st,
                       (const double *)nullptr, z, reinterpret_cast<const TS *>(h.T.get()), D, h.n,
                       (int)h.P, h.nparts, done, f);
    return grid;
}

template <class TS, class TC>
static unsigned launch_fused(const L1GSPlan &h, double *z, const TC *D, hipStream_t st,
                             const CGFuse &f) {
    const int *done = reinterpret_cast<const int *>(reinterpret_cast<const char *>(f.st) +
                                                    offsetof(CGDev, done));
    if (h.cache == HS_CACHE_PACKED_SYM)
        return launch_symt<TS, TC, FORM_CG>(h, (const double *)nullptr, z, D, st, done, f);
    if (h.P == 64 && h.variant > 0) {
        const TS *T = reinterpret_cast<const TS *>(h.T.get());
        const dim3 grid(grid_for(h.nparts, 1)), block(64);
        if (h.variant == 1)
            hipLaunchKernelGGL((k_sweep_ring64<TS, TC, 8, 12, FORM_CG>), grid, block, 0, st,
                               (const double *)nullptr, z, T, D, h.n, h.nparts, done, f);
        else
            hipLaunchKernelGGL((k_sweep_ring64<TS, TC, 16, 24, FORM_CG>), grid, block, 0, st,
                               (const double *)nullptr, z, T, D, h.n, h.nparts, done, f);
        return (unsigned)h.nparts;
    }
    if (h.P <= 4) return launch_seg_fused<4, TS, TC>(h, z, D, st, done, f);
    if (h.P <= 8) return launch_seg_fused<8, TS, TC>(h, z, D, st, done, f);
    if (h.P <= 16) return launch_seg_fused<16, TS, TC>(h, z, D, st, done, f);
    if (h.P <= 32) return launch_seg_fused<32, TS, TC>(h, z, D, st, done, f);
    return launch_seg_fused<64, TS, TC>(h, z, D, st, done, f);
}

bool l1gs_can_fuse(const L1GSPlan &h) {
    return (h.cache == HS_CACHE_PACKED || h.cache == HS_CACHE_PACKED_SYM) &&
           h.sweep == HS_SWEEP_SYMMETRIC && h.P <= 64 && h.n > 0;
}

bool l1gs_can_smooth_fused(const L1GSPlan &h) {
    return h.cache != HS_CACHE_VIEW && h.P <= 64 && h.n > 0;
}

void l1gs_smooth_fused(const L1GSPlan &h, const SmoothFuse &sf, hipStream_t st) {
    if (h.prec == HS_PREC_FP64) launch_smooth<double, double>(h, h.D.get(), st, sf);
    else if (h.prec == HS_PREC_FP32) launch_smooth<float, float>(h, h.Df.get(), st, sf);
    else launch_smooth<__hip_bfloat16, float>(h, h.Df.get(), st, sf);
    check_launch();
}

unsigned l1gs_apply_fused(const L1GSPlan &h, double *z, hipStream_t st, const CGFuse &f) {
    unsigned count;
    if (h.prec == HS_PREC_FP64) count = launch_fused<double, double>(h, z, h.D.get(), st, f);
    else if (h.prec == HS_PREC_FP32) count = launch_fused<float, float>(h, z, h.Df.get(), st, f);
    else count = launch_fused<__hip_bfloat16, float>(h, z, h.Df.get(), st, f);
    check_launch();
    return count;
}

template <int WMAX>
static void launch_view_ring_w(const L1GSPlan &h, const double *x, double *y, const int64_t *rp,
                               const int64_t *ri, const double *rv, hipStream_t st,
                               const int *done) {
    const dim3 grid(grid_for(h.nparts, 1)), block(64);
    const SellView<double> A = h.vsell.view();
    if (h.sweep == HS_SWEEP_FORWARD)
        hipLaunchKernelGGL((k_sweep_view_ring<MODE_FWD, WMAX>), grid, block, 0, st, x, y, A, rp,
                           ri, rv, h.D.get(), h.n, h.nparts, done);
    else if (h.sweep == HS_SWEEP_BACKWARD)
        hipLaunchKernelGGL((k_sweep_view_ring<MODE_BWD, WMAX>), grid, block, 0, st, x, y, A, rp,
                           ri, rv, h.D.get(), h.n, h.nparts, done);
    else
        hipLaunchKernelGGL((k_sweep_view_ring<MODE_SYM, WMAX>), grid, block, 0, st, x, y, A, rp,
                           ri, rv, h.D.get(), h.n, h.nparts, done);
}

// the register budget follows the widest slice: 8 entries (5- and 7-point
// stencils), 16, or 32 (wider slices take the general kernel wave by wave)
static void launch_view_ring(const L1GSPlan &h, const double *x, double *y, const int64_t *rp,
                             const int64_t *ri, const double *rv, hipStream_t st,
                             const int *done) {
    if (h.vsell_w <= 8) launch_view_ring_w<8>(h, x, y, rp, ri, rv, st, done);
    else if (h.vsell_w <= 16) launch_view_ring_w<16>(h, x, y, rp, ri, rv, st, done);
    else launch_view_ring_w<32>(h, x, y, rp, ri, rv, st, done);
}

template <int R, bool RING>
static void launch_view_wide_r(const L1GSPlan &h, const double *x, double *y, const int64_t *rp,
                               const int64_t *ri, const double *rv, hipStream_t st,
                               const int *done) {
    const dim3 grid(grid_for(h.nparts, 1)), block(64);
    const SellView<double> A = h.vsell.view();
    if (h.sweep == HS_SWEEP_FORWARD)
        hipLaunchKernelGGL((k_sweep_view_wide<R, MODE_FWD, RING>), grid, block, 0, st, x, y, A,
                           rp, ri, rv, h.D.get(), h.n, h.P, h.nparts, done);
    else if (h.sweep == HS_SWEEP_BACKWARD)
        hipLaunchKernelGGL((k_sweep_view_wide<R, MODE_BWD, RING>), grid, block, 0, st, x, y, A,
                           rp, ri, rv, h.D.get(), h.n, h.P, h.nparts, done);
    else
        hipLaunchKernelGGL((k_sweep_view_wide<R, MODE_SYM, RING>), grid, block, 0, st, x, y, A,
                           rp, ri, rv, h.D.get(), h.n, h.P, h.nparts, done);
}

// partsize 65..256: the ring kernel over the sliced-ELL rows (built at
// partsizes 64..256), else the queue form throughout
template <int R>
static void launch_view_wide(const L1GSPlan &h, const double *x, double *y, const int64_t *rp,
                             const int64_t *ri, const double *rv, hipStream_t st,
                             const int *done) {
    if (h.vsell.n == h.n && h.vsell.nent > 0)
        launch_view_wide_r<R, true>(h, x, y, rp, ri, rv, st, done);
    else launch_view_wide_r<R, false>(h, x, y, rp, ri, rv, st, done);
}

// Device-side application y = M^-1 x (x, y device pointers; may alias).
void l1gs_apply_device(const L1GSPlan &h, const double *x, double *y, hipStream_t st,
                       const int *done) {
    if (h.n == 0) return;
    if (h.cache == HS_CACHE_VIEW && h.P == 64 && h.vsell.n == h.n) {
        const int64_t *rp = h.rowwise ? h.vptr.get() : h.wptr.get();
        const int64_t *ri = h.rowwise ? h.vind.get() : h.wind.get();
        const double *rv = h.rowwise ? h.vval.get() : h.wval.get();
        launch_view_ring(h, x, y, rp, ri, rv, st, done);
    } else if (h.cache == HS_CACHE_VIEW && h.P > 64 && h.P <= 256) {
        const int64_t *rp = h.rowwise ? h.vptr.get() : h.wptr.get();
        const int64_t *ri = h.rowwise ? h.vind.get() : h.wind.get();
        const double *rv = h.rowwise ? h.vval.get() : h.wval.get();
        if (h.P <= 128) launch_view_wide<2>(h, x, y, rp, ri, rv, st, done);
        else if (h.P <= 192) launch_view_wide<3>(h, x, y, rp, ri, rv, st, done);
        else launch_view_wide<4>(h, x, y, rp, ri, rv, st, done);
    } else if (h.cache == HS_CACHE_VIEW && h.P <= 64) {
        // rows: the matrix itself when it is read row-wise, else its transpose
        const int64_t *rp = h.rowwise ? h.vptr.get() : h.wptr.get();
        const int64_t *ri = h.rowwise ? h.vind.get() : h.wind.get();
        const double *rv = h.rowwise ? h.vval.get() : h.wval.get();
        const dim3 grid(grid_for(h.nparts, 1)), block(64);
        if (h.sweep == HS_SWEEP_FORWARD)
            hipLaunchKernelGGL(k_sweep_view_wave<MODE_FWD>, grid, block, 0, st, x, y, rp, ri, rv,
                               h.D.get(), h.n, h.P, h.nparts, done);
        else if (h.sweep == HS_SWEEP_BACKWARD)
            hipLaunchKernelGGL(k_sweep_view_wave<MODE_BWD>, grid, block, 0, st, x, y, rp, ri, rv,
                               h.D.get(), h.n, h.P, h.nparts, done);
        else
            hipLaunchKernelGGL(k_sweep_view_wave<MODE_SYM>, grid, block, 0, st, x, y, rp, ri, rv,
                               h.D.get(), h.n, h.P, h.nparts, done);
    } else if (h.cache == HS_CACHE_VIEW) {
        unsigned grid = grid_for(h.nparts, 64);
        const int64_t *dg = h.vdiag.get();
        if (h.sweep == HS_SWEEP_FORWARD)
            hipLaunchKernelGGL(k_sweep_view<MODE_FWD>, dim3(grid), dim3(64), 0, st, x, y,
                               h.vptr.get(), h.vind.get(), h.vval.get(), dg, h.rowwise, h.D.get(),
                               h.n, h.P, h.nparts, done);
        else if (h.sweep == HS_SWEEP_BACKWARD)
            hipLaunchKernelGGL(k_sweep_view<MODE_BWD>, dim3(grid), dim3(64), 0, st, x, y,
                               h.vptr.get(), h.vind.get(), h.vval.get(), dg, h.rowwise, h.D.get(),
                               h.n, h.P, h.nparts, done);
        else
            hipLaunchKernelGGL(k_sweep_view<MODE_SYM>, dim3(grid), dim3(64), 0, st, x, y,
                               h.vptr.get(), h.vind.get(), h.vval.get(), dg, h.rowwise, h.D.get(),
                               h.n, h.P, h.nparts, done);
    } else if (h.prec == HS_PREC_FP64) {
        launch_packed<double, double>(h, x, y, h.D.get(), st, done);
    } else if (h.prec == HS_PREC_FP32) {
        launch_packed<float, float>(h, x, y, h.Df.get(), st, done);
    } else {
        launch_packed<__hip_bfloat16, float>(h, x, y, h.Df.get(), st, done);
    }
    check_launch();
}

// D_Dl1 of a non-symmetric CSC: the reference scans every column for row i
// (:1135-1216), visiting row i's entries in ascending column order.  The same
// sums come from the rows of the CSR transpose (a stable sort by row keeps
// ascending columns inside each row), in O(nnz) instead of O(N^2), on the
// device.
static void nonsym_D(L1GSPlan &h, int64_t nnz, const DBuf<int64_t> &ptr, const DBuf<int64_t> &ind,
                     const DBuf<double> &val, double eta, hipStream_t st, bool keep_rows) {
    const int64_t n = h.n;
    DBuf<int64_t> trp, tci;
    DBuf<double> tcv;
    csc_transpose_device(n, nnz, ptr.get(), ind.get(), val.get(), trp, tci, tcv, st);
    hipLaunchKernelGGL((k_setup_rows<double>), dim3(grid_for(n, 256)), dim3(256), 0, st, n, h.P,
                       trp.get(), tci.get(), tcv.get(), eta, h.D.get(), (double *)nullptr,
                       h.sweep, 0, 0);
    check_launch();
    HS_CHECK_HIP(hipStreamSynchronize(st));
    if (keep_rows) {  // the MatrixViewCache's row-wise copy
        h.wptr = std::move(trp);
        h.wind = std::move(tci);
        h.wval = std::move(tcv);
    }
}

int64_t l1gs_rows(const L1GSPlan &h) { return h.n; }
int l1gs_precision(const L1GSPlan &h) { return h.prec; }
uint64_t l1gs_serial(const L1GSPlan &h) { return h.serial; }

// The sliced-ELL copy of the rows the MatrixViewCache ring kernels read at
// partsize 64 to 256 (the matrix itself when read row-wise, else its
// transpose), and its widest slice.
static void build_view_sell(L1GSPlan &h, int64_t nnz, hipStream_t st) {
    if (h.P < 64 || h.P > 256 || h.n == 0) return;
    const int64_t *rp = h.rowwise ? h.vptr.get() : h.wptr.get();
    const int64_t *ri = h.rowwise ? h.vind.get() : h.wind.get();
    const double *rv = h.rowwise ? h.vval.get() : h.wval.get();
    sell_build_device(h.vsell, h.n, nnz, rp, ri, rv, st);
    std::vector<int64_t> sp(h.vsell.nslices + 1);
    HS_CHECK_HIP(hipMemcpy(sp.data(), h.vsell.sptr.get(), sp.size() * 8, hipMemcpyDeviceToHost));
    int64_t wmax = 0;
    for (int64_t k = 0; k < h.vsell.nslices; ++k) wmax = std::max(wmax, (sp[k + 1] - sp[k]) / 64);
    h.vsell_w = (int)std::min<int64_t>(wmax, 1 << 20);
}

template <class TS>
static void build_packed(L1GSPlan &h, int64_t nnz, const DBuf<int64_t> &ptr,
                         const DBuf<int64_t> &ind, const DBuf<double> &val, double eta,
                         hipStream_t st) {
    const int want_L = h.sweep != HS_SWEEP_BACKWARD;
    const int want_U = h.sweep != HS_SWEEP_FORWARD && h.cache != HS_CACHE_PACKED_SYM;
    const int mode = h.tri_mode();
    const int64_t elems = h.tri_elems();
    h.T.alloc(elems * sizeof(TS));
    if (elems) HS_CHECK_HIP(hipMemsetAsync(h.T.get(), 0, h.T.bytes(), st));
    TS *T = reinterpret_cast<TS *>(h.T.get());
    unsigned grid = grid_for(h.n, 256);
    if (h.rowwise) {
        // one pass: D_Dl1 and both triangles (the reference launches its
        // packing kernel once per triangle and recomputes D_Dl1, :805-837)
        hipLaunchKernelGGL((k_setup_rows<TS>), dim3(grid), dim3(256), 0, st, h.n, h.P, ptr.get(),
                           ind.get(), val.get(), eta, h.D.get(), T, mode, want_L, want_U);
        check_launch();
        return;
    }
    nonsym_D(h, nnz, ptr, ind, val, eta, st, false);
    hipLaunchKernelGGL((k_pack_cols<TS>), dim3(grid), dim3(256), 0, st, h.n, h.P, ptr.get(),
                       ind.get(), val.get(), T, mode, want_L, want_U);
    check_launch();
}

}  // namespace hs

struct hs_l1gs : hs::L1GSPlan {};

using namespace hs;

extern "C" {

int hs_l1gs_build(int64_t n, const int64_t *ptr, const int64_t *ind, const double *val, int fmt,
                  int is_sym, int64_t partsize, double eta, int sweep, int cache, int precision,
                  int64_t threads, int64_t blocks, void *stream, hs_l1gs_t *out) {
    return guard([&] {
        HS_REQUIRE(out != nullptr, HS_EINVAL, "out handle pointer is NULL");
        *out = nullptr;
        HS_REQUIRE(partsize > 0, HS_EINVAL, "partsize must be greater than 0");
        HS_REQUIRE(threads > 0, HS_EINVAL, "`threads` must be greater than 0");
        HS_REQUIRE(blocks > 0, HS_EINVAL, "`blocks` must be greater than 0");
        HS_REQUIRE(n >= 0, HS_EINVAL, "matrix size must be non-negative");
        HS_REQUIRE(fmt == HS_FMT_CSC || fmt == HS_FMT_CSR, HS_EINVAL, "unknown matrix format");
        HS_REQUIRE(sweep >= 0 && sweep <= 2, HS_EINVAL, "unknown sweep");
        HS_REQUIRE(cache == HS_CACHE_PACKED || cache == HS_CACHE_VIEW ||
                       cache == HS_CACHE_PACKED_SYM,
                   HS_EINVAL, "unknown cache strategy");
        HS_REQUIRE(cache != HS_CACHE_PACKED_SYM ||
                       (sweep == HS_SWEEP_SYMMETRIC && is_sym && partsize <= 64),
                   HS_EUNSUPPORTED,
                   "the one-triangle packed cache needs isSymA = true, SymmetricSweep and "
                   "partsize <= 64");
        HS_REQUIRE(precision >= HS_PREC_FP64 && precision <= HS_PREC_BF16, HS_EINVAL,
                   "unknown precision");
        HS_REQUIRE(!(cache == HS_CACHE_VIEW && precision != HS_PREC_FP64), HS_EUNSUPPORTED,
                   "MatrixViewCache is provided in FP64 only");
        HS_REQUIRE(n == 0 || (ptr && ind && val), HS_EINVAL, "matrix arrays are NULL");
        hipStream_t st = as_stream(stream);
        std::unique_ptr<hs_l1gs> h(new hs_l1gs());
        h->n = n;
        h->P = partsize;
        h->sweep = sweep;
        h->cache = cache;
        h->prec = precision;
        h->fmt = fmt;
        h->sym = is_sym ? 1 : 0;
        h->rowwise = (fmt == HS_FMT_CSR) || is_sym;
        // _blockpartitioning, GPU (:353-365)
        h->nchunks = blocks;
        h->nparts = (int64_t)std::ceil((double)n / (double)partsize);
        int64_t cs = (int64_t)std::ceil((double)h->nparts / (double)blocks);
        h->chunksize = cs <= threads ? cs : threads;
        int64_t last = n - (h->nparts - 1) * partsize;
        h->S = (partsize * (partsize - 1) * (h->nparts - 1)) / 2 + (last * (last - 1)) / 2;
        if (n == 0) h->S = 0;

        const int64_t nnz = n > 0 ? ptr[n] - 1 : 0;
        HS_REQUIRE(nnz >= 0, HS_EINVAL, "invalid pointer array");
        DBuf<int64_t> dptr(n + 1), dind(nnz);
        DBuf<double> dval(nnz);
        if (n > 0) {
            HS_CHECK_HIP(hipMemcpyAsync(dptr.get(), ptr, (n + 1) * 8, hipMemcpyHostToDevice, st));
            if (nnz) {
                HS_CHECK_HIP(hipMemcpyAsync(dind.get(), ind, nnz * 8, hipMemcpyHostToDevice, st));
                HS_CHECK_HIP(hipMemcpyAsync(dval.get(), val, nnz * 8, hipMemcpyHostToDevice, st));
            }
            // every kernel below indexes through these arrays
            validate_sparse_device(n, nnz, dptr.get(), dind.get(), st);
        }
        DBuf<int64_t> ddiag;
        if (!h->rowwise && n > 0) {
            // DiagonalIndices: SingularException(col) on a missing or zero diagonal
            ddiag.alloc(n);
            DBuf<unsigned long long> bad(1);
            unsigned long long init = ~0ull, got = 0;
            HS_CHECK_HIP(hipMemcpyAsync(bad.get(), &init, 8, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_diag_indices, dim3(grid_for(n, 256)), dim3(256), 0, st, n,
                               dptr.get(), dind.get(), dval.get(), ddiag.get(), bad.get());
            check_launch();
            HS_CHECK_HIP(hipMemcpyAsync(&got, bad.get(), 8, hipMemcpyDeviceToHost, st));
            HS_CHECK_HIP(hipStreamSynchronize(st));
            if (got != ~0ull) {
                set_error(HS_ESINGULAR, "SingularException(" + std::to_string(got) + ")",
                          (int64_t)got);
                throw Error{HS_ESINGULAR};
            }
        }
        h->D.alloc(n);
        if (n > 0) {
            if (cache != HS_CACHE_VIEW) {
                if (precision == HS_PREC_FP64)
                    build_packed<double>(*h, nnz, dptr, dind, dval, eta, st);
                else if (precision == HS_PREC_FP32)
                    build_packed<float>(*h, nnz, dptr, dind, dval, eta, st);
                else
                    build_packed<__hip_bfloat16>(*h, nnz, dptr, dind, dval, eta, st);
            } else {
                // MatrixViewCache: D_Dl1 only; keep the matrix on the device
                if (h->rowwise) {
                    hipLaunchKernelGGL((k_setup_rows<double>), dim3(grid_for(n, 256)), dim3(256),
                                       0, st, n, h->P, dptr.get(), dind.get(), dval.get(), eta,
                                       h->D.get(), (double *)nullptr, h->sweep, 0, 0);
                    check_launch();
                } else {
                    nonsym_D(*h, nnz, dptr, dind, dval, eta, st, true);
                }
                // the view kernels read 0-based diagonal positions
                h->vptr = std::move(dptr);
                h->vind = std::move(dind);
                h->vval = std::move(dval);
                h->vdiag = std::move(ddiag);
                build_view_sell(*h, nnz, st);
            }
            if (precision != HS_PREC_FP64) {
                h->Df.alloc(n);
                hipLaunchKernelGGL((k_convert<float>), dim3(grid_for(n, 256)), dim3(256), 0, st, n,
                                   h->D.get(), h->Df.get());
                check_launch();
            }
        }
        HS_CHECK_HIP(hipStreamSynchronize(st));
        *out = h.release();
    });
}

int hs_l1gs_apply(hs_l1gs_t h, double *y, const double *x, int loc, void *stream, int flags) {
    return guard([&] {
        HS_REQUIRE(h != nullptr, HS_EINVAL, "NULL preconditioner handle");
        HS_REQUIRE(h->n == 0 || (x && y), HS_EINVAL, "NULL vector");
        hipStream_t st = as_stream(stream);
        if (h->n == 0) return;
        if (loc == HS_LOC_DEVICE) {
            l1gs_apply_device(*h, x, y, st, nullptr);
            finish(st, flags);
        } else {
            // host vectors: y .= x; adapt; sweep; copyto!(y, _y)  (:367-384)
            size_t bytes = (size_t)h->n * 8;
            h->stage.ensure(h->n);
            HS_CHECK_HIP(hipMemcpyAsync(h->stage.a.get(), x, bytes, hipMemcpyHostToDevice, st));
            l1gs_apply_device(*h, h->stage.a.get(), h->stage.a.get(), st, nullptr);
            HS_CHECK_HIP(hipMemcpyAsync(y, h->stage.a.get(), bytes, hipMemcpyDeviceToHost, st));
            HS_CHECK_HIP(hipStreamSynchronize(st));
        }
    });
}

int hs_l1gs_info(hs_l1gs_t h, int64_t *out) {
    return guard([&] {
        HS_REQUIRE(h && out, HS_EINVAL, "NULL argument");
        out[0] = h->P;
        out[1] = h->nparts;
        out[2] = h->nchunks;
        out[3] = h->chunksize;
        out[4] = h->n;
        out[5] = h->S;
        out[6] = h->sweep;
        out[7] = h->cache;
    });
}

int hs_l1gs_get(hs_l1gs_t h, int what, double *out) {
    return guard([&] {
        HS_REQUIRE(h && out, HS_EINVAL, "NULL argument");
        const int64_t n = h->n, P = h->P;
        if (what == HS_GET_D_DL1) {
            if (n) HS_CHECK_HIP(hipMemcpy(out, h->D.get(), n * 8, hipMemcpyDeviceToHost));
            return;
        }
        HS_REQUIRE(what == HS_GET_SL || what == HS_GET_SU, HS_EINVAL, "unknown field");
        HS_REQUIRE(h->cache != HS_CACHE_VIEW, HS_EINVAL,
                   "packed buffers exist only for PackedBufferCache");
        const bool has = (what == HS_GET_SL) ? h->sweep != HS_SWEEP_BACKWARD
                                             : h->sweep != HS_SWEEP_FORWARD;
        HS_REQUIRE(has, HS_EINVAL,
                   what == HS_GET_SL ? "this sweep has no lower buffer"
                                     : "this sweep has no upper buffer");
        const int64_t elems = h->tri_elems();
        const bool one_tri = h->cache == HS_CACHE_PACKED_SYM;
        std::vector<double> internal(elems);
        const size_t es = h->elem();
        std::vector<uint8_t> raw(elems * es);
        if (elems) HS_CHECK_HIP(hipMemcpy(raw.data(), h->T.get(), raw.size(), hipMemcpyDeviceToHost));
        for (int64_t t = 0; t < elems; ++t) {
            if (h->prec == HS_PREC_FP64) {
                double v;
                std::memcpy(&v, &raw[t * 8], 8);
                internal[t] = v;
            } else if (h->prec == HS_PREC_FP32) {
                float v;
                std::memcpy(&v, &raw[t * 4], 4);
                internal[t] = v;
            } else {
                uint16_t b;
                std::memcpy(&b, &raw[t * 2], 2);
                uint32_t w = (uint32_t)b << 16;
                float v;
                std::memcpy(&v, &w, 4);
                internal[t] = v;
            }
        }
        // device block layout -> reference row-major buffers (:1264, :1398)
        for (int64_t k = 0; k < h->nparts; ++k) {
            int64_t s = k * P;
            int64_t Pa = (n - s < P) ? n - s : P;
            int64_t boff = k * ((P * (P - 1)) / 2);
            int64_t lb, ub;
            block_bases(k, P, Pa, h->tri_mode(), lb, ub);
            for (int64_t i = 0; i < Pa; ++i) {
                if (what == HS_GET_SL) {
                    for (int64_t j = 0; j < i; ++j)
                        out[boff + (i * (i - 1)) / 2 + j] = internal[lb + colL_off(j, Pa) + (i - j - 1)];
                } else {
                    int64_t row_start = i * Pa - (i * (i + 1)) / 2;
                    for (int64_t j = i + 1; j < Pa; ++j)  // U(i,j) = L(j,i) with one triangle
                        out[boff + row_start + (j - i - 1)] =
                            one_tri ? internal[lb + colL_off(i, Pa) + (j - i - 1)]
                                    : internal[ub + colU_off(j, Pa) + i];
                }
            }
        }
    });
}

int hs_l1gs_set_variant(hs_l1gs_t h, int variant) {
    return guard([&] {
        HS_REQUIRE(h, HS_EINVAL, "NULL preconditioner handle");
        HS_REQUIRE(variant >= 0 && variant <= 2, HS_EINVAL, "unknown kernel variant");
        h->variant = variant;
    });
}

// ---------------------------------------------------------------------------
// persisted state: the built preconditioner as a file, reloaded without the
// matrix.  Layout (little-endian): a fixed header, then the arrays in header
// order, each exactly as stored on the device; the header ends with an
// FNV-1a 64 checksum of everything after it.
// ---------------------------------------------------------------------------
namespace {

constexpr char kMagic[8] = {'H', 'S', 'L', '1', 'G', 'S', '0', '2'};

struct FileHeader {
    char magic[8];
    int64_t n, P, nparts, nchunks, chunksize, S;
    int32_t sweep, cache, prec, fmt, sym, rowwise, variant, pad;
    int64_t nbytes[7];  // D, Df, T, vptr, vind, vdiag, vval
    uint64_t checksum;  // FNV-1a 64 of the header before this field, then of the payload
};

uint64_t fnv1a(uint64_t h, const uint8_t *p, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

uint64_t header_hash(const FileHeader &hd) {
    return fnv1a(0xcbf29ce484222325ULL, reinterpret_cast<const uint8_t *>(&hd),
                 offsetof(FileHeader, checksum));
}

// Largest row count and packed-buffer size a file may describe: n * 8 and the
// buffer byte counts stay far from int64 overflow.
constexpr int64_t kMaxRows = int64_t(1) << 40;
constexpr int64_t kMaxElems = int64_t(1) << 50;

// Everything the loaded plan's kernels index by is recomputed from (n, P, sweep,
// cache, prec, fmt, sym) exactly as hs_l1gs_build derives it, and the header
// must agree with it: the partitioning, S, the layout switches and every array
// size.  Returns an empty string or what is wrong.
std::string check_header(const FileHeader &hd) {
    const int64_t n = hd.n, P = hd.P;
    if (n < 0 || n > kMaxRows || P <= 0 || P > kMaxRows) return "rows or partsize out of range";
    if (hd.sweep < 0 || hd.sweep > 2) return "unknown sweep";
    if (hd.cache != HS_CACHE_PACKED && hd.cache != HS_CACHE_VIEW && hd.cache != HS_CACHE_PACKED_SYM)
        return "unknown cache";
    if (hd.prec < HS_PREC_FP64 || hd.prec > HS_PREC_BF16) return "unknown precision";
    if (hd.fmt != HS_FMT_CSC && hd.fmt != HS_FMT_CSR) return "unknown matrix format";
    if (hd.sym != 0 && hd.sym != 1) return "bad isSymA flag";
    if (hd.rowwise != ((hd.fmt == HS_FMT_CSR || hd.sym) ? 1 : 0)) return "bad row-wise flag";
    if (hd.variant < 0 || hd.variant > 2 || hd.pad != 0) return "bad kernel variant";
    if (hd.cache == HS_CACHE_PACKED_SYM &&
        !(hd.sweep == HS_SWEEP_SYMMETRIC && hd.sym == 1 && P <= 64))
        return "one-triangle cache without isSymA, SymmetricSweep and partsize <= 64";
    if (hd.cache == HS_CACHE_VIEW && hd.prec != HS_PREC_FP64)
        return "MatrixViewCache in a precision other than FP64";
    const int64_t nparts = n > 0 ? (n + P - 1) / P : 0;
    if (hd.nparts != nparts) return "partition count";
    if (hd.nchunks < 1 || hd.chunksize < 0) return "chunking";
    if (nparts > 0 && (hd.chunksize < 1 || hd.chunksize > (nparts + hd.nchunks - 1) / hd.nchunks))
        return "chunking";
    if (nparts == 0 && hd.chunksize != 0) return "chunking";
    // S = P(P-1)(nparts-1)/2 + Pl(Pl-1)/2, Pl the last partition's size (:699-701)
    __int128 S = 0;
    if (n > 0) {
        const int64_t last = n - (nparts - 1) * P;
        S = (__int128)P * (P - 1) * (nparts - 1) / 2 + (__int128)last * (last - 1) / 2;
    }
    if (S > kMaxElems || hd.S != (int64_t)S) return "packed size";
    const int64_t elem = hd.prec == HS_PREC_FP64 ? 8 : (hd.prec == HS_PREC_FP32 ? 4 : 2);
    const int64_t tri = hd.cache == HS_CACHE_VIEW
                            ? 0
                            : (hd.cache == HS_CACHE_PACKED && hd.sweep == HS_SWEEP_SYMMETRIC
                                   ? 2 * hd.S
                                   : hd.S);
    if (hd.nbytes[0] != n * 8 || hd.nbytes[1] != (hd.prec == HS_PREC_FP64 ? 0 : n * 4) ||
        hd.nbytes[2] != tri * elem)
        return "array sizes";
    if (hd.cache != HS_CACHE_VIEW) {
        for (int k = 3; k < 7; ++k)
            if (hd.nbytes[k] != 0) return "array sizes";
        return "";
    }
    // MatrixViewCache: the matrix as given (1-based pointers, indices, values)
    // and, for a non-symmetric CSC, its diagonal positions
    const int64_t nnzb = hd.nbytes[4];
    if (n > 0 && (nnzb < 0 || nnzb % 8 != 0 || nnzb / 8 > kMaxElems || hd.nbytes[6] != nnzb ||
                  hd.nbytes[3] != (n + 1) * 8 || hd.nbytes[5] != (hd.rowwise ? 0 : n * 8)))
        return "matrix sizes";
    if (n == 0) {
        for (int k = 3; k < 7; ++k)
            if (hd.nbytes[k] != 0) return "matrix sizes";
    }
    return "";
}

struct FileCloser {
    void operator()(std::FILE *f) const {
        if (f) std::fclose(f);
    }
};

}  // namespace

int hs_l1gs_save(hs_l1gs_t h, const char *path) {
    return guard([&] {
        HS_REQUIRE(h && path, HS_EINVAL, "NULL argument");
        FileHeader hd{};
        std::memcpy(hd.magic, kMagic, 8);
        hd.n = h->n;
        hd.P = h->P;
        hd.nparts = h->nparts;
        hd.nchunks = h->nchunks;
        hd.chunksize = h->chunksize;
        hd.S = h->S;
        hd.sweep = h->sweep;
        hd.cache = h->cache;
        hd.prec = h->prec;
        hd.fmt = h->fmt;
        hd.sym = h->sym;
        hd.rowwise = h->rowwise;
        hd.variant = h->variant;
        const void *dev[7] = {h->D.get(), h->Df.get(), h->T.get(), h->vptr.get(),
                              h->vind.get(), h->vdiag.get(), h->vval.get()};
        const size_t sizes[7] = {h->D.bytes(), h->Df.bytes(), h->T.bytes(), h->vptr.bytes(),
                                 h->vind.bytes(), h->vdiag.bytes(), h->vval.bytes()};
        std::vector<std::vector<uint8_t>> host(7);
        for (int k = 0; k < 7; ++k) hd.nbytes[k] = (int64_t)sizes[k];
        uint64_t ck = header_hash(hd);
        for (int k = 0; k < 7; ++k) {
            host[k].resize(sizes[k]);
            if (sizes[k])
                HS_CHECK_HIP(hipMemcpy(host[k].data(), dev[k], sizes[k], hipMemcpyDeviceToHost));
            ck = fnv1a(ck, host[k].data(), sizes[k]);
        }
        hd.checksum = ck;
        std::unique_ptr<std::FILE, FileCloser> f(std::fopen(path, "wb"));
        HS_REQUIRE(f, HS_EINVAL, std::string("cannot open for writing: ") + path);
        bool ok = std::fwrite(&hd, sizeof hd, 1, f.get()) == 1;
        for (int k = 0; k < 7 && ok; ++k)
            ok = sizes[k] == 0 || std::fwrite(host[k].data(), 1, sizes[k], f.get()) == sizes[k];
        ok = ok && std::fflush(f.get()) == 0;
        HS_REQUIRE(ok, HS_EINVAL, std::string("write failed: ") + path);
    });
}

int hs_l1gs_load(const char *path, void *stream, hs_l1gs_t *out) {
    return guard([&] {
        HS_REQUIRE(path && out, HS_EINVAL, "NULL argument");
        *out = nullptr;
        std::unique_ptr<std::FILE, FileCloser> f(std::fopen(path, "rb"));
        HS_REQUIRE(f, HS_EINVAL, std::string("cannot open: ") + path);
        FileHeader hd{};
        HS_REQUIRE(std::fread(&hd, sizeof hd, 1, f.get()) == 1, HS_EINVAL,
                   "not a saved L1GS preconditioner (short header)");
        HS_REQUIRE(std::memcmp(hd.magic, kMagic, 8) == 0, HS_EINVAL,
                   "not a saved L1GS preconditioner (bad magic)");
        const int64_t n = hd.n;
        const std::string bad = check_header(hd);
        HS_REQUIRE(bad.empty(), HS_EINVAL, "corrupt L1GS file (" + bad + ")");
        std::vector<std::vector<uint8_t>> host(7);
        uint64_t ck = header_hash(hd);
        for (int k = 0; k < 7; ++k) {
            host[k].resize((size_t)hd.nbytes[k]);
            HS_REQUIRE(hd.nbytes[k] == 0 ||
                           std::fread(host[k].data(), 1, host[k].size(), f.get()) == host[k].size(),
                       HS_EINVAL, "truncated L1GS file");
            ck = fnv1a(ck, host[k].data(), host[k].size());
        }
        HS_REQUIRE(ck == hd.checksum, HS_EINVAL, "corrupt L1GS file (checksum mismatch)");
        hipStream_t st = as_stream(stream);
        std::unique_ptr<hs_l1gs> h(new hs_l1gs());
        h->n = n;
        h->P = hd.P;
        h->nparts = hd.nparts;
        h->nchunks = hd.nchunks;
        h->chunksize = hd.chunksize;
        h->S = hd.S;
        h->sweep = hd.sweep;
        h->cache = hd.cache;
        h->prec = hd.prec;
        h->fmt = hd.fmt;
        h->sym = hd.sym;
        h->rowwise = hd.rowwise;
        h->variant = hd.variant;
        auto up = [&](auto &buf, int k) {
            using T = std::remove_reference_t<decltype(*buf.get())>;
            buf.alloc((size_t)hd.nbytes[k] / sizeof(T));
            if (hd.nbytes[k])
                HS_CHECK_HIP(hipMemcpyAsync(buf.get(), host[k].data(), (size_t)hd.nbytes[k],
                                            hipMemcpyHostToDevice, st));
        };
        up(h->D, 0);
        up(h->Df, 1);
        up(h->T, 2);
        up(h->vptr, 3);
        up(h->vind, 4);
        up(h->vdiag, 5);
        up(h->vval, 6);
        if (h->cache == HS_CACHE_VIEW && n > 0) {
            // the view kernels index through the stored matrix
            const int64_t nnz = (int64_t)(hd.nbytes[4] / 8);
            validate_sparse_device(n, nnz, h->vptr.get(), h->vind.get(), st);
            if (!h->rowwise) {
                // diagonal positions recomputed from the validated matrix rather
                // than trusted from the file; the wave kernel's row-wise copy is
                // rebuilt, not stored
                DBuf<unsigned long long> bad(1);
                unsigned long long init = ~0ull, got = 0;
                HS_CHECK_HIP(hipMemcpyAsync(bad.get(), &init, 8, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(k_diag_indices, dim3(grid_for(n, 256)), dim3(256), 0, st, n,
                                   h->vptr.get(), h->vind.get(), h->vval.get(), h->vdiag.get(),
                                   bad.get());
                check_launch();
                HS_CHECK_HIP(hipMemcpyAsync(&got, bad.get(), 8, hipMemcpyDeviceToHost, st));
                HS_CHECK_HIP(hipStreamSynchronize(st));
                HS_REQUIRE(got == ~0ull, HS_EINVAL, "corrupt L1GS file (matrix diagonal)");
                csc_transpose_device(n, nnz, h->vptr.get(), h->vind.get(), h->vval.get(), h->wptr,
                                     h->wind, h->wval, st);
            }
            // the ring kernels' sliced-ELL rows are rebuilt, not stored
            build_view_sell(*h, nnz, st);
        }
        HS_CHECK_HIP(hipStreamSynchronize(st));  // the host buffers go out of scope
        *out = h.release();
    });
}

int hs_l1gs_destroy(hs_l1gs_t h) {
    return guard([&] { delete h; });
}

}  // extern "C"
