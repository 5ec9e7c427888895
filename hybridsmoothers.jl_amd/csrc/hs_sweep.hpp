// hs_sweep.hpp -- device building blocks of the partitioned L1GS sweeps,
// shared by the single-system kernels (hs_l1gs.hip) and the batched engine
// (hs_batch.hip): packed-block layout, the D_Dl1 rule, segment broadcasts,
// correctly rounded division from a reciprocal, and the per-partition
// right-looking recurrences.  The layout and recurrence are described at the
// top of hs_l1gs.hip.
#pragma once

#include "hs_plans.hpp"

#include <cmath>
#include <type_traits>

namespace hs {

// MODE_SYMT: symmetric sweep of a symmetric matrix from the lower triangle
// alone (inside a partition U = L^T), stored like a forward sweep's buffer.
enum { MODE_FWD = 0, MODE_BWD = 1, MODE_SYM = 2, MODE_SYMT = 3 };

// ---------------------------------------------------------------------------
// layout helpers
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int64_t colL_off(int64_t j, int64_t Pa) {
    return j * (Pa - 1) - (j * (j - 1)) / 2;
}
__host__ __device__ __forceinline__ int64_t colU_off(int64_t j, int64_t Pa) {
    return (Pa * (Pa - 1) - j * (j + 1)) / 2;
}
// elements per partition block of the device buffer
__host__ __device__ __forceinline__ int64_t blk_stride(int64_t P, int mode) {
    return mode == MODE_SYM ? P * (P - 1) : (P * (P - 1)) / 2;
}
// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// elements of a segment's LDS copy of L in the one-triangle symmetric sweep:
// PM rows of PM + 1 (sweep_full_symt); for PM = 64 the half-size copy of
// sweep_full_symt64, 32 lines of 65 and a 64-element area for dead stores
__host__ __device__ constexpr int symt_lds(int PM) { return PM == 64 ? 32 * 65 + 64 : PM * (PM + 1); }
// start of the lower and upper triangles of partition k (Pa rows)
__host__ __device__ __forceinline__ void block_bases(int64_t k, int64_t P, int64_t Pa, int mode,
                                                     int64_t &lb, int64_t &ub) {
    lb = k * blk_stride(P, mode);
    ub = (mode == MODE_SYM) ? lb + (Pa * (Pa - 1)) / 2 : lb;
}

__host__ __device__ __forceinline__ double dl1_rule(double a, double d, double eta) {
    // _compute_D_Dl1_kernel! (:1016-1022): a_ii + (a_ii >= eta*d ? 0 : d/2)
    double star = (a >= eta * d) ? 0.0 : d / 2.0;
    return a + star;
}


// Aligned full-wave chunk loads of the packed triangles, which an apply
// streams once from HBM: nontemporal (the `nt` cache policy).  A plain read
// of 32 KB blocks, one wave each, runs at 5.7 TB/s on this chip with ordinary
// loads and at 6.2 TB/s with nontemporal ones.  Only for loads that touch
// each line once: the column-by-column loads of the direct kernels overlap
// the previous column's lines, which an `nt` load does not keep in cache
// (measured: those kernels slow down, up to 2x).
template <class TS>
__device__ __forceinline__ TS ld_stream(const TS *p) {
    if constexpr (std::is_same_v<TS, double> || std::is_same_v<TS, float>) {
        return __builtin_nontemporal_load(p);
    } else {
        static_assert(sizeof(TS) == 2, "16-bit storage type");
        const unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
        TS r;
        __builtin_memcpy(&r, &b, 2);
        return r;
    }
}

// ---------------------------------------------------------------------------
// sweep kernels
// ---------------------------------------------------------------------------
// Broadcast of lane j of every PM-lane segment to the whole segment.
// PM = 64: v_readlane into SGPRs; PM = 32: readlanes of lanes j and j+32 and a
// per-half select; smaller segments: ds_bpermute (__shfl with width PM).
template <int PM>
__device__ __forceinline__ double seg_bcast(double v, int j) {
    if constexpr (PM == 64) {
        int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
        int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
        return __hiloint2double(hi, lo);
    } else if constexpr (PM == 32) {
        int lo0 = __builtin_amdgcn_readlane(__double2loint(v), j);
        int hi0 = __builtin_amdgcn_readlane(__double2hiint(v), j);
        int lo1 = __builtin_amdgcn_readlane(__double2loint(v), j + 32);
        int hi1 = __builtin_amdgcn_readlane(__double2hiint(v), j + 32);
        bool up = (threadIdx.x & 32) != 0;
        return __hiloint2double(up ? hi1 : hi0, up ? lo1 : lo0);
    } else {
        return __shfl(v, j, PM);
    }
}
template <int PM>
__device__ __forceinline__ float seg_bcast(float v, int j) {
    if constexpr (PM == 64) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
    } else if constexpr (PM == 32) {
        int a = __builtin_amdgcn_readlane(__float_as_int(v), j);
        int b = __builtin_amdgcn_readlane(__float_as_int(v), j + 32);
        return __int_as_float((threadIdx.x & 32) ? b : a);
    } else {
        return __shfl(v, j, PM);
    }
}

// v_writelane_b32 (no clang builtin exists; this binds the LLVM intrinsic).
extern "C" __device__ int hs_writelane_i32(int value, int lane, int old)
    __asm("llvm.amdgcn.writelane.i32");

// Broadcast lane j of every segment and record the broadcast value in lane j
// of `keep` (v_writelane from the SGPR copy: no per-step lane masks).
template <int PM>
__device__ __forceinline__ double bcast_keep(double v, int j, double &keep) {
    if constexpr (PM == 64) {
        int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
        int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
        int klo = hs_writelane_i32(lo, j, __double2loint(keep));
        int khi = hs_writelane_i32(hi, j, __double2hiint(keep));
        keep = __hiloint2double(khi, klo);
        return __hiloint2double(hi, lo);
    } else if constexpr (PM == 32) {
        int lo0 = __builtin_amdgcn_readlane(__double2loint(v), j);
        int hi0 = __builtin_amdgcn_readlane(__double2hiint(v), j);
        int lo1 = __builtin_amdgcn_readlane(__double2loint(v), j + 32);
        int hi1 = __builtin_amdgcn_readlane(__double2hiint(v), j + 32);
        int klo = hs_writelane_i32(lo0, j, __double2loint(keep));
        int khi = hs_writelane_i32(hi0, j, __double2hiint(keep));
        klo = hs_writelane_i32(lo1, j + 32, klo);
        khi = hs_writelane_i32(hi1, j + 32, khi);
        keep = __hiloint2double(khi, klo);
        bool up = (threadIdx.x & 32) != 0;
        return __hiloint2double(up ? hi1 : hi0, up ? lo1 : lo0);
    } else {
        const int lane = threadIdx.x % PM;
        keep = (lane == j) ? v : keep;
        return __shfl(v, j, PM);
    }
}
template <int PM>
__device__ __forceinline__ float bcast_keep(float v, int j, float &keep) {
    if constexpr (PM == 64) {
        int b = __builtin_amdgcn_readlane(__float_as_int(v), j);
        keep = __int_as_float(hs_writelane_i32(b, j, __float_as_int(keep)));
        return __int_as_float(b);
    } else if constexpr (PM == 32) {
        int a = __builtin_amdgcn_readlane(__float_as_int(v), j);
        int b = __builtin_amdgcn_readlane(__float_as_int(v), j + 32);
        int k = hs_writelane_i32(a, j, __float_as_int(keep));
        keep = __int_as_float(hs_writelane_i32(b, j + 32, k));
        return __int_as_float((threadIdx.x & 32) ? b : a);
    } else {
        const int lane = threadIdx.x % PM;
        keep = (lane == j) ? v : keep;
        return __shfl(v, j, PM);
    }
}

// Broadcast for segments of at most 32 lanes (several segments per wave):
// ds_bpermute from lane j of the lane's own segment -- the step is the
// instruction's immediate offset from base4 = 4 * (first lane of the segment)
// -- and lane j's value kept by v_cmp_eq + v_cndmask on VCC, emitted together
// so the compiler neither hoists 31 comparisons nor holds their masks.  Eight
// VALU operations per recurrence step instead of thirteen with
// readlane/writelane, which for 32-lane segments also need two moves and a
// select (one scalar operand per VALU instruction on this chip).
template <int J>
__device__ __forceinline__ int keep_lane_i32(int keep, int v, int slane) {
    asm("v_cmp_eq_u32 vcc, %3, %2\n\tv_cndmask_b32 %0, %0, %1, vcc"
        : "+v"(keep)
        : "v"(v), "v"(slane), "n"(J)
        : "vcc");
    return keep;
}
template <int J>
__device__ __forceinline__ float bcast_keep_perm(float v, float &keep, int base4, int slane) {
    keep = __int_as_float(keep_lane_i32<J>(__float_as_int(keep), __float_as_int(v), slane));
    return __int_as_float(__builtin_amdgcn_ds_bpermute(base4 + 4 * J, __float_as_int(v)));
}
template <int J>
__device__ __forceinline__ double bcast_keep_perm(double v, double &keep, int base4, int slane) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    keep = __hiloint2double(keep_lane_i32<J>(__double2hiint(keep), hi, slane),
                            keep_lane_i32<J>(__double2loint(keep), lo, slane));
    return __hiloint2double(__builtin_amdgcn_ds_bpermute(base4 + 4 * J, hi),
                            __builtin_amdgcn_ds_bpermute(base4 + 4 * J, lo));
}
// The full-partition recurrences' broadcast: SGPR readlane/writelane for
// 64-lane segments, ds_bpermute for narrower ones.  `lane` is the lane's index
// in its segment; base4 = 4 * (its segment's first lane in the wave).
template <int PM, int J, class TC>
__device__ __forceinline__ TC bcast_step(TC v, TC &keep, int base4, int lane) {
    if constexpr (PM == 64) return bcast_keep<64>(v, J, keep);
    else return bcast_keep_perm<J>(v, keep, base4, lane);
}
// index of the calling lane's PM-lane segment within its wave
template <int PM>
__device__ __forceinline__ int seg_of_wave() { return (int)((threadIdx.x & 63) / PM); }

// base4 of the calling lane, opaque to the optimiser: per-step addresses
// base4 + 4j must fold into the instruction offset, not be precomputed.
__device__ __forceinline__ int seg_base4(int PM) {
    int b = (int)((threadIdx.x & 63) & ~(PM - 1)) * 4;
    asm volatile("" : "+v"(b));
    return b;
}

// Correctly rounded z/d from the correctly rounded reciprocal r = 1/d:
// q0 = z*r, e = z - q0*d (exact with an FMA), q = q0 + e*r rounds to z/d
// (Markstein's theorem: with r within half an ulp of 1/d and q0 within one
// ulp of z/d, the correction step gives the correctly rounded quotient,
// provided nothing overflows or underflows -- e is then exact).  The sweep
// checks |d| in [2^-400, 2^400] once per partition (else true division); the
// theorem's remaining conditions hold when z/d is well inside the normal
// range, |z/d| in [2^-500, 2^500] (then z, z*r and e are normal too).  Outside
// it the result can differ from z / d: an overflowing quotient gives NaN
// instead of +-inf, a quotient in the subnormal range can be an ulp off, and a
// zero remainder z = -0 gives +0.  The same code is compiled for the host
// (libhs_check.so) and tested there against IEEE division.
__host__ __device__ __forceinline__ double quot_fast(double z, double d, double r) {
    double q0 = z * r;
    double e = __builtin_fma(-q0, d, z);
    return __builtin_fma(e, r, q0);
}
__host__ __device__ __forceinline__ float quot_fast(float z, float d, float r) {
    float q0 = z * r;
    float e = __builtin_fmaf(-q0, d, z);
    return __builtin_fmaf(e, r, q0);
}
template <bool FAST, class TC>
__device__ __forceinline__ TC quot(TC z, TC d, TC r) {
    if constexpr (FAST) return quot_fast(z, d, r);
    else return z / d;
}
__host__ __device__ __forceinline__ bool safe_divisor(double d) {
    double a = fabs(d);
    return a >= 0x1p-400 && a <= 0x1p400;
}
__host__ __device__ __forceinline__ bool safe_divisor(float d) {
    float a = fabsf(d);
    return a >= 0x1p-60f && a <= 0x1p60f;
}

// One row of a recurrence: its input x, the products L(i,j) y_j added one by
// one, and the quotient (x - sum) / d.  FP64 is the reference's arithmetic bit
// for bit: the products are summed unfused in order and the quotient is
// correctly rounded.  FP32 (the reduced-precision smoothers, which have no
// reference arithmetic to reproduce) keeps the remainder rem = x - sum, takes
// each term off with one FMA, and multiplies by the reciprocal of d: two VALU
// operations per step instead of five.  FAST = false divides (unsafe d).
template <class TC>
struct RowAcc;
template <>
struct RowAcc<double> {
    double x, acc;
    __device__ __forceinline__ RowAcc() : x(0.0), acc(0.0) {}
    __device__ __forceinline__ explicit RowAcc(double xr) : x(xr), acc(0.0) {}
    template <bool FAST = true>
    __device__ __forceinline__ double q(double d, double r) const {
        return quot<FAST>(x - acc, d, r);
    }
    __device__ __forceinline__ void add(double l, double y) { acc = acc + l * y; }
};
template <>
struct RowAcc<float> {
    float rem;
    __device__ __forceinline__ RowAcc() : rem(0.0f) {}
    __device__ __forceinline__ explicit RowAcc(float xr) : rem(xr) {}
    template <bool FAST = true>
    __device__ __forceinline__ float q(float d, float r) const {
        if constexpr (FAST) return rem * r;
        else return rem / d;
    }
    __device__ __forceinline__ void add(float l, float y) { rem = __builtin_fmaf(-l, y, rem); }
};

// Any partition shape (last, short partition; P below the segment width):
// steps guarded by the actual size Pa.
template <int PM, int MODE, bool FAST, class TS, class TC>
__device__ __forceinline__ TC sweep_generic(TC xr, TC dr, TC rr, int lane, int Pa, bool live,
                                            const TS *__restrict__ Lk,
                                            const TS *__restrict__ Uk) {
    TC yr = xr;
    if constexpr (MODE != MODE_BWD) {
        RowAcc<TC> row(xr);
        TC yfin = (TC)0;
#pragma unroll
        for (int j = 0; j < PM; ++j) {
            if (j < Pa) {
                TC q = row.template q<FAST>(dr, rr);
                yfin = (lane == j) ? q : yfin;
                if (j < Pa - 1) {
                    TC yj = seg_bcast<PM>(q, j);
                    bool has = (lane > j) && live;
                    TC l = has ? load_as<TC>(Lk[colL_off(j, Pa) + (lane - j - 1)]) : (TC)0;
                    if (lane > j) row.add(l, yj);
                }
            }
        }
        yr = yfin;
    }
    if constexpr (MODE == MODE_SYM || MODE == MODE_SYMT) xr = dr * yr;  // y .= D_Dl1 .* y (:897)
    if constexpr (MODE != MODE_FWD) {
        RowAcc<TC> row(xr);
        TC yfin = (TC)0;
#pragma unroll
        for (int jj = 0; jj < PM; ++jj) {
            const int j = PM - 1 - jj;
            if (j < Pa) {
                TC q = row.template q<FAST>(dr, rr);
                yfin = (lane == j) ? q : yfin;
                if (j > 0) {
                    TC yj = seg_bcast<PM>(q, j);
                    TC u = (TC)0;
                    if (lane < j) {
                        // U(i, j) = L(j, i): row j of column i of the lower triangle
                        u = (MODE == MODE_SYMT)
                                ? load_as<TC>(Lk[colL_off(lane, Pa) + (j - lane - 1)])
                                : load_as<TC>(Uk[colU_off(j, Pa) + lane]);
                    }
                    if (lane < j) row.add(u, yj);
                }
            }
        }
        yr = yfin;
    }
    return yr;
}

// Full partitions (Pa == PM): straight-line, branch-free recurrence.  The
// column loads of both passes form one stream that runs RING columns ahead
// of the step consuming them, so every wave keeps RING*64 values in flight
// while it computes and the compiler can wait on exact load counts.
template <int PM, int MODE, int RING, class TS, class TC>
__device__ __forceinline__ TC sweep_full(TC xr, TC dr, TC rr, int lane,
                                         const TS *__restrict__ Lk, const TS *__restrict__ Uk) {
    constexpr int NL = (MODE != MODE_BWD) ? PM - 1 : 0;  // forward columns 0..PM-2
    constexpr int NU = (MODE != MODE_FWD) ? PM - 1 : 0;  // backward columns PM-1..1
    constexpr int NC = NL + NU;
    constexpr int R = RING < NC ? RING : NC;
    // Unconditional loads: lanes whose row is already final (i <= j forward,
    // i >= j backward) read a neighbouring in-block value into an accumulator
    // that is never used again.  Indices stay inside the triangle: forward
    // column j reads colL(j) + lane - j - 1 >= 0 for j >= 1 (clamped for j = 0);
    // backward column j reads colU(j) + lane, clamped to the column for the
    // last stored columns (small j), whose tail would pass the triangle's end.
    auto load_col = [&](int c) -> TS {
        if (c < NL) {
            const int j = c;
            const int off = (j == 0) ? (lane > 0 ? lane - 1 : 0) : lane - j - 1;
            return Lk[colL_off(j, PM) + off];
        }
        const int j = PM - 1 - (c - NL);
        constexpr int64_t UEND = (PM * (PM - 1)) / 2;
        const bool tail = colU_off(j, PM) + PM > UEND;
        const int off = tail ? (lane < j - 1 ? lane : j - 1) : lane;
        return Uk[colU_off(j, PM) + off];
    };
    TS ring[R > 0 ? R : 1];
#pragma unroll
    for (int c = 0; c < R; ++c) ring[c] = load_col(c);
    const int base4 = seg_base4(PM);

    TC yr = xr;
    if constexpr (MODE != MODE_BWD) {
        RowAcc<TC> row(xr);
        TC yfin = (TC)0;
        static_for<0, PM - 1>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const TS v = ring[j % R];
            if constexpr (j + R < NC) ring[j % R] = load_col(j + R);
            TC q = row.q(dr, rr);
            TC yj = bcast_step<PM, j>(q, yfin, base4, lane);
            row.add(load_as<TC>(v), yj);  // dead for lanes <= j
        });
        TC q = row.q(dr, rr);
        yr = (lane == PM - 1) ? q : yfin;
    }
    if constexpr (MODE == MODE_SYM) xr = dr * yr;  // y .= D_Dl1 .* y  (:897)
    if constexpr (MODE != MODE_FWD) {
        RowAcc<TC> row(xr);
        TC yfin = (TC)0;
        static_for<0, PM - 1>([&](auto JJ) {
            constexpr int jj = decltype(JJ)::value;
            constexpr int j = PM - 1 - jj;
            constexpr int c = NL + jj;
            const TS v = ring[c % R];
            if constexpr (c + R < NC) ring[c % R] = load_col(c + R);
            TC q = row.q(dr, rr);
            TC yj = bcast_step<PM, j>(q, yfin, base4, lane);
            row.add(load_as<TC>(v), yj);  // dead for lanes >= j
        });
        TC q = row.q(dr, rr);
        yr = (lane == 0) ? q : yfin;
    }
    return yr;
}


// Full partitions of the one-triangle symmetric sweep.  The forward pass is
// the streamed right-looking recurrence over the columns of L; each lane also
// stores the value it loads into its own row of a copy of L in the segment's
// LDS area: lane i writes L(i, j) at i*(PM+1) + j at step j.  Lanes i <= j
// hold a dead value and write it into their row's unused upper part, so the
// store needs no lane mask; the row stride PM + 1 spreads a step's stores
// over distinct banks.  The backward pass needs U(i, j) = L(j, i) in lane i
// at step j: row j of the copy, j*(PM+1) + i, one contiguous read, prefetched
// a few steps ahead.  The products and their order are those of
// sweep_full<PM, MODE_SYM>, so the results are bit-identical to the
// two-triangle sweep of the same (symmetric) matrix.
template <int PM, int RING, class TS, class TC>
__device__ __forceinline__ TC sweep_full_symt(TC xr, TC dr, TC rr, int lane,
                                              const TS *__restrict__ Lk, TS *lds) {
    constexpr int NL = PM - 1;
    constexpr int R = RING < NL ? RING : NL;
    constexpr int RS = PM + 1;  // row stride of the copy
    auto load_col = [&](int j) -> TS {
        const int off = (j == 0) ? (lane > 0 ? lane - 1 : 0) : lane - j - 1;
        return Lk[colL_off(j, PM) + off];
    };
    TS *const wrow = lds + lane * RS;  // this lane's row
    const TS *const rcol = lds + lane;  // this lane's column of every row
    const int base4 = seg_base4(PM);
    TS ring[R];
#pragma unroll
    for (int c = 0; c < R; ++c) ring[c] = load_col(c);
    RowAcc<TC> row(xr);
    TC yfin = (TC)0;
    static_for<0, NL>([&](auto J) {
        constexpr int j = decltype(J)::value;
        // keep each step's load, LDS store and arithmetic together: hoisted
        // stores would wait for their loads long before the step needs them
        __builtin_amdgcn_sched_barrier(0);
        const TS v = ring[j % R];
        if constexpr (j + R < NL) ring[j % R] = load_col(j + R);
        wrow[j] = v;
        TC q = row.q(dr, rr);
        TC yj = bcast_step<PM, j>(q, yfin, base4, lane);
        row.add(load_as<TC>(v), yj);  // dead for lanes <= j
    });
    TC q = row.q(dr, rr);
    TC yr = (lane == PM - 1) ? q : yfin;
    const TC xb = dr * yr;  // y .= D_Dl1 .* y  (:897)
    // the copy is complete; other lanes' rows are read below, RB steps ahead
    // (LDS latency is short: a few rows in flight hide it)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    constexpr int RB = 4 < NL ? 4 : NL;
    TS rb[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) rb[c] = rcol[(PM - 1 - c) * RS];
    row = RowAcc<TC>(xb);
    yfin = (TC)0;
    static_for<0, NL>([&](auto JJ) {
        constexpr int jj = decltype(JJ)::value;
        constexpr int j = PM - 1 - jj;
        const TS v = rb[jj % RB];
        if constexpr (jj + RB < NL) rb[jj % RB] = rcol[(PM - 1 - (jj + RB)) * RS];
        TC qb = row.q(dr, rr);
        TC yj = bcast_step<PM, j>(qb, yfin, base4, lane);
        row.add(load_as<TC>(v), yj);  // dead for lanes >= j
    });
    q = row.q(dr, rr);
    return (lane == 0) ? q : yfin;
}

// PM = 64 form of sweep_full_symt with half the LDS (2144 elements instead of
// 4160), so that an FP64 wave needs 17 KB.  Row r of L has r entries; rows r
// and 63 - r share line m = min(r, 63 - r) of 65 elements: row m (m < 32) at
// positions [0, m), row 63 - m at [m + 1, 64).  Lane r stores L(r, j) at
// wbase + j with wbase = 65 r (r < 32) or 65 (63 - r) + 64 - r (r >= 32).  A
// lane whose row is complete (lane j from step j on) would store into the
// other row of its line, so at step j lane j's base moves (v_writelane, one
// operation per step) to G - j, G = 32 * 65: the dead stores of step s then go
// to G + s - r, distinct for the dead lanes r <= s and inside [G, G + 63).
// (wbase is kept in bytes.)
// The backward pass reads row j at a step-constant offset from lds + lane:
// 65 j (j < 32) or 65 (63 - j) + 64 - j (j >= 32).  Same loads, products and
// order as sweep_full_symt: bit-identical results.
template <int RING, class TS, class TC>
__device__ __forceinline__ TC sweep_full_symt64(TC xr, TC dr, TC rr, int lane,
                                                const TS *__restrict__ Lk, TS *lds) {
    constexpr int PM = 64, NL = PM - 1;
    constexpr int R = RING < NL ? RING : NL;
    constexpr int RS = 65, G = 32 * RS;
    auto load_col = [&](int j) -> TS {
        const int off = (j == 0) ? (lane > 0 ? lane - 1 : 0) : lane - j - 1;
        return Lk[colL_off(j, PM) + off];
    };
    // byte offsets, so that a step's store address is wbase plus an immediate
    constexpr int ES = (int)sizeof(TS);
    int wbase = (lane < 32 ? lane * RS : (63 - lane) * RS + 64 - lane) * ES;
    asm volatile("" : "+v"(wbase));
    char *const lbytes = reinterpret_cast<char *>(lds);
    const TS *const rrow = lds + lane;
    const int base4 = seg_base4(PM);
    TS ring[R];
#pragma unroll
    for (int c = 0; c < R; ++c) ring[c] = load_col(c);
    RowAcc<TC> row(xr);
    TC yfin = (TC)0;
    static_for<0, NL>([&](auto J) {
        constexpr int j = decltype(J)::value;
        __builtin_amdgcn_sched_barrier(0);
        const TS v = ring[j % R];
        if constexpr (j + R < NL) ring[j % R] = load_col(j + R);
        wbase = hs_writelane_i32((G - j) * ES, j, wbase);
        *reinterpret_cast<TS *>(lbytes + wbase + j * ES) = v;
        TC q = row.q(dr, rr);
        TC yj = bcast_step<PM, j>(q, yfin, base4, lane);
        row.add(load_as<TC>(v), yj);  // dead for lanes <= j
    });
    TC q = row.q(dr, rr);
    TC yr = (lane == PM - 1) ? q : yfin;
    const TC xb = dr * yr;  // y .= D_Dl1 .* y  (:897)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    auto row_at = [](int j) { return j < 32 ? j * RS : (63 - j) * RS + 64 - j; };
    constexpr int RB = 4;
    TS rb[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) rb[c] = rrow[row_at(PM - 1 - c)];
    row = RowAcc<TC>(xb);
    yfin = (TC)0;
    static_for<0, NL>([&](auto JJ) {
        constexpr int jj = decltype(JJ)::value;
        constexpr int j = PM - 1 - jj;
        const TS v = rb[jj % RB];
        if constexpr (jj + RB < NL) rb[jj % RB] = rrow[row_at(PM - 1 - (jj + RB))];
        TC qb = row.q(dr, rr);
        TC yj = bcast_step<PM, j>(qb, yfin, base4, lane);
        row.add(load_as<TC>(v), yj);  // dead for lanes >= j
    });
    q = row.q(dr, rr);
    return (lane == 0) ? q : yfin;
}



// Staged form of the one-triangle symmetric sweep, for a wave whose 64 / PM
// segments hold consecutive full partitions: their lower-triangle blocks are
// one contiguous run of WE = 32 (PM - 1) values in the packed buffer.  The
// wave copies that run into LDS unchanged, with aligned full-wave nontemporal
// loads (lane l of chunk c takes value 64 c + l), and both passes read the
// copy: forward column j of a segment's block is contiguous (one value per
// lane), and the backward pass reads U(i, j) = L(j, i) at colL(i) + (j - i - 1)
// -- one fixed offset per lane plus the step, so an immediate offset per step.
// All loads are issued first; each chunk is written to LDS just before the
// first forward read that needs it, so the first steps start while later
// chunks are still in flight.  Lanes with no term at a step read an in-block
// value into an accumulator that is no longer used (as in sweep_full).  The
// loads, products and their order are those of sweep_full_symt, so the
// results are bit-identical.  `wl`: the wave's LDS area (>= NCH * 64 values),
// `wsrc`: the packed block of the wave's first partition.
template <int PM>
struct SymtStage {
    static constexpr int SEGS = 64 / PM;
    static constexpr int BLK = PM * (PM - 1) / 2;
    static constexpr int WE = SEGS * BLK;
    static constexpr int colL(int j) { return j * (PM - 1) - j * (j - 1) / 2; }
    // chunks of cw values that must be in LDS before forward column j is read
    // (all segments)
    static constexpr int need(int j, int cw) {
        const int hi = (SEGS - 1) * BLK + colL(j + 1) - 1;  // highest value read
        return hi / cw + 1;
    }
};

template <int PM, class TS, class TC>
__device__ __forceinline__ TC sweep_symt_staged(TC xr, TC dr, TC rr, int lane, int seg,
                                                const TS *__restrict__ wsrc, TS *wl) {
    using St = SymtStage<PM>;
    constexpr int NL = PM - 1, WE = St::WE, BLK = St::BLK;
    // each lane loads V consecutive values as one word W: BF16 four to an
    // 8-byte word, so that a load moves 512 B per wave as in FP64 (the 2-byte
    // loads took the 128^3 BF16 apply 46-47 us, 8-byte ones 44; FP32 measured
    // slower with 8-byte words and keeps one value per lane).  A wave's
    // blocks start word-aligned: WE and the wave's first block offset are
    // multiples of V, the buffers 256-byte aligned.
    using W = std::conditional_t<sizeof(TS) == 2, uint64_t,
                                 std::conditional_t<sizeof(TS) == 4, uint32_t, uint64_t>>;
    constexpr int V = (int)(sizeof(W) / sizeof(TS)), CW = 64 * V, NCH = (WE + CW - 1) / CW;
    static_assert(WE % V == 0, "a wave's blocks are whole words");
    constexpr int RB = NL < 4 ? NL : 4;  // reads in flight, at most one per column
    static_assert(St::need(NL - 1, CW) == NCH, "the forward pass reads every chunk");
    const int l64 = (int)(threadIdx.x & 63);
    W chunk[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int idx = CW * c + V * l64;
        chunk[c] = __builtin_nontemporal_load(
            reinterpret_cast<const W *>(wsrc + (idx < WE ? idx : WE - V)));
    }
    auto lds_fence = [] { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); };
    // chunks [A, B) to LDS (compile-time range)
    auto write_chunks = [&](auto A, auto B) {
        static_for<decltype(A)::value, decltype(B)::value>([&](auto C) {
            constexpr int c = decltype(C)::value;
            const int idx = CW * c + V * l64;
            if (idx < WE) *reinterpret_cast<W *>(wl + idx) = chunk[c];
        });
    };
    // per-lane bases, opaque so that per-step addresses fold into immediate offsets
    int fb = seg * BLK + lane - 1;  // forward column j: fb + colL(j) - j (j >= 1)
    int bb = seg * BLK + St::colL(lane) - lane - 1;  // backward row j: bb + j
    asm volatile("" : "+v"(fb), "+v"(bb));
    const int f0 = seg * BLK + (lane > 0 ? lane - 1 : 0);  // forward column 0
    auto fwd_read = [&](auto J) -> TS {
        constexpr int j = decltype(J)::value;
        if constexpr (j == 0) return wl[f0];
        else return wl[fb + (St::colL(j) - j)];
    };
    const int base4 = seg_base4(PM);

    write_chunks(std::integral_constant<int, 0>{}, std::integral_constant<int, St::need(RB - 1, CW)>{});
    lds_fence();
    TS rb[RB];
    static_for<0, RB>([&](auto C) { rb[decltype(C)::value] = fwd_read(C); });
    RowAcc<TC> row(xr);
    TC yfin = (TC)0;
    static_for<0, NL>([&](auto J) {
        constexpr int j = decltype(J)::value;
        __builtin_amdgcn_sched_barrier(0);
        const TS v = rb[j % RB];
        if constexpr (j + RB < NL) {
            constexpr int a = St::need(j + RB - 1, CW), b = St::need(j + RB, CW);
            if constexpr (b > a) {
                write_chunks(std::integral_constant<int, a>{}, std::integral_constant<int, b>{});
                lds_fence();
            }
            rb[j % RB] = fwd_read(std::integral_constant<int, j + RB>{});
        }
        TC q = row.q(dr, rr);
        TC yj = bcast_step<PM, j>(q, yfin, base4, lane);
        row.add(load_as<TC>(v), yj);  // dead for lanes <= j
    });
    TC q = row.q(dr, rr);
    TC yr = (lane == PM - 1) ? q : yfin;
    const TC xb = dr * yr;  // y .= D_Dl1 .* y  (:897)
    static_for<0, RB>([&](auto C) { rb[decltype(C)::value] = wl[bb + (NL - decltype(C)::value)]; });
    row = RowAcc<TC>(xb);
    yfin = (TC)0;
    static_for<0, NL>([&](auto JJ) {
        constexpr int jj = decltype(JJ)::value;
        constexpr int j = PM - 1 - jj;
        __builtin_amdgcn_sched_barrier(0);
        const TS v = rb[jj % RB];
        if constexpr (jj + RB < NL) rb[jj % RB] = wl[bb + (NL - (jj + RB))];
        TC qb = row.q(dr, rr);
        TC yj = bcast_step<PM, j>(qb, yfin, base4, lane);
        row.add(load_as<TC>(v), yj);  // dead for lanes >= j
    });
    q = row.q(dr, rr);
    return (lane == 0) ? q : yfin;
}

// One partition per PM-lane segment; returns x_i * y_i of the lane's row
// (r_i * z_i in the CG forms) or 0.  `input(i)` supplies the sweep's input
// for row i (and may update vectors of that row).  Every lane of the wave
// must call it together: the full-partition recurrence moves values between
// the segments of a wave (v_readlane / v_writelane, which ignore the
// execution mask), so the choice of recurrence is made once per wave, and a
// segment past the last partition (k >= nparts) runs the last partition's
// recurrence on dead lanes -- its loads stay inside the buffers and it stores
// nothing.  MODE_SYMT takes `lds`: this segment's symt_lds(PM) elements.
// PAD (P == PM): the last partition is stored as a full block (the rows past
// n have no couplings), and is swept as a full partition whose missing rows
// are identity rows (d = 1, x = 0): they stay exactly 0 and add exact zero
// terms to the real rows, whose results are therefore those of the short
// partition's own recurrence.
// `output(i, y_i)` stores the lane's result (sweep_seg_run: y[i] = y_i).
// LSTR: LDS elements per segment.  With symt_lds(PM) every MODE_SYMT form
// fits; with the block size PM (PM - 1) / 2 only the staged one, and a wave
// that cannot stage its blocks takes the register form (same results).
template <int PM, int MODE, class TS, class TC, bool PAD, int LSTR = symt_lds(PM), class In,
          class Out>
__device__ __forceinline__ double sweep_seg_run_io(In &&input, Out &&output,
                                                   const TS *__restrict__ T,
                                                   const TC *__restrict__ D, int64_t n, int P,
                                                   int64_t nparts, int64_t k, int lane,
                                                   TS *lds = nullptr) {
    const bool valid = k < nparts;
    const int64_t kk = valid ? k : nparts - 1;
    const int64_t s = kk * (int64_t)P;
    const int64_t rem = n - s;
    const int Pr = (int)(rem < P ? rem : P);  // rows of this partition
    const int Pa = PAD ? PM : Pr;             // rows of its stored block
    int64_t lb, ub;
    block_bases(kk, P, Pa, MODE, lb, ub);
    const bool live = valid && lane < Pr;
    const TS *Lk = T + lb;
    const TS *Uk = T + ub;

    double xin = 0.0;
    if (live) xin = input(s + lane);
    const TC xr = (TC)xin;
    const TC dr = live ? D[s + lane] : (TC)1;
    const TC rr = (TC)1 / dr;
    TC yr;
    if (__any(!safe_divisor(dr))) {
        yr = sweep_generic<PM, MODE, false, TS, TC>(xr, dr, rr, lane, Pa, live, Lk, Uk);
    } else if (MODE == MODE_SYMT && P == PM &&
               __all(valid && (PAD || Pr == PM) && (k % SymtStage<PM>::SEGS) == seg_of_wave<PM>())) {
        // every segment of the wave holds a full partition, the wave's
        // partitions consecutive: stage their blocks (LDS areas of a wave's
        // segments are contiguous, symt_lds(PM) values each)
        if constexpr (MODE == MODE_SYMT) {
            const int sw = seg_of_wave<PM>();
            yr = sweep_symt_staged<PM, TS, TC>(xr, dr, rr, lane, sw, T + (kk - sw) * SymtStage<PM>::BLK,
                                               lds - sw * LSTR);
        }
    } else if (__all(Pa == PM) && (MODE != MODE_SYMT || LSTR >= symt_lds(PM))) {
        if constexpr (MODE == MODE_SYMT && PM == 64)
            yr = sweep_full_symt64<32, TS, TC>(xr, dr, rr, lane, Lk, lds);
        else if constexpr (MODE == MODE_SYMT)
            yr = sweep_full_symt<PM, 32, TS, TC>(xr, dr, rr, lane, Lk, lds);
        else
            yr = sweep_full<PM, MODE, 16, TS, TC>(xr, dr, rr, lane, Lk, Uk);
    } else {
        yr = sweep_generic<PM, MODE, true, TS, TC>(xr, dr, rr, lane, Pa, live, Lk, Uk);
    }
    if (live) output(s + lane, (double)yr);
    return live ? xin * (double)yr : 0.0;
}

template <int PM, int MODE, class TS, class TC, bool PAD, class In>
__device__ __forceinline__ double sweep_seg_run(In &&input, double *y, const TS *__restrict__ T,
                                                const TC *__restrict__ D, int64_t n, int P,
                                                int64_t nparts, int64_t k, int lane,
                                                TS *lds = nullptr) {
    return sweep_seg_run_io<PM, MODE, TS, TC, PAD>(
        input, [&](int64_t i, double v) { y[i] = v; }, T, D, n, P, nparts, k, lane, lds);
}

// Forms of the single-system sweep kernels.
enum { FORM_PLAIN = 0, FORM_CG = 1, FORM_SMOOTH = 2 };

// The plain form (input x), the fused CG form (input r - alpha q, with the
// CG vector updates of cg_prologue) and the fused smoother form.
template <int PM, int MODE, class TS, class TC, int FORM, int LSTR = symt_lds(PM)>
__device__ __forceinline__ double sweep_seg_body(const double *x, double *y,
                                                 const TS *__restrict__ T,
                                                 const TC *__restrict__ D, int64_t n, int P,
                                                 int64_t nparts, int64_t k, int lane,
                                                 const CGFuse &f, const SmoothFuse &sf,
                                                 TS *lds = nullptr) {
    if constexpr (FORM == FORM_CG)
        return sweep_seg_run_io<PM, MODE, TS, TC, false, LSTR>(
            [&](int64_t i) { return cg_prologue(f, i); },
            [&](int64_t i, double z) {
                if constexpr (std::is_same_v<TC, float>) {
                    if (f.z32) {
                        f.z32[i] = (float)z;  // z is an FP32 number: exact
                        return;
                    }
                }
                y[i] = z;
            },
            T, D, n, P, nparts, k, lane, lds);
    else if constexpr (FORM == FORM_SMOOTH)
        return sweep_seg_run_io<PM, MODE, TS, TC, false, LSTR>(
            [&](int64_t i) { return row_dot(sf.A, sf.x, i) - sf.b[i]; },
            [&](int64_t i, double z) { sf.xout[i] = sf.x[i] - z; }, T, D, n, P, nparts, k, lane,
            lds);
    else
        return sweep_seg_run_io<PM, MODE, TS, TC, false, LSTR>(
            [&](int64_t i) { return x[i]; }, [&](int64_t i, double v) { y[i] = v; }, T, D, n, P,
            nparts, k, lane, lds);
}

}  // namespace hs
