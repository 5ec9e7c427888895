// hs_sparse.hpp -- device CSR operator, SpMV building blocks and deterministic
// reductions shared by the Chebyshev smoothers and the CG driver.
#pragma once

#include "hs_common.hpp"

#include <vector>

namespace hs {

// ---------------------------------------------------------------------------
// The device operator: sliced ELL with 64-row slices.  Slice s holds rows
// [64s, 64s+64); its entries are stored column-major -- entry k of row 64s+l
// at sptr[s] + 64k + l -- padded to the longest row of the slice with
// (column = the row itself, value = 0).  One wave reads one slice with fully
// coalesced loads, and every row is still summed in its CSR order (padding
// comes last and adds zero), so results equal a row-by-row CSR product.
// When every stored column is within 32,767 of its row (banded and stencil
// matrices up to that bandwidth), the products read the column as a 16-bit
// offset from the row (col16: 2 bytes per entry instead of 4).  Otherwise,
// when in every slice the k-th entries of the 64 rows span fewer than 65,536
// columns (stencils and banded matrices of wider bandwidth, e.g. the 192^3
// grid's +-36,864 couplings), as a 16-bit delta from a 32-bit base per
// (slice, k) (kbase, cdelta).  The 32-bit columns are kept for the read-back
// and for matrices where neither fits.
// The operator is built on the device (hs_setup.hip); the row lengths are
// kept so that the CSR can be read back (csr_download).
// ---------------------------------------------------------------------------
struct CGWork;                 // CG workspace (hs_pcg.hip), owned by the matrix
void cgwork_free(CGWork *w);

template <class T>
struct SellView {
    const int64_t *sptr;
    const int32_t *col;
    const T *val;
    int64_t n;
    const int16_t *col16 = nullptr;  // column - row, when every entry fits
    // otherwise, when the k-th entries of every slice span fewer than 65,536
    // columns: entry p's column is kbase[p >> 6] + cdelta[p] (slices start at
    // multiples of 64 entries, so p >> 6 numbers the (slice, k) groups)
    const int32_t *kbase = nullptr;
    const uint16_t *cdelta = nullptr;
    const int32_t *rowlen = nullptr;  // stored entries of each row (the rest is padding)
};

// column of stored entry p of row i in the 16-bit formats (kbase view or offset from the row)
template <class T>
__device__ __forceinline__ int64_t sell_col16(const SellView<T> &A, int64_t i, int64_t p) {
    return A.col16 ? i + A.col16[p] : (int64_t)A.kbase[p >> 6] + A.cdelta[p];
}

struct CSR {
    uint64_t serial = next_serial();
    int64_t n = 0, nnz = 0, nslices = 0, nent = 0;
    DBuf<int64_t> sptr;
    DBuf<int32_t> col;
    DBuf<double> val;
    DBuf<float> val32;
    DBuf<int32_t> rowlen;
    DBuf<int16_t> col16;
    DBuf<int32_t> kbase;   // with cdelta: the 16-bit format of matrices col16 cannot hold
    DBuf<uint16_t> cdelta;
    CGWork *cg = nullptr;

    CSR() = default;
    CSR(const CSR &) = delete;
    CSR &operator=(const CSR &) = delete;
    ~CSR() {
        if (cg) cgwork_free(cg);
    }
    void ensure_fp32(hipStream_t st);
    SellView<double> view() const {
        return {sptr.get(), col.get(), val.get(), n, col16.get(), kbase.get(), cdelta.get(),
                rowlen.get()};
    }
    SellView<float> view32() const {
        return {sptr.get(), col.get(), val32.get(), n, col16.get(), kbase.get(), cdelta.get(),
                rowlen.get()};
    }
    template <class T> SellView<T> view_as() const {
        if constexpr (sizeof(T) == 8) return view();
        else return view32();
    }
    // bytes per stored column as the products read it: 2 (offset from the row,
    // or delta from the (slice, k) base) or 4
    int column_bytes() const { return (col16.get() || kbase.get()) ? 2 : 4; }
    // algorithmic bytes of one product: values, column indices, slice offsets
    // (and the (slice, k) bases)
    int64_t bytes(int vbytes) const {
        return nent * (vbytes + column_bytes()) + (nslices + 1) * 8 +
               (kbase.get() ? (nent / 64) * 4 : 0);
    }
};

// Device setup (hs_setup.hip).  Arrays are 1-based Int64 as SparseMatrixCSC /
// SparseMatrixCSR hold them.
// HS_EINVAL unless ptr1 starts at 1, never decreases and ends at nnz + 1,
// and every index is in [1, n] (device arrays; checked before any other
// kernel reads them).
void validate_sparse_device(int64_t n, int64_t nnz, const int64_t *ptr1, const int64_t *ind1,
                            hipStream_t st);
// CSR of a device CSC (stable by rows: columns ascend inside every row).
void csc_transpose_device(int64_t n, int64_t nnz, const int64_t *colptr1, const int64_t *rowval1,
                          const double *nzval, DBuf<int64_t> &rp1, DBuf<int64_t> &ci1,
                          DBuf<double> &cv, hipStream_t st);
// The sliced-ELL operator of a device CSR.
void sell_build_device(CSR &A, int64_t n, int64_t nnz, const int64_t *rp1, const int64_t *ci1,
                       const double *cv, hipStream_t st);
// The operator of host CSC (fmt HS_FMT_CSC) or CSR arrays: upload, validate,
// transpose if needed, pack.
void csr_create_device(CSR &A, int64_t n, const int64_t *ptr, const int64_t *ind,
                       const double *val, int fmt, hipStream_t st);
// The operator's CSR back on the host (1-based; rowptr1 has n + 1 entries).
void csr_download(const CSR &A, int64_t *rowptr1, int64_t *colval1, double *nzval);

// ---------------------------------------------------------------------------
// row products, summed in row order
// ---------------------------------------------------------------------------
// Folds the stored entries of row i in order (padding last):
// acc = f(acc, value, gather(column)) entry by entry.  The entries are taken
// eight at a time: all column indices and values of a batch are loaded first,
// with no branch around the loads (past the slice width, the same in the whole
// wave, the slice's last entry is loaded again and not folded), then the
// batch's gathers are issued together, then the terms are folded in row
// order.  The loop form (load, gather, fold, one entry after the other) left a
// wave one dependent load chain per entry.  The column format is the same
// for the whole matrix, so the branch on it is uniform.
template <int CF, class T, class R, class G, class F>
__device__ __forceinline__ R row_fold_t(const SellView<T> &A, int64_t i, R acc, G &&gather,
                                        F &&f) {
    constexpr int B = 8;
    const int64_t s = i >> 6;
    const int64_t s0 = A.sptr[s];
    const int64_t p0 = s0 + (i & 63);
    const int w = (int)((A.sptr[s + 1] - s0) >> 6);  // slice width, same in the wave
    for (int k0 = 0; k0 < w; k0 += B) {
        int64_t c[B];
        T v[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int64_t p = p0 + 64 * (int64_t)((k0 + k < w) ? k0 + k : w - 1);
            if constexpr (CF == 1) c[k] = i + A.col16[p];
            else if constexpr (CF == 2) c[k] = (int64_t)A.kbase[p >> 6] + A.cdelta[p];
            else c[k] = (int64_t)A.col[p];
            v[k] = A.val[p];
        }
        decltype(gather(c[0])) g[B];
#pragma unroll
        for (int k = 0; k < B; ++k) g[k] = gather(c[k]);
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (k0 + k < w) acc = f(acc, v[k], g[k]);
    }
    return acc;
}
template <class T, class R, class G, class F>
__device__ __forceinline__ R row_fold(const SellView<T> &A, int64_t i, R acc, G &&gather,
                                      F &&f) {
    if (A.col16) return row_fold_t<1>(A, i, acc, gather, f);
    if (A.kbase) return row_fold_t<2>(A, i, acc, gather, f);
    return row_fold_t<0>(A, i, acc, gather, f);
}

template <class T, class TX>
__device__ __forceinline__ T row_dot(const SellView<T> &A, const TX *__restrict__ x, int64_t i) {
    return row_fold(
        A, i, (T)0, [&](int64_t c) { return (T)x[c]; },
        [](T acc, T a, T xc) { return acc + a * xc; });
}

// r_i - a_1 x_1 - a_2 x_2 - ... term by term (5-argument mul! with alpha = -1, beta = 1)
template <class T, class TX>
__device__ __forceinline__ T row_sub(const SellView<T> &A, const TX *__restrict__ x, int64_t i,
                                     T r) {
    return row_fold(
        A, i, r, [&](int64_t c) { return (T)x[c]; }, [](T acc, T a, T xc) { return acc - a * xc; });
}

// sum_j |a_ij| in row order
template <class T>
__device__ __forceinline__ double row_abs_sum(const SellView<T> &A, int64_t i) {
    const int64_t s = i >> 6;
    const int l = (int)(i & 63);
    const int64_t e = A.sptr[s + 1];
    double acc = 0.0;
    for (int64_t p = A.sptr[s] + l; p < e; p += 64) acc += fabs((double)A.val[p]);
    return acc;
}

// ---------------------------------------------------------------------------
// deterministic reductions (fixed tree; grid result by the last block)
// ---------------------------------------------------------------------------
constexpr int RED_BLOCK = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over a 256-thread block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *smem4) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) smem4[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = ((smem4[0] + smem4[1]) + smem4[2]) + smem4[3];
    __syncthreads();
    return s;
}

// Grid-wide deterministic sums are done in two launches: every block stores
// its partial (plain store), then one block adds the partials in index order
// (finalize_sum).  Kernel boundaries give the visibility; a "last block
// finishes" scheme would need an agent-scope release in every block, which on
// this chip writes back the XCD's L2 each time.
// Sum over a block of up to 1024 threads (blockDim a multiple of 64); the
// result is valid in thread 0.
__device__ __forceinline__ double block_sum_wide(double v, double *smem16) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) smem16[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < nw; ++k) s += smem16[k];
    __syncthreads();
    return s;
}

constexpr int FIN_BLOCK = 1024;  // threads of the finalize kernels

// The dispatcher hands workgroup b of a launch to XCD b mod 8.  xcd_tile(b, G)
// numbers the G tiles so that each XCD works through one contiguous range of
// them (XCD x gets tiles [x q + min(x, r), ...) with q = G / 8, r = G mod 8):
// a stencil row's neighbours in the next and previous planes were then loaded
// by the same XCD, into the same L2.  A bijection of [0, G).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned G) {
    const unsigned x = b & 7u, q = G >> 3, r = G & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Thread t adds partials t, t + blockDim, t + 2 blockDim, ... one after the
// other; the loads of 32 consecutive terms are issued before their additions
// (same order, same result), so the loop is not one load latency per term
// (32,768 partials over 1,024 threads: one round of loads).
__device__ __forceinline__ double finalize_sum(const double *__restrict__ partials, unsigned count,
                                              unsigned stride, double *smem16) {
    constexpr unsigned U = 32;
    const unsigned step = blockDim.x;
    double v = 0.0;
    unsigned b = threadIdx.x;
    for (; b + (U - 1) * step < count; b += U * step) {
        double t[U];
#pragma unroll
        for (unsigned k = 0; k < U; ++k) t[k] = partials[(size_t)(b + k * step) * stride];
#pragma unroll
        for (unsigned k = 0; k < U; ++k) v += t[k];
    }
    for (; b < count; b += step) v += partials[(size_t)b * stride];
    return block_sum_wide(v, smem16);  // valid in thread 0
}

}  // namespace hs

struct hs_csr : hs::CSR {};
