// hs_sparse.hpp -- device CSR operator, SpMV building blocks and deterministic
// reductions shared by the Chebyshev smoothers and the CG driver.
#pragma once

#include "hs_common.hpp"

#include <vector>

namespace hs {

// ---------------------------------------------------------------------------
// CSR on the device: 0-based int32 column indices, int64 row pointers.
// Values in FP64 (always) and optionally FP32 (reduced-precision smoothers).
// ---------------------------------------------------------------------------
struct CGWork;                 // CG workspace (hs_pcg.hip), owned by the matrix
void cgwork_free(CGWork *w);

struct CSR {
    uint64_t serial = next_serial();
    int64_t n = 0, nnz = 0;
    DBuf<int64_t> rowptr;
    DBuf<int32_t> col;
    DBuf<double> val;
    DBuf<float> val32;
    CGWork *cg = nullptr;

    CSR() = default;
    CSR(const CSR &) = delete;
    CSR &operator=(const CSR &) = delete;
    ~CSR() {
        if (cg) cgwork_free(cg);
    }
    void ensure_fp32(hipStream_t st);
};

// Host CSR (1-based Int64, as SparseMatrixCSR stores it) from 1-based CSC by a
// stable counting sort on rows: columns ascend inside every row.
void csc_to_csr_host(int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                     std::vector<int64_t> &rp, std::vector<int64_t> &ci, std::vector<double> &cv);

// Uploads a 1-based CSR (host) as a device CSR.
void csr_upload(CSR &A, int64_t n, const int64_t *rp1, const int64_t *ci1, const double *cv,
                hipStream_t st);

// ---------------------------------------------------------------------------
// row dot product of a CSR row with a gathered vector, summed in row order
// ---------------------------------------------------------------------------
template <class T, class TX>
__device__ __forceinline__ T row_dot(const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                     const T *__restrict__ v, const TX *__restrict__ x, int64_t i) {
    T acc = (T)0;
    const int64_t p1 = rp[i + 1];
    for (int64_t p = rp[i]; p < p1; ++p) acc = acc + v[p] * (T)x[ci[p]];
    return acc;
}

// r_i - a_1 x_1 - a_2 x_2 - ... term by term (5-argument mul! with alpha = -1, beta = 1)
template <class T, class TX>
__device__ __forceinline__ T row_sub(const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                     const T *__restrict__ v, const TX *__restrict__ x, int64_t i,
                                     T r) {
    const int64_t p1 = rp[i + 1];
    for (int64_t p = rp[i]; p < p1; ++p) r = r - v[p] * (T)x[ci[p]];
    return r;
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
__device__ __forceinline__ double finalize_sum(const double *__restrict__ partials, unsigned count,
                                              unsigned stride, double *smem4) {
    double v = 0.0;
    for (unsigned b = threadIdx.x; b < count; b += blockDim.x) v += partials[(size_t)b * stride];
    return block_sum(v, smem4);  // valid in thread 0
}

}  // namespace hs

struct hs_csr : hs::CSR {};
