// hs_setup.hip -- setup of the sparse operands on the device: validation of
// the caller's 1-based arrays, the CSC -> CSR transpose and the sliced-ELL
// operator, so that a build does O(nnz) host work only to copy its input.
//
// The transpose is the SparseMatrixCSR(A) of src/chebyshev.jl:144 and the
// row-wise view of a non-symmetric CSC that the O(nnz) l1 diagonal needs
// (the reference scans all columns per row, :1135-1216).  It is a stable
// radix sort of the entries, taken in CSC order, by row: inside each row the
// columns then ascend, exactly the order of a serial counting sort, so every
// later row sum is added in the order the reference adds it.
#include "hs_sparse.hpp"

#include <hipcub/hipcub.hpp>

namespace hs {

namespace {

constexpr int BLK = 256;

// bad bit 1: pointer array not starting at 1, decreasing or not ending at
// nnz + 1; bit 2: an index outside [1, n].
__global__ void k_check_ptr(int64_t n, int64_t nnz, const int64_t *__restrict__ ptr1,
                            unsigned *__restrict__ bad) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const bool ok = (c > 0 || ptr1[0] == 1) && ptr1[c + 1] >= ptr1[c] &&
                    (c < n - 1 || ptr1[n] == nnz + 1);
    if (!ok) atomicOr(bad, 1u);
}

__global__ void k_check_ind(int64_t n, int64_t nnz, const int64_t *__restrict__ ind1,
                            unsigned *__restrict__ bad) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz && (ind1[p] < 1 || ind1[p] > n)) atomicOr(bad, 2u);
}

// sort keys (0-based rows) and values (entry positions) of a CSC
__global__ void k_row_keys(int64_t nnz, const int64_t *__restrict__ rowval1,
                           int32_t *__restrict__ key, int32_t *__restrict__ pos) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) {
        key[p] = (int32_t)(rowval1[p] - 1);
        pos[p] = (int32_t)p;
    }
}

// column of every entry of a CSC, one thread per column
__global__ void k_entry_cols(int64_t n, const int64_t *__restrict__ colptr1,
                             int32_t *__restrict__ colof) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int64_t p = colptr1[c] - 1; p < colptr1[c + 1] - 1; ++p) colof[p] = (int32_t)c;
}

// entries in row order: 1-based column indices and values
__global__ void k_csr_gather(int64_t nnz, const int32_t *__restrict__ pos,
                             const int32_t *__restrict__ colof, const double *__restrict__ nzval,
                             int64_t *__restrict__ ci1, double *__restrict__ cv) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nnz) {
        const int32_t p = pos[q];
        ci1[q] = (int64_t)colof[p] + 1;
        cv[q] = nzval[p];
    }
}

// 1-based row pointer from the sorted keys: thread q (0..nnz) sets the start
// of the rows after entry q-1's row up to entry q's row (to row n for q = nnz)
__global__ void k_row_starts(int64_t n, int64_t nnz, const int32_t *__restrict__ key,
                             int64_t *__restrict__ rp1) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > nnz) return;
    const int64_t lo = (q == 0) ? 0 : (int64_t)key[q - 1] + 1;
    const int64_t hi = (q == nnz) ? n : (int64_t)key[q];
    for (int64_t r = lo; r <= hi; ++r) rp1[r] = q + 1;
}

// slice widths: one wave per 64-row slice; ent[s] = 64 * (longest row)
__global__ __launch_bounds__(BLK) void k_sell_width(int64_t n, int64_t nslices,
                                                    const int64_t *__restrict__ rp1,
                                                    int64_t *__restrict__ ent,
                                                    int32_t *__restrict__ rowlen) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t len = 0;
    if (i < n) {
        len = rp1[i + 1] - rp1[i];
        rowlen[i] = (int32_t)len;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(len, off, 64);
        len = o > len ? o : len;
    }
    const int64_t s = i >> 6;
    if ((threadIdx.x & 63) == 0 && s < nslices) ent[s] = 64 * len;
}

// slice s, lane l = row 64 s + l: entry k at sptr[s] + 64 k + l, padded with
// (own row, 0.0); rows past n (the last slice) pad with (n - 1, 0.0).  Also
// the 16-bit offsets column - row; *wide is set when one does not fit.
__global__ __launch_bounds__(BLK) void k_sell_pack(int64_t n, int64_t nslices,
                                                   const int64_t *__restrict__ sptr,
                                                   const int64_t *__restrict__ rp1,
                                                   const int64_t *__restrict__ ci1,
                                                   const double *__restrict__ cv,
                                                   int32_t *__restrict__ col,
                                                   double *__restrict__ val,
                                                   int16_t *__restrict__ col16,
                                                   int *__restrict__ wide) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = i >> 6;
    if (s >= nslices) return;
    const int l = (int)(i & 63);
    const int64_t w = (sptr[s + 1] - sptr[s]) >> 6;
    const int64_t p0 = i < n ? rp1[i] - 1 : 0;
    const int64_t len = i < n ? rp1[i + 1] - rp1[i] : 0;
    const int32_t pad = (int32_t)(i < n ? i : n - 1);
    bool fits = true;
    for (int64_t k = 0; k < w; ++k) {
        const int64_t q = sptr[s] + 64 * k + l;
        const int32_t c = k < len ? (int32_t)(ci1[p0 + k] - 1) : pad;
        col[q] = c;
        val[q] = k < len ? cv[p0 + k] : 0.0;
        const int64_t d = (int64_t)c - i;
        fits = fits && d >= -32767 && d <= 32767;
        col16[q] = (int16_t)(fits ? d : 0);
    }
    if (!fits) atomicOr(wide, 1);
}

// The (slice, k) bases of the 16-bit delta format: the 64 entries of group
// g = q >> 6 (slice s, position k) are one wave's; base = their smallest
// column, delta = column - base; *wide2 is set when a group spans more than
// 65,535 columns.
__global__ __launch_bounds__(BLK) void k_sell_kbase(int64_t nent, const int32_t *__restrict__ col,
                                                    int32_t *__restrict__ kbase,
                                                    uint16_t *__restrict__ cdelta,
                                                    int *__restrict__ wide2) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nent) return;  // nent is a multiple of 64: whole waves
    const int32_t c = col[q];
    int32_t lo = c, hi = c;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t a = __shfl_xor(lo, off, 64), b = __shfl_xor(hi, off, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((q & 63) == 0) kbase[q >> 6] = lo;
    const int64_t d = (int64_t)c - lo;
    cdelta[q] = (uint16_t)(d <= 65535 ? d : 0);
    if ((q & 63) == 0 && (int64_t)hi - lo > 65535) atomicOr(wide2, 1);
}

}  // namespace

void validate_sparse_device(int64_t n, int64_t nnz, const int64_t *ptr1, const int64_t *ind1,
                            hipStream_t st) {
    if (n == 0) return;
    DBuf<unsigned> bad(1);
    unsigned h = 0;
    HS_CHECK_HIP(hipMemcpyAsync(bad.get(), &h, sizeof h, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_check_ptr, dim3(grid_for(n, BLK)), dim3(BLK), 0, st, n, nnz, ptr1,
                       bad.get());
    if (nnz)
        hipLaunchKernelGGL(k_check_ind, dim3(grid_for(nnz, BLK)), dim3(BLK), 0, st, n, nnz, ind1,
                           bad.get());
    check_launch();
    HS_CHECK_HIP(hipMemcpyAsync(&h, bad.get(), sizeof h, hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    HS_REQUIRE(!(h & 1u), HS_EINVAL,
               "invalid sparse matrix: the pointer array must start at 1, not decrease and end "
               "at nnz + 1");
    HS_REQUIRE(!(h & 2u), HS_EINVAL, "invalid sparse matrix: an index is outside 1:n");
}

void csc_transpose_device(int64_t n, int64_t nnz, const int64_t *colptr1, const int64_t *rowval1,
                          const double *nzval, DBuf<int64_t> &rp1, DBuf<int64_t> &ci1,
                          DBuf<double> &cv, hipStream_t st) {
    HS_REQUIRE(n < (int64_t)INT32_MAX && nnz < (int64_t)INT32_MAX, HS_EUNSUPPORTED,
               "the device transpose takes fewer than 2^31-1 rows and entries");
    rp1.alloc(n + 1);
    ci1.alloc(nnz);
    cv.alloc(nnz);
    if (nnz > 0) {
        DBuf<int32_t> key(nnz), key2(nnz), pos(nnz), pos2(nnz), colof(nnz);
        hipLaunchKernelGGL(k_row_keys, dim3(grid_for(nnz, BLK)), dim3(BLK), 0, st, nnz, rowval1,
                           key.get(), pos.get());
        hipLaunchKernelGGL(k_entry_cols, dim3(grid_for(n, BLK)), dim3(BLK), 0, st, n, colptr1,
                           colof.get());
        check_launch();
        int bits = 1;
        while (bits < 31 && (int64_t{1} << bits) < n) ++bits;
        size_t tmp_bytes = 0;
        HS_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.get(), key2.get(),
                                                        pos.get(), pos2.get(), (int)nnz, 0, bits,
                                                        st));
        DBuf<uint8_t> tmp(tmp_bytes > 0 ? tmp_bytes : 1);
        HS_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tmp_bytes, key.get(),
                                                        key2.get(), pos.get(), pos2.get(),
                                                        (int)nnz, 0, bits, st));
        hipLaunchKernelGGL(k_csr_gather, dim3(grid_for(nnz, BLK)), dim3(BLK), 0, st, nnz,
                           pos2.get(), colof.get(), nzval, ci1.get(), cv.get());
        hipLaunchKernelGGL(k_row_starts, dim3(grid_for(nnz + 1, BLK)), dim3(BLK), 0, st, n, nnz,
                           key2.get(), rp1.get());
        check_launch();
        HS_CHECK_HIP(hipStreamSynchronize(st));  // the temporaries go out of scope
    } else {
        hipLaunchKernelGGL(k_row_starts, dim3(1), dim3(BLK), 0, st, n, nnz, (const int32_t *)nullptr,
                           rp1.get());
        check_launch();
    }
}

void sell_build_device(CSR &A, int64_t n, int64_t nnz, const int64_t *rp1, const int64_t *ci1,
                       const double *cv, hipStream_t st) {
    HS_REQUIRE(n < (int64_t)INT32_MAX, HS_EUNSUPPORTED, "more than 2^31-1 rows");
    A.n = n;
    A.nnz = nnz;
    A.nslices = (n + 63) / 64;
    if (n == 0) return;
    const int64_t rows = A.nslices * 64;
    DBuf<int64_t> ent(A.nslices + 1);
    A.rowlen.alloc(n);
    HS_CHECK_HIP(hipMemsetAsync(ent.get(), 0, ent.bytes(), st));
    hipLaunchKernelGGL(k_sell_width, dim3(grid_for(rows, BLK)), dim3(BLK), 0, st, n, A.nslices,
                       rp1, ent.get(), A.rowlen.get());
    check_launch();
    A.sptr.alloc(A.nslices + 1);
    size_t tmp_bytes = 0;
    HS_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ent.get(), A.sptr.get(),
                                                  (int)(A.nslices + 1), st));
    DBuf<uint8_t> tmp(tmp_bytes > 0 ? tmp_bytes : 1);
    HS_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tmp_bytes, ent.get(), A.sptr.get(),
                                                  (int)(A.nslices + 1), st));
    HS_CHECK_HIP(hipMemcpyAsync(&A.nent, A.sptr.get() + A.nslices, sizeof(int64_t),
                                hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    A.col.alloc(A.nent);
    A.val.alloc(A.nent);
    A.col16.alloc(A.nent);
    DBuf<int> wide(1);
    HS_CHECK_HIP(hipMemsetAsync(wide.get(), 0, sizeof(int), st));
    hipLaunchKernelGGL(k_sell_pack, dim3(grid_for(rows, BLK)), dim3(BLK), 0, st, n, A.nslices,
                       A.sptr.get(), rp1, ci1, cv, A.col.get(), A.val.get(), A.col16.get(),
                       wide.get());
    check_launch();
    int h_wide = 0;
    HS_CHECK_HIP(hipMemcpyAsync(&h_wide, wide.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));  // the caller's inputs may go out of scope
    if (!h_wide) return;
    A.col16.release();  // some column is farther than 32,767 from its row
    // the (slice, k)-base format, if every group spans fewer than 65,536 columns
    A.kbase.alloc(A.nent / 64);
    A.cdelta.alloc(A.nent);
    HS_CHECK_HIP(hipMemsetAsync(wide.get(), 0, sizeof(int), st));
    hipLaunchKernelGGL(k_sell_kbase, dim3(grid_for(A.nent, BLK)), dim3(BLK), 0, st, A.nent,
                       A.col.get(), A.kbase.get(), A.cdelta.get(), wide.get());
    check_launch();
    HS_CHECK_HIP(hipMemcpyAsync(&h_wide, wide.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    if (h_wide) {  // 32-bit columns
        A.kbase.release();
        A.cdelta.release();
    }
}

void csr_create_device(CSR &A, int64_t n, const int64_t *ptr, const int64_t *ind,
                       const double *val, int fmt, hipStream_t st) {
    HS_REQUIRE(n < (int64_t)INT32_MAX, HS_EUNSUPPORTED, "more than 2^31-1 rows");
    const int64_t nnz = n > 0 ? ptr[n] - 1 : 0;
    HS_REQUIRE(nnz >= 0, HS_EINVAL, "invalid pointer array");
    A.n = n;
    A.nnz = nnz;
    if (n == 0) return;
    DBuf<int64_t> dptr(n + 1), dind(nnz);
    DBuf<double> dval(nnz);
    HS_CHECK_HIP(hipMemcpyAsync(dptr.get(), ptr, (n + 1) * 8, hipMemcpyHostToDevice, st));
    if (nnz) {
        HS_CHECK_HIP(hipMemcpyAsync(dind.get(), ind, nnz * 8, hipMemcpyHostToDevice, st));
        HS_CHECK_HIP(hipMemcpyAsync(dval.get(), val, nnz * 8, hipMemcpyHostToDevice, st));
    }
    validate_sparse_device(n, nnz, dptr.get(), dind.get(), st);
    if (fmt == HS_FMT_CSR) {
        sell_build_device(A, n, nnz, dptr.get(), dind.get(), dval.get(), st);
        return;
    }
    DBuf<int64_t> rp1, ci1;
    DBuf<double> cv;
    csc_transpose_device(n, nnz, dptr.get(), dind.get(), dval.get(), rp1, ci1, cv, st);
    sell_build_device(A, n, nnz, rp1.get(), ci1.get(), cv.get(), st);
}

void csr_download(const CSR &A, int64_t *rowptr1, int64_t *colval1, double *nzval) {
    const int64_t n = A.n;
    rowptr1[0] = 1;
    if (n == 0) return;
    std::vector<int64_t> sp(A.nslices + 1);
    std::vector<int32_t> rl(n), col(A.nent);
    std::vector<double> val(A.nent);
    HS_CHECK_HIP(hipMemcpy(sp.data(), A.sptr.get(), sp.size() * 8, hipMemcpyDeviceToHost));
    HS_CHECK_HIP(hipMemcpy(rl.data(), A.rowlen.get(), rl.size() * 4, hipMemcpyDeviceToHost));
    if (A.nent) {
        HS_CHECK_HIP(hipMemcpy(col.data(), A.col.get(), col.size() * 4, hipMemcpyDeviceToHost));
        HS_CHECK_HIP(hipMemcpy(val.data(), A.val.get(), val.size() * 8, hipMemcpyDeviceToHost));
    }
    int64_t q = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t base = sp[i >> 6] + (i & 63);
        for (int64_t k = 0; k < rl[i]; ++k) {
            colval1[q] = (int64_t)col[base + 64 * k] + 1;
            nzval[q] = val[base + 64 * k];
            ++q;
        }
        rowptr1[i + 1] = q + 1;
    }
}

}  // namespace hs
