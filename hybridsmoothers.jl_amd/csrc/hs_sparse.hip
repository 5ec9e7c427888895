// hs_sparse.hip -- device CSR operator (the SparseMatrixCSR of
// src/chebyshev.jl:144 and the system matrix of the CG driver) and its SpMV.
#include "hs_sparse.hpp"

#include <algorithm>
#include <memory>

namespace hs {

void csc_to_csr_host(int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                     std::vector<int64_t> &rp, std::vector<int64_t> &ci, std::vector<double> &cv) {
    const int64_t nnz = colptr[n] - 1;
    rp.assign(n + 1, 0);
    ci.resize(nnz);
    cv.resize(nnz);
    for (int64_t p = 0; p < nnz; ++p) rp[rowval[p]] += 1;
    rp[0] = 1;
    for (int64_t i = 1; i <= n; ++i) rp[i] += rp[i - 1];
    std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
    for (int64_t c = 1; c <= n; ++c)
        for (int64_t p = colptr[c - 1]; p <= colptr[c] - 1; ++p) {
            int64_t q = cur[rowval[p - 1] - 1]++;
            ci[q - 1] = c;
            cv[q - 1] = nzval[p - 1];
        }
}

template <class TD>
__global__ void k_to_type(int64_t n, const double *__restrict__ s, TD *__restrict__ d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (TD)s[i];
}

void csr_upload(CSR &A, int64_t n, const int64_t *rp1, const int64_t *ci1, const double *cv,
                hipStream_t st) {
    HS_REQUIRE(n < (int64_t)INT32_MAX, HS_EUNSUPPORTED, "more than 2^31-1 rows");
    const int64_t nnz = rp1[n] - 1;
    A.n = n;
    A.nnz = nnz;
    A.nslices = (n + 63) / 64;
    A.h_rp.resize(n + 1);
    A.h_ci.resize(nnz);
    A.h_cv.assign(cv, cv + nnz);
    for (int64_t i = 0; i <= n; ++i) A.h_rp[i] = rp1[i] - 1;
    for (int64_t p = 0; p < nnz; ++p) A.h_ci[p] = (int32_t)(ci1[p] - 1);
    // slice widths and offsets
    std::vector<int64_t> sp(A.nslices + 1, 0);
    for (int64_t s = 0; s < A.nslices; ++s) {
        int64_t w = 0;
        for (int64_t i = 64 * s; i < std::min<int64_t>(64 * s + 64, n); ++i)
            w = std::max<int64_t>(w, A.h_rp[i + 1] - A.h_rp[i]);
        sp[s + 1] = sp[s] + 64 * w;
    }
    A.nent = sp[A.nslices];
    std::vector<int32_t> col(A.nent);
    std::vector<double> val(A.nent);
    for (int64_t s = 0; s < A.nslices; ++s) {
        const int64_t w = (sp[s + 1] - sp[s]) / 64;
        for (int l = 0; l < 64; ++l) {
            const int64_t i = 64 * s + l;
            const int64_t len = i < n ? A.h_rp[i + 1] - A.h_rp[i] : 0;
            for (int64_t k = 0; k < w; ++k) {
                const int64_t q = sp[s] + 64 * k + l;
                if (k < len) {
                    col[q] = A.h_ci[A.h_rp[i] + k];
                    val[q] = A.h_cv[A.h_rp[i] + k];
                } else {  // padding: the row's own column, value zero
                    col[q] = (int32_t)(i < n ? i : n - 1);
                    val[q] = 0.0;
                }
            }
        }
    }
    A.sptr.alloc(A.nslices + 1);
    A.col.alloc(A.nent);
    A.val.alloc(A.nent);
    HS_CHECK_HIP(hipMemcpyAsync(A.sptr.get(), sp.data(), (A.nslices + 1) * 8,
                                hipMemcpyHostToDevice, st));
    if (A.nent) {
        HS_CHECK_HIP(hipMemcpyAsync(A.col.get(), col.data(), A.nent * 4, hipMemcpyHostToDevice, st));
        HS_CHECK_HIP(hipMemcpyAsync(A.val.get(), val.data(), A.nent * 8, hipMemcpyHostToDevice, st));
    }
    HS_CHECK_HIP(hipStreamSynchronize(st));  // the host staging vectors go out of scope
}

void CSR::ensure_fp32(hipStream_t st) {
    if (val32.get() || nent == 0) return;
    val32.alloc(nent);
    hipLaunchKernelGGL((k_to_type<float>), dim3(grid_for(nent, 256)), dim3(256), 0, st, nent,
                       val.get(), val32.get());
    check_launch();
}

// y = A x, one thread per row, sequential row sums (mul! of SparseMatricesCSR)
__global__ void k_spmv(SellView<double> A, const double *__restrict__ x, double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n) y[i] = row_dot(A, x, i);
}

void spmv_device(const CSR &A, const double *x, double *y, hipStream_t st) {
    if (A.n == 0) return;
    hipLaunchKernelGGL(k_spmv, dim3(grid_for(A.n, 256)), dim3(256), 0, st, A.view(), x, y);
    check_launch();
}

}  // namespace hs

using namespace hs;

extern "C" {

int hs_csr_create(int64_t n, const int64_t *ptr, const int64_t *ind, const double *val, int fmt,
                  void *stream, hs_csr_t *out) {
    return guard([&] {
        HS_REQUIRE(out, HS_EINVAL, "out handle pointer is NULL");
        *out = nullptr;
        HS_REQUIRE(n >= 0, HS_EINVAL, "matrix size must be non-negative");
        HS_REQUIRE(fmt == HS_FMT_CSC || fmt == HS_FMT_CSR, HS_EINVAL, "unknown matrix format");
        HS_REQUIRE(n == 0 || (ptr && ind && val), HS_EINVAL, "matrix arrays are NULL");
        hipStream_t st = as_stream(stream);
        std::unique_ptr<hs_csr> A(new hs_csr());
        if (n > 0) {
            if (fmt == HS_FMT_CSR) {
                csr_upload(*A, n, ptr, ind, val, st);
            } else {
                std::vector<int64_t> rp, ci;
                std::vector<double> cv;
                csc_to_csr_host(n, ptr, ind, val, rp, ci, cv);
                csr_upload(*A, n, rp.data(), ci.data(), cv.data(), st);
            }
        }
        *out = A.release();
    });
}

int hs_csr_spmv(hs_csr_t A, double *y, const double *x, int loc, void *stream, int flags) {
    return guard([&] {
        HS_REQUIRE(A, HS_EINVAL, "NULL matrix handle");
        HS_REQUIRE(A->n == 0 || (x && y), HS_EINVAL, "NULL vector");
        HS_REQUIRE(x != y, HS_EINVAL, "y and x must not alias");
        hipStream_t st = as_stream(stream);
        if (A->n == 0) return;
        if (loc == HS_LOC_DEVICE) {
            spmv_device(*A, x, y, st);
            finish(st, flags);
            return;
        }
        DBuf<double> dx(A->n), dy(A->n);
        HS_CHECK_HIP(hipMemcpyAsync(dx.get(), x, A->n * 8, hipMemcpyHostToDevice, st));
        spmv_device(*A, dx.get(), dy.get(), st);
        HS_CHECK_HIP(hipMemcpyAsync(y, dy.get(), A->n * 8, hipMemcpyDeviceToHost, st));
        HS_CHECK_HIP(hipStreamSynchronize(st));
    });
}

int hs_csr_info(hs_csr_t A, int64_t *out) {
    return guard([&] {
        HS_REQUIRE(A && out, HS_EINVAL, "NULL argument");
        out[0] = A->n;
        out[1] = A->nnz;
    });
}

int hs_csr_get(hs_csr_t A, int64_t *rowptr, int64_t *colval, double *nzval) {
    return guard([&] {
        HS_REQUIRE(A && rowptr && (A->nnz == 0 || (colval && nzval)), HS_EINVAL, "NULL argument");
        if (A->n == 0) {
            rowptr[0] = 1;
            return;
        }
        for (int64_t i = 0; i <= A->n; ++i) rowptr[i] = A->h_rp[i] + 1;
        for (int64_t p = 0; p < A->nnz; ++p) {
            colval[p] = (int64_t)A->h_ci[p] + 1;
            nzval[p] = A->h_cv[p];
        }
    });
}

int hs_csr_destroy(hs_csr_t A) {
    return guard([&] { delete A; });
}

}  // extern "C"
