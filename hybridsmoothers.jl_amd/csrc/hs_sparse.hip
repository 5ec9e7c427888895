// hs_sparse.hip -- device CSR operator (the SparseMatrixCSR of
// src/chebyshev.jl:144 and the system matrix of the CG driver) and its SpMV.
#include "hs_sparse.hpp"

#include <algorithm>
#include <memory>

namespace hs {

template <class TD>
__global__ void k_to_type(int64_t n, const double *__restrict__ s, TD *__restrict__ d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (TD)s[i];
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
        csr_create_device(*A, n, ptr, ind, val, fmt, st);
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

int hs_csr_column_bytes(hs_csr_t A, int64_t *bytes) {
    return guard([&] {
        HS_REQUIRE(A && bytes, HS_EINVAL, "NULL argument");
        *bytes = A->column_bytes();
    });
}

int hs_csr_get(hs_csr_t A, int64_t *rowptr, int64_t *colval, double *nzval) {
    return guard([&] {
        HS_REQUIRE(A && rowptr && (A->nnz == 0 || (colval && nzval)), HS_EINVAL, "NULL argument");
        csr_download(*A, rowptr, colval, nzval);
    });
}

int hs_csr_destroy(hs_csr_t A) {
    return guard([&] { delete A; });
}

}  // extern "C"
