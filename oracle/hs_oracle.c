/*
 * hs_oracle.c -- CPU oracle for the HybridSmoothers.jl smoother hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithms of
 * the reference package (HybridSmoothers.jl v0.1.4) so that the HIP library can
 * be checked against them.  Only the test-suite, the smoke check and the CPU
 * baseline leg of the benchmark may load it; the product library never does.
 *
 * Conventions follow the reference exactly: sparse inputs are Julia-style
 * 1-based Int64 CSC (colptr, rowval, nzval) or CSR (rowptr, colval, nzval);
 * partitions, work-items and packed offsets use the reference's formulas.
 * Every sum is accumulated in the same order as the reference, and the file is
 * compiled with -ffp-contract=off so that a*b+c is never fused (Julia does not
 * fuse `acc += a*b` either).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   partitioning         closed forms, reference test fields
 *   diag_indices         reference SingularException cases, brute force
 *   d_dl1 (all paths)    reference known answers (N=9, eta cases, non-sym A2),
 *                        closed forms on stencils, literal O(N^2) vs O(nnz)
 *   pack lower/upper     reference known answers (P=2, P=3), dense brute force
 *   sweeps               reference known answers, scipy triangular solves
 *                        (partsize=N), dense block brute force
 *   l1_rowsum            reference known answer [3,4,4,4,3]
 *   spectral radius      dense eigvals of M^-1 A (power-iteration limit)
 *   cheb1_astar          closed form k=1 (1/3), tabulated values
 *   cheb smooth          dense polynomial restatement, reference properties
 *   pcg                  scipy/dense solve; iteration counts vs Krylov.jl are
 *                        PARITY UNPINNED (third-party, not importable here)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

/* ---------------------------------------------------------------------- */
/* Partitioning: l1_gauss_seidel.jl:340-365                                 */
/* out = {partsize, nparts, nchunks, chunksize}                            */
/* ---------------------------------------------------------------------- */
static i64 ceil_div_float(i64 a, i64 b)
{
    /* Julia: convert(Ti, a / b |> ceil) with a Float64 quotient */
    return (i64)ceil((double)a / (double)b);
}

void hs_or_partition_cpu(i64 n, i64 partsize, i64 chunks, i64 *out)
{
    i64 nparts = ceil_div_float(n, partsize);
    i64 nchunks = chunks;
    i64 chunksize = (nparts + nchunks - 1) / nchunks; /* cld */
    out[0] = partsize;
    out[1] = nparts;
    out[2] = nchunks;
    out[3] = chunksize;
}

void hs_or_partition_gpu(i64 n, i64 partsize, i64 threads, i64 blocks, i64 *out)
{
    i64 nchunks = blocks;
    i64 nparts = ceil_div_float(n, partsize);
    i64 chunksize = ceil_div_float(nparts, nchunks);
    if (chunksize > threads) chunksize = threads;
    out[0] = partsize;
    out[1] = nparts;
    out[2] = nchunks;
    out[3] = chunksize;
}

/* Partition k (1-based) -> rows [start, end], l1_gauss_seidel.jl:947-958 */
static void part_bounds(i64 n, i64 partsize, i64 k, i64 *start, i64 *end)
{
    *start = (k - 1) * partsize + 1;
    i64 e = *start + partsize - 1;
    *end = e < n ? e : n;
}

/*
 * Work-item traversal of DiagonalPartsIterator (l1_gauss_seidel.jl:913-945)
 * under the KA.CPU mapping: nchunks workgroups of chunksize work-items; each
 * workgroup runs on one thread, work-item g (1-based) visits partitions
 * g, g+stride, ... with stride = nchunks*chunksize.
 */
typedef void (*part_fn)(void *ctx, i64 k, i64 start, i64 end);

static void for_each_partition(i64 n, i64 partsize, i64 nparts, i64 nchunks,
                               i64 chunksize, int nthreads, part_fn fn, void *ctx)
{
    i64 stride = nchunks * chunksize;
    i64 w;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (w = 0; w < nchunks; ++w) {
        for (i64 t = 0; t < chunksize; ++t) {
            i64 g = w * chunksize + t + 1;
            for (i64 k = g; k <= nparts; k += stride) {
                i64 s, e;
                part_bounds(n, partsize, k, &s, &e);
                fn(ctx, k, s, e);
            }
        }
    }
    (void)nthreads;
}

/* ---------------------------------------------------------------------- */
/* DiagonalIndices: l1_gauss_seidel.jl:422-436                              */
/* Returns 0 on success, else the 1-based column of the first missing or    */
/* zero diagonal (the reference throws SingularException(col)).             */
/* ---------------------------------------------------------------------- */
i64 hs_or_diag_indices(i64 n, const i64 *colptr, const i64 *rowval,
                       const double *nzval, i64 *diag)
{
    for (i64 col = 1; col <= n; ++col) {
        i64 r1 = colptr[col - 1];
        i64 r2 = colptr[col] - 1;
        /* searchsortedfirst(rowval, col, r1, r2): first p in [r1,r2] with rowval[p] >= col */
        i64 lo = r1, hi = r2 + 1;
        while (lo < hi) {
            i64 mid = lo + ((hi - lo) >> 1);
            if (rowval[mid - 1] < col) lo = mid + 1; else hi = mid;
        }
        r1 = lo;
        if (r1 > r2 || rowval[r1 - 1] != col || nzval[r1 - 1] == 0.0) return col;
        diag[col - 1] = r1;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Diagonal / off-partition scans: l1_gauss_seidel.jl:1101-1242             */
/* ---------------------------------------------------------------------- */
/* _diag_offpart_rowwise (:1101-1129): CSR, or symmetric CSC (row i = col i) */
static void diag_offpart_rowwise(const i64 *ptr, const i64 *ind, const double *val,
                                 i64 idx, i64 ps, i64 pe, double *b, double *d)
{
    double bb = 0.0, dd = 0.0;
    for (i64 p = ptr[idx - 1]; p <= ptr[idx] - 1; ++p) {
        i64 j = ind[p - 1];
        double v = val[p - 1];
        if (j == idx) bb = v;
        else if (j < ps || j > pe) dd += fabs(v);
    }
    *b = bb;
    *d = dd;
}

/* binary search of row `idx` in rowval[lo..hi] (1-based positions); 0 if absent */
static i64 bsearch_row(const i64 *rowval, i64 lo, i64 hi, i64 idx)
{
    i64 left = lo, right = hi;
    while (left <= right) {
        i64 mid = (left + right) >> 1;
        i64 row = rowval[mid - 1];
        if (row < idx) left = mid + 1;
        else if (row > idx) right = mid - 1;
        else return mid;
    }
    return 0;
}

/* _diag_offpart, non-symmetric CSC (:1135-1216): literal O(N) scan per row */
static void diag_offpart_csc_literal(i64 n, const i64 *colptr, const i64 *rowval,
                                     const double *nzval, const i64 *diag, i64 idx,
                                     i64 ps, i64 pe, double *b, double *d)
{
    double dd = 0.0;
    *b = nzval[diag[idx - 1] - 1];
    for (i64 col = 1; col <= idx - 1; ++col) {
        if (!(col < ps || col > pe)) continue;
        i64 p_lo = diag[col - 1] + 1, p_hi = colptr[col] - 1;
        if (p_lo > p_hi) continue;
        if (rowval[p_lo - 1] <= idx && rowval[p_hi - 1] >= idx) {
            i64 m = bsearch_row(rowval, p_lo, p_hi, idx);
            if (m) dd += fabs(nzval[m - 1]);
        }
    }
    for (i64 col = idx + 1; col <= n; ++col) {
        if (!(col < ps || col > pe)) continue;
        i64 p_lo = colptr[col - 1], p_hi = diag[col - 1] - 1;
        if (p_lo > p_hi) continue;
        if (rowval[p_lo - 1] <= idx && rowval[p_hi - 1] >= idx) {
            i64 m = bsearch_row(rowval, p_lo, p_hi, idx);
            if (m) dd += fabs(nzval[m - 1]);
        }
    }
    *d = dd;
}

/* D_Dl1 rule of _compute_D_Dl1_kernel! (:1016-1022) */
static double dl1_rule(double a, double d, double eta)
{
    double star = (a >= eta * d) ? 0.0 : d / 2.0;
    return a + star;
}

enum { FMT_CSC = 0, FMT_CSR = 1 };

typedef struct {
    i64 n;
    const i64 *ptr, *ind;
    const double *val;
    int fmt, sym;
    const i64 *diag;    /* non-sym CSC only (1-based positions) */
    const i64 *tptr;    /* O(nnz) variant: CSR transpose of the CSC (1-based) */
    const i64 *tind;
    const double *tval;
    double eta;
    double *D;
} dl1_ctx;

static void dl1_part_literal(void *vc, i64 k, i64 s, i64 e)
{
    dl1_ctx *c = (dl1_ctx *)vc;
    (void)k;
    for (i64 i = s; i <= e; ++i) {
        double a, d;
        if (c->fmt == FMT_CSR || c->sym)
            diag_offpart_rowwise(c->ptr, c->ind, c->val, i, s, e, &a, &d);
        else
            diag_offpart_csc_literal(c->n, c->ptr, c->ind, c->val, c->diag, i, s, e, &a, &d);
        c->D[i - 1] = dl1_rule(a, d, c->eta);
    }
}

/*
 * D_Dl1 for every row.  fmt: 0 CSC / 1 CSR; sym: isSymA.  For non-symmetric
 * CSC `diag` must hold DiagonalIndices; this path is the reference's literal
 * O(N^2) scan and is meant for small N.
 */
void hs_or_d_dl1(i64 n, const i64 *ptr, const i64 *ind, const double *val, int fmt,
                 int sym, const i64 *diag, i64 partsize, i64 nparts, i64 nchunks,
                 i64 chunksize, double eta, double *D, int nthreads)
{
    dl1_ctx c = {n, ptr, ind, val, fmt, sym, diag, 0, 0, 0, eta, D};
    for_each_partition(n, partsize, nparts, nchunks, chunksize, nthreads, dl1_part_literal, &c);
}

/*
 * CSR transpose of a 1-based CSC by a stable counting sort on the row index:
 * within each row the entries come out in ascending column order, the order
 * in which the literal scan above visits them.
 */
void hs_or_csc_to_csr(i64 n, i64 m, const i64 *colptr, const i64 *rowval,
                      const double *nzval, i64 *rowptr, i64 *colval, double *outval)
{
    /* n rows, m columns */
    i64 nnz = colptr[m] - 1;
    for (i64 i = 0; i <= n; ++i) rowptr[i] = 0;
    for (i64 p = 0; p < nnz; ++p) rowptr[rowval[p]] += 1;
    rowptr[0] = 1;
    for (i64 i = 1; i <= n; ++i) rowptr[i] += rowptr[i - 1];
    i64 *cursor = (i64 *)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
    for (i64 i = 0; i < n; ++i) cursor[i] = rowptr[i];
    for (i64 col = 1; col <= m; ++col) {
        for (i64 p = colptr[col - 1]; p <= colptr[col] - 1; ++p) {
            i64 r = rowval[p - 1];
            i64 q = cursor[r - 1]++;
            colval[q - 1] = col;
            outval[q - 1] = nzval[p - 1];
        }
    }
    free(cursor);
}

static void dl1_part_fast(void *vc, i64 k, i64 s, i64 e)
{
    dl1_ctx *c = (dl1_ctx *)vc;
    (void)k;
    for (i64 i = s; i <= e; ++i) {
        double a, d;
        diag_offpart_rowwise(c->tptr, c->tind, c->tval, i, s, e, &a, &d);
        /* the reference reads a_ii through DiagonalIndices (:1148) */
        a = c->val[c->diag[i - 1] - 1];
        c->D[i - 1] = dl1_rule(a, d, c->eta);
    }
}

/* O(nnz) restatement of the non-symmetric CSC D_Dl1 (bit-equal to the literal scan) */
void hs_or_d_dl1_nonsym_fast(i64 n, const i64 *colptr, const i64 *rowval,
                             const double *nzval, const i64 *diag, i64 partsize,
                             i64 nparts, i64 nchunks, i64 chunksize, double eta,
                             double *D, int nthreads)
{
    i64 nnz = colptr[n] - 1;
    i64 *rp = (i64 *)malloc(sizeof(i64) * (size_t)(n + 1));
    i64 *ci = (i64 *)malloc(sizeof(i64) * (size_t)(nnz > 0 ? nnz : 1));
    double *cv = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    hs_or_csc_to_csr(n, n, colptr, rowval, nzval, rp, ci, cv);
    dl1_ctx c = {n, colptr, rowval, nzval, FMT_CSC, 0, diag, rp, ci, cv, eta, D};
    for_each_partition(n, partsize, nparts, nchunks, chunksize, nthreads, dl1_part_fast, &c);
    free(rp);
    free(ci);
    free(cv);
}

/* ---------------------------------------------------------------------- */
/* Packing: l1_gauss_seidel.jl:1248-1519 and buffer size :698-701          */
/* ---------------------------------------------------------------------- */
i64 hs_or_packed_size(i64 n, i64 partsize, i64 nparts)
{
    i64 last = n - (nparts - 1) * partsize;
    return (partsize * (partsize - 1) * (nparts - 1)) / 2 + (last * (last - 1)) / 2;
}

typedef struct {
    i64 n, P;
    const i64 *ptr, *ind;
    const double *val;
    int fmt, sym, upper;
    double *buf;
} pack_ctx;

static void pack_part(void *vc, i64 k, i64 s, i64 e)
{
    pack_ctx *c = (pack_ctx *)vc;
    i64 P = c->P, Pa = e - s + 1;
    i64 block_off = (k - 1) * ((P * (P - 1)) / 2);
    int rowwise = (c->fmt == FMT_CSR) || c->sym;
    if (!c->upper) {
        if (rowwise) { /* _pack_strict_lower_rowwise! :1248-1277 */
            for (i64 i = s; i <= e; ++i) {
                i64 li = i - s + 1;
                i64 row_off = ((li - 1) * (li - 2)) / 2;
                for (i64 p = c->ptr[i - 1]; p <= c->ptr[i] - 1; ++p) {
                    i64 j = c->ind[p - 1];
                    if (j >= s && j < i) c->buf[block_off + row_off + (j - s + 1) - 1] = c->val[p - 1];
                }
            }
        } else { /* non-symmetric CSC lower :1279-1324 */
            for (i64 col = s; col <= e - 1; ++col) {
                i64 lj = col - s + 1;
                i64 p = c->ptr[col - 1], pend = c->ptr[col] - 1;
                while (p <= pend && c->ind[p - 1] <= col) ++p;
                while (p <= pend) {
                    i64 i = c->ind[p - 1];
                    if (i > e) break;
                    i64 li = i - s + 1;
                    i64 row_off = ((li - 1) * (li - 2)) / 2;
                    c->buf[block_off + row_off + lj - 1] = c->val[p - 1];
                    ++p;
                }
            }
        }
    } else {
        if (rowwise) { /* _pack_strict_upper_rowwise! :1380-1413 */
            for (i64 i = s; i <= e; ++i) {
                i64 li = i - s + 1;
                i64 row_start = (li - 1) * Pa - ((li - 1) * li) / 2;
                for (i64 p = c->ptr[i - 1]; p <= c->ptr[i] - 1; ++p) {
                    i64 j = c->ind[p - 1];
                    if (j > i && j <= e) {
                        i64 lj = j - s + 1;
                        c->buf[block_off + row_start + (lj - li) - 1] = c->val[p - 1];
                    }
                }
            }
        } else { /* non-symmetric CSC upper :1415-1469 */
            for (i64 col = s + 1; col <= e; ++col) {
                i64 lj = col - s + 1;
                for (i64 p = c->ptr[col - 1]; p <= c->ptr[col] - 1; ++p) {
                    i64 i = c->ind[p - 1];
                    if (i < col) {
                        if (i < s) continue;
                        i64 li = i - s + 1;
                        i64 row_start = (li - 1) * Pa - ((li - 1) * li) / 2;
                        c->buf[block_off + row_start + (lj - li) - 1] = c->val[p - 1];
                    } else {
                        break;
                    }
                }
            }
        }
    }
}

/* Packs the strict lower (upper=0) or upper (upper=1) in-partition triangles.
 * `buf` must be zero-filled by the caller (the reference allocates zeros). */
void hs_or_pack(i64 n, const i64 *ptr, const i64 *ind, const double *val, int fmt,
                int sym, i64 partsize, i64 nparts, i64 nchunks, i64 chunksize,
                int upper, double *buf, int nthreads)
{
    pack_ctx c = {n, partsize, ptr, ind, val, fmt, sym, upper, buf};
    for_each_partition(n, partsize, nparts, nchunks, chunksize, nthreads, pack_part, &c);
}

/* ---------------------------------------------------------------------- */
/* Sweeps: _apply_sweep_kernel! (:1068-1095), packed accumulation           */
/* (:1547-1585) and MatrixView accumulation (:1588-1737)                    */
/* ---------------------------------------------------------------------- */
enum { CACHE_PACKED = 0, CACHE_VIEW = 1 };

typedef struct {
    i64 n, P;
    double *y;
    const double *D;
    int upper, cache;
    const double *buf;                    /* packed */
    const i64 *ptr, *ind;                 /* view */
    const double *val;
    int fmt, sym;
    const i64 *diag;
} sweep_ctx;

static double acc_packed_lower(const sweep_ctx *c, i64 k, i64 s, i64 i)
{
    i64 P = c->P;
    i64 block_off = (k - 1) * ((P * (P - 1)) / 2);
    i64 li = i - s + 1;
    i64 row_off = ((li - 1) * (li - 2)) / 2;
    double acc = 0.0;
    for (i64 lj = 1; lj <= li - 1; ++lj)
        acc += c->buf[block_off + row_off + lj - 1] * c->y[s + lj - 1 - 1];
    return acc;
}

static double acc_packed_upper(const sweep_ctx *c, i64 k, i64 s, i64 e, i64 i)
{
    i64 P = c->P, Pa = e - s + 1;
    i64 block_off = (k - 1) * ((P * (P - 1)) / 2);
    i64 li = i - s + 1;
    i64 row_start = (li - 1) * Pa - ((li - 1) * li) / 2;
    double acc = 0.0;
    i64 num_upper = Pa - li;
    for (i64 off = 1; off <= num_upper; ++off) {
        i64 gj = i + off;
        if (gj > e) break;
        acc += c->buf[block_off + row_start + off - 1] * c->y[gj - 1];
    }
    return acc;
}

/* _accumulate_triangular_csc_binsearch (:1616-1650) */
static double acc_csc_binsearch(const sweep_ctx *c, i64 i, i64 col_lo, i64 col_hi, int upper)
{
    double acc = 0.0;
    for (i64 col = col_lo; col <= col_hi; ++col) {
        i64 di = c->diag[col - 1];
        i64 p_lo, p_hi;
        if (!upper) { p_lo = di + 1; p_hi = c->ptr[col] - 1; }
        else { p_lo = c->ptr[col - 1]; p_hi = di - 1; }
        if (p_lo <= p_hi && c->ind[p_lo - 1] <= i && c->ind[p_hi - 1] >= i) {
            i64 m = bsearch_row(c->ind, p_lo, p_hi, i);
            if (m) acc += c->val[m - 1] * c->y[col - 1];
        }
    }
    return acc;
}

static double acc_view(const sweep_ctx *c, i64 s, i64 e, i64 i)
{
    int rowwise = (c->fmt == FMT_CSR) || c->sym;
    if (rowwise) {
        double acc = 0.0;
        for (i64 p = c->ptr[i - 1]; p <= c->ptr[i] - 1; ++p) {
            i64 j = c->ind[p - 1];
            if (!c->upper) {
                if (s <= j && j < i) acc += c->val[p - 1] * c->y[j - 1];
            } else {
                if (i < j && j <= e) acc += c->val[p - 1] * c->y[j - 1];
            }
        }
        return acc;
    }
    if (!c->upper) {
        i64 hi = (i - 1 < e) ? i - 1 : e;
        return acc_csc_binsearch(c, i, s, hi, 0);
    }
    return acc_csc_binsearch(c, i, i + 1, e, 1);
}

static void sweep_part(void *vc, i64 k, i64 s, i64 e)
{
    sweep_ctx *c = (sweep_ctx *)vc;
    if (!c->upper) {
        for (i64 i = s; i <= e; ++i) {
            double acc = c->cache == CACHE_PACKED ? acc_packed_lower(c, k, s, i) : acc_view(c, s, e, i);
            c->y[i - 1] = (c->y[i - 1] - acc) / c->D[i - 1];
        }
    } else {
        for (i64 i = e; i >= s; --i) {
            double acc = c->cache == CACHE_PACKED ? acc_packed_upper(c, k, s, e, i) : acc_view(c, s, e, i);
            c->y[i - 1] = (c->y[i - 1] - acc) / c->D[i - 1];
        }
    }
}

/* One in-place partitioned sweep with a packed triangle buffer. */
void hs_or_sweep_packed(i64 n, double *y, const double *buf, int upper, const double *D,
                        i64 partsize, i64 nparts, i64 nchunks, i64 chunksize, int nthreads)
{
    sweep_ctx c;
    memset(&c, 0, sizeof c);
    c.n = n; c.P = partsize; c.y = y; c.D = D; c.upper = upper; c.cache = CACHE_PACKED; c.buf = buf;
    for_each_partition(n, partsize, nparts, nchunks, chunksize, nthreads, sweep_part, &c);
}

/* One in-place partitioned sweep reading the matrix (MatrixViewCache). */
void hs_or_sweep_view(i64 n, double *y, const i64 *ptr, const i64 *ind, const double *val,
                      int fmt, int sym, const i64 *diag, int upper, const double *D,
                      i64 partsize, i64 nparts, i64 nchunks, i64 chunksize, int nthreads)
{
    sweep_ctx c;
    memset(&c, 0, sizeof c);
    c.n = n; c.P = partsize; c.y = y; c.D = D; c.upper = upper; c.cache = CACHE_VIEW;
    c.ptr = ptr; c.ind = ind; c.val = val; c.fmt = fmt; c.sym = sym; c.diag = diag;
    for_each_partition(n, partsize, nparts, nchunks, chunksize, nthreads, sweep_part, &c);
}

/* y .= D_Dl1 .* y between the two passes of the symmetric sweep (:897) */
void hs_or_scale(i64 n, double *y, const double *D, int nthreads)
{
    i64 i;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (i = 0; i < n; ++i) y[i] = D[i] * y[i];
    (void)nthreads;
}

/* ---------------------------------------------------------------------- */
/* Chebyshev smoothers: chebyshev.jl                                        */
/* ---------------------------------------------------------------------- */
/* _l1_rowsum (:27-36): l1[row] += |a| over CSC columns in ascending order */
void hs_or_l1_rowsum(i64 n, i64 m, const i64 *colptr, const i64 *rowval,
                     const double *nzval, double *l1)
{
    for (i64 i = 0; i < n; ++i) l1[i] = 0.0;
    for (i64 col = 1; col <= m; ++col)
        for (i64 p = colptr[col - 1]; p <= colptr[col] - 1; ++p)
            l1[rowval[p - 1] - 1] += fabs(nzval[p - 1]);
}

/* CSR SpMV y = A x, sequential row sums (SparseMatricesCSR mul!) */
void hs_or_spmv(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                const double *x, double *y, int nthreads)
{
    i64 i;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (i = 0; i < n; ++i) {
        double acc = 0.0;
        for (i64 p = rowptr[i]; p <= rowptr[i + 1] - 1; ++p) acc += val[p - 1] * x[colval[p - 1] - 1];
        y[i] = acc;
    }
    (void)nthreads;
}

/* r <- r - A z, term by term in row order (5-argument mul! with alpha=-1, beta=1) */
void hs_or_spmv_sub(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                    const double *z, double *r, int nthreads)
{
    i64 i;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (i = 0; i < n; ++i) {
        double acc = r[i];
        for (i64 p = rowptr[i]; p <= rowptr[i + 1] - 1; ++p) acc -= val[p - 1] * z[colval[p - 1] - 1];
        r[i] = acc;
    }
    (void)nthreads;
}

static double norm2_seq(i64 n, const double *w)
{
    double s = 0.0;
    for (i64 i = 0; i < n; ++i) s += w[i] * w[i];
    return sqrt(s);
}

/* _l1jacobi_spectral_radius (:40-54) */
double hs_or_spectral_radius(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                             const double *l1inv, int power_iters, int nthreads)
{
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double v0 = 1.0 / sqrt((double)n);
    for (i64 i = 0; i < n; ++i) v[i] = v0;
    double rho = 1.0;
    for (int it = 0; it < power_iters; ++it) {
        hs_or_spmv(n, rowptr, colval, val, v, w, nthreads);
        for (i64 i = 0; i < n; ++i) w[i] *= l1inv[i];
        rho = norm2_seq(n, w);
        if (rho == 0.0) break;
        for (i64 i = 0; i < n; ++i) v[i] = w[i] / rho;
    }
    free(v);
    free(w);
    return rho;
}

/* _cheb1_astar (:81-89): bisection on x = sqrt(a*) */
double hs_or_cheb1_astar(int k)
{
    double lo = 1e-10, hi = 1.0 - 1e-10;
    for (int it = 0; it < 60; ++it) {
        double mid = (lo + hi) / 2.0;
        double x = mid;
        double phi = 8.0 * k * pow(1.0 - x * x, 2.0 * k) + x * (pow(1.0 - x, 4.0 * k) - pow(1.0 + x, 4.0 * k));
        if (phi > 0) lo = mid; else hi = mid;
    }
    double m = (lo + hi) / 2.0;
    return m * m;
}

/* ChebyshevFourth smooth! (:158-183).  r, z are the smoother's workspaces. */
void hs_or_cheb4_smooth(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                        const double *l1inv, double rho, int degree, double *x,
                        const double *b, double *r, double *z, int nthreads)
{
    hs_or_spmv(n, rowptr, colval, val, x, r, nthreads);
    for (i64 i = 0; i < n; ++i) r[i] = b[i] - r[i];
    for (i64 i = 0; i < n; ++i) z[i] = 0.0;
    for (int k = 1; k <= degree; ++k) {
        double ak = (double)(2 * k - 3) / (double)(2 * k + 1);
        double bk = (double)(8 * k - 4) / ((double)(2 * k + 1) * rho);
        for (i64 i = 0; i < n; ++i) {
            double zi = ak * z[i] + bk * l1inv[i] * r[i];
            z[i] = zi;
            x[i] += zi;
        }
        hs_or_spmv_sub(n, rowptr, colval, val, z, r, nthreads);
    }
}

/* ChebyshevFirst smooth! (:278-322) */
void hs_or_cheb1_smooth(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                        const double *l1inv, double rho, double a, int degree, double *x,
                        const double *b, double *r, double *z, double *tmp, int nthreads)
{
    double inv_rho = 1.0 / rho;
    hs_or_spmv(n, rowptr, colval, val, x, tmp, nthreads);
    double c = 2.0 / (1.0 + a);
    for (i64 i = 0; i < n; ++i) {
        double ri = l1inv[i] * (b[i] - tmp[i]) * inv_rho;
        r[i] = ri;
        z[i] = c * ri;
    }
    double rs = (1.0 - a) / (1.0 + a);
    for (int k = 1; k <= degree; ++k) {
        for (i64 i = 0; i < n; ++i) x[i] += z[i];
        if (k < degree) {
            hs_or_spmv(n, rowptr, colval, val, z, tmp, nthreads);
            for (i64 i = 0; i < n; ++i) r[i] -= l1inv[i] * tmp[i] * inv_rho;
            double rsn = 1.0 / (2.0 * (1.0 + a) / (1.0 - a) - rs);
            double c1 = rsn * rs, c2 = 4.0 * rsn / (1.0 - a);
            for (i64 i = 0; i < n; ++i) z[i] = c1 * z[i] + c2 * r[i];
            rs = rsn;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Preconditioned CG in the form of Krylov.jl `cg` with ldiv=true, x0 = 0:   */
/* stop when sqrt(r'z) <= atol + rtol*sqrt(r0'z0).  (Third-party; not in   */
/* the reference tree: iteration-count parity against it is unpinned.)     */
/* ---------------------------------------------------------------------- */
typedef void (*prec_fn)(void *ctx, double *z, const double *r);

typedef struct {
    /* L1GS packed symmetric sweep */
    i64 n, partsize, nparts, nchunks, chunksize;
    const double *D, *SL, *SU;
    int sweep; /* 0 fwd, 1 bwd, 2 sym */
    int nthreads;
} l1gs_prec;

static double dot_seq(i64 n, const double *a, const double *b)
{
    double s = 0.0;
    for (i64 i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* ldiv!(y, P, x) for a packed L1GS plan (:386-396 then :854-906) */
void hs_or_l1gs_ldiv(i64 n, double *y, const double *x, const double *D, const double *SL,
                     const double *SU, int sweep, i64 partsize, i64 nparts, i64 nchunks,
                     i64 chunksize, int nthreads)
{
    memcpy(y, x, sizeof(double) * (size_t)n);
    if (sweep == 0 || sweep == 2)
        hs_or_sweep_packed(n, y, SL, 0, D, partsize, nparts, nchunks, chunksize, nthreads);
    if (sweep == 2) hs_or_scale(n, y, D, nthreads);
    if (sweep == 1 || sweep == 2)
        hs_or_sweep_packed(n, y, SU, 1, D, partsize, nparts, nchunks, chunksize, nthreads);
}

static void l1gs_prec_apply(void *vc, double *z, const double *r)
{
    l1gs_prec *p = (l1gs_prec *)vc;
    hs_or_l1gs_ldiv(p->n, z, r, p->D, p->SL, p->SU, p->sweep, p->partsize, p->nparts,
                    p->nchunks, p->chunksize, p->nthreads);
}

typedef struct {
    i64 n;
    const i64 *rowptr, *colval;
    const double *val, *l1inv;
    double rho, a;
    int kind, degree, nthreads;
    double *r, *zz, *tmp;
} cheb_prec;

/* Chebyshev as a left preconditioner: z = 0; smooth!(z, S, r) */
static void cheb_prec_apply(void *vc, double *z, const double *r)
{
    cheb_prec *p = (cheb_prec *)vc;
    for (i64 i = 0; i < p->n; ++i) z[i] = 0.0;
    if (p->kind == 4)
        hs_or_cheb4_smooth(p->n, p->rowptr, p->colval, p->val, p->l1inv, p->rho, p->degree, z, r,
                           p->r, p->zz, p->nthreads);
    else
        hs_or_cheb1_smooth(p->n, p->rowptr, p->colval, p->val, p->l1inv, p->rho, p->a, p->degree,
                           z, r, p->r, p->zz, p->tmp, p->nthreads);
}

static void ident_apply(void *vc, double *z, const double *r)
{
    i64 n = *(const i64 *)vc;
    memcpy(z, r, sizeof(double) * (size_t)n);
}

/* stats = {iterations, final sqrt(r'z), initial sqrt(r'z), status(0 solved,1 tired,2 breakdown)} */
static void pcg_core(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                     const double *b, double *x, double rtol, double atol, i64 itmax,
                     prec_fn M, void *mctx, int nthreads, double *stats)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *z = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    double *q = (double *)malloc(sizeof(double) * (size_t)n);
    for (i64 i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; }
    M(mctx, z, r);
    memcpy(p, z, sizeof(double) * (size_t)n);
    double gamma = dot_seq(n, r, z);
    double rnorm = sqrt(gamma);
    double rnorm0 = rnorm;
    i64 iter = 0;
    int status = 0;
    if (gamma == 0.0) {
        stats[0] = 0; stats[1] = 0; stats[2] = 0; stats[3] = 0;
        free(r); free(z); free(p); free(q);
        return;
    }
    double eps = atol + rtol * rnorm;
    int solved = rnorm <= eps;
    while (!solved && iter < itmax) {
        hs_or_spmv(n, rowptr, colval, val, p, q, nthreads);
        double pq = dot_seq(n, p, q);
        if (pq <= 0.0) { status = 2; break; }
        double alpha = gamma / pq;
        for (i64 i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (i64 i = 0; i < n; ++i) r[i] -= alpha * q[i];
        M(mctx, z, r);
        double gnext = dot_seq(n, r, z);
        rnorm = sqrt(gnext);
        solved = (rnorm <= eps) || (rnorm + 1.0 <= 1.0);
        if (!solved) {
            double beta = gnext / gamma;
            gamma = gnext;
            for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        iter += 1;
    }
    if (!solved && status == 0) status = 1;
    stats[0] = (double)iter;
    stats[1] = rnorm;
    stats[2] = rnorm0;
    stats[3] = status;
    free(r); free(z); free(p); free(q);
}

void hs_or_pcg_l1gs(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                    const double *b, double *x, double rtol, double atol, i64 itmax,
                    const double *D, const double *SL, const double *SU, int sweep,
                    i64 partsize, i64 nparts, i64 nchunks, i64 chunksize, int nthreads,
                    double *stats)
{
    l1gs_prec pc = {n, partsize, nparts, nchunks, chunksize, D, SL, SU, sweep, nthreads};
    pcg_core(n, rowptr, colval, val, b, x, rtol, atol, itmax, l1gs_prec_apply, &pc, nthreads, stats);
}

void hs_or_pcg_cheb(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                    const double *b, double *x, double rtol, double atol, i64 itmax,
                    const double *l1inv, double rho, double a, int kind, int degree,
                    int nthreads, double *stats)
{
    double *w1 = (double *)malloc(sizeof(double) * (size_t)n);
    double *w2 = (double *)malloc(sizeof(double) * (size_t)n);
    double *w3 = (double *)malloc(sizeof(double) * (size_t)n);
    cheb_prec pc = {n, rowptr, colval, val, l1inv, rho, a, kind, degree, nthreads, w1, w2, w3};
    pcg_core(n, rowptr, colval, val, b, x, rtol, atol, itmax, cheb_prec_apply, &pc, nthreads, stats);
    free(w1); free(w2); free(w3);
}

void hs_or_pcg_none(i64 n, const i64 *rowptr, const i64 *colval, const double *val,
                    const double *b, double *x, double rtol, double atol, i64 itmax,
                    int nthreads, double *stats)
{
    i64 nn = n;
    pcg_core(n, rowptr, colval, val, b, x, rtol, atol, itmax, ident_apply, &nn, nthreads, stats);
}

int hs_or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
