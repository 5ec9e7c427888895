/*
 * hsmooth.h -- C ABI of libhsmooth.so, the MI355X (gfx950) smoother library.
 *
 * Plain C types only (int64_t, double, void*).  Sparse inputs are host arrays
 * in the layout Julia's SparseMatrixCSC{Float64,Int64} / SparseMatrixCSR holds:
 * 1-based Int64 pointer and index arrays, Float64 values.  Vector arguments
 * are host or device pointers as selected by `loc`.  Every call is synchronous
 * unless HS_FLAG_ASYNC is passed, matching the reference's `synchronize` after
 * each launch.  On failure a negative status is returned and hs_last_error()
 * describes it; HS_ESINGULAR stores the offending 1-based column in
 * hs_last_error_aux() (Julia maps it to LinearAlgebra.SingularException(col)).
 *
 * Reference interfaces replaced (HybridSmoothers.jl v0.1.4, file:line):
 *   hs_l1gs_build   -> (builder::L1GSPrecBuilder)(A, partsize; isSymA, eta, sweep,
 *                      cache_strategy)  src/l1_gauss_seidel.jl:287-338, with
 *                      _blockpartitioning (GPU) :353-365, DiagonalIndices :422-436,
 *                      _make_sweep_plan :626-843, _precompute_and_pack_kernel! :1028-1066
 *   hs_l1gs_apply   -> LinearSolve.ldiv!(y, P, x)  :367-384 (+ _apply_sweep! :846-906)
 *   hs_l1gs_info    -> P.partitioning fields  :118-124
 *   hs_l1gs_get     -> P.sweep.{op,lop,uop}.D_DL1, .L.SLbuffer, .U.SUbuffer  :487-518
 *   hs_cheb_build   -> _build_smoother(ChebyshevFourth|ChebyshevFirst, A)
 *                      src/chebyshev.jl:141-156, :258-276
 *   hs_cheb_smooth  -> smooth!(x, s, b)  src/chebyshev.jl:158-183, :278-322
 *   hs_cheb_get     -> s.l1inv, s.rho, s.a, s.r, s.z, s.tmp  src/chebyshev.jl:121-129, :224-234
 *   hs_csr_*        -> the SparseMatrixCSR operator of chebyshev.jl:144 / the
 *                      LinearProblem matrix the Krylov solver multiplies by
 *   hs_pcg_solve    -> solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P)
 *                      (README.md:27-29, test/test_preconditioners.jl:93)
 */
#ifndef HSMOOTH_H
#define HSMOOTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HS_OK 0
#define HS_EINVAL (-1)         /* bad argument: Julia ErrorException */
#define HS_ESINGULAR (-2)      /* missing/zero diagonal: SingularException(aux) */
#define HS_EDEVICE (-3)        /* HIP runtime error */
#define HS_ENOMEM (-4)         /* device allocation failed */
#define HS_ENODEVICE (-5)      /* backend not functional (no usable GPU) */
#define HS_EUNSUPPORTED (-6)   /* combination not provided by this library */
#define HS_EAGAIN (-7)         /* batched engine full right now: retry after collecting */

/* enums */
#define HS_FMT_CSC 0
#define HS_FMT_CSR 1
#define HS_SWEEP_FORWARD 0
#define HS_SWEEP_BACKWARD 1
#define HS_SWEEP_SYMMETRIC 2
#define HS_CACHE_PACKED 0      /* PackedBufferCache */
#define HS_CACHE_VIEW 1        /* MatrixViewCache */
#define HS_CACHE_PACKED_SYM 2  /* packed, lower triangle only, U = L^T: symmetric A (isSymA or
                                  a symmetric CSR), SymmetricSweep, partsize <= 64; no
                                  counterpart in the reference (its PackedBufferCache stores
                                  both triangles, :788-799); same results, half the bytes */
#define HS_PREC_FP64 0         /* FP64 values, FP64 arithmetic */
#define HS_PREC_FP32 1         /* FP32 values and packed buffers, FP32 accumulation */
#define HS_PREC_BF16 2         /* BF16 off-diagonals, FP32 diagonal and accumulation */
#define HS_LOC_HOST 0
#define HS_LOC_DEVICE 1
#define HS_FLAG_ASYNC 1        /* do not synchronise the stream before returning */

#define HS_GET_D_DL1 0
#define HS_GET_SL 1            /* strict lower, reference packed layout */
#define HS_GET_SU 2            /* strict upper, reference packed layout */

typedef struct hs_l1gs *hs_l1gs_t;
typedef struct hs_csr *hs_csr_t;
typedef struct hs_cheb *hs_cheb_t;

/* ---- library ---------------------------------------------------------- */
/* Selects the device and checks that it is a usable GPU (KA `functional`). */
int hs_init(int device);
const char *hs_last_error(void);
int64_t hs_last_error_aux(void);
/* Bytes of device memory held by live handles of this process. */
int64_t hs_device_bytes(void);

/* ---- L1 Gauss-Seidel -------------------------------------------------- */
/*
 * Builds the preconditioner on the device.  n: rows (square matrix);
 * ptr/ind/val: 1-based CSC (fmt=HS_FMT_CSC) or CSR (HS_FMT_CSR) host arrays;
 * is_sym: isSymA (ignored for CSR, whose rows are read directly);
 * threads/blocks: GPUConfig, used for the partitioning fields exactly as
 * _blockpartitioning does (results never depend on them).
 */
int hs_l1gs_build(int64_t n, const int64_t *ptr, const int64_t *ind, const double *val,
                  int fmt, int is_sym, int64_t partsize, double eta, int sweep, int cache,
                  int precision, int64_t threads, int64_t blocks, void *stream,
                  hs_l1gs_t *out);
/* y = M^-1 x.  x and y may alias.  loc: HS_LOC_HOST or HS_LOC_DEVICE. */
int hs_l1gs_apply(hs_l1gs_t h, double *y, const double *x, int loc, void *stream, int flags);
/* out[0..7] = partsize, nparts, nchunks, chunksize, n, packed_size, sweep, cache */
int hs_l1gs_info(hs_l1gs_t h, int64_t *out);
/* Copies D_DL1 (n) or SL/SU (packed_size) in the reference layout to host `out`. */
int hs_l1gs_get(hs_l1gs_t h, int what, double *out);
/* Kernel variant of the partsize-64 symmetric sweep (results are identical):
 * 0 column loads straight into registers; 1 and 2 chunks streamed through an
 * LDS ring of 8 / 16 chunks, 12 / 24 steps ahead (2 is the default).  Used by
 * the autotuner.  No counterpart in the reference (one kernel per sweep). */
int hs_l1gs_set_variant(hs_l1gs_t h, int variant);
/* Persisted state (no counterpart in the reference, which rebuilds; for
 * services that evict and re-admit systems): writes the built preconditioner --
 * partitioning, D_Dl1, packed triangles and (MatrixViewCache) the device
 * matrix, exactly as stored on the device -- to `path` with a checksum, and
 * recreates it on the device without the matrix or a rebuild.  Loading a
 * truncated, corrupted or foreign file fails with HS_EINVAL. */
int hs_l1gs_save(hs_l1gs_t h, const char *path);
int hs_l1gs_load(const char *path, void *stream, hs_l1gs_t *out);
int hs_l1gs_destroy(hs_l1gs_t h);

/* ---- device sparse operator (CSR) ------------------------------------- */
/* Device CSR of A from 1-based CSC or CSR host arrays (CSC is transposed). */
int hs_csr_create(int64_t n, const int64_t *ptr, const int64_t *ind, const double *val,
                  int fmt, void *stream, hs_csr_t *out);
/* y = A x on device or host vectors. */
int hs_csr_spmv(hs_csr_t A, double *y, const double *x, int loc, void *stream, int flags);
/* out[0..1] = n, nnz */
int hs_csr_info(hs_csr_t A, int64_t *out);
/* Copies 1-based rowptr (n+1), colval (nnz) and nzval (nnz) to host arrays. */
int hs_csr_get(hs_csr_t A, int64_t *rowptr, int64_t *colval, double *nzval);
/* Bytes the products read per stored column: 2 when every column lies within
 * 32,767 of its row (16-bit offsets), else 4.  No counterpart in the
 * reference (a layout diagnostic: the operator's algorithmic bytes depend on it). */
int hs_csr_column_bytes(hs_csr_t A, int64_t *bytes);
int hs_csr_destroy(hs_csr_t A);

/* ---- Chebyshev smoothers --------------------------------------------- */
/* kind: 4 (ChebyshevFourth) or 1 (ChebyshevFirst).  Power iteration in FP64
 * on the device.  precision: HS_PREC_FP64 or HS_PREC_FP32 (values of the
 * smoother's SpMV and recurrences; the spectral estimate stays FP64). */
int hs_cheb_build(int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                  int kind, int degree, int power_iters, int precision, void *stream,
                  hs_cheb_t *out);
/* smooth!(x, s, b): x updated in place. */
int hs_cheb_smooth(hs_cheb_t s, double *x, const double *b, int loc, void *stream, int flags);
/* z = 0; smooth!(z, s, r)  -- the smoother used as a left preconditioner. */
int hs_cheb_apply(hs_cheb_t s, double *z, const double *r, int loc, void *stream, int flags);
/* out = {rho, a*, degree, kind, n} */
int hs_cheb_info(hs_cheb_t s, double *out);
/* what: 0 l1inv, 1 r, 2 z, 3 tmp (n doubles each); r and z as the last
 * hs_cheb_smooth or hs_cheb_apply left them (hs_pcg_solve uses them as its
 * preconditioner's workspace) */
int hs_cheb_get(hs_cheb_t s, int what, double *out);
hs_csr_t hs_cheb_matrix(hs_cheb_t s);
int hs_cheb_destroy(hs_cheb_t s);

/* ---- AMG smoother glue -------------------------------------------------- */
/*
 * AMG.smooth!(x, s::HybridSmoother, b) (ext/HybridSmoothersAMGExt.jl:29-35):
 * `iters` times { r = A x - b; z = M^-1 r; x -= z } with the L1GS
 * preconditioner P (any sweep and layout) and the device matrix A; the
 * reference runs 2 iterations.  x is updated in place; loc as elsewhere.
 */
int hs_l1gs_smooth(hs_l1gs_t P, hs_csr_t A, double *x, const double *b, int iters, int loc,
                   void *stream, int flags);
/*
 * No counterpart in the reference.  hs_l1gs_smooth on device vectors that
 * allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): its residual and correction vectors live in a fixed device
 * array of the library (2 x 2,097,152 doubles, 32 MiB per GPU, held until
 * the process ends), so systems of more than 2,097,152 rows are refused with
 * HS_EINVAL.  x and b are device pointers.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other threads, other streams)
 * race and must be serialised by the caller.  Same kernels and results as
 * hs_l1gs_smooth.
 */
int hs_l1gs_smooth_noalloc(hs_l1gs_t P, hs_csr_t A, double *x, const double *b, int iters,
                           void *stream, int flags);

/* ---- preconditioned CG ------------------------------------------------ */
#define HS_PCG_NONE 0
#define HS_PCG_L1GS 1
#define HS_PCG_CHEB 2

typedef struct {
    int64_t iters;        /* CG iterations taken */
    int64_t status;       /* 0 solved, 1 iteration limit, 2 breakdown (p'Ap <= 0) */
    double rnorm0;        /* sqrt(r0' z0) */
    double rnorm;         /* final sqrt(r' z) */
    double true_relres;   /* ||b - A x|| / ||b||, FP64 on the device */
    double seconds;       /* wall time of the solve on the host clock */
} hs_pcg_stats;

/*
 * CG with x0 = 0 and left preconditioner `prec` (an hs_l1gs_t or hs_cheb_t,
 * as given by prec_kind), stopping when sqrt(r'z) <= atol + rtol*sqrt(r0'z0)
 * (Krylov.jl `cg` rule) or after itmax iterations.  FP64 throughout the outer
 * iteration whatever the preconditioner's precision.
 */
int hs_pcg_solve(hs_csr_t A, int prec_kind, void *prec, const double *b, double *x,
                 double rtol, double atol, int64_t itmax, int loc, void *stream,
                 hs_pcg_stats *stats);
/*
 * hs_pcg_solve that also meets a stated true residual (no counterpart in the
 * reference, whose solve stops on the preconditioned norm alone): at each
 * stop of the rule above the FP64 true relative residual ||b - A x|| / ||b||
 * is computed on the device, and while it exceeds true_rtol the same CG
 * recurrence continues (no restart) with its threshold scaled down by
 * true_rtol / (true residual) * 0.95, until it is met or itmax is reached.
 * true_rtol = 0 is hs_pcg_solve.
 */
int hs_pcg_solve_true(hs_csr_t A, int prec_kind, void *prec, const double *b, double *x,
                      double rtol, double atol, int64_t itmax, double true_rtol, int loc,
                      void *stream, hs_pcg_stats *stats);

/* ---- workload generator -------------------------------------------------- */
/*
 * Host generator of the serving workload (no counterpart in the reference;
 * its tests build 1D/2D stencils in test/test_matrices.jl:5-19): 5-point
 * (nz = 1) or 7-point variable-coefficient Poisson on nx x ny x nz, face
 * coefficients exp(sigma * N(0,1)) from a counter-based generator keyed by
 * `seed`, Dirichlet faces in the diagonal.  Writes 1-based CSR rowptr (n+1),
 * colval and nzval (nnz); the matrix is symmetric, so these are also its CSC
 * arrays.  With colval == NULL only *nnz_out is set.
 */
int hs_problem_lognormal(int64_t nx, int64_t ny, int64_t nz, double sigma, uint64_t seed,
                         int64_t *rowptr, int64_t *colval, double *nzval, int64_t *nnz_out);

/* ---- batched persistent PCG (serving engine) ---------------------------- */
/*
 * No counterpart in the reference, whose users solve one system per
 * `solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P)` call (README.md:27-29).
 * Many SPD systems of uneven size stay resident on the device and advance
 * together: every graph launch runs `passes` CG iterations of every active
 * system, each preconditioned by its own symmetric L1GS (isSymA, eta, the
 * engine's partsize and precision) or its own l1-Jacobi Chebyshev smoother of
 * the 4th kind (the engine's degree and precision; rho from `power_iters`
 * power-iteration steps, run inside the engine's passes before the system's
 * first CG step).  Each system's arithmetic is that of hs_pcg_solve with that
 * preconditioner (x0 = 0, FP64 CG, Krylov.jl's rule), so its result does not
 * depend on the other systems in the batch.
 */
typedef struct hs_batch *hs_batch_t;

typedef struct {
    int64_t arena_bytes;    /* device memory for resident systems (matrix, smoother, vectors) */
    int64_t max_systems;    /* resident systems at once */
    int64_t staging_bytes;  /* pinned host staging for uploads; the largest system must fit */
    int64_t partsize;       /* 16, 32 or 64 */
    int precision;          /* smoother storage: HS_PREC_FP64, HS_PREC_FP32 or HS_PREC_BF16 */
    int passes;             /* CG iterations per step (one graph launch) */
    int grid;               /* workgroups of the persistent sweep kernel; 0 = as many as are
                               resident at once (the SpMV uses one per 256-row tile) */
    double eta;             /* the l1 rule's eta (reference default 1.5) */
    int triangles;          /* packed triangles stored per partition: 1 the lower only (U = L^T,
                               half the bytes; partsize 16, 32 and 64 in FP32/BF16),
                               2 both, 0 automatic (1 where possible) */
    int smoother;           /* 0 or HS_PCG_L1GS: L1GS (the fields above); HS_PCG_CHEB:
                               ChebyshevFourth(degree) as built by hs_cheb_build (kind 4,
                               precision FP64 or FP32; partsize, grid, eta and triangles
                               are not used) */
    int degree;             /* Chebyshev degree, 2 to 16 */
    int power_iters;        /* Chebyshev power-iteration steps for rho (the reference's 50) */
} hs_batch_config;

typedef struct {
    int64_t id;             /* as returned by hs_batch_submit */
    int64_t n;
    int64_t iters;          /* CG iterations */
    int64_t status;         /* 0 solved, 1 iteration limit, 2 breakdown */
    double rnorm0, rnorm;   /* sqrt(r0'z0), final sqrt(r'z) */
    double true_relres;     /* ||b - A x|| / ||b||, FP64 on the device */
    double t_submit, t_start, t_done;  /* hs_batch_clock() at submit, first step, retirement */
} hs_batch_result;

int hs_batch_create(const hs_batch_config *cfg, hs_batch_t *out);
/* Arena bytes a system with this 1-based pointer array (n+1 entries) occupies. */
int hs_batch_footprint(hs_batch_t e, int64_t n, const int64_t *ptr, int64_t *bytes);
/*
 * Admits one symmetric positive definite system: 1-based CSR arrays (or the
 * CSC arrays of the symmetric matrix, whose columns are its rows), host b.
 * The upload and the preconditioner build run on the engine's copy stream,
 * overlapping the steps.  x_out (host, n doubles) receives the solution when
 * the system is collected and must stay valid until then.  Returns HS_EAGAIN,
 * changing nothing, when no slot or no arena block is free.
 */
int hs_batch_submit(hs_batch_t e, int64_t n, const int64_t *ptr, const int64_t *ind,
                    const double *val, const double *b, double *x_out, double rtol,
                    double atol, int64_t itmax, int64_t *id);
/* Activates uploaded systems and enqueues one step; returns without waiting. */
int hs_batch_step(hs_batch_t e);
/* Waits for the last step, retires finished systems (x copied to x_out, their
 * memory freed) and returns up to `max` results not yet reported. */
int hs_batch_collect(hs_batch_t e, hs_batch_result *out, int64_t max, int64_t *count);
/* out[0..11] = active, uploading, free slots, free arena bytes, arena bytes,
 * largest free arena block, steps, retired, results waiting, the workgroups
 * of the SpMV (in the last step: the smallest of 1024, 8192, 65536 covering
 * the tile table) and sweep tile kernels (0 with the Chebyshev smoother, whose
 * kernels run on the SpMV's grid), and the packed triangles stored per
 * partition (1: the lower one only -- U = L^T -- whenever the sweep's LDS copy
 * of L fits: partsize 16, partsize 32 and 64 in FP32/BF16; 2: both; 0 with
 * the Chebyshev smoother) */
int hs_batch_info(hs_batch_t e, int64_t *out);
/* Copies a vector of a resident (submitted, not yet collected) system to host
 * `out` after the last step: what = 0 x, 1 r, 2 z, 3 b, 4 q (n doubles each),
 * 5 D_Dl1 (n, widened to FP64; L1GS), 6 the two partials of each 256-row
 * tile, 7 l1inv (n, FP64; Chebyshev), 8 rho (one value; Chebyshev, final once
 * the power iteration has ended), 9 the bytes per stored column of its upload
 * (one value: 2 for 16-bit offsets from the row, 4 for 32-bit columns).  For
 * tests and diagnostics. */
int hs_batch_peek(hs_batch_t e, int64_t id, int what, double *out);
/* 1 while the last step is still running on the device, 0 once it has
 * finished (or none is pending), negative on a device error.  Lets a caller
 * admit new systems exactly while the device is busy. */
int hs_batch_busy(hs_batch_t e);
/* Seconds since the engine was created (the clock of hs_batch_result). */
double hs_batch_clock(hs_batch_t e);
int hs_batch_destroy(hs_batch_t e);

#ifdef __cplusplus
}
#endif

#endif /* HSMOOTH_H */
