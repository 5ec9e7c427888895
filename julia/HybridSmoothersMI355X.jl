"""
    HybridSmoothersMI355X

The MI355X (gfx950) device path of HybridSmoothers.jl v0.1.4, reached through
the C ABI of `libhsmooth.so` (`include/hsmooth.h`) with `ccall`.

Loaded next to the reference package,

    using HybridSmoothers, AMDGPU, HybridSmoothersMI355X

it adds methods to the reference's own types for the `ROCBackend`, so the
reference API is unchanged for its callers:

    P = L1GSPrecBuilder(ROCBackend(); threads = 256, blocks = 20)(A, partsize;
            isSymA = true, η = 1.5, sweep = SymmetricSweep(), cache_strategy = PackedBufferCache())
    P = L1GSPrecBuilder(ROCBackend(); threads = 256, blocks = 20)(Symmetric(A), partsize)
    LinearSolve.ldiv!(y, P, x)                 # Vector or ROCVector
    y = P \\ x
    Vector(P.sweep.lop.D_DL1); P.sweep.lop.L.SLbuffer; P.partitioning.nparts
    S = HybridSmoothers._build_smoother(ChebyshevFourth(; degree = 3, backend = ROCBackend()), A)
    HybridSmoothers.smooth!(x, S, b); S.l1inv; S.ρ; S.A
    sol = solve(LinearProblem(A, b), KrylovJL_CG(); Pl = P)   # one ldiv! per iteration

Methods replaced (src/l1_gauss_seidel.jl, src/chebyshev.jl of the reference):
* `(builder::L1GSPrecBuilder)(A, partsize; isSymA, η, sweep, cache_strategy)`  :287-296
* `(builder::L1GSPrecBuilder)(A::Symmetric, partsize; η, sweep, cache_strategy)`  :305-313
* `LinearSolve.ldiv!(y, P, x)`  :367-396;  `P \\ x`  :398-407
* `_build_smoother(::ChebyshevFourth / ::ChebyshevFirst, A)`  chebyshev.jl:141-156, :258-276
* `smooth!(x, s, b)`  chebyshev.jl:158-183, :278-322
* `AMG.smooth!(x, s::HybridSmoother, b)`  ext/HybridSmoothersAMGExt.jl:29-35 (as `amg_smooth!`)
Additions without a reference counterpart: the whole PCG solve on the device
(`pcg_solve`), the allocation-free smoother round, persisted preconditioners,
the batched engine (`BatchEngine`).
"""
module HybridSmoothersMI355X

using SparseArrays: SparseArrays, SparseMatrixCSC
using SparseMatricesCSR: SparseMatricesCSR, SparseMatrixCSR
using LinearAlgebra: LinearAlgebra, Symmetric, SingularException
import LinearSolve
import AMDGPU
using AMDGPU: ROCBackend, ROCVector
import HybridSmoothers as HS
import HybridSmoothers: L1GSPrecBuilder, GPUConfig, BlockPartitioning, ForwardSweep,
    BackwardSweep, SymmetricSweep, PackedBufferCache, MatrixViewCache, ChebyshevFourth,
    ChebyshevFirst
import Base: \

export PackedSymmetricCache, DeviceMatrix, pcg_solve, amg_smooth!, amg_smooth_noalloc!,
    save_preconditioner, load_preconditioner, BatchEngine

const libhsmooth = get(ENV, "HSMOOTH_LIB",
    joinpath(@__DIR__, "..", "hybridsmoothers.jl_amd", "hybridsmoothers", "libhsmooth.so"))

# include/hsmooth.h constants
const HS_OK = Cint(0)
const HS_ESINGULAR = Cint(-2)
const HS_EAGAIN = Cint(-7)
const FMT_CSC = Cint(0)
const FMT_CSR = Cint(1)
const SWEEP_FORWARD = Cint(0)
const SWEEP_BACKWARD = Cint(1)
const SWEEP_SYMMETRIC = Cint(2)
const CACHE_PACKED = Cint(0)
const CACHE_VIEW = Cint(1)
const CACHE_PACKED_SYM = Cint(2)
const LOC_HOST = Cint(0)
const LOC_DEVICE = Cint(1)
const GET_D_DL1 = Cint(0)
const GET_SL = Cint(1)
const GET_SU = Cint(2)
const PCG_NONE = Cint(0)
const PCG_L1GS = Cint(1)
const PCG_CHEB = Cint(2)
const PRECISION = Dict(:fp64 => Cint(0), :fp32 => Cint(1), :bf16 => Cint(2))

"""
    PackedSymmetricCache()

Cache strategy of this library (no reference counterpart): the strict lower
triangle of each partition alone, `U = Lᵀ` read back from it, for symmetric `A`
(`isSymA = true` or a `Symmetric` wrapper), `SymmetricSweep`, partsize ≤ 64.
Results equal `PackedBufferCache` bit for bit from half the packed bytes.
"""
struct PackedSymmetricCache <: HS.AbstractCacheStrategy end

sweepcode(::ForwardSweep) = SWEEP_FORWARD
sweepcode(::BackwardSweep) = SWEEP_BACKWARD
sweepcode(::SymmetricSweep) = SWEEP_SYMMETRIC
cachecode(::PackedBufferCache) = CACHE_PACKED
cachecode(::MatrixViewCache) = CACHE_VIEW
cachecode(::PackedSymmetricCache) = CACHE_PACKED_SYM

# Status -> exception, as the reference raises them: `error(...)` with the
# library's message, `SingularException(col)` for a missing or zero diagonal.
function check(st::Cint)
    st == HS_OK && return nothing
    msg = unsafe_string(ccall((:hs_last_error, libhsmooth), Cstring, ()))
    if st == HS_ESINGULAR
        col = ccall((:hs_last_error_aux, libhsmooth), Int64, ())
        throw(SingularException(Int(col)))
    end
    error(msg)
end

stream_handle() = AMDGPU.stream().stream  # the task's HIP stream (hipStream_t)

# the HIP device current on this thread (as AMDGPU.jl selected it), asked of
# the HIP runtime itself so that no device numbering convention is assumed
function current_hip_device()
    dev = Ref{Cint}(0)
    ccall((:hipGetDevice, "libamdhip64"), Cint, (Ref{Cint},), dev) == 0 ||
        error("hipGetDevice failed")
    return dev[]
end

function __init__()
    AMDGPU.device()  # initialises AMDGPU.jl's device selection
    check(ccall((:hs_init, libhsmooth), Cint, (Cint,), current_hip_device()))
end

# the arrays of a SparseMatrixCSC{Float64,Int64} / SparseMatrixCSR{1,Float64,Int64}
sparse_arrays(A::SparseMatrixCSC{Float64, Int64}) = (A.colptr, A.rowval, A.nzval, FMT_CSC)
sparse_arrays(A::SparseMatrixCSR{1, Float64, Int64}) =
    (HS.getrowptr(A), HS.colvals(A), A.nzval, FMT_CSR)

# ---------------------------------------------------------------------------
# L1 Gauss-Seidel
# ---------------------------------------------------------------------------
"""
The device preconditioner behind `L1GSPrecBuilder{ROCBackend}`.  Fields the
reference's tests read are materialised from the device in the reference's
layout: `P.partitioning` (a `BlockPartitioning`), `P.sweep.op.D_DL1`,
`P.sweep.{lop,uop}.D_DL1`, `P.sweep.op.L.SLbuffer`, `P.sweep.op.U.SUbuffer`,
`P.sweep.lop.L.SLbuffer`, `P.sweep.uop.U.SUbuffer` (all `Vector{Float64}`).
"""
mutable struct MI355XL1GSPreconditioner
    handle::Ptr{Cvoid}
    n::Int
    packed_size::Int
    sweep_kind::Cint
    cache::Cint
    partitioning::BlockPartitioning{Int64, ROCBackend}
    function MI355XL1GSPreconditioner(h::Ptr{Cvoid}, backend::ROCBackend)
        info = Vector{Int64}(undef, 8)
        check(ccall((:hs_l1gs_info, libhsmooth), Cint, (Ptr{Cvoid}, Ptr{Int64}), h, info))
        part = BlockPartitioning(info[1], info[2], info[3], info[4], backend)
        P = new(h, Int(info[5]), Int(info[6]), Cint(info[7]), Cint(info[8]), part)
        finalizer(P) do p
            ccall((:hs_l1gs_destroy, libhsmooth), Cint, (Ptr{Cvoid},), getfield(p, :handle))
        end
        return P
    end
end

function build_l1gs(builder::L1GSPrecBuilder{<:ROCBackend, <:GPUConfig}, A, partsize::Integer,
                    isSymA::Bool, η, sweep::HS.AbstractSweep,
                    cache_strategy::HS.AbstractCacheStrategy, precision::Symbol)
    ptr, ind, val, fmt = sparse_arrays(A)
    n = size(A, 1)
    size(A, 2) == n || error("the matrix must be square")
    cfg = builder.device_config
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve ptr ind val begin
        check(ccall((:hs_l1gs_build, libhsmooth), Cint,
            (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Cint, Cint, Int64, Cdouble, Cint, Cint,
             Cint, Int64, Int64, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
            n, ptr, ind, val, fmt, Cint(isSymA), Int64(partsize), Float64(η), sweepcode(sweep),
            cachecode(cache_strategy), PRECISION[precision], Int64(cfg.threads),
            Int64(cfg.blocks), stream_handle(), out))
    end
    return MI355XL1GSPreconditioner(out[], builder.backend)
end

# (builder::L1GSPrecBuilder)(A, partsize; isSymA, η, sweep, cache_strategy)  (:287-296);
# `precision = :fp32 / :bf16` selects the reduced-precision smoother (an addition).
function (builder::L1GSPrecBuilder{<:ROCBackend, <:GPUConfig})(
    A::Union{SparseMatrixCSC{Float64, Int64}, SparseMatrixCSR{1, Float64, Int64}},
    partsize::Integer = builder.device_config.blocks;
    isSymA::Bool = false,
    η = 1.5,
    sweep::HS.AbstractSweep = SymmetricSweep(),
    cache_strategy::HS.AbstractCacheStrategy = PackedBufferCache(),
    precision::Symbol = :fp64,
)
    return build_l1gs(builder, A, partsize, isSymA, η, sweep, cache_strategy, precision)
end

# (builder::L1GSPrecBuilder)(A::Symmetric, partsize; η, sweep, cache_strategy)  (:305-313):
# the full matrix from the stored triangle (get_data, :604-610) with isSymA = true
function (builder::L1GSPrecBuilder{<:ROCBackend, <:GPUConfig})(
    A::Symmetric{Float64, <:SparseMatrixCSC{Float64, Int64}},
    partsize::Integer = builder.device_config.blocks;
    η = 1.5,
    sweep::HS.AbstractSweep = SymmetricSweep(),
    cache_strategy::HS.AbstractCacheStrategy = PackedBufferCache(),
    precision::Symbol = :fp64,
)
    return build_l1gs(builder, HS.get_data(A), partsize, true, η, sweep, cache_strategy,
                      precision)
end

# LinearSolve.ldiv!(y, P, x): overwrites y, leaves x unchanged, returns nothing (:367-396)
function LinearSolve.ldiv!(y::Vector{Float64}, P::MI355XL1GSPreconditioner, x::Vector{Float64})
    length(x) == length(y) == getfield(P, :n) || throw(DimensionMismatch("vector length"))
    GC.@preserve x y begin
        check(ccall((:hs_l1gs_apply, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(P, :handle), pointer(y), pointer(x), LOC_HOST, C_NULL, Cint(0)))
    end
    return nothing
end
function LinearSolve.ldiv!(y::ROCVector{Float64}, P::MI355XL1GSPreconditioner,
                           x::ROCVector{Float64})
    length(x) == length(y) == getfield(P, :n) || throw(DimensionMismatch("vector length"))
    GC.@preserve x y begin
        check(ccall((:hs_l1gs_apply, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(P, :handle), pointer(y), pointer(x), LOC_DEVICE, stream_handle(), Cint(0)))
    end
    return nothing
end
# any other vector type: y .= x on the host, sweep, copy back (the generic method, :367-384)
function LinearSolve.ldiv!(y::AbstractVector, P::MI355XL1GSPreconditioner, x::AbstractVector)
    w = Vector{Float64}(x)
    LinearSolve.ldiv!(w, P, w)
    copyto!(y, w)
    return nothing
end

# P \ x allocates y like x (:398-407)
function (\)(P::MI355XL1GSPreconditioner, x::AbstractVector)
    y = similar(x, Float64)
    LinearSolve.ldiv!(y, P, x)
    return y
end

function download_l1gs(P::MI355XL1GSPreconditioner, what::Cint, len::Integer)
    out = Vector{Float64}(undef, len)
    check(ccall((:hs_l1gs_get, libhsmooth), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                getfield(P, :handle), what, out))
    return out
end

# P.sweep.op / .lop / .uop and their D_DL1, L.SLbuffer, U.SUbuffer (:487-518)
struct SweepView
    P::MI355XL1GSPreconditioner
end
struct OperatorView
    P::MI355XL1GSPreconditioner
    side::Symbol  # :L (lower operator) or :U (upper operator)
end
struct LowerCacheView
    P::MI355XL1GSPreconditioner
end
struct UpperCacheView
    P::MI355XL1GSPreconditioner
end

function Base.getproperty(P::MI355XL1GSPreconditioner, s::Symbol)
    s === :sweep && return SweepView(P)
    return getfield(P, s)
end
Base.propertynames(::MI355XL1GSPreconditioner) = (:partitioning, :sweep)

function Base.getproperty(v::SweepView, s::Symbol)
    P = getfield(v, :P)
    kind = getfield(P, :sweep_kind)
    if s === :op
        kind == SWEEP_SYMMETRIC && error("a SymmetricSweep plan has `lop` and `uop`")
        return OperatorView(P, kind == SWEEP_FORWARD ? :L : :U)
    elseif s === :lop || s === :uop
        kind == SWEEP_SYMMETRIC || error("a one-sided sweep plan has `op`")
        return OperatorView(P, s === :lop ? :L : :U)
    end
    return getfield(v, s)
end

function Base.getproperty(o::OperatorView, s::Symbol)
    P = getfield(o, :P)
    side = getfield(o, :side)
    s === :D_DL1 && return download_l1gs(P, GET_D_DL1, getfield(P, :n))
    s === :L && side === :L && return LowerCacheView(P)
    s === :U && side === :U && return UpperCacheView(P)
    return getfield(o, s)
end

function Base.getproperty(v::LowerCacheView, s::Symbol)
    P = getfield(v, :P)
    s === :SLbuffer && return download_l1gs(P, GET_SL, getfield(P, :packed_size))
    return getfield(v, s)
end
function Base.getproperty(v::UpperCacheView, s::Symbol)
    P = getfield(v, :P)
    s === :SUbuffer && return download_l1gs(P, GET_SU, getfield(P, :packed_size))
    return getfield(v, s)
end

"""
    save_preconditioner(P, path); load_preconditioner(path; backend = ROCBackend())

A built preconditioner written to a file (partitioning, D_DL1, packed
triangles as stored on the device, checksummed) and recreated without the
matrix.  No reference counterpart (the reference rebuilds).
"""
function save_preconditioner(P::MI355XL1GSPreconditioner, path::AbstractString)
    check(ccall((:hs_l1gs_save, libhsmooth), Cint, (Ptr{Cvoid}, Cstring),
                getfield(P, :handle), path))
    return nothing
end
function load_preconditioner(path::AbstractString; backend::ROCBackend = ROCBackend())
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hs_l1gs_load, libhsmooth), Cint, (Cstring, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                path, stream_handle(), out))
    return MI355XL1GSPreconditioner(out[], backend)
end

# ---------------------------------------------------------------------------
# the device operator (SparseMatrixCSR of chebyshev.jl:144; the CG matrix)
# ---------------------------------------------------------------------------
mutable struct DeviceMatrix
    handle::Ptr{Cvoid}
    n::Int
    owned::Bool
    function DeviceMatrix(h::Ptr{Cvoid}, n::Integer, owned::Bool)
        M = new(h, Int(n), owned)
        owned && finalizer(M) do m
            ccall((:hs_csr_destroy, libhsmooth), Cint, (Ptr{Cvoid},), m.handle)
        end
        return M
    end
end

function DeviceMatrix(A::Union{SparseMatrixCSC{Float64, Int64}, SparseMatrixCSR{1, Float64, Int64}})
    ptr, ind, val, fmt = sparse_arrays(A)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve ptr ind val begin
        check(ccall((:hs_csr_create, libhsmooth), Cint,
            (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
            size(A, 1), ptr, ind, val, fmt, stream_handle(), out))
    end
    return DeviceMatrix(out[], size(A, 1), true)
end

# the operator back as a SparseMatrixCSR{1} (1-based)
function SparseMatricesCSR.SparseMatrixCSR(M::DeviceMatrix)
    info = Vector{Int64}(undef, 2)
    check(ccall((:hs_csr_info, libhsmooth), Cint, (Ptr{Cvoid}, Ptr{Int64}), M.handle, info))
    n, nnz = info[1], info[2]
    rowptr = Vector{Int64}(undef, n + 1)
    colval = Vector{Int64}(undef, nnz)
    nzval = Vector{Float64}(undef, nnz)
    check(ccall((:hs_csr_get, libhsmooth), Cint,
        (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}), M.handle, rowptr, colval, nzval))
    return SparseMatrixCSR{1}(n, n, rowptr, colval, nzval)
end

function LinearAlgebra.mul!(y::Vector{Float64}, M::DeviceMatrix, x::Vector{Float64})
    GC.@preserve x y begin
        check(ccall((:hs_csr_spmv, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            M.handle, pointer(y), pointer(x), LOC_HOST, C_NULL, Cint(0)))
    end
    return y
end
function LinearAlgebra.mul!(y::ROCVector{Float64}, M::DeviceMatrix, x::ROCVector{Float64})
    GC.@preserve x y begin
        check(ccall((:hs_csr_spmv, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            M.handle, pointer(y), pointer(x), LOC_DEVICE, stream_handle(), Cint(0)))
    end
    return y
end

# ---------------------------------------------------------------------------
# Chebyshev smoothers (src/chebyshev.jl)
# ---------------------------------------------------------------------------
"""
The device smoother behind `_build_smoother(ChebyshevFourth/First(; backend = ROCBackend()), A)`.
Fields of the reference's smoother structs (chebyshev.jl:121-129, :224-234) are
properties: `s.A` (a `SparseMatrixCSR`), `s.l1inv`, `s.ρ`, `s.a` (first kind),
`s.r`, `s.z`, `s.tmp` (as the last `smooth!` left them), `s.degree`, `s.backend`.
"""
mutable struct MI355XChebyshevSmoother
    handle::Ptr{Cvoid}
    n::Int
    kind::Int
    degree::Int
    backend::ROCBackend
end

function build_cheb(backend::ROCBackend, kind::Integer, degree::Integer,
                    A::SparseMatrixCSC{Float64, Int64}, precision::Symbol, power_iters::Integer)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve A begin
        check(ccall((:hs_cheb_build, libhsmooth), Cint,
            (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Cint, Cint, Cint, Cint, Ptr{Cvoid},
             Ref{Ptr{Cvoid}}),
            size(A, 1), A.colptr, A.rowval, A.nzval, Cint(kind), Cint(degree),
            Cint(power_iters), PRECISION[precision], stream_handle(), out))
    end
    s = MI355XChebyshevSmoother(out[], size(A, 1), Int(kind), Int(degree), backend)
    finalizer(s) do x
        ccall((:hs_cheb_destroy, libhsmooth), Cint, (Ptr{Cvoid},), getfield(x, :handle))
    end
    return s
end

# _build_smoother(config, A) (chebyshev.jl:141-156, :258-276); ρ from 50 FP64
# power-iteration steps on the device, as the reference's _l1jacobi_spectral_radius
HS._build_smoother(config::ChebyshevFourth{<:ROCBackend}, A::SparseMatrixCSC{Float64, Int64};
                   precision::Symbol = :fp64, power_iters::Integer = 50) =
    build_cheb(config.backend, 4, config.degree, A, precision, power_iters)
HS._build_smoother(config::ChebyshevFirst{<:ROCBackend}, A::SparseMatrixCSC{Float64, Int64};
                   precision::Symbol = :fp64, power_iters::Integer = 50) =
    build_cheb(config.backend, 1, config.degree, A, precision, power_iters)

function cheb_info(s::MI355XChebyshevSmoother)
    out = Vector{Float64}(undef, 5)
    check(ccall((:hs_cheb_info, libhsmooth), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                getfield(s, :handle), out))
    return out  # rho, a*, degree, kind, n
end
function cheb_vector(s::MI355XChebyshevSmoother, what::Integer)
    out = Vector{Float64}(undef, getfield(s, :n))
    check(ccall((:hs_cheb_get, libhsmooth), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                getfield(s, :handle), Cint(what), out))
    return out
end

function Base.getproperty(s::MI355XChebyshevSmoother, f::Symbol)
    (f === :ρ || f === :rho) && return cheb_info(s)[1]
    f === :a && return cheb_info(s)[2]
    f === :l1inv && return cheb_vector(s, 0)
    f === :r && return cheb_vector(s, 1)
    f === :z && return cheb_vector(s, 2)
    f === :tmp && return cheb_vector(s, 3)
    if f === :A
        h = ccall((:hs_cheb_matrix, libhsmooth), Ptr{Cvoid}, (Ptr{Cvoid},), getfield(s, :handle))
        return SparseMatrixCSR(DeviceMatrix(h, getfield(s, :n), false))  # owned by s
    end
    return getfield(s, f)
end

# smooth!(x, s, b): x updated in place (chebyshev.jl:158-183, :278-322)
function HS.smooth!(x::Vector{Float64}, s::MI355XChebyshevSmoother, b::Vector{Float64})
    GC.@preserve x b begin
        check(ccall((:hs_cheb_smooth, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(s, :handle), pointer(x), pointer(b), LOC_HOST, C_NULL, Cint(0)))
    end
    return x
end
function HS.smooth!(x::ROCVector{Float64}, s::MI355XChebyshevSmoother, b::ROCVector{Float64})
    GC.@preserve x b begin
        check(ccall((:hs_cheb_smooth, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(s, :handle), pointer(x), pointer(b), LOC_DEVICE, stream_handle(), Cint(0)))
    end
    return x
end

# the smoother as a Krylov left preconditioner: ldiv!(z, S, r) := (z .= 0; smooth!(z, S, r))
function LinearSolve.ldiv!(z::Vector{Float64}, s::MI355XChebyshevSmoother, r::Vector{Float64})
    GC.@preserve z r begin
        check(ccall((:hs_cheb_apply, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(s, :handle), pointer(z), pointer(r), LOC_HOST, C_NULL, Cint(0)))
    end
    return nothing
end
function LinearSolve.ldiv!(z::ROCVector{Float64}, s::MI355XChebyshevSmoother,
                           r::ROCVector{Float64})
    GC.@preserve z r begin
        check(ccall((:hs_cheb_apply, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(s, :handle), pointer(z), pointer(r), LOC_DEVICE, stream_handle(), Cint(0)))
    end
    return nothing
end

# ---------------------------------------------------------------------------
# the whole PCG solve on the device (solve(LinearProblem(A, b), KrylovJL_CG(); Pl))
# ---------------------------------------------------------------------------
"Statistics of one device solve (`hs_pcg_stats`)."
struct PCGStats
    iters::Int64
    status::Int64       # 0 solved, 1 iteration limit, 2 breakdown
    rnorm0::Cdouble     # sqrt(r0'z0)
    rnorm::Cdouble      # final sqrt(r'z)
    true_relres::Cdouble
    seconds::Cdouble
end

prec_args(::Nothing) = (PCG_NONE, C_NULL)
prec_args(P::MI355XL1GSPreconditioner) = (PCG_L1GS, getfield(P, :handle))
prec_args(S::MI355XChebyshevSmoother) = (PCG_CHEB, getfield(S, :handle))

"""
    pcg_solve(M::DeviceMatrix, b; Pl = nothing, reltol = 1e-8, abstol = 0.0, maxiters = -1,
              true_reltol = 0.0)

FP64 CG from x0 = 0 with Krylov.jl's stopping rule sqrt(r'z) ≤ abstol + reltol·sqrt(r0'z0),
the preconditioner applied on the device each iteration; returns `(x, stats)` with the
FP64 true relative residual in `stats`.  `true_reltol > 0`: the recurrence continues past
the rule's stop until ‖b - A x‖/‖b‖ ≤ true_reltol (checked on the device at each stop).
"""
function pcg_solve(M::DeviceMatrix, b::Vector{Float64}; Pl = nothing, reltol = 1e-8,
                   abstol = 0.0, maxiters::Integer = -1, true_reltol = 0.0)
    kind, ph = prec_args(Pl)
    x = similar(b)
    st = Ref{PCGStats}()
    GC.@preserve b x Pl begin
        check(ccall((:hs_pcg_solve_true, libhsmooth), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Int64,
             Cdouble, Cint, Ptr{Cvoid}, Ref{PCGStats}),
            M.handle, kind, ph, pointer(b), pointer(x), Float64(reltol), Float64(abstol),
            Int64(maxiters), Float64(true_reltol), LOC_HOST, C_NULL, st))
    end
    return x, st[]
end
function pcg_solve(M::DeviceMatrix, b::ROCVector{Float64}; Pl = nothing, reltol = 1e-8,
                   abstol = 0.0, maxiters::Integer = -1, true_reltol = 0.0)
    kind, ph = prec_args(Pl)
    x = similar(b)
    st = Ref{PCGStats}()
    GC.@preserve b x Pl begin
        check(ccall((:hs_pcg_solve_true, libhsmooth), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Int64,
             Cdouble, Cint, Ptr{Cvoid}, Ref{PCGStats}),
            M.handle, kind, ph, pointer(b), pointer(x), Float64(reltol), Float64(abstol),
            Int64(maxiters), Float64(true_reltol), LOC_DEVICE, stream_handle(), st))
    end
    return x, st[]
end

# ---------------------------------------------------------------------------
# AMG smoother glue (ext/HybridSmoothersAMGExt.jl:15-35)
# ---------------------------------------------------------------------------
"""
    amg_smooth!(x, M, P, b; iters = 2)

`AMG.smooth!(x, s::HybridSmoother, b)` with `s.A` as a `DeviceMatrix` and `s.s` an
L1GS preconditioner: `iters` rounds of x -= P⁻¹(A x - b), one fused launch per round for
packed triangles with partsize ≤ 64.
"""
function amg_smooth!(x::Vector{Float64}, M::DeviceMatrix, P::MI355XL1GSPreconditioner,
                     b::Vector{Float64}; iters::Integer = 2)
    GC.@preserve x b begin
        check(ccall((:hs_l1gs_smooth, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cint, Ptr{Cvoid}, Cint),
            getfield(P, :handle), M.handle, pointer(x), pointer(b), Cint(iters), LOC_HOST,
            C_NULL, Cint(0)))
    end
    return x
end
function amg_smooth!(x::ROCVector{Float64}, M::DeviceMatrix, P::MI355XL1GSPreconditioner,
                     b::ROCVector{Float64}; iters::Integer = 2)
    GC.@preserve x b begin
        check(ccall((:hs_l1gs_smooth, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cint, Ptr{Cvoid}, Cint),
            getfield(P, :handle), M.handle, pointer(x), pointer(b), Cint(iters), LOC_DEVICE,
            stream_handle(), Cint(0)))
    end
    return x
end

"""
    amg_smooth_noalloc!(x, M, P, b; iters = 2)

`amg_smooth!` on device vectors that allocates no device memory (for code under
HIP graph capture): the residual and correction live in the library's fixed
array of 2 × 2,097,152 values per GPU, so larger systems are refused; calls that
overlap on one GPU race.  No reference counterpart.
"""
function amg_smooth_noalloc!(x::ROCVector{Float64}, M::DeviceMatrix,
                             P::MI355XL1GSPreconditioner, b::ROCVector{Float64};
                             iters::Integer = 2)
    GC.@preserve x b begin
        check(ccall((:hs_l1gs_smooth_noalloc, libhsmooth), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}, Cint),
            getfield(P, :handle), M.handle, pointer(x), pointer(b), Cint(iters),
            stream_handle(), Cint(0)))
    end
    return x
end

# ---------------------------------------------------------------------------
# batched persistent PCG and serving (no reference counterpart)
# ---------------------------------------------------------------------------
"Engine configuration (`hs_batch_config`)."
struct BatchConfig
    arena_bytes::Int64
    max_systems::Int64
    staging_bytes::Int64
    partsize::Int64
    precision::Cint
    passes::Cint
    grid::Cint
    eta::Cdouble
    triangles::Cint
    smoother::Cint
    degree::Cint
    power_iters::Cint
end

"One collected solve (`hs_batch_result`)."
struct BatchResult
    id::Int64
    n::Int64
    iters::Int64
    status::Int64
    rnorm0::Cdouble
    rnorm::Cdouble
    true_relres::Cdouble
    t_submit::Cdouble
    t_start::Cdouble
    t_done::Cdouble
end

"""
    BatchEngine(; partsize = 32, precision = :fp32, arena_bytes = 32 << 30, max_systems = 1024,
                passes = 8, smoother = :l1gs, degree = 2, power_iters = 50)

Many SPD systems resident on the device and solved together, each with its own
L1GS (symmetric, isSymA) or ChebyshevFourth(degree) preconditioner and the
arithmetic of `pcg_solve`.  `submit!` returns an id, or `nothing` when the
engine is full right now; `step!` runs `passes` iterations of every active
system; `collect!` returns finished solves (their x written to the vector given
at submission, which must stay alive until then).
"""
mutable struct BatchEngine
    handle::Ptr{Cvoid}
    outputs::Dict{Int64, Vector{Float64}}  # x vectors of resident systems, kept alive
end

function BatchEngine(; partsize::Integer = 32, precision::Symbol = :fp32,
                     arena_bytes::Integer = Int64(32) << 30, max_systems::Integer = 1024,
                     staging_bytes::Integer = Int64(256) << 20, passes::Integer = 8,
                     grid::Integer = 0, eta = 1.5, triangles::Integer = 0,
                     smoother::Symbol = :l1gs, degree::Integer = 2, power_iters::Integer = 50)
    cfg = Ref(BatchConfig(arena_bytes, max_systems, staging_bytes, partsize,
                          PRECISION[precision], passes, grid, eta, triangles,
                          smoother === :cheb ? PCG_CHEB : PCG_L1GS, degree, power_iters))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hs_batch_create, libhsmooth), Cint, (Ref{BatchConfig}, Ref{Ptr{Cvoid}}),
                cfg, out))
    e = BatchEngine(out[], Dict{Int64, Vector{Float64}}())
    finalizer(e) do x
        ccall((:hs_batch_destroy, libhsmooth), Cint, (Ptr{Cvoid},), x.handle)
    end
    return e
end

# A symmetric positive definite, so the arrays of its CSC are those of its CSR
function submit!(e::BatchEngine, A::SparseMatrixCSC{Float64, Int64}, b::Vector{Float64};
                 rtol = 1e-8, atol = 0.0, itmax::Integer = -1)
    n = size(A, 1)
    x = Vector{Float64}(undef, n)
    id = Ref{Int64}(-1)
    st = GC.@preserve A b x ccall((:hs_batch_submit, libhsmooth), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Cdouble, Cdouble, Int64, Ref{Int64}),
        e.handle, n, A.colptr, A.rowval, A.nzval, b, x, Float64(rtol), Float64(atol),
        Int64(itmax), id)
    st == HS_EAGAIN && return nothing
    check(st)
    e.outputs[id[]] = x
    return id[]
end

step!(e::BatchEngine) = check(ccall((:hs_batch_step, libhsmooth), Cint, (Ptr{Cvoid},), e.handle))

"Finished solves as (result, x) pairs."
function collect!(e::BatchEngine; max::Integer = 1024)
    out = Vector{BatchResult}(undef, max)
    count = Ref{Int64}(0)
    check(ccall((:hs_batch_collect, libhsmooth), Cint,
        (Ptr{Cvoid}, Ptr{BatchResult}, Int64, Ref{Int64}), e.handle, out, Int64(max), count))
    return [(r, pop!(e.outputs, r.id)) for r in out[1:count[]]]
end

function info(e::BatchEngine)
    out = Vector{Int64}(undef, 12)
    check(ccall((:hs_batch_info, libhsmooth), Cint, (Ptr{Cvoid}, Ptr{Int64}), e.handle, out))
    return out
end

end # module
