# OxiBLASMI355X -- drop-in GPU backend for OxiBLAS.jl's level-3 hot path.
#
# Keeps the reference's API (gemm!, lu, lu!, LU <: Factorization, ldiv!,
# ldiv, \, solve, solve_transposed) and calls liboxiblas_mi355x.so, the
# hand-written HIP/CDNA4 kernels, through `ccall` on the C ABI declared in
# include/oxiblas_mi355x.h.  Matrices live in device memory as `DeviceMatrix`
# (column-major Float64 with a leading dimension, as the reference's
# Matrix{Float64}); host `Matrix{Float64}` arguments are uploaded, computed on
# the GPU and copied back, so existing call sites keep working unchanged.
module OxiBLASMI355X

using LinearAlgebra
import LinearAlgebra: ldiv!, issuccess, Factorization, SingularException
import ChainRulesCore
import ChainRulesCore: rrule, NoTangent, unthunk, @thunk

export DeviceMatrix, gemm!, lu, lu!, ldiv, solve, solve_transposed, mixed_solve, equilibrate

const LIB = get(ENV, "OXIBLAS_MI355X_LIB",
                normpath(joinpath(@__DIR__, "..", "..", "..", "liboxiblas_mi355x.so")))

# ---------------------------------------------------------------------------
# handle (one per device, like the reference's GemmWorkspace pool it replaces:
# src/blas/gemm.jl:107-153)
# ---------------------------------------------------------------------------
const _HANDLES = Dict{Int,Ptr{Cvoid}}()
const _LOCK = ReentrantLock()

function handle(device::Integer=0)
    lock(_LOCK) do
        get!(_HANDLES, Int(device)) do
            h = Ref{Ptr{Cvoid}}(C_NULL)
            rc = ccall((:oxi_create, LIB), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, h)
            rc == 0 || error("oxi_create failed ($rc)")
            h[]
        end
    end
end

lasterror(h) = unsafe_string(ccall((:oxi_last_error, LIB), Cstring, (Ptr{Cvoid},), h))

# Return codes: negative = invalid argument (nothing written), positive = runtime failure.
function _check(h, rc, what)
    rc == 0 && return nothing
    rc < 0 && throw(ArgumentError("$what: invalid argument $(-rc)"))
    error("$what: $(lasterror(h)) (code $rc)")
end

# ---------------------------------------------------------------------------
# device matrices
# ---------------------------------------------------------------------------
mutable struct DeviceMatrix <: AbstractMatrix{Float64}
    ptr::Ptr{Float64}
    m::Int
    n::Int
    ld::Int
    owner::Bool
    function DeviceMatrix(m::Integer, n::Integer; ld::Integer=max(1, m))
        h = handle()
        p = Ref{Ptr{Cvoid}}(C_NULL)
        _check(h, ccall((:oxi_malloc, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, UInt64), h, p, 8 * ld * n), "malloc")
        D = new(Ptr{Float64}(p[]), m, n, ld, true)
        finalizer(D) do d
            d.owner && d.ptr != C_NULL && ccall((:oxi_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), handle(), d.ptr)
        end
        D
    end
    DeviceMatrix(ptr::Ptr{Float64}, m, n, ld) = new(ptr, m, n, ld, false)   # a view, not owned
end

Base.size(D::DeviceMatrix) = (D.m, D.n)
Base.getindex(::DeviceMatrix, ::Int, ::Int) = error("DeviceMatrix: copy to the host with Array(D)")

function DeviceMatrix(A::StridedMatrix{Float64})
    m, n = size(A)
    D = DeviceMatrix(m, n)
    H = Matrix{Float64}(A)   # contiguous host copy
    h = handle()
    _check(h, ccall((:oxi_memcpy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, UInt64, Cint),
                   h, D.ptr, H, 8 * m * n, 0), "memcpy")
    D
end

function Base.Array(D::DeviceMatrix)
    H = Matrix{Float64}(undef, D.ld, D.n)
    h = handle()
    _check(h, ccall((:oxi_memcpy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, UInt64, Cint),
                   h, H, D.ptr, 8 * D.ld * D.n, 1), "memcpy")
    H[1:D.m, :]
end

# contiguous sub-block view D[rows, cols]
function Base.view(D::DeviceMatrix, rows::UnitRange{Int}, cols::UnitRange{Int})
    DeviceMatrix(D.ptr + 8 * ((first(rows) - 1) + (first(cols) - 1) * D.ld), length(rows), length(cols), D.ld)
end

_trans(t) = t in (:N, 'N') ? 'N' : t in (:T, 'T') ? 'T' : t in (:C, 'C') ? 'C' :
            throw(ArgumentError("transpose must be N, T or C"))   # src/blas/gemm.jl:156-161

# ---------------------------------------------------------------------------
# gemm!  (src/blas/gemm.jl:625-703)
# ---------------------------------------------------------------------------
"""
    gemm!(C, A, B, α=1, β=0; transA=:N, transB=:N)

C ← α op(A) op(B) + β C on the MI355X (FP64 MFMA).  β=0 never reads C; α=0 or
K=0 never reads A/B; an operand overlapping C is snapshotted first.
"""
function gemm!(C::DeviceMatrix, A::DeviceMatrix, B::DeviceMatrix, α=1.0, β=0.0; transA=:N, transB=:N)
    ta, tb = _trans(transA), _trans(transB)
    am, ak = ta == 'N' ? (A.m, A.n) : (A.n, A.m)
    bk, bn = tb == 'N' ? (B.m, B.n) : (B.n, B.m)
    ak == bk && (C.m, C.n) == (am, bn) ||
        throw(DimensionMismatch("GEMM: C$(size(C)), A$(size(A)), B$(size(B))"))
    h = handle()
    rc = ccall((:oxi_dgemm, LIB), Cint,
               (Ptr{Cvoid}, Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                Float64, Ptr{Float64}, Int64),
               h, ta, tb, C.m, C.n, ak, Float64(α), A.ptr, A.ld, B.ptr, B.ld, Float64(β), C.ptr, C.ld)
    _check(h, rc, "gemm!")
    C
end

function gemm!(C::StridedMatrix{Float64}, A::StridedMatrix{Float64}, B::StridedMatrix{Float64}, α=1.0, β=0.0;
               transA=:N, transB=:N)
    Cd = DeviceMatrix(C)
    gemm!(Cd, DeviceMatrix(A), DeviceMatrix(B), α, β; transA, transB)
    copyto!(C, Array(Cd))
end

# ---------------------------------------------------------------------------
# LU  (src/api/linalg_api.jl:145-218, src/lapack/factorizations.jl:5-77)
# ---------------------------------------------------------------------------
"""
    LU <: Factorization{Float64}

Dense LU with row pivoting whose factors and pivots stay on the GPU:
`F.L*F.U == A[F.p,:]`.
"""
mutable struct LU <: Factorization{Float64}
    factors::DeviceMatrix
    ipiv::Ptr{Int64}          # device, 1-based global rows
    info::Int
end

Base.size(F::LU) = size(F.factors)
issuccess(F::LU) = F.info == 0

function _pivots(F::LU)
    q = min(size(F)...)
    v = Vector{Int64}(undef, q)
    h = handle()
    _check(h, ccall((:oxi_memcpy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, UInt64, Cint),
                   h, v, getfield(F, :ipiv), 8 * q, 1), "memcpy")
    v
end

function Base.getproperty(F::LU, s::Symbol)
    s === :ipiv && return _pivots(F)
    if s === :p
        m = getfield(F, :factors).m
        perm = collect(1:m)
        for (j, p) in enumerate(_pivots(F))
            perm[j], perm[p] = perm[p], perm[j]
        end
        return perm
    end
    if s === :L || s === :U
        a = Array(getfield(F, :factors)); m, n = size(a)
        return s === :L ? (tril(a[:, 1:min(m, n)], -1) + Matrix(1.0I, m, min(m, n))) : triu(a[1:min(m, n), :])
    end
    getfield(F, s)
end
Base.propertynames(::LU, private::Bool=false) = (:L, :U, :p, (private ? (:factors, :ipiv, :info) : ())...)

function _getrf!(A::DeviceMatrix; check::Bool=true, blocksize::Integer=0)
    h = handle()
    q = min(A.m, A.n)
    piv = Ref{Ptr{Cvoid}}(C_NULL)
    _check(h, ccall((:oxi_malloc, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, UInt64), h, piv, 8 * max(q, 1)), "malloc")
    info = Ref{Int64}(0)
    rc = ccall((:oxi_dgetrf, LIB), Cint,
               (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Int64}),
               h, A.m, A.n, A.ptr, A.ld, Ptr{Int64}(piv[]), blocksize, info)
    _check(h, rc, "lu")
    F = LU(A, Ptr{Int64}(piv[]), Int(info[]))
    finalizer(f -> ccall((:oxi_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), handle(), getfield(f, :ipiv)), F)
    check && info[] != 0 && throw(SingularException(Int(info[])))
    F
end

"""
    lu(A; check=true, blocksize=0) -> LU

Partial-pivoted blocked LU of a copy of A (OxiBLAS.lu, src/api/linalg_api.jl:217).
"""
lu(A::StridedMatrix{Float64}; check::Bool=true, blocksize::Integer=0) = _getrf!(DeviceMatrix(A); check, blocksize)
lu(A::DeviceMatrix; check::Bool=true, blocksize::Integer=0) = _getrf!(_copy(A); check, blocksize)

"""
    lu!(A::DeviceMatrix; check=true) -> LU

In-place LU: A becomes `F.factors` (OxiBLAS.lu!, src/api/linalg_api.jl:218).
"""
lu!(A::DeviceMatrix; check::Bool=true, blocksize::Integer=0) = _getrf!(A; check, blocksize)

function _copy(A::DeviceMatrix)
    D = DeviceMatrix(A.m, A.n)
    # D <- 1*A + 0*D through gemm! with an identity would cost a GEMM; copy on the device instead
    h = handle()
    for j in 0:A.n-1
        _check(h, ccall((:oxi_memcpy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, UInt64, Cint),
                       h, D.ptr + 8 * j * D.ld, A.ptr + 8 * j * A.ld, 8 * A.m, 2), "memcpy")
    end
    D
end

function _ready(F::LU)
    m, n = size(F)
    m == n || throw(DimensionMismatch("matrix is not square"))
    F.info == 0 || throw(SingularException(F.info))      # src/api/linalg_api.jl:163-167
end

function _getrs!(F::LU, B::DeviceMatrix, trans::Char)
    _ready(F)
    B.m == F.factors.m || throw(DimensionMismatch("factorization has $(F.factors.m) rows, right-hand side has $(B.m)"))
    h = handle()
    rc = ccall((:oxi_dgetrs, LIB), Cint,
               (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Float64}, Int64),
               h, trans, F.factors.m, B.n, F.factors.ptr, F.factors.ld, getfield(F, :ipiv), B.ptr, B.ld)
    _check(h, rc, "ldiv!")
    B
end

# ldiv!(F, B) and its transpose/adjoint (src/api/linalg_api.jl:168-184)
ldiv!(F::LU, B::DeviceMatrix) = _getrs!(F, B, 'N')
ldiv!(F::LinearAlgebra.TransposeFactorization{Float64,LU}, B::DeviceMatrix) = _getrs!(parent(F), B, 'T')
ldiv!(F::LinearAlgebra.AdjointFactorization{Float64,LU}, B::DeviceMatrix) = _getrs!(parent(F), B, 'C')
function ldiv!(F::Union{LU,LinearAlgebra.TransposeFactorization{Float64,LU},LinearAlgebra.AdjointFactorization{Float64,LU}},
               B::StridedVecOrMat{Float64})
    Bd = DeviceMatrix(reshape(B, size(B, 1), :))
    ldiv!(F, Bd)
    copyto!(B, vec(Array(Bd)))
end

"""
    ldiv(F, B) -> X

A new solution of F \\ B (OxiBLAS.ldiv, src/api/linalg_api_solve.jl:188-199).
"""
ldiv(F, B::StridedVecOrMat{Float64}) = ldiv!(F, copy(B))
ldiv(F, B::DeviceMatrix) = ldiv!(F, _copy(B))
Base.:\(F::LU, B::StridedVecOrMat{Float64}) = ldiv(F, B)
Base.:\(F::LU, B::DeviceMatrix) = ldiv(F, B)

solve(F::LU, b) = F \ b                                                 # src/lapack/factorizations.jl:177-193
function solve_transposed(F::LU, b; trans=:C)                           # src/lapack/expert_refinement.jl:2-23
    t = _trans(trans)
    t == 'N' && return F \ b
    ldiv(t == 'T' ? transpose(F) : adjoint(F), b)
end

function LinearAlgebra.logabsdet(F::LU)                                 # src/api/linalg_api.jl:185-191
    _ready(F) === nothing
    d = diag(Array(F.factors)); p = F.p
    s = prod(sign, d)
    cycles = 0; seen = falses(length(p))
    for i in eachindex(p)
        seen[i] && continue
        cycles += 1; j = i
        while !seen[j]; seen[j] = true; j = p[j]; end
    end
    (sum(log ∘ abs, d), isodd(length(p) - cycles) ? -s : s)
end
LinearAlgebra.det(F::LU) = F.info != 0 ? 0.0 : (la = logabsdet(F); la[2] * exp(la[1]))

# ---------------------------------------------------------------------------
# adjoints: x = F \ b reuses the stored factors (B̄ = A⁻ᵀ X̄ by the transposed
# solve, Ā = -B̄ Xᵀ by gemm!), and gemm!'s product rule
# ---------------------------------------------------------------------------
function rrule(::typeof(\), F::LU, B::StridedVecOrMat{Float64})
    X = F \ B
    function ldiv_pullback(X̄)
        B̄ = solve_transposed(F, Matrix{Float64}(reshape(unthunk(X̄), size(B, 1), :)); trans=:T)
        Ā = zeros(size(F)...)
        gemm!(Ā, B̄, reshape(X, size(B, 1), :), -1.0, 0.0; transB=:T)
        (NoTangent(), ChainRulesCore.Tangent{LU}(; factors=Ā), reshape(B̄, size(B)))
    end
    X, ldiv_pullback
end

# x = A \ B on a plain matrix: one factorization kept in HBM, one fused
# pullback call (transposed solve + rank-nrhs GEMM) through oxi_dgesv_adjoint.
function rrule(::typeof(\), A::StridedMatrix{Float64}, B::StridedVecOrMat{Float64})
    F = lu(A)
    n, k = size(A, 1), size(B, 2)
    Xd = DeviceMatrix(Matrix{Float64}(reshape(B, n, k)))
    _getrs!(F, Xd, 'N')
    X = Array(Xd)
    function ldiv_pullback(X̄)
        X̄d = DeviceMatrix(Matrix{Float64}(reshape(unthunk(X̄), n, k)))
        B̄d = DeviceMatrix(n, k); Ād = DeviceMatrix(n, n)
        h = handle()
        rc = ccall((:oxi_dgesv_adjoint, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Float64}, Int64, Ptr{Float64},
                    Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
                   h, n, k, F.factors.ptr, F.factors.ld, getfield(F, :ipiv), Xd.ptr, Xd.ld, X̄d.ptr, X̄d.ld,
                   B̄d.ptr, B̄d.ld, Ād.ptr, Ād.ld)
        _check(h, rc, "ldiv pullback")
        (NoTangent(), Array(Ād), reshape(Array(B̄d), size(B)))
    end
    reshape(X, size(B)), ldiv_pullback
end

function rrule(::typeof(*), A::StridedMatrix{Float64}, B::StridedMatrix{Float64})
    C = zeros(size(A, 1), size(B, 2)); gemm!(C, A, B)
    function mul_pullback(C̄)
        c = Matrix{Float64}(unthunk(C̄))
        Ā = @thunk(gemm!(zeros(size(A)), c, B, 1.0, 0.0; transB=:T))
        B̄ = @thunk(gemm!(zeros(size(B)), A, c, 1.0, 0.0; transA=:T))
        (NoTangent(), Ā, B̄)
    end
    C, mul_pullback
end

# ---------------------------------------------------------------------------
# mixed precision (src/lapack/extended_lapack.jl:175-226) and equilibration
# (src/lapack/dense_utilities.jl:191-212)
# ---------------------------------------------------------------------------
struct IRInfo
    iterations::Int64
    inner_iterations::Int64
    converged::Int32
    fallback::Int32
    reason::Int32
    solver::Int32
    backward_error::Float64
    normwise_backward_error::Float64
    residual_norm::Float64
    factor_ms::Float64
    refine_ms::Float64
end
IRInfo() = IRInfo(0, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0)

# low type -> OXI_PREC_*: Float32 (the reference's default), Float16, and the
# names of types Base lacks
function _prec(lowtype)
    lowtype === Float32 && return Cint(3)
    lowtype === Float16 && return Cint(1)
    name = lowercase(string(lowtype))
    occursin("bfloat16", name) && return Cint(0)
    (occursin("float8", name) || occursin("e4m3", name)) && return Cint(2)
    throw(ArgumentError("low precision type must be Float32, Float16, BFloat16 or Float8_E4M3FN"))
end
const _REASONS = (:converged, :iteration_limit, :nonfinite_iteration, :working_factor_fallback)

"""
    mixed_solve(A, b; maxiter=10, rtol=1e-13, lowtype=Float32, solver=:ir, gmres_restart=0)

Radix equilibration, a low-precision LU on the GPU (FP32 panel; trailing
products in `lowtype`) and FP64 refinement until the componentwise backward
error is <= rtol; `solver=:gmres` runs GMRES-IR, `:auto` switches to it when
classical refinement stalls.  Returns the reference's NamedTuple.
"""
function mixed_solve(A::StridedMatrix{Float64}, b::StridedVecOrMat{Float64}; maxiter::Integer=10,
                     rtol::Real=1e-13, lowtype=Float32, solver::Symbol=:ir, gmres_restart::Integer=0,
                     factorization::Symbol=:lu, pivot::Symbol=:partial)
    factorization === :lu || throw(ArgumentError("the GPU mixed solve implements factorization=:lu"))
    n = size(A, 1); size(A, 2) == n || throw(DimensionMismatch("mixed factorization dimensions differ"))
    B = reshape(b, size(b, 1), :); k = size(B, 2)
    Ad = DeviceMatrix(Matrix{Float64}(A)); Bd = DeviceMatrix(Matrix{Float64}(B)); Xd = DeviceMatrix(n, k)
    info = Ref(IRInfo())
    h = handle()
    rc = ccall((:oxi_dgesv_mixed, LIB), Cint,
               (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                Int32, Float64, Int32, Int32, Ref{IRInfo}),
               h, _prec(lowtype), Cint(solver === :ir ? 0 : solver === :gmres ? 1 : 2), n, k, Ad.ptr, Ad.ld, Bd.ptr,
               Bd.ld, Xd.ptr, Xd.ld, Int32(maxiter), Float64(rtol), Int32(gmres_restart),
               Int32(pivot === :rbt ? 1 : 0), info)
    rc == 6 && throw(ArgumentError("finite mixed input required"))
    _check(h, rc, "mixed_solve")
    r = info[]
    X = Array(Xd)
    (; x=ndims(b) == 1 ? vec(X) : X, converged=r.converged != 0, iterations=Int(r.iterations),
     residual=r.residual_norm, backward_error=r.backward_error, metric=:componentwise_backward_error,
     low_type=lowtype, fallback=r.fallback != 0, factorization=:lu, finite=r.reason != 2,
     reason=_REASONS[r.reason + 1], gmres_iterations=Int(r.inner_iterations))
end

"""
    equilibrate(A; radix=false) -> (; matrix, row_scale, column_scale)
"""
function equilibrate(A::StridedMatrix{Float64}; radix::Bool=false)
    m, n = size(A)
    Ad = DeviceMatrix(Matrix{Float64}(A)); Ed = DeviceMatrix(m, n)
    r = DeviceMatrix(max(m, 1), 1); c = DeviceMatrix(max(n, 1), 1)
    h = handle()
    rc = ccall((:oxi_dequilibrate, LIB), Cint,
               (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Cint),
               h, m, n, Ad.ptr, Ad.ld, Ed.ptr, Ed.ld, r.ptr, c.ptr, Cint(radix))
    _check(h, rc, "equilibrate")
    (; matrix=Array(Ed), row_scale=vec(Array(r))[1:m], column_scale=vec(Array(c))[1:n])
end

end # module
