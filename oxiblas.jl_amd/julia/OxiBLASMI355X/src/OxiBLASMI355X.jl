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
import LinearAlgebra: ldiv!, issuccess, Factorization, SingularException, PosDefException
import ChainRulesCore
import ChainRulesCore: rrule, NoTangent, unthunk, @thunk

export DeviceMatrix, gemm!, lu, lu!, ldiv, solve, solve_transposed, mixed_solve, equilibrate,
       syrk!, herk!, gemm_batched!, gemm_strided_batched!, Cholesky, cholesky, cholesky!, cholesky_factor,
       native_inverse, backward_error, inverse_norm_estimate, reciprocal_condition, refine_solve, expert_solve

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
function lu(A::DeviceMatrix; check::Bool=true, blocksize::Integer=0)
    # the copy is written by the factorization's change of layout (oxi_dgetrf_copy)
    h = handle()
    F = DeviceMatrix(A.m, A.n)
    q = min(A.m, A.n)
    piv = Ref{Ptr{Cvoid}}(C_NULL)
    _check(h, ccall((:oxi_malloc, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, UInt64), h, piv, 8 * max(q, 1)), "malloc")
    info = Ref{Int64}(0)
    rc = ccall((:oxi_dgetrf_copy, LIB), Cint,
               (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Int64}),
               h, A.m, A.n, A.ptr, A.ld, F.ptr, F.ld, Ptr{Int64}(piv[]), blocksize, info)
    _check(h, rc, "lu")
    Fl = LU(F, Ptr{Int64}(piv[]), Int(info[]))
    finalizer(f -> ccall((:oxi_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), handle(), getfield(f, :ipiv)), Fl)
    check && info[] != 0 && throw(SingularException(Int(info[])))
    Fl
end

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

# ldiv!(F, B) and its transpose/adjoint (src/api/linalg_api.jl:168-184); a
# complex right-hand side on the real factors is solved as its real and
# imaginary parts (_apply_split!, :67-78)
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
function ldiv!(F::Union{LU,LinearAlgebra.TransposeFactorization{Float64,LU},LinearAlgebra.AdjointFactorization{Float64,LU}},
               B::StridedVecOrMat{ComplexF64})
    re = ldiv!(F, Matrix{Float64}(real(reshape(B, size(B, 1), :))))
    im_ = ldiv!(F, Matrix{Float64}(imag(reshape(B, size(B, 1), :))))
    copyto!(B, vec(complex.(re, im_)))
end

ldiv(F, B::StridedVecOrMat{Float64}) = ldiv!(F, copy(B))
ldiv(F, B::StridedVecOrMat{ComplexF64}) = ldiv!(F, copy(B))
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
    m, n = size(F); m == n || throw(DimensionMismatch("matrix is not square"))
    la = Ref{Float64}(0); sg = Ref{Float64}(0); h = handle()
    rc = ccall((:oxi_dlogabsdet, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ref{Float64}, Ref{Float64}),
               h, n, F.factors.ptr, F.factors.ld, getfield(F, :ipiv), la, sg)
    _check(h, rc, "logabsdet")
    (la[], sg[])          # a singular factor gives (-Inf, 0.0), no exception
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
classical refinement stalls.  `pivot=:rbt` (random butterflies, no
pivoting) and `pivot=:tournament` (tournament pivoting per 32-column leaf)
have no reference counterpart.  Returns the reference's NamedTuple.
"""
function mixed_solve(A::StridedMatrix{Float64}, b::StridedVecOrMat{Float64}; maxiter::Integer=10,
                     rtol::Real=1e-13, lowtype=Float32, solver::Symbol=:ir, gmres_restart::Integer=0,
                     factorization::Symbol=:lu, pivot::Symbol=:partial)
    factorization === :cholesky && return _mixed_cholesky(A, b; maxiter, rtol, lowtype)
    factorization === :lu || throw(ArgumentError("the GPU mixed solve implements factorization=:lu and :cholesky"))
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
               Int32(pivot === :rbt ? 1 : pivot === :tournament ? 2 : 0), info)
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

# ---------------------------------------------------------------------------
# what is built on the LU factors (src/lapack/dense_utilities.jl:14-36,
# src/api/linalg_api.jl:1101-1143)
# ---------------------------------------------------------------------------
function Base.inv(F::LU)                                                 # native_inverse from the factors
    _ready(F)
    n = size(F, 1); X = DeviceMatrix(n, n); h = handle()
    rc = ccall((:oxi_dgetri, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Float64}, Int64),
               h, n, F.factors.ptr, F.factors.ld, getfield(F, :ipiv), X.ptr, X.ld)
    rc == 4 && throw(SingularException(F.info))
    _check(h, rc, "inv")
    Array(X)
end
native_inverse(A::StridedMatrix{Float64}) = inv(lu(A))
LinearAlgebra.det(A::DeviceMatrix) = det(lu(A; check=false))
LinearAlgebra.logabsdet(A::DeviceMatrix) = logabsdet(lu(A; check=false))

# ---------------------------------------------------------------------------
# syrk!/herk! (src/blas/level3.jl:779-781), batched GEMM (src/blas/batched.jl)
# ---------------------------------------------------------------------------
function _rankk!(C::DeviceMatrix, A::DeviceMatrix, α, β, uplo, t)
    n = C.m; C.n == n || throw(DimensionMismatch("C must be square"))
    k = t == 'N' ? A.n : A.m
    (t == 'N' ? A.m : A.n) == n || throw(DimensionMismatch("rank-k dimensions differ"))
    h = handle()
    rc = ccall((:oxi_dsyrk, LIB), Cint,
               (Ptr{Cvoid}, Cchar, Cchar, Int64, Int64, Float64, Ptr{Float64}, Int64, Float64, Ptr{Float64}, Int64),
               h, uplo === :U ? 'U' : 'L', t, n, k, Float64(α), A.ptr, A.ld, Float64(β), C.ptr, C.ld)
    _check(h, rc, "syrk!")
    C
end
function syrk!(C::DeviceMatrix, A::DeviceMatrix, α=1.0, β=0.0; uplo=:U, trans=:N)
    t = _trans(trans); t == 'C' && throw(ArgumentError("symmetric rank-k accepts N or T"))
    _rankk!(C, A, α, β, uplo, t)
end
function herk!(C::DeviceMatrix, A::DeviceMatrix, α=1.0, β=0.0; uplo=:U, trans=:N)
    t = _trans(trans); t == 'T' && throw(ArgumentError("Hermitian rank-k accepts N or C"))
    _rankk!(C, A, α, β, uplo, t)
end
for f in (:syrk!, :herk!)
    @eval function $f(C::StridedMatrix{Float64}, A::StridedMatrix{Float64}, α=1.0, β=0.0; kwargs...)
        Cd = DeviceMatrix(C); $f(Cd, DeviceMatrix(A), α, β; kwargs...); copyto!(C, Array(Cd))
    end
end

function gemm_strided_batched!(C::Array{Float64,3}, A::Array{Float64,3}, B::Array{Float64,3}, α=1.0, β=0.0;
                               transA=:N, transB=:N)
    ta, tb = _trans(transA), _trans(transB)
    size(C, 3) == size(A, 3) == size(B, 3) || throw(DimensionMismatch("batch counts differ"))
    am, ak = ta == 'N' ? (size(A, 1), size(A, 2)) : (size(A, 2), size(A, 1))
    bk, bn = tb == 'N' ? (size(B, 1), size(B, 2)) : (size(B, 2), size(B, 1))
    ak == bk && (size(C, 1), size(C, 2)) == (am, bn) || throw(DimensionMismatch("batched GEMM dimensions differ"))
    nb = size(C, 3)
    up(X) = DeviceMatrix(reshape(X, size(X, 1), :))            # the batch as one device block
    Ad, Bd, Cd = up(A), up(B), up(C)
    h = handle()
    rc = ccall((:oxi_dgemm_strided_batched, LIB), Cint,
               (Ptr{Cvoid}, Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Int64,
                Int64, Float64, Ptr{Float64}, Int64, Int64, Int64),
               h, ta, tb, size(C, 1), size(C, 2), ak, Float64(α), Ad.ptr, max(1, size(A, 1)), size(A, 1) * size(A, 2),
               Bd.ptr, max(1, size(B, 1)), size(B, 1) * size(B, 2), Float64(β), Cd.ptr, max(1, size(C, 1)),
               size(C, 1) * size(C, 2), nb)
    _check(h, rc, "gemm_strided_batched!")
    copyto!(C, reshape(Array(Cd), size(C)))
end

function gemm_batched!(Cs::Vector{DeviceMatrix}, As::Vector{DeviceMatrix}, Bs::Vector{DeviceMatrix}, α=1.0, β=0.0;
                       transA=:N, transB=:N)
    ta, tb = _trans(transA), _trans(transB)
    length(Cs) == length(As) == length(Bs) || throw(DimensionMismatch("batch counts differ"))
    groups = Dict{NTuple{6,Int},Vector{Int}}()
    for i in eachindex(Cs)
        A, B, C = As[i], Bs[i], Cs[i]
        ak = ta == 'N' ? A.n : A.m
        push!(get!(groups, (C.m, C.n, ak, A.ld, B.ld, C.ld), Int[]), i)
    end
    h = handle()
    for ((m, n, k, lda, ldb, ldc), idx) in groups                 # one launch per shape
        pa = [As[i].ptr for i in idx]; pb = [Bs[i].ptr for i in idx]; pc = [Cs[i].ptr for i in idx]
        rc = ccall((:oxi_dgemm_batched, LIB), Cint,
                   (Ptr{Cvoid}, Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Ptr{Float64}}, Int64, Ptr{Ptr{Float64}},
                    Int64, Float64, Ptr{Ptr{Float64}}, Int64, Int64),
                   h, ta, tb, m, n, k, Float64(α), pa, lda, pb, ldb, Float64(β), pc, ldc, length(idx))
        _check(h, rc, "gemm_batched!")
    end
    Cs
end

# ---------------------------------------------------------------------------
# Cholesky (src/lapack/factorizations.jl:206-300, src/api/linalg_api.jl:221-296)
# ---------------------------------------------------------------------------
mutable struct Cholesky <: Factorization{Float64}
    factors::DeviceMatrix      # L in the lower triangle, zero above
    info::Int
end
Base.size(F::Cholesky) = size(F.factors)
issuccess(F::Cholesky) = F.info == 0
function Base.getproperty(F::Cholesky, s::Symbol)
    (s === :L || s === :UL) && return LowerTriangular(Array(getfield(F, :factors)))
    s === :U && return UpperTriangular(Matrix(transpose(Array(getfield(F, :factors)))))
    s === :uplo && return 'L'
    getfield(F, s)
end
function _potrf!(A::DeviceMatrix, check::Bool)
    A.m == A.n || throw(DimensionMismatch("matrix is not square"))
    info = Ref{Int64}(0); h = handle()
    rc = ccall((:oxi_dpotrf, LIB), Cint, (Ptr{Cvoid}, Cchar, Int64, Ptr{Float64}, Int64, Ref{Int64}),
               h, 'L', A.n, A.ptr, A.ld, info)
    _check(h, rc, "cholesky")
    check && info[] != 0 && throw(PosDefException(Int(info[])))
    Cholesky(A, Int(info[]))
end
function cholesky(A::StridedMatrix{Float64}; check::Bool=true)
    LinearAlgebra.checksquare(A)
    if !ishermitian(A)
        check && throw(PosDefException(-1))
        return Cholesky(DeviceMatrix(A), -1)
    end
    _potrf!(DeviceMatrix(A), check)
end
cholesky(A::DeviceMatrix; check::Bool=true) = _potrf!(_copy(A), check)
cholesky!(A::DeviceMatrix; check::Bool=true) = _potrf!(A, check)
cholesky_factor(A::StridedMatrix{Float64}) = _potrf!(DeviceMatrix(A), true)
function ldiv!(F::Cholesky, B::DeviceMatrix)
    F.info == 0 || throw(PosDefException(F.info))
    B.m == F.factors.m || throw(DimensionMismatch("Cholesky solve dimensions differ"))
    h = handle()
    rc = ccall((:oxi_dpotrs, LIB), Cint, (Ptr{Cvoid}, Cchar, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
               h, 'L', F.factors.m, B.n, F.factors.ptr, F.factors.ld, B.ptr, B.ld)
    _check(h, rc, "ldiv!")
    B
end
function ldiv!(F::Cholesky, B::StridedVecOrMat{Float64})
    Bd = DeviceMatrix(reshape(B, size(B, 1), :)); ldiv!(F, Bd); copyto!(B, vec(Array(Bd)))
end
Base.:\(F::Cholesky, B::StridedVecOrMat{Float64}) = ldiv!(F, copy(B))
solve(F::Cholesky, b) = F \ b
function LinearAlgebra.logabsdet(F::Cholesky)                          # linalg_api.jl:247-251
    F.info == 0 || throw(PosDefException(F.info))
    (2 * sum(log, diag(Array(F.factors))), 1.0)
end

# ---------------------------------------------------------------------------
# expert drivers and estimates (src/lapack/expert_refinement.jl,
# src/lapack/advanced_factorizations.jl:591-645, dense_utilities.jl:214-228)
# ---------------------------------------------------------------------------
struct RefineInfo
    iterations::Int64
    converged::Int32
    reserved::Int32
    backward_error::Float64
    residual_norm::Float64
    rcond_estimate::Float64
    forward_error_estimate::Float64
end
RefineInfo() = RefineInfo(0, 0, 0, 0.0, 0.0, 0.0, 0.0)

function backward_error(A::StridedMatrix{Float64}, x::StridedVecOrMat{Float64}, b::StridedVecOrMat{Float64})
    X = reshape(x, size(x, 1), :); B = reshape(b, size(b, 1), :)
    size(A, 2) == size(X, 1) && size(B) == (size(A, 1), size(X, 2)) || throw(DimensionMismatch("residual dimensions differ"))
    Ad, Xd, Bd = DeviceMatrix(Matrix(A)), DeviceMatrix(Matrix(X)), DeviceMatrix(Matrix(B))
    Rd = DeviceMatrix(size(B)...); e = Ref{Float64}(0); h = handle()
    rc = ccall((:oxi_dbackward_error, LIB), Cint,
               (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                Ptr{Float64}, Int64, Ref{Float64}),
               h, size(A, 1), size(A, 2), size(B, 2), Ad.ptr, Ad.ld, Xd.ptr, Xd.ld, Bd.ptr, Bd.ld, Rd.ptr, Rd.ld, e)
    _check(h, rc, "backward_error")
    R = Array(Rd)
    (; error=e[], residual=ndims(b) == 1 ? vec(R) : R)
end

function inverse_norm_estimate(F::LU, ::Type{Float64}, n::Integer; maxiter=8)
    n == size(F, 1) || throw(DimensionMismatch("n differs from the factorization"))
    est = Ref{Float64}(0); h = handle()
    rc = ccall((:oxi_dinverse_norm_estimate, LIB), Cint,
               (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Int32, Ref{Float64}),
               h, n, F.factors.ptr, F.factors.ld, getfield(F, :ipiv), Int32(maxiter), est)
    _check(h, rc, "inverse_norm_estimate")
    est[]
end
function reciprocal_condition(A::StridedMatrix{Float64}; maxiter=8)
    n = LinearAlgebra.checksquare(A); n == 0 && return 1.0
    F = lu(A; check=false); F.info != 0 && return 0.0
    v = opnorm(A, 1) * inverse_norm_estimate(F, Float64, n; maxiter)
    iszero(v) ? 1.0 : min(1.0, 1 / v)
end

function refine_solve(A::StridedMatrix{Float64}, F::LU, b::StridedVecOrMat{Float64}; maxiter=8, rtol=8eps())
    n = LinearAlgebra.checksquare(A); B = reshape(b, n, :)
    Ad, Bd, Xd = DeviceMatrix(Matrix(A)), DeviceMatrix(Matrix(B)), DeviceMatrix(n, size(B, 2))
    info = Ref(RefineInfo()); h = handle()
    rc = ccall((:oxi_drefine_solve, LIB), Cint,
               (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Float64}, Int64,
                Ptr{Float64}, Int64, Int32, Float64, Ref{RefineInfo}),
               h, n, size(B, 2), Ad.ptr, Ad.ld, F.factors.ptr, F.factors.ld, getfield(F, :ipiv), Bd.ptr, Bd.ld, Xd.ptr,
               Xd.ld, Int32(maxiter), Float64(rtol), info)
    rc == 4 && throw(SingularException(F.info))
    _check(h, rc, "refine_solve")
    r = info[]; X = Array(Xd)
    (; x=ndims(b) == 1 ? vec(X) : X, converged=r.converged != 0, iterations=Int(r.iterations),
     backward_error=r.backward_error, residual_norm=r.residual_norm, rcond_estimate=r.rcond_estimate,
     forward_error_estimate=r.forward_error_estimate)
end

function expert_solve(A::StridedMatrix{Float64}, b::StridedVecOrMat{Float64}; maxiter=8, rtol=8eps(),
                      equilibrate=true, pivot=:partial)
    pivot === :partial || throw(ArgumentError("expert_solve implements pivot=:partial"))
    n = LinearAlgebra.checksquare(A); B = reshape(b, n, :)
    Ad, Bd, Xd = DeviceMatrix(Matrix(A)), DeviceMatrix(Matrix(B)), DeviceMatrix(n, size(B, 2))
    LUd = DeviceMatrix(n, n); rows = DeviceMatrix(max(n, 1), 1); cols = DeviceMatrix(max(n, 1), 1)
    h = handle(); piv = Ref{Ptr{Cvoid}}(C_NULL)
    _check(h, ccall((:oxi_malloc, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, UInt64), h, piv, 8 * max(n, 1)), "malloc")
    info = Ref(RefineInfo())
    rc = ccall((:oxi_dexpert_solve, LIB), Cint,
               (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32, Float64,
                Int32, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ref{RefineInfo}),
               h, n, size(B, 2), Ad.ptr, Ad.ld, Bd.ptr, Bd.ld, Xd.ptr, Xd.ld, Int32(maxiter), Float64(rtol),
               Int32(equilibrate), LUd.ptr, LUd.ld, Ptr{Int64}(piv[]), rows.ptr, cols.ptr, info)
    if rc == 4
        m = match(r"U\((\d+),", lasterror(h))
        throw(SingularException(m === nothing ? 0 : parse(Int, m.captures[1])))
    end
    _check(h, rc, "expert_solve")
    F = LU(LUd, Ptr{Int64}(piv[]), 0)
    finalizer(f -> ccall((:oxi_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), handle(), getfield(f, :ipiv)), F)
    r = info[]; X = Array(Xd)
    (; x=ndims(b) == 1 ? vec(X) : X, converged=r.converged != 0, iterations=Int(r.iterations),
     backward_error=r.backward_error, row_scale=vec(Array(rows))[1:n], column_scale=vec(Array(cols))[1:n], factor=F)
end

# mixed_solve(A, b; factorization=:cholesky): oxi_dposv_mixed
function _mixed_cholesky(A, b; maxiter, rtol, lowtype)
    lowtype === Float32 || throw(ArgumentError("the Cholesky mixed solve takes Float32 low factors"))
    n = LinearAlgebra.checksquare(A); B = reshape(b, n, :); k = size(B, 2)
    Ad, Bd, Xd = DeviceMatrix(Matrix{Float64}(A)), DeviceMatrix(Matrix{Float64}(B)), DeviceMatrix(n, k)
    info = Ref(IRInfo()); h = handle()
    rc = ccall((:oxi_dposv_mixed, LIB), Cint,
               (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32,
                Float64, Ref{IRInfo}),
               h, Cint(3), n, k, Ad.ptr, Ad.ld, Bd.ptr, Bd.ld, Xd.ptr, Xd.ld, Int32(maxiter), Float64(rtol), info)
    rc == 6 && throw(ArgumentError("finite mixed input required"))
    if rc == 8
        m = match(r"A\((\d+),", lasterror(h))
        throw(PosDefException(m === nothing ? 0 : parse(Int, m.captures[1])))
    end
    _check(h, rc, "mixed_solve")
    r = info[]; X = Array(Xd)
    (; x=ndims(b) == 1 ? vec(X) : X, converged=r.converged != 0, iterations=Int(r.iterations),
     residual=r.residual_norm, backward_error=r.backward_error, metric=:componentwise_backward_error,
     low_type=Float32, fallback=r.fallback != 0, factorization=:cholesky, finite=r.reason != 2,
     reason=_REASONS[r.reason + 1])
end

end # module
