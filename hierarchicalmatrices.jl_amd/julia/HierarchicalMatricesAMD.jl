"""
    HierarchicalMatricesAMD

Julia front-end of libhmx (C ABI in include/hmx.h), implemented by `ccall`:

  * `DeviceKernelMatrix{T}` (T = Float64 or Float32): the constructor and
    product methods of `KernelMatrix` (src/KernelMatrix.jl:5-47) on the GPU,
    for the named kernels (every form, gradients) or any Julia f (stored and
    hybrid forms; also from a reference-built `KernelMatrix`).
  * `DeviceHierarchicalMatrix`: a `HierarchicalMatrix` with `LowRankMatrix` and
    `Matrix` leaves (src/HierarchicalMatrix.jl) flattened into a device plan
    (`hm_plan_from_blocks`), with `mul!` and the diagonal scalings `rmul!` /
    `lmul!` (src/HierarchicalMatrix.jl:15-16).
  * `DeviceEvenBarycentric`: EvenBarycentricMatrix{Float64|Float32} leaves (src/BarycentricMatrix.jl:5-46)
    in one device plan, with their `mul!` (src/algebra.jl:168-239).
  * `device_lowrank`: LowRankMatrix `+`, `-`, `*`, `'*` with recompression
    (src/LowRankMatrix.jl:95-126) on the device.
  * `kernel_matvec(f, x, y, b, a, bb, c, d)`: `KernelMatrix(f, x, y, a, bb, c, d) * b`
    with a ChainRulesCore `rrule` whose pullback gives the cotangents of x, y
    and b on the device (`kernelmatrix_pullback`).

Vectors may be host vectors (staged through device buffers) or device pointers
(e.g. `pointer` of an AMDGPU.jl `ROCVector`), through `mul_device!`.
"""
module HierarchicalMatricesAMD

using LinearAlgebra
import HierarchicalMatrices: HierarchicalMatrix, LowRankMatrix, KernelMatrix, blocksize,
                             chebyshevpoints
import ChainRulesCore

const libhmx = joinpath(@__DIR__, "..", "hmx", "libhmx.so")
const libhip = "libamdhip64"

export DeviceKernelMatrix, DeviceHierarchicalMatrix, DeviceEvenBarycentric, device_lowrank, cauchykernel, coulombkernel,
       coulombprimekernel, logkernel, mul_device!, kernel_matvec, kernelmatrix_pullback

# kernels of examples/Kernel.jl:34-37, known to the device code by id
@inline cauchykernel(x::T, y::T) where T = inv(x - y)
@inline coulombkernel(x::T, y::T) where T = inv((x - y)^2)
@inline coulombprimekernel(x::T, y::T) where T = inv((x - y)^3)
@inline logkernel(x::T, y::T) where T = log(abs(x - y))

kernel_id(::typeof(cauchykernel)) = Cint(0)
kernel_id(::typeof(coulombkernel)) = Cint(1)
kernel_id(::typeof(coulombprimekernel)) = Cint(2)
kernel_id(::typeof(logkernel)) = Cint(3)
# any other f is a host function (HM_USER): the device builds the tree and the
# bases U, V, and f's values at the dense leaves and far node pairs are
# handed over by set_leaf_data! (stored and hybrid forms only)
const HM_USER = Cint(4)
kernel_id(f) = HM_USER

# plan and product flags of include/hmx.h
const HM_PLAN_MATRIX_FREE = Cuint(1)
const HM_PLAN_HYBRID = Cuint(8)
const HM_MV_FAR_F32 = Cuint(8)
const HM_MV_FAR_BF16X3 = Cuint(16)
const HM_MV_MFMA = Cuint(32)
const HM_MV_NO_MFMA = Cuint(64)
const HM_MV_FAR_BF16X6 = Cuint(128)

function check(st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:hm_last_error, libhmx), Cstring, ()))
    throw(ErrorException("libhmx error $st: $msg"))
end

BLOCKRANK(::Type{Float64}) = Int(ccall((:hm_blockrank, libhmx), Cint, (Cint,), 52))
BLOCKRANK(::Type{Float32}) = Int(ccall((:hm_blockrank, libhmx), Cint, (Cint,), 23))
BLOCKSIZE(T) = 4BLOCKRANK(T)

destroy!(h) = (h.plan != C_NULL && ccall((:hm_plan_destroy, libhmx), Cint, (Ptr{Cvoid},), h.plan); h.plan = C_NULL)

mutable struct DeviceKernelMatrix{T<:Union{Float32,Float64}} <: AbstractMatrix{T}
    plan::Ptr{Cvoid}
    m::Int
    n::Int
    function DeviceKernelMatrix{T}(plan, m, n) where T
        H = new{T}(plan, m, n)
        finalizer(destroy!, H)
        H
    end
end

Base.size(H::DeviceKernelMatrix) = (H.m, H.n)

mutable struct DeviceHierarchicalMatrix <: AbstractMatrix{Float64}
    plan::Ptr{Cvoid}
    m::Int
    n::Int
    function DeviceHierarchicalMatrix(plan, m, n)
        H = new(plan, m, n)
        finalizer(destroy!, H)
        H
    end
end

Base.size(H::DeviceHierarchicalMatrix) = (H.m, H.n)

"""
    DeviceKernelMatrix(f, x, y, a, b, c, d; blocksize, rank, mode=:stored)

Same arguments as `KernelMatrix(f, x, y, a, b, c, d)` (src/KernelMatrix.jl:47):
x and y non-increasing `Vector{T}` (T = Float64 or Float32), intervals given as
(upper, lower). `mode` is `:stored`, `:hybrid` or `:free`. A Float32 matrix
splits at Float32 midpoints and forms and applies every factor in Float32, as
the generic reference code does for T = Float32.
"""
function DeviceKernelMatrix(f::Function, x::Vector{T}, y::Vector{T}, a::T, b::T, c::T, d::T;
                            blocksize::Int = BLOCKSIZE(T), rank::Int = BLOCKRANK(T),
                            mode::Symbol = :stored) where T<:Union{Float32,Float64}
    flags = mode === :free ? HM_PLAN_MATRIX_FREE : mode === :hybrid ? HM_PLAN_HYBRID : Cuint(0)
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x y begin
        # ccall needs a constant function name: one call per element type
        st = if T === Float64
            ccall((:hm_plan_from_points, libhmx), Cint,
                  (Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Float64, Float64,
                   Float64, Cint, Cint, Cuint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                  kernel_id(f), x, length(x), y, length(y), a, b, c, d, blocksize, rank, flags,
                  C_NULL, plan)
        else
            ccall((:hm_plan_from_points_f32, libhmx), Cint,
                  (Cint, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32, Float32, Float32,
                   Float32, Cint, Cint, Cuint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                  kernel_id(f), x, length(x), y, length(y), a, b, c, d, blocksize, rank, flags,
                  C_NULL, plan)
        end
        check(st)
    end
    H = DeviceKernelMatrix{T}(plan[], length(x), length(y))
    if kernel_id(f) == HM_USER
        mode === :free && throw(ArgumentError("a host function f has no matrix-free form"))
        set_leaf_data!(H, user_leaf_data(f, H, x, y, rank)...)
    end
    H
end

# leaf list of a plan in traversal order: (code, depth, row0, col0, m, n) and
# (a, b, c, d) per leaf (hm_plan_blocks)
function plan_blocks(H)
    cnt = Ref{Int64}(0)
    check(ccall((:hm_plan_blocks, libhmx), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Float64}, Int64,
                                                  Ref{Int64}), H.plan, C_NULL, C_NULL, 0, cnt))
    ints = Vector{Int64}(undef, 6cnt[])
    dbls = Vector{Float64}(undef, 4cnt[])
    check(ccall((:hm_plan_blocks, libhmx), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Float64}, Int64,
                                                  Ref{Int64}), H.plan, ints, dbls, cnt[], cnt))
    reshape(ints, 6, :), reshape(dbls, 4, :)
end

# f at every dense leaf's (x_i, y_j) and at every far leaf's mapped node pairs
# F[m, n] = f(t_m, s_n), t = (a+b)/2 + (b-a)/2 x_hat in T, as
# BarycentricPoly2D forms them (src/BarycentricMatrix.jl:147-178)
function user_leaf_data(f, H::DeviceKernelMatrix{T}, x::Vector{T}, y::Vector{T}, r) where T
    ints, dbls = plan_blocks(H)
    A, F = T[], T[]
    xh = chebyshevpoints(T, r)
    for t in 1:size(ints, 2)
        code, i0, j0, m, n = ints[1, t], ints[3, t], ints[4, t], ints[5, t], ints[6, t]
        (m == 0 || n == 0) && continue
        if code == 3
            for j in 1:n, i in 1:m
                push!(A, f(x[i0 + i], y[j0 + j]))
            end
        else
            a, b, c, d = T(dbls[1, t]), T(dbls[2, t]), T(dbls[3, t]), T(dbls[4, t])
            tx = [T(0.5) * (a + b) + T(0.5) * (b - a) * xh[p] for p in 1:r]
            ty = [T(0.5) * (c + d) + T(0.5) * (d - c) * xh[q] for q in 1:r]
            for q in 1:r, p in 1:r
                push!(F, f(tx[p], ty[q]))
            end
        end
    end
    A, F
end

"""
    set_leaf_data!(H, A, F)

Dense leaves and far-field cores of a DeviceKernelMatrix of a host function
(`hm_plan_set_leaf_data`): A every dense leaf (column-major, traversal order),
F every far leaf's r x r core.
"""
function set_leaf_data!(H::DeviceKernelMatrix{T}, A::Vector{T}, F::Vector{T}) where T
    GC.@preserve A F begin
        st = if T === Float64
            ccall((:hm_plan_set_leaf_data, libhmx), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid}),
                  H.plan, A, length(A), F, length(F), C_NULL)
        else
            ccall((:hm_plan_set_leaf_data_f32, libhmx), Cint,
                  (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Ptr{Cvoid}),
                  H.plan, A, length(A), F, length(F), C_NULL)
        end
        check(st)
    end
    H
end

# the leaves of a reference-built KernelMatrix in its traversal order
# (src/KernelMatrix.jl:17-45): codes 1 KernelMatrix, 2 BarycentricMatrix2D,
# 3 Matrix (@hierarchical KernelMatrix BarycentricMatrix2D Matrix)
function kernel_leaves!(out, K::KernelMatrix)
    M, N = blocksize(K)
    for m = 1:M, n = 1:N
        code = K.assigned[m, n]
        if code == 1
            kernel_leaves!(out, K.KernelMatrixblocks[m, n])
        elseif code == 2
            push!(out, K.BarycentricMatrix2Dblocks[m, n])
        elseif code == 3
            push!(out, K.Matrixblocks[m, n])
        end
    end
    out
end

"""
    DeviceKernelMatrix(K::KernelMatrix{T}, x, y, a, b, c, d; blocksize, rank, mode=:stored)

A KernelMatrix the reference built from an arbitrary f (src/KernelMatrix.jl:
47-116) on the device: the plan rebuilds the same tree from the points and
intervals K was built from, forms U and V (which do not depend on f) on the
device, and takes K's own dense leaves and far cores B.F. The leaf sizes are
checked against the plan's, leaf for leaf.
"""
function DeviceKernelMatrix(K::KernelMatrix{T}, x::Vector{T}, y::Vector{T}, a::T, b::T, c::T,
                            d::T; blocksize::Int = BLOCKSIZE(T), rank::Int = BLOCKRANK(T),
                            mode::Symbol = :stored) where T<:Union{Float32,Float64}
    mode === :free && throw(ArgumentError("a host function f has no matrix-free form"))
    flags = mode === :hybrid ? HM_PLAN_HYBRID : Cuint(0)
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x y begin
        st = if T === Float64
            ccall((:hm_plan_from_points, libhmx), Cint,
                  (Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Float64, Float64,
                   Float64, Cint, Cint, Cuint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                  HM_USER, x, length(x), y, length(y), a, b, c, d, blocksize, rank, flags,
                  C_NULL, plan)
        else
            ccall((:hm_plan_from_points_f32, libhmx), Cint,
                  (Cint, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32, Float32, Float32,
                   Float32, Cint, Cint, Cuint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                  HM_USER, x, length(x), y, length(y), a, b, c, d, blocksize, rank, flags,
                  C_NULL, plan)
        end
        check(st)
    end
    H = DeviceKernelMatrix{T}(plan[], length(x), length(y))
    ints, _ = plan_blocks(H)
    L = kernel_leaves!(Any[], K)
    length(L) == size(ints, 2) || throw(ArgumentError("K's tree differs from the plan's"))
    A, F = T[], T[]
    for (t, leaf) in enumerate(L)
        size(leaf) == (ints[5, t], ints[6, t]) || throw(ArgumentError("leaf $t differs in size"))
        (ints[5, t] == 0 || ints[6, t] == 0) && continue
        if ints[1, t] == 3
            append!(A, vec(Matrix{T}(leaf)))
        else
            size(leaf.B.F) == (rank, rank) || throw(ArgumentError("leaf $t: rank differs"))
            append!(F, vec(leaf.B.F))
        end
    end
    set_leaf_data!(H, A, F)
end

"""
    mul_device!(u_dev, H, v_dev, istart, jstart; INCX=1, INCY=1, beta=1, k=1, ldu=0, ldv=0, flags=0)

u[istart + (i-1)INCY + q ldu] = beta*u[...] + (H v[:, q])_i for q < k on device
pointers (1-based offsets as mul!(u, H, v, istart, jstart), src/KernelMatrix.jl:17).
`flags` takes the HM_MV_* values (reduced-precision far field, MFMA passes).
"""
function mul_device!(u::Ptr{T}, H::Union{DeviceKernelMatrix{T},DeviceHierarchicalMatrix},
                     v::Ptr{T}, istart::Int, jstart::Int; INCX::Int = 1, INCY::Int = 1,
                     beta::Int = 1, k::Int = 1, ldu::Int = 0, ldv::Int = 0, flags::Cuint = Cuint(0),
                     stream::Ptr{Cvoid} = C_NULL) where T
    st = if T === Float64
        ccall((:hm_matvec, libhmx), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Int64, Int64, Int64, Int64,
               Int64, Cint, Cuint, Ptr{Cvoid}),
              H.plan, u, v, istart - 1, jstart - 1, INCY, INCX, k, ldu, ldv, beta, flags, stream)
    else
        ccall((:hm_matvec_f32, libhmx), Cint,
              (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Int64, Int64, Int64, Int64, Int64, Int64,
               Int64, Cint, Cuint, Ptr{Cvoid}),
              H.plan, u, v, istart - 1, jstart - 1, INCY, INCX, k, ldu, ldv, beta, flags, stream)
    end
    check(st)
    u
end

# host vectors: staged through device buffers
function devbuf(nbytes)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:hipMalloc, libhip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, nbytes)
    st == 0 || throw(OutOfMemoryError())
    p[]
end
devfree(p) = ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), p)
h2d(dst, src::Vector{T}) where T = ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), dst, src, sizeof(T) * length(src), 1)
d2h(dst::Vector{T}, src) where T = ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), dst, src, sizeof(T) * length(dst), 2)

function staged_mul!(u::Vector{T}, H, v::AbstractVector{T}, istart, jstart, INCX, INCY) where T
    vv = Vector{T}(v)
    du, dv = devbuf(sizeof(T) * length(u)), devbuf(sizeof(T) * length(vv))
    try
        h2d(du, u); h2d(dv, vv)
        mul_device!(Ptr{T}(du), H, Ptr{T}(dv), istart, jstart; INCX = INCX, INCY = INCY)
        d2h(u, du)
    finally
        devfree(du); devfree(dv)
    end
    u
end

# mul!(u, H, v, istart, jstart): u += H v (src/KernelMatrix.jl:17-45)
LinearAlgebra.mul!(u::Vector{T}, H::DeviceKernelMatrix{T}, v::AbstractVector{T}, istart::Int,
                   jstart::Int) where T = staged_mul!(u, H, v, istart, jstart, 1, 1)

# LinearAlgebra.mul!(u, H, v) = mul!(u, H, v, 1, 1), src/KernelMatrix.jl:14
LinearAlgebra.mul!(u::Vector{T}, H::DeviceKernelMatrix{T}, v::AbstractVector{T}) where T =
    mul!(u, H, v, 1, 1)

# H * x, src/KernelMatrix.jl:5-8
Base.:*(H::DeviceKernelMatrix{T}, x::AbstractVector{T}) where T = mul!(zeros(T, size(H, 1)), H, x)

# ------------------------------------------------- differentiation

"""
    kernelmatrix_pullback(K, b, ū) -> (x̄, ȳ, b̄)

Cotangents of u = K(x, y) b for the output cotangent ū (host vectors):
b̄ = Kᵀ ū and ȳ_j = b_j ∂(Kᵀ ū)_j/∂y_j in one adjoint pass
(`hm_matvec_t_grad`), x̄_i = ū_i ∂(K b)_i/∂x_i (`hm_matvec_grad_x`): the exact
derivatives of the computed product (near field regenerated with the
kernel's derivative, far field through its Chebyshev-form basis).
"""
function kernelmatrix_pullback(K::DeviceKernelMatrix{T}, b::AbstractVector{T},
                               ū::AbstractVector{T}) where T
    m, n = size(K)
    bb, uu = Vector{T}(b), Vector{T}(ū)
    db, dr = devbuf(sizeof(T) * n), devbuf(sizeof(T) * m)
    dgb, dgy, dgx = devbuf(sizeof(T) * n), devbuf(sizeof(T) * n), devbuf(sizeof(T) * m)
    b̄, ȳ, x̄ = Vector{T}(undef, n), Vector{T}(undef, n), Vector{T}(undef, m)
    try
        h2d(db, bb); h2d(dr, uu)
        if T === Float64
            check(ccall((:hm_matvec_t_grad, libhmx), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64,
                         Int64, Int64, Cint, Ptr{Cvoid}),
                        K.plan, dgb, dgy, dr, db, 1, 1, 1, 1, 0, C_NULL))
            check(ccall((:hm_matvec_grad_x, libhmx), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Int64,
                         Int64, Cint, Ptr{Cvoid}),
                        K.plan, dgx, db, dr, 0, 0, 1, 1, 1, 0, C_NULL))
        else
            check(ccall((:hm_matvec_t_grad_f32, libhmx), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64,
                         Int64, Int64, Cint, Ptr{Cvoid}),
                        K.plan, dgb, dgy, dr, db, 1, 1, 1, 1, 0, C_NULL))
            check(ccall((:hm_matvec_grad_x_f32, libhmx), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Int64,
                         Int64, Cint, Ptr{Cvoid}),
                        K.plan, dgx, db, dr, 0, 0, 1, 1, 1, 0, C_NULL))
        end
        d2h(b̄, dgb); d2h(ȳ, dgy); d2h(x̄, dgx)
    finally
        devfree(db); devfree(dr); devfree(dgb); devfree(dgy); devfree(dgx)
    end
    x̄, ȳ, b̄
end

"""
    kernel_matvec(f, x, y, b, a, bb, c, d; kw...)

`DeviceKernelMatrix(f, x, y, a, bb, c, d; kw...) * b`, differentiable in x, y
and b (the ChainRulesCore `rrule` below); the intervals only select the
tree, so they get no tangent.
"""
kernel_matvec(f, x::Vector{T}, y::Vector{T}, b::AbstractVector{T}, a::T, bb::T, c::T, d::T;
              kw...) where T = DeviceKernelMatrix(f, x, y, a, bb, c, d; kw...) * b

function ChainRulesCore.rrule(::typeof(kernel_matvec), f, x::Vector{T}, y::Vector{T},
                              b::AbstractVector{T}, a::T, bb::T, c::T, d::T; kw...) where T
    K = DeviceKernelMatrix(f, x, y, a, bb, c, d; kw...)
    u = K * b
    function kernel_matvec_pullback(ū)
        x̄, ȳ, b̄ = kernelmatrix_pullback(K, b, ChainRulesCore.unthunk(ū))
        no = ChainRulesCore.NoTangent()
        return (no, no, x̄, ȳ, b̄, no, no, no, no)
    end
    return u, kernel_matvec_pullback
end

# ------------------------------------------------- HierarchicalMatrix leaves

# the leaves in the traversal order of src/HierarchicalMatrix.jl:24-52 with
# their absolute (0-based) offsets
function leaves!(out, H::HierarchicalMatrix, i0, j0)
    M, N = blocksize(H)
    p = 0
    for m = 1:M
        q = 0
        for n = 1:N
            code = H.assigned[m, n]
            if code == 1
                leaves!(out, H.HierarchicalMatrixblocks[m, n], i0 + p, j0 + q)
            elseif code == 2
                push!(out, (H.LowRankMatrixblocks[m, n], i0 + p, j0 + q))
            elseif code == 3
                push!(out, (H.Matrixblocks[m, n], i0 + p, j0 + q))
            end
            q += blocksize(H, 1, n, 2)
        end
        p += blocksize(H, m, N, 1)
    end
    out
end

"""
    DeviceHierarchicalMatrix(H::HierarchicalMatrix{Float64})

Flattens H (LowRankMatrix and Matrix leaves, nested HierarchicalMatrix nodes)
into one device plan (`hm_plan_from_blocks`); LowRank ranks up to 64.
"""
function DeviceHierarchicalMatrix(H::HierarchicalMatrix{Float64})
    L = leaves!(Any[], H, 0, 0)
    nl = length(L)
    kind = Vector{Int32}(undef, nl)
    row0, col0, m, n, rk = (Vector{Int64}(undef, nl) for _ in 1:5)
    A, U, S, V = (fill(Ptr{Float64}(C_NULL), nl) for _ in 1:4)
    keep = Any[]
    for (t, (c, i0, j0)) in enumerate(L)
        row0[t], col0[t] = i0, j0
        m[t], n[t] = size(c)
        if c isa LowRankMatrix
            kind[t] = 2
            rk[t] = length(c.Σ.diag)
            u, s, v = Matrix{Float64}(c.U), Vector{Float64}(c.Σ.diag), Matrix{Float64}(c.V)
            push!(keep, u, s, v)
            U[t], S[t], V[t] = pointer(u), pointer(s), pointer(v)
        else
            kind[t] = 3
            rk[t] = 0
            a = Matrix{Float64}(c)
            push!(keep, a)
            A[t] = pointer(a)
        end
    end
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep begin
        check(ccall((:hm_plan_from_blocks, libhmx), Cint,
                    (Int64, Int64, Int64, Ptr{Int32}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                     Ptr{Int64}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}},
                     Ptr{Ptr{Float64}}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                    size(H, 1), size(H, 2), nl, kind, row0, col0, m, n, rk, A, U, S, V, C_NULL, plan))
    end
    DeviceHierarchicalMatrix(plan[], size(H, 1), size(H, 2))
end

# mul!(y, H, x, istart, jstart, INCX, INCY), src/HierarchicalMatrix.jl:24-52
LinearAlgebra.mul!(y::Vector{Float64}, H::DeviceHierarchicalMatrix, x::AbstractVector{Float64},
                   istart::Int, jstart::Int, INCX::Int = 1, INCY::Int = 1) =
    staged_mul!(y, H, x, istart, jstart, INCX, INCY)

LinearAlgebra.mul!(y::Vector{Float64}, H::DeviceHierarchicalMatrix, x::AbstractVector{Float64}) =
    mul!(y, H, x, 1, 1)

Base.:*(H::DeviceHierarchicalMatrix, x::AbstractVector{Float64}) = mul!(zeros(size(H, 1)), H, x)

# rmul!(H, Diagonal(b)) and lmul!(Diagonal(b), H), src/HierarchicalMatrix.jl:15-16,
# applied to the device factors in place (hm_plan_scale)
function device_scale!(H::DeviceHierarchicalMatrix, b::Vector{Float64}, side::Integer)
    db = devbuf(8length(b))
    try
        h2d(db, b)
        check(ccall((:hm_plan_scale, libhmx), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                    H.plan, side, db, 0, C_NULL))
        ccall((:hipDeviceSynchronize, libhip), Cint, ())
    finally
        devfree(db)
    end
    H
end
LinearAlgebra.rmul!(H::DeviceHierarchicalMatrix, b::Diagonal{Float64}) = device_scale!(H, b.diag, 0)
LinearAlgebra.lmul!(b::Diagonal{Float64}, H::DeviceHierarchicalMatrix) = device_scale!(H, b.diag, 1)

"""
    DeviceEvenBarycentric(leaves, nrows, ncols)

A device plan of `EvenBarycentricMatrix{T}` leaves, T = Float64 or Float32
(src/BarycentricMatrix.jl:5-46), placed at 0-based (row0, col0) in an
nrows x ncols matrix: `leaves` is a vector of `(row0, col0, B)`, all of one T.
F comes from each B (the caller's f, evaluated by the reference's
constructor); W is formed on the device in T. A single leaf at (0, 0) gives
the leaf's own `mul!(u, B, v, istart, jstart)`.
"""
mutable struct DeviceEvenBarycentric{T<:Union{Float64,Float32}}
    plan::Ptr{Cvoid}
    m::Int
    n::Int
end
Base.size(E::DeviceEvenBarycentric) = (E.m, E.n)
Base.eltype(::DeviceEvenBarycentric{T}) where {T} = T

for (T, create, matvec) in ((Float64, :hm_even_create, :hm_even_matvec),
                            (Float32, :hm_even_create_f32, :hm_even_matvec_f32))
    @eval begin
        function even_create(::Type{$T}, nrows, ncols, nl, row0, col0, abcd, rank, Fp, plan)
            ccall(($(QuoteNode(create)), libhmx), Cint,
                  (Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Cint,
                   Ptr{Ptr{$T}}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                  nrows, ncols, nl, row0, col0, abcd, rank, Fp, C_NULL, plan)
        end
        function even_matvec(E::DeviceEvenBarycentric{$T}, u::Ptr{$T}, v::Ptr{$T}, ioff, joff,
                             k, ldu, ldv, beta, stream)
            ccall(($(QuoteNode(matvec)), libhmx), Cint,
                  (Ptr{Cvoid}, Ptr{$T}, Ptr{$T}, Int64, Int64, Int64, Int64, Int64, Int64,
                   Int64, Cint, Ptr{Cvoid}),
                  E.plan, u, v, ioff, joff, 1, 1, k, ldu, ldv, beta, stream)
        end
    end
end

function DeviceEvenBarycentric(leaves::AbstractVector, nrows::Integer, ncols::Integer)
    nl = length(leaves)
    nl > 0 || throw(ArgumentError("no leaves"))
    T = eltype(leaves[1][3].F)
    T in (Float64, Float32) || throw(ArgumentError("element type must be Float64 or Float32"))
    rank = length(leaves[1][3].x)
    row0 = Int64[l[1] for l in leaves]
    col0 = Int64[l[2] for l in leaves]
    abcd = Int64[]
    Fs = Matrix{T}[]
    for (_, _, B) in leaves
        length(B.x) == rank || throw(ArgumentError("all leaves must have one rank"))
        eltype(B.F) == T || throw(ArgumentError("all leaves must have one element type"))
        append!(abcd, (B.a, B.b, B.c, B.d))
        push!(Fs, Matrix{T}(B.F))
    end
    Fp = [pointer(F) for F in Fs]
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve Fs begin
        check(even_create(T, nrows, ncols, nl, row0, col0, abcd, rank, Fp, plan))
    end
    E = DeviceEvenBarycentric{T}(plan[], nrows, ncols)
    finalizer(e -> (e.plan != C_NULL && ccall((:hm_even_destroy, libhmx), Cint, (Ptr{Cvoid},), e.plan);
                    e.plan = C_NULL), E)
    E
end

# mul!(u, B, v, istart, jstart), src/algebra.jl:168-239, for every leaf at once:
# row I pairs with column J when (istart-1+I) + (jstart-1+J) is even
function mul_device!(u::Ptr{T}, E::DeviceEvenBarycentric{T}, v::Ptr{T}, istart::Int,
                     jstart::Int; beta::Int = 1, k::Int = 1, ldu::Int = 0, ldv::Int = 0,
                     stream::Ptr{Cvoid} = C_NULL) where {T}
    check(even_matvec(E, u, v, istart - 1, jstart - 1, k, ldu, ldv, beta, stream))
    u
end

function LinearAlgebra.mul!(u::Vector{T}, E::DeviceEvenBarycentric{T}, v::AbstractVector{T},
                            istart::Int = 1, jstart::Int = 1) where {T}
    vv = Vector{T}(v)
    du, dv = devbuf(sizeof(T) * length(u)), devbuf(sizeof(T) * length(vv))
    try
        h2d(du, u); h2d(dv, vv)
        mul_device!(Ptr{T}(du), E, Ptr{T}(dv), istart, jstart)
        d2h(u, du)
    finally
        devfree(du); devfree(dv)
    end
    u
end

# LowRankMatrix arithmetic with recompression (src/LowRankMatrix.jl:82-158) on
# the device: hm_lowrank_op over any number of items; ops of include/hmx.h
const HM_LR_ADD, HM_LR_SUB, HM_LR_MUL, HM_LR_ADJMUL, HM_LR_DMUL, HM_LR_MULD, HM_LR_SVDTRUNC = 0:6

"""
    device_lowrank(op, L1, L2) -> LowRankMatrix{Float64}

`L1 + L2` (op = HM_LR_ADD), `L1 - L2` (HM_LR_SUB), `L1 * L2` (HM_LR_MUL) or
`L1' * L2` (HM_LR_ADJMUL) of two LowRankMatrix{Float64}, computed on the GPU
(thin QR, SVD of the core, getrank truncation, as the reference's methods).
"""
function device_lowrank(op::Integer, L1::LowRankMatrix{Float64}, L2::LowRankMatrix{Float64})
    U1, S1, V1 = Matrix(L1.U), Vector(L1.Σ.diag), Matrix(L1.V)
    U2, S2, V2 = Matrix(L2.U), Vector(L2.Σ.diag), Matrix(L2.V)
    r1, r2 = length(S1), length(S2)
    m, n, k = op == HM_LR_ADJMUL ? (size(V1, 1), size(V2, 1), size(U1, 1)) :
              op == HM_LR_MUL ? (size(U1, 1), size(V2, 1), size(V1, 1)) :
              (size(U1, 1), size(V1, 1), 0)
    rmax = op in (HM_LR_ADD, HM_LR_SUB) ? r1 + r2 : min(r1, r2)
    host = (U1, S1, V1, U2, S2, V2)
    dins = [devbuf(sizeof(a)) for a in host]
    douts = [devbuf(8m * rmax), devbuf(8rmax), devbuf(8n * rmax)]
    rank = Int64[0]
    try
        for (d, a) in zip(dins, host)
            h2d(d, vec(a))
        end
        ins = Ptr{Cvoid}[dins..., C_NULL]
        check(ccall((:hm_lowrank_op, libhmx), Cint,
                    (Cint, Int64, Ptr{Int64}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Int64}, Ptr{Cvoid}),
                    op, 1, Int64[m, n, k, r1, r2], ins, douts, rank, C_NULL))
        r = rank[1]
        U, S, V = zeros(m * rmax), zeros(rmax), zeros(n * rmax)
        d2h(U, douts[1]); d2h(S, douts[2]); d2h(V, douts[3])
        return LowRankMatrix(reshape(U, m, rmax)[:, 1:r], Diagonal(S[1:r]), reshape(V, n, rmax)[:, 1:r])
    finally
        foreach(devfree, dins)
        foreach(devfree, douts)
    end
end

end # module
