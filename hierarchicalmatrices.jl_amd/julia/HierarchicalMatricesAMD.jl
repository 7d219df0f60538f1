"""
    HierarchicalMatricesAMD

Julia front-end of libhmx (C ABI in include/hmx.h): a device-backed
`KernelMatrix` with the constructor and product methods of HierarchicalMatrices.jl
(src/KernelMatrix.jl:5-47), implemented by `ccall`.

Vectors may be host `Vector{Float64}` (staged through device buffers) or any
array whose `pointer` is a device address (e.g. an AMDGPU.jl `ROCVector`), via
`mul_device!`.
"""
module HierarchicalMatricesAMD

using LinearAlgebra

const libhmx = joinpath(@__DIR__, "..", "hmx", "libhmx.so")
const libhip = "libamdhip64"

export DeviceKernelMatrix, cauchykernel, coulombkernel, coulombprimekernel, logkernel,
       mul_device!

# kernels of examples/Kernel.jl:34-37, known to the device code by id
@inline cauchykernel(x::T, y::T) where T = inv(x - y)
@inline coulombkernel(x::T, y::T) where T = inv((x - y)^2)
@inline coulombprimekernel(x::T, y::T) where T = inv((x - y)^3)
@inline logkernel(x::T, y::T) where T = log(abs(x - y))

kernel_id(::typeof(cauchykernel)) = Cint(0)
kernel_id(::typeof(coulombkernel)) = Cint(1)
kernel_id(::typeof(coulombprimekernel)) = Cint(2)
kernel_id(::typeof(logkernel)) = Cint(3)
kernel_id(f) = throw(ArgumentError("libhmx knows cauchykernel, coulombkernel, coulombprimekernel and logkernel; got $f"))

const HM_PLAN_MATRIX_FREE = Cuint(1)
const HM_PLAN_HYBRID = Cuint(8)

function check(st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:hm_last_error, libhmx), Cstring, ()))
    throw(ErrorException("libhmx error $st: $msg"))
end

BLOCKRANK(::Type{Float64}) = Int(ccall((:hm_blockrank, libhmx), Cint, (Cint,), 52))
BLOCKRANK(::Type{Float32}) = Int(ccall((:hm_blockrank, libhmx), Cint, (Cint,), 23))
BLOCKSIZE(T) = 4BLOCKRANK(T)

mutable struct DeviceKernelMatrix <: AbstractMatrix{Float64}
    plan::Ptr{Cvoid}
    m::Int
    n::Int
    function DeviceKernelMatrix(plan, m, n)
        H = new(plan, m, n)
        finalizer(h -> (h.plan != C_NULL && ccall((:hm_plan_destroy, libhmx), Cint, (Ptr{Cvoid},), h.plan); h.plan = C_NULL), H)
        H
    end
end

Base.size(H::DeviceKernelMatrix) = (H.m, H.n)

"""
    DeviceKernelMatrix(f, x, y, a, b, c, d; blocksize, rank, mode=:stored)

Same arguments as `KernelMatrix(f, x, y, a, b, c, d)` (src/KernelMatrix.jl:47):
x and y non-increasing, intervals given as (upper, lower). `mode` is `:stored`,
`:hybrid` or `:free`.
"""
function DeviceKernelMatrix(f::Function, x::Vector{Float64}, y::Vector{Float64},
                            a::Float64, b::Float64, c::Float64, d::Float64;
                            blocksize::Int = BLOCKSIZE(Float64), rank::Int = BLOCKRANK(Float64),
                            mode::Symbol = :stored)
    flags = mode === :free ? HM_PLAN_MATRIX_FREE : mode === :hybrid ? HM_PLAN_HYBRID : Cuint(0)
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x y begin
        check(ccall((:hm_plan_from_points, libhmx), Cint,
                    (Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Float64, Float64,
                     Float64, Cint, Cint, Cuint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                    kernel_id(f), x, length(x), y, length(y), a, b, c, d, blocksize, rank,
                    flags, C_NULL, plan))
    end
    DeviceKernelMatrix(plan[], length(x), length(y))
end

"""
    mul_device!(u_dev, H, v_dev, istart, jstart; INCX=1, INCY=1, beta=1)

u[istart + (i-1)INCY] = beta*u[...] + (H v)_i on device pointers (1-based offsets,
as mul!(u, H, v, istart, jstart), src/KernelMatrix.jl:17).
"""
function mul_device!(u::Ptr{Float64}, H::DeviceKernelMatrix, v::Ptr{Float64}, istart::Int,
                     jstart::Int; INCX::Int = 1, INCY::Int = 1, beta::Int = 1,
                     stream::Ptr{Cvoid} = C_NULL)
    check(ccall((:hm_matvec, libhmx), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Int64, Int64, Int64,
                 Int64, Int64, Cint, Cuint, Ptr{Cvoid}),
                H.plan, u, v, istart - 1, jstart - 1, INCY, INCX, 1, 0, 0, beta, 0, stream))
    u
end

# host vectors: staged through device buffers
function devbuf(n)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:hipMalloc, libhip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, 8n)
    st == 0 || throw(OutOfMemoryError())
    Ptr{Float64}(p[])
end
devfree(p) = ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), p)
h2d(dst, src::Vector{Float64}) = ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), dst, src, 8length(src), 1)
d2h(dst::Vector{Float64}, src) = ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), dst, src, 8length(dst), 2)

# mul!(u, H, v, istart, jstart): u += H v (src/KernelMatrix.jl:17-45)
function LinearAlgebra.mul!(u::Vector{Float64}, H::DeviceKernelMatrix, v::AbstractVector{Float64},
                            istart::Int, jstart::Int)
    vv = Vector{Float64}(v)
    du, dv = devbuf(length(u)), devbuf(length(vv))
    try
        h2d(du, u); h2d(dv, vv)
        mul_device!(du, H, dv, istart, jstart)
        d2h(u, du)
    finally
        devfree(du); devfree(dv)
    end
    u
end

# LinearAlgebra.mul!(u, H, v) = mul!(u, H, v, 1, 1), src/KernelMatrix.jl:14
LinearAlgebra.mul!(u::Vector{Float64}, H::DeviceKernelMatrix, v::AbstractVector{Float64}) =
    mul!(u, H, v, 1, 1)

# H * x, src/KernelMatrix.jl:5-8
Base.:*(H::DeviceKernelMatrix, x::AbstractVector{Float64}) = mul!(zeros(Float64, size(H, 1)), H, x)

end # module
