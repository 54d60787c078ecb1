# Status codes, handles and the far-field kernel matrix (include/nca_hip.h, top part).

const NCA_OK, NCA_ERR_ARGUMENT, NCA_ERR_DIMENSION, NCA_ERR_HIP, NCA_ERR_SINGULAR, NCA_ERR_LIMIT = 0, 1, 2, 3, 4, 5
const NCA_OP_N, NCA_OP_T, NCA_OP_C = 0, 1, 2
const NCA_PREC_FP64, NCA_PREC_MIXED = 0, 1

# A nonzero status becomes the exception the reference raises for the same case.
function check(status::Int32)
    status == NCA_OK && return nothing
    msg = unsafe_string(ccall((:nca_last_error, libnca_hip), Cstring, ()))
    status == NCA_ERR_DIMENSION && throw(DimensionMismatch(msg))
    status == NCA_ERR_ARGUMENT && throw(ArgumentError(msg))
    return error(msg)
end

struct HelmholtzC          # nca_helmholtz: G(x, y) = cis(sign k r) / (4 pi r), zero at r = 0
    wavenumber::Float64
    sign::Int32
    reserved::Int32
end

mutable struct Context     # one device and its stream
    ptr::Ptr{Cvoid}
    function Context(device::Integer=0)
        r = Ref{Ptr{Cvoid}}()
        check(ccall((:nca_ctx_create, libnca_hip), Int32, (Int32, Ptr{Ptr{Cvoid}}), device, r))
        c = new(r[])
        finalizer(c -> ccall((:nca_ctx_destroy, libnca_hip), Int32, (Ptr{Cvoid},), c.ptr), c)
        return c
    end
end

synchronize(c::Context) = check(ccall((:nca_ctx_synchronize, libnca_hip), Int32, (Ptr{Cvoid},), c.ptr))

function stream(c::Context)
    s = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:nca_ctx_stream, libnca_hip), Int32, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), c.ptr, s))
    return s[]
end

mutable struct Points      # device copy of a point set
    ptr::Ptr{Cvoid}
    n::Int64
    function Points(ctx::Context, pts::AbstractVector{SVector{3,Float64}})
        v = Vector{SVector{3,Float64}}(pts)
        r = Ref{Ptr{Cvoid}}()
        GC.@preserve v check(ccall((:nca_points_create, libnca_hip), Int32,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Ptr{Cvoid}}), ctx.ptr, Ptr{Float64}(pointer(v)), length(v), r))
        p = new(r[], length(v))
        finalizer(p -> ccall((:nca_points_destroy, libnca_hip), Int32, (Ptr{Cvoid},), p.ptr), p)
        return p
    end
end

# The far-field kernel matrix whose type routes the seams to this extension.
struct HIPKernelMatrix{T} <: AdaptiveCrossApproximation.AbstractKernelMatrix{T}
    ctx::Context
    kernel::HelmholtzC
    test::Points
    trial::Points
    m::Int
    n::Int
end

function HIPKernelMatrix(wavenumber::Real, testpoints, trialpoints; sign::Integer=1, ctx::Context=Context())
    return HIPKernelMatrix{ComplexF64}(ctx, HelmholtzC(Float64(wavenumber), Int32(sign), Int32(0)),
        Points(ctx, testpoints), Points(ctx, trialpoints), length(testpoints), length(trialpoints))
end

Base.size(A::HIPKernelMatrix) = (A.m, A.n)

# 1-based CSR of a collection of index vectors
csr(v) = (Int64.(NCA.ptrfromcounts(length.(v))), Int64.(reduce(vcat, v; init=Int[])))
