# Seam 3: assemble_couplingstore (src/couplingmatrices.jl:11-53). The blocks
# S_i = A[testpivots[t], trialpivots[s]] of every far pair stay on the device
# in the CSR order of the test far data; the store keeps the handle and the H^2
# operator borrows it.

mutable struct HIPCoupling
    ptr::Ptr{Cvoid}
end

function _pivotcsr(pivots)
    return csr([isassigned(pivots, i) ? pivots[i] : Int[] for i in eachindex(pivots)])
end

function _couplingstore(A::HIPKernelMatrix{ComplexF64}, testpivots, trialpivots, fp::Vector{Int64},
                        fn::Vector{Int64})
    tpp, tpi = _pivotcsr(testpivots)
    spp, spi = _pivotcsr(trialpivots)
    r = Ref{Ptr{Cvoid}}()
    k = Ref(A.kernel)
    GC.@preserve tpp tpi spp spi fp fn check(ccall((:nca_coupling_assemble, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Int64,
         Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Cvoid}}),
        A.ctx.ptr, k, A.test.ptr, A.trial.ptr, length(fp) - 1, fp, fn, length(spp) - 1, tpp, tpi, spp, spi, r))
    c = HIPCoupling(r[])
    finalizer(c -> ccall((:nca_coupling_destroy, libnca_hip), Int32, (Ptr{Cvoid},), c.ptr), c)
    return c
end

function assemble_couplingstore(A::HIPKernelMatrix{ComplexF64}, testpivots, trialpivots,
                                testfardata::FarData, trialfardata::FarData; ntasks=nothing, pbar=nothing)
    fp, fn = Int64.(NCA.farptr(testfardata)), Int64.(NCA.farnodes(testfardata))
    c = _couplingstore(A, testpivots, trialpivots, fp, fn)
    return CouplingStore{ComplexF64}(CouplingTraversalPlan(fp, fn), c)   # the blocks field holds the handle
end

# The blocks back on the host (tests, diagnostics): 0-based complex offsets and entries.
function download(c::HIPCoupling, ctx::Context)
    nb, ne = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:nca_coupling_size, libnca_hip), Int32, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), c.ptr, nb, ne))
    off = zeros(Int64, nb[] + 1)
    blocks = Vector{ComplexF64}(undef, max(ne[], 1))
    check(ccall((:nca_coupling_download, libnca_hip), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{ComplexF64}),
        ctx.ptr, c.ptr, off, blocks))
    return off, blocks[1:ne[]]
end
