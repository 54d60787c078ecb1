# Factors kept on the device: a whole side of BatchedBottomUp in one call per
# side (one launch per level, rows built on the device from the level below),
# TopDown on the device, and the operator packed from the factors without a
# host round trip. The reference's GPU path downloads Z after every level
# (compresslevel.jl:131); these entry points keep it on the device.

mutable struct HIPFactors                 # nca_factors: skeleton factors X by block id 1..nids
    ptr::Ptr{Cvoid}
    ctx::Context
    nids::Int
    function HIPFactors(ctx::Context, nids::Integer)
        r = Ref{Ptr{Cvoid}}()
        check(ccall((:nca_factors_create, libnca_hip), Int32, (Ptr{Cvoid}, Int64, Ptr{Ptr{Cvoid}}), ctx.ptr, nids, r))
        f = new(r[], ctx, nids)
        finalizer(f -> ccall((:nca_factors_destroy, libnca_hip), Int32, (Ptr{Cvoid},), f.ptr), f)
        return f
    end
end

# All factors, level by level in call order (each block's X column-major).
function download(f::HIPFactors)
    n = Ref{Int64}(0)
    check(ccall((:nca_factors_size, libnca_hip), Int32, (Ptr{Cvoid}, Ref{Int64}), f.ptr, n))
    X = Vector{ComplexF64}(undef, max(n[], 1))
    check(ccall((:nca_factors_download, libnca_hip), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}),
        f.ctx.ptr, f.ptr, X))
    return X[1:n[]]
end

# Nodes of every level as the drivers take them (1-based CSR), deepest first
# (BottomUp) or root first (TopDown).
function levelcsr(tree::TreeC, level::AbstractVector{<:Integer}; deepestfirst::Bool)
    levs = sort(unique(level); rev=deepestfirst)
    ids = [findall(==(l), level) for l in levs]
    return Int64[1; 1 .+ cumsum(length.(ids))], Int64.(reduce(vcat, ids; init=Int[]))
end

# testbases / trialbases(::BatchedBottomUp) for one side (batchedbottomup.jl:246-447),
# all levels on the device; transposed = the trial side. Returns ranks, rows of
# every block and the row pivots of the blocks in level order.
function hipbottomup!(A::HIPKernelMatrix, f::HIPFactors, self::TreeC, levptr, levids, fpptr, fpidx;
                      maxrank::Integer=50, tol=1e-4, iaca::Bool=false, transposed::Bool=false)
    selfpts, otherpts = transposed ? (A.trial, A.test) : (A.test, A.trial)
    nn = self.nnodes
    ranks, mrows, indep = zeros(Int32, nn), zeros(Int64, nn), zeros(Int64, nn)
    fpp, fpi = Int64.(fpptr), Int64.(fpidx)
    pivcap = sum(min.(diff(fpp), maxrank); init=0)
    piv = Vector{Int64}(undef, max(pivcap, 1))
    k = Ref(A.kernel)
    GC.@preserve self levptr levids fpp fpi check(ccall((:nca_bottomup_side, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
         Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Float64, Int32,
         Int32, Ptr{Cvoid}, Ptr{Int32}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}),
        A.ctx.ptr, k, selfpts.ptr, otherpts.ptr, transposed, nn, self.childptr, self.children, self.lo, self.hi,
        self.perm, length(levptr) - 1, levptr, levids, fpp, fpi, maxrank, tol, 1, iaca, f.ptr, ranks, mrows,
        pivcap, piv, indep))
    return ranks, mrows, piv[1:sum(max.(ranks, 0); init=0)]
end

# testbases / trialbases(::TopDown) for one side (topdown.jl:51-206) on the
# device: ranks (-1: no far field), rows of every factor, and the row (cluster)
# and column (far) pivots of the clusters with pivots in level order.
function hiptopdown!(A::HIPKernelMatrix, f::HIPFactors, self::TreeC, other::TreeC, farptr, farnodes, levptr,
                     levids; maxrank::Integer=50, tol=1e-4, transposed::Bool=false)
    selfpts, otherpts = transposed ? (A.trial, A.test) : (A.test, A.trial)
    nn = self.nnodes
    ranks, mrows, indep = zeros(Int32, nn), zeros(Int64, nn), zeros(Int64, nn)
    fp, fn = Int64.(farptr), Int64.(farnodes)
    pivcap = sum(min.(self.nvalues, maxrank); init=0)
    piv, cpiv = Vector{Int64}(undef, max(pivcap, 1)), Vector{Int64}(undef, max(pivcap, 1))
    k = Ref(A.kernel)
    GC.@preserve self other fp fn levptr levids check(ccall((:nca_topdown_side, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
         Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
         Int64, Ptr{Int64}, Ptr{Int64}, Int64, Float64, Ptr{Cvoid}, Ptr{Int32}, Ptr{Int64}, Int64, Ptr{Int64},
         Ptr{Int64}, Ptr{Int64}),
        A.ctx.ptr, k, selfpts.ptr, otherpts.ptr, transposed, nn, self.parent, self.childptr, self.children,
        self.lo, self.hi, self.perm, other.nnodes, other.lo, other.hi, other.perm, fp, fn, length(levptr) - 1,
        levptr, levids, maxrank, tol, f.ptr, ranks, mrows, pivcap, piv, cpiv, indep))
    n = sum(max.(ranks, 0); init=0)
    return ranks, mrows, piv[1:n], cpiv[1:n]
end

# The operator with its basis and transfer stores packed on the device from
# resident factors (the core's plans and offsets as for HIPOperator; its block
# arrays are not read).
function HIPOperator(A::HIPKernelMatrix, core, partition, testf::HIPFactors, trialf::HIPFactors;
                     precision::Integer=NCA_PREC_FP64)
    d, keep = h2desc(core, partition)
    r = Ref{Ptr{Cvoid}}()
    k = Ref(A.kernel)
    GC.@preserve keep check(ccall((:nca_h2_create_resident, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{H2DescC}, Ptr{Cvoid}, Int32, Ptr{Cvoid},
         Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
        A.ctx.ptr, k, A.test.ptr, A.trial.ptr, Ref(d), core.couplingmatrices.blocks.ptr, precision, testf.ptr,
        trialf.ptr, r))
    push!(keep, testf, trialf)
    return _finalize(HIPOperator(r[], keep))
end

# Launch choices saved from autotune! (a persisted tuning table), applied
# without timing; devicename keys such a table.
tuning_apply!(op::HIPOperator, t::HIPTuning) =
    check(ccall((:nca_h2_tuning_apply, libnca_hip), Int32, (Ptr{Cvoid}, Ref{HIPTuning}), op.ptr, t))

function devicename(ctx::Context)
    buf = Vector{UInt8}(undef, 256)
    check(ccall((:nca_ctx_device_name, libnca_hip), Int32, (Ptr{Cvoid}, Ptr{UInt8}, Int64), ctx.ptr, buf,
        length(buf)))
    return String(buf[1:(findfirst(==(0x00), buf) - 1)])
end
