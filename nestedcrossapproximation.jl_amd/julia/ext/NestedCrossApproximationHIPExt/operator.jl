# Seam 4: the device operator behind LinearMaps._unsafe_mul!
# (src/nca/petrovgalerkinnca.jl:415-425, stages :431-589). nca_h2_create takes
# the core's plans as the Julia structs hold them (1-based) plus the packed
# blocks; the near field is passed as its CSR plan and evaluated on the device
# (or stored on the device, store_near!).

struct H2DescC                       # nca_h2_desc, field order of include/nca_hip.h
    M::Int64; N::Int64; nn_test::Int64; nn_trial::Int64
    test_perm::Ptr{Int64}; test_lo::Ptr{Int64}; test_hi::Ptr{Int64}
    trial_perm::Ptr{Int64}; trial_lo::Ptr{Int64}; trial_hi::Ptr{Int64}
    test_cptr::Ptr{Int64}; trial_cptr::Ptr{Int64}
    nbasis_test::Int64; nbasis_trial::Int64
    basis_test_nodes::Ptr{Int64}; basis_test_off::Ptr{Int64}
    basis_trial_nodes::Ptr{Int64}; basis_trial_off::Ptr{Int64}
    basis_test::Ptr{ComplexF64}; basis_trial::Ptr{ComplexF64}
    nlev_test::Int64; nlev_trial::Int64
    tt_levptr::Ptr{Int64}; tt_levidcs::Ptr{Int64}; tt_chdptr::Ptr{Int64}; tt_chds::Ptr{Int64}; tt_off::Ptr{Int64}
    st_levptr::Ptr{Int64}; st_levidcs::Ptr{Int64}; st_chdptr::Ptr{Int64}; st_chds::Ptr{Int64}; st_off::Ptr{Int64}
    tt_blocks::Ptr{ComplexF64}; st_blocks::Ptr{ComplexF64}
    cp_ptr::Ptr{Int64}; cp_nodes::Ptr{Int64}
    nearptr::Ptr{Int64}; nearnodes::Ptr{Int64}
end

packed(blocks) = (reduce(vcat, vec.(blocks); init=ComplexF64[]), Int64[0; cumsum(length.(blocks))])

# values(tree, node) as ranges of one permutation (H2Trees keeps them so)
function layout(tree)
    perm = reduce(vcat, (H2Trees.values(tree, l) for l in H2Trees.leaves(tree)); init=Int[])
    pos = invperm(perm)
    lo = [isempty(H2Trees.values(tree, n)) ? 1 : minimum(pos[H2Trees.values(tree, n)])
          for n in 1:H2Trees.numberofnodes(tree)]
    hi = [lo[n] + length(H2Trees.values(tree, n)) - 1 for n in eachindex(lo)]
    return Int64.(perm), Int64.(lo), Int64.(hi)
end

mutable struct HIPOperator
    ptr::Ptr{Cvoid}
    keep::Vector{Any}                # host arrays and the coupling handle the operator borrows
end

function _finalize(op::HIPOperator)
    finalizer(o -> ccall((:nca_h2_destroy, libnca_hip), Int32, (Ptr{Cvoid},), o.ptr), op)
    return op
end

# The nca_h2_desc of a core, and the host arrays it points into.
function h2desc(core, partition)
    tp, tl, th = layout(H2Trees.testtree(core.tree))
    sp, sl, sh = layout(H2Trees.trialtree(core.tree))
    tb, tbo = packed(core.nestedtestbases.blocks)
    sb, sbo = packed(core.nestedtrialbases.blocks)
    ttb, tto = packed(core.testtransfermatrices.blocks)
    stb, sto = packed(core.trialtransfermatrices.blocks)
    tplan, splan = core.testtransfermatrices.plan, core.trialtransfermatrices.plan
    cplan = core.couplingmatrices.plan
    nptr, nnodes = Int64.(partition.nearptr), Int64.(partition.nearnodes)
    keep = Any[tp, tl, th, sp, sl, sh, tb, tbo, sb, sbo, ttb, tto, stb, sto, tplan, splan, cplan,
               core.disaggregationplan, core.aggregationplan, nptr, nnodes, core.couplingmatrices.blocks]
    d = H2DescC(length(tp), length(sp), length(tl), length(sl), pointer(tp), pointer(tl), pointer(th),
        pointer(sp), pointer(sl), pointer(sh),
        pointer(core.disaggregationplan.globalidxtocoeffptr), pointer(core.aggregationplan.globalidxtocoeffptr),
        length(core.nestedtestbases.plan.nodes), length(core.nestedtrialbases.plan.nodes),
        pointer(core.nestedtestbases.plan.nodes), pointer(tbo), pointer(core.nestedtrialbases.plan.nodes),
        pointer(sbo), pointer(tb), pointer(sb), length(tplan.leveltoidcsptr) - 1, length(splan.leveltoidcsptr) - 1,
        pointer(tplan.leveltoidcsptr), pointer(tplan.levelidcs), pointer(tplan.idxtochdsptr),
        pointer(tplan.chdsidcs), pointer(tto),
        pointer(splan.leveltoidcsptr), pointer(splan.levelidcs), pointer(splan.idxtochdsptr),
        pointer(splan.chdsidcs), pointer(sto),
        pointer(ttb), pointer(stb), pointer(cplan.nodetofarnodesptr), pointer(cplan.farnodes),
        pointer(nptr), pointer(nnodes))
    return d, keep
end

function HIPOperator(A::HIPKernelMatrix, core, partition; precision::Integer=NCA_PREC_FP64)
    d, keep = h2desc(core, partition)
    r = Ref{Ptr{Cvoid}}()
    k = Ref(A.kernel)
    GC.@preserve keep check(ccall((:nca_h2_create, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{H2DescC}, Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}),
        A.ctx.ptr, k, A.test.ptr, A.trial.ptr, Ref(d), core.couplingmatrices.blocks.ptr, precision, r))
    return _finalize(HIPOperator(r[], keep))
end

# FP64 (0) or MIXED (1): c32 operators and source coefficients on f32 MFMA,
# FP64 destination coefficients, distance and phase k r.
setprecision!(op::HIPOperator, precision::Integer) =
    check(ccall((:nca_h2_set_precision, libnca_hip), Int32, (Ptr{Cvoid}, Int32), op.ptr, precision))

# The near field stored on the device (the reference CPU path's layout) for
# forward products instead of evaluated; false drops it.
store_near!(op::HIPOperator, on::Bool=true) =
    check(ccall((:nca_h2_store_near, libnca_hip), Int32, (Ptr{Cvoid}, Int32), op.ptr, on))

# op: 0 = A, 1 = transpose(A), 2 = adjoint(A); y = alpha op(A) x + beta y, host arrays
function hipmul!(y::AbstractVecOrMat{ComplexF64}, op::HIPOperator, opcode::Integer,
                 x::AbstractVecOrMat{ComplexF64}, alpha=true, beta=false)
    a = ComplexF64(alpha)
    b = ComplexF64(beta)
    GC.@preserve y x check(ccall((:nca_h2_matvec, libnca_hip), Int32,
        (Ptr{Cvoid}, Int32, Int64, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Int64, Ref{ComplexF64},
         Ref{ComplexF64}),
        op.ptr, opcode, size(x, 2), x, stride(x, 2), y, stride(y, 2), a, b))
    return y
end

# Device arrays (e.g. ROCArray pointers): enqueued on the context stream.
struct HIPMatrix
    ptr::Ptr{Cvoid}                  # complex FP64, column-major
    ld::Int64
end

function hipmul_device!(y::HIPMatrix, op::HIPOperator, opcode::Integer, x::HIPMatrix, nrhs::Integer,
                        alpha=true, beta=false; noalloc::Bool=false)
    a = [real(ComplexF64(alpha)), imag(ComplexF64(alpha))]
    b = [real(ComplexF64(beta)), imag(ComplexF64(beta))]
    # noalloc: the library's static device scratch instead of the operator's
    # workspace, for callers that cannot allocate (graph capture, solver loops)
    GC.@preserve a b begin
        if noalloc
            check(ccall((:nca_h2_matvec_device_noalloc, libnca_hip), Int32,
                (Ptr{Cvoid}, Int32, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}),
                op.ptr, opcode, nrhs, x.ptr, x.ld, y.ptr, y.ld, a, b))
        else
            check(ccall((:nca_h2_matvec_device, libnca_hip), Int32,
                (Ptr{Cvoid}, Int32, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}),
                op.ptr, opcode, nrhs, x.ptr, x.ld, y.ptr, y.ld, a, b))
        end
    end
    return y
end

# One device sweep captured once and replayed (fixed device buffers). The
# graph must be dropped before its operator; a product that grows the
# workspace, autotune! or store_near! makes later launches throw.
mutable struct HIPGraph
    ptr::Ptr{Cvoid}
    op::HIPOperator
end

function hipgraph(op::HIPOperator, opcode::Integer, x::HIPMatrix, y::HIPMatrix, nrhs::Integer)
    g = Ref{Ptr{Cvoid}}(C_NULL)
    a = [1.0, 0.0]
    b = [0.0, 0.0]
    GC.@preserve a b check(ccall((:nca_h2_graph_create, libnca_hip), Int32,
        (Ptr{Cvoid}, Int32, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
        op.ptr, opcode, nrhs, x.ptr, x.ld, y.ptr, y.ld, a, b, g))
    gr = HIPGraph(g[], op)
    finalizer(h -> ccall((:nca_h2_graph_destroy, libnca_hip), Int32, (Ptr{Cvoid},), h.ptr), gr)
    return gr
end

launch!(g::HIPGraph) = check(ccall((:nca_h2_graph_launch, libnca_hip), Int32, (Ptr{Cvoid},), g.ptr))

# Time the launch choices on this operator once and keep the fastest.
struct HIPTuning
    near_k3::Int32; leaf_order::Int32; near_xcd::Int32; couple_k3::Int32
    couple_order::Int32; couple_xcd::Int32; candidates::Int32; near_stagger::Int32
    near_ms::Float64; couple_ms::Float64; near_ms_start::Float64; couple_ms_start::Float64
end

function autotune!(op::HIPOperator; reps::Integer=3)
    t = Ref{HIPTuning}()
    check(ccall((:nca_h2_autotune, libnca_hip), Int32, (Ptr{Cvoid}, Int32, Ref{HIPTuning}), op.ptr, reps, t))
    return t[]
end

# The GPU core: PetrovGalerkinNCACore's fields (src/nca/petrovgalerkinnca.jl:43-101)
# plus the device operator; a PetrovGalerkinNCA over it routes mul! here.
struct HIPCore{T,C}
    core::C
    op::HIPOperator
end
HIPCore(core, op::HIPOperator) = HIPCore{ComplexF64,typeof(core)}(core, op)
Base.getproperty(c::HIPCore, s::Symbol) = s in (:core, :op) ? getfield(c, s) : getproperty(getfield(c, :core), s)

# The LinearMaps entry point (the busy flag and ConcurrentWorkspaceError stay in
# withworkspace, petrovgalerkinnca.jl:293-300; products of several workspaces
# over one operator are serialized by the operator's lock in the library).
LinearMaps._unsafe_mul!(y::AbstractVecOrMat, A::NCA.PetrovGalerkinNCA{<:Any,<:HIPCore}, x::AbstractVecOrMat) =
    NCA.withworkspace(() -> hipmul!(y, A.core.op, NCA_OP_N, x), A)
LinearMaps._unsafe_mul!(y::AbstractVecOrMat,
                        A::LinearMaps.TransposeMap{<:Any,<:NCA.PetrovGalerkinNCA{<:Any,<:HIPCore}},
                        x::AbstractVecOrMat) =
    NCA.withworkspace(() -> hipmul!(y, A.lmap.core.op, NCA_OP_T, x), A.lmap)
LinearMaps._unsafe_mul!(y::AbstractVecOrMat,
                        A::LinearMaps.AdjointMap{<:Any,<:NCA.PetrovGalerkinNCA{<:Any,<:HIPCore}},
                        x::AbstractVecOrMat) =
    NCA.withworkspace(() -> hipmul!(y, A.lmap.core.op, NCA_OP_C, x), A.lmap)
