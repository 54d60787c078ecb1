# Directional (wideband) operators: with isnear = isnearwideband(k) the
# reference builds DirectionalData and the directional stores
# (DirBasisTraversalPlan, DirTransferTraversalPlan, DirCouplingTraversalPlan,
# src/nca/abstractnca.jl:5-9,41-48,78-82); they are passed as they are to
# nca_dh2_create, and the handle is an ordinary HIPOperator. The coupling store
# is assembled over direction indices: the CSR is diridxtofarptr of the test
# data with the trial direction index of every block (sdiridcs of
# assemble_couplingstore(::DirectionalData), src/couplingmatrices.jl:55-111).

struct DH2DescC                      # nca_dh2_desc, field order of include/nca_hip.h
    M::Int64; N::Int64; nn_test::Int64; nn_trial::Int64; ne_test::Int64; ne_trial::Int64
    test_perm::Ptr{Int64}; test_lo::Ptr{Int64}; test_hi::Ptr{Int64}
    trial_perm::Ptr{Int64}; trial_lo::Ptr{Int64}; trial_hi::Ptr{Int64}
    test_dirptr::Ptr{Int64}; trial_dirptr::Ptr{Int64}
    test_cptr::Ptr{Int64}; trial_cptr::Ptr{Int64}
    nbnodes_test::Int64; nbnodes_trial::Int64
    bt_nodes::Ptr{Int64}; bt_dirptr::Ptr{Int64}; bt_diridx::Ptr{Int64}; bt_off::Ptr{Int64}
    bs_nodes::Ptr{Int64}; bs_dirptr::Ptr{Int64}; bs_diridx::Ptr{Int64}; bs_off::Ptr{Int64}
    bt_blocks::Ptr{ComplexF64}; bs_blocks::Ptr{ComplexF64}
    nlev_test::Int64; nlev_trial::Int64
    tt_levptr::Ptr{Int64}; tt_levidcs::Ptr{Int64}; tt_chdptr::Ptr{Int64}; tt_chds::Ptr{Int64}; tt_off::Ptr{Int64}
    st_levptr::Ptr{Int64}; st_levidcs::Ptr{Int64}; st_chdptr::Ptr{Int64}; st_chds::Ptr{Int64}; st_off::Ptr{Int64}
    tt_blocks::Ptr{ComplexF64}; st_blocks::Ptr{ComplexF64}
    cp_nodeptr::Ptr{Int64}; cp_tidcs::Ptr{Int64}; cp_sidcs::Ptr{Int64}
    nearptr::Ptr{Int64}; nearnodes::Ptr{Int64}
end

function HIPDirectionalOperator(A::HIPKernelMatrix, core, testdata, trialdata, partition;
                                precision::Integer=NCA_PREC_FP64)
    tp, tl, th = layout(H2Trees.testtree(core.tree))
    sp, sl, sh = layout(H2Trees.trialtree(core.tree))
    B, S = core.nestedtestbases, core.nestedtrialbases
    E, F = core.testtransfermatrices, core.trialtransfermatrices
    C = core.couplingmatrices
    tb, tbo = packed(B.blocks)
    sb, sbo = packed(S.blocks)
    ttb, tto = packed(E.blocks)
    stb, sto = packed(F.blocks)
    nptr, nnodes = Int64.(partition.nearptr), Int64.(partition.nearnodes)
    keep = Any[tp, tl, th, sp, sl, sh, tb, tbo, sb, sbo, ttb, tto, stb, sto, B.plan, S.plan, E.plan, F.plan,
               C.plan, testdata, trialdata, core.disaggregationplan, core.aggregationplan, nptr, nnodes, C.blocks]
    d = DH2DescC(length(tp), length(sp), length(tl), length(sl), length(testdata.dirs), length(trialdata.dirs),
        pointer(tp), pointer(tl), pointer(th), pointer(sp), pointer(sl), pointer(sh),
        pointer(testdata.nodetodirsptr), pointer(trialdata.nodetodirsptr),
        pointer(core.disaggregationplan.globalidxtocoeffptr), pointer(core.aggregationplan.globalidxtocoeffptr),
        length(B.plan.nodes), length(S.plan.nodes),
        pointer(B.plan.nodes), pointer(B.plan.idxtodiridxptr), pointer(B.plan.diridx), pointer(tbo),
        pointer(S.plan.nodes), pointer(S.plan.idxtodiridxptr), pointer(S.plan.diridx), pointer(sbo),
        pointer(tb), pointer(sb), length(E.plan.leveltoidcsptr) - 1, length(F.plan.leveltoidcsptr) - 1,
        pointer(E.plan.leveltoidcsptr), pointer(E.plan.levelidcs), pointer(E.plan.idxtochdsptr),
        pointer(E.plan.chdsidcs), pointer(tto),
        pointer(F.plan.leveltoidcsptr), pointer(F.plan.levelidcs), pointer(F.plan.idxtochdsptr),
        pointer(F.plan.chdsidcs), pointer(sto),
        pointer(ttb), pointer(stb), pointer(C.plan.nodetodiridxptr), pointer(C.plan.tidcs), pointer(C.plan.sidcs),
        pointer(nptr), pointer(nnodes))
    r = Ref{Ptr{Cvoid}}()
    k = Ref(A.kernel)
    GC.@preserve keep check(ccall((:nca_dh2_create, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{DH2DescC}, Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}),
        A.ctx.ptr, k, A.test.ptr, A.trial.ptr, Ref(d), C.blocks.ptr, precision, r))
    return _finalize(HIPOperator(r[], keep))
end
