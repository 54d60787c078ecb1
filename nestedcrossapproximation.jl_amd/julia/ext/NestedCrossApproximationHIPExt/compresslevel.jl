# Seam 1: assemble_level (src/compressors/batchedbottomup.jl:101-116) returns a
# device-resident level buffer, whose type selects the methods of seam 2,
# compresslevel_test!/compresslevel_trial! (batchedbottomup.jl:145-244):
# batched ACA (or IACA) and the skeleton solve with the pivots kept on the
# device (the CUDA precedent is ext/NestedCrossApproximationCUDAExt/compresslevel.jl).

mutable struct HIPLevelStack          # (sum m) x maxrank complex FP64 on the device, padded columns
    ptr::Ptr{Cvoid}
    ctx::Context
    rowoffsets::Vector{Int}
    ncols::Vector{Int}
    maxrank::Int
end

function assemble_level(A::HIPKernelMatrix{ComplexF64}, rowidcs::AbstractVector{<:AbstractVector{Int}},
                        colidcs::AbstractVector{<:AbstractVector{Int}}; maxrank::Int, ntasks=nothing,
                        transposed::Bool=false)
    length(rowidcs) == length(colidcs) ||
        throw(DimensionMismatch("row and column index collections must have equal length"))
    rp, ri = csr(rowidcs)
    cp, ci = csr(colidcs)
    r = Ref{Ptr{Cvoid}}()
    k = Ref(A.kernel)
    GC.@preserve rp ri cp ci check(ccall((:nca_level_assemble, libnca_hip), Int32,
        (Ptr{Cvoid}, Ref{HelmholtzC}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
         Ptr{Int64}, Int64, Int32, Ptr{Ptr{Cvoid}}),
        A.ctx.ptr, k, A.test.ptr, A.trial.ptr, length(rowidcs), rp, ri, cp, ci, maxrank, transposed, r))
    st = HIPLevelStack(r[], A.ctx, [0; cumsum(length.(rowidcs))], length.(colidcs), maxrank)
    finalizer(s -> ccall((:nca_stack_destroy, libnca_hip), Int32, (Ptr{Cvoid},), s.ptr), st)
    return st
end

_isiaca(cfg) = cfg isa AdaptiveCrossApproximation.IACA
_tolerance(cfg) = cfg.convergence.tol

# ACA(FullPivoting(), FullPivoting(), FNormEstimator(tol)) with the sampled stop
# (batchedbottomup.jl:118-125), or the IACA (geometric column order, rows by
# maximum value), then X_b = L_b P_b^{-1} for every block (the single-Z form of
# compresslevel.jl:14-25). Returns ranks, local row pivots and the factors.
function _compress(cfg, st::HIPLevelStack)
    nb, mr = length(st.ncols), st.maxrank
    m = diff(st.rowoffsets)
    ranks = zeros(Int32, nb)
    rp = zeros(Int64, mr, nb)
    cp = zeros(Int64, mr, nb)
    xoff = zeros(Int64, nb + 1)
    indep = zeros(Int64, nb)
    cap = sum(m .* min.(m, st.ncols, mr); init=0)
    X = Vector{ComplexF64}(undef, max(cap, 1))
    tol = Float64(_tolerance(cfg))
    status = GC.@preserve ranks rp cp xoff X indep begin
        if _isiaca(cfg)
            ccall((:nca_compress_level_iaca, libnca_hip), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Int32}, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64},
                 Ptr{ComplexF64}, Ptr{Int64}),
                st.ctx.ptr, st.ptr, tol, ranks, rp, cp, cap, xoff, X, indep)
        else
            ccall((:nca_compress_level, libnca_hip), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Int32, Ptr{Int32}, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64},
                 Ptr{ComplexF64}, Ptr{Int64}),
                st.ctx.ptr, st.ptr, tol, 1, ranks, rp, cp, cap, xoff, X, indep)
        end
    end
    if status == NCA_ERR_SINGULAR                     # _crossfactorization's error (utils.jl:38-80)
        b = findfirst(>=(0), indep)
        throw(SingularCrossMatrixError(:nested_basis, Int(ranks[b]), Int(indep[b])))
    end
    check(status)
    return ranks, rp, xoff, X, m
end

# Test side: basis = X (m x r), transfer = X[child rows, :] (|tau_c| x r).
function compresslevel_test!(cfg, st::HIPLevelStack, subtree, activeclusters, tvalues, childranges,
                             nodechildren, tpivots, bases, transfer, maxrank::Int, pbar; scheduler=nothing)
    ranks, rp, xoff, X, m = _compress(cfg, st)
    for (i, t) in enumerate(activeclusters)
        r = Int(ranks[i])
        tpivots[t] = tvalues[i][rp[1:r, i]]
        Xi = reshape(view(X, (xoff[i] + 1):xoff[i + 1]), m[i], r)
        if H2Trees.isleaf(subtree, t)
            bases[t] = Matrix(Xi)
        else
            transfer[t] = [Matrix(Xi[cr, :]) for cr in childranges[i]]
        end
        NCA.next!(pbar)
    end
    return nothing
end

# Trial side: the blocks were assembled transposed (transposed = true), so
# basis = transpose(X) and transfer = transpose(X[child rows, :]), no conjugation.
function compresslevel_trial!(cfg, st::HIPLevelStack, subtree, activeclusters, svalues, childranges,
                              nodechildren, spivots, bases, transfer, maxrank::Int, pbar; scheduler=nothing)
    ranks, rp, xoff, X, m = _compress(cfg, st)
    for (i, s) in enumerate(activeclusters)
        r = Int(ranks[i])
        spivots[s] = svalues[i][rp[1:r, i]]
        Xi = reshape(view(X, (xoff[i] + 1):xoff[i + 1]), m[i], r)
        if H2Trees.isleaf(subtree, s)
            bases[s] = Matrix(transpose(Xi))
        else
            transfer[s] = [Matrix(transpose(Xi[cr, :])) for cr in childranges[i]]
        end
        NCA.next!(pbar)
    end
    return nothing
end
