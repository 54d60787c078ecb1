# The partition and Phase 1 on the device. The reference takes the partition
# from interactions(tree, isnear) and the Phase-1 pivots from
# TreeMimicryPivoting, both in AdaptiveCrossApproximation.jl; these compute
# them on the device and hand the results to the seams above.

# Cluster tree as the library takes it: 1-based child CSR, centers 3 x nn,
# radii, value ranges lo/hi of one permutation, parents (0 at a root).
struct TreeC
    nnodes::Int64
    childptr::Vector{Int64}
    children::Vector{Int64}
    center::Vector{Float64}
    radius::Vector{Float64}
    lo::Vector{Int64}
    hi::Vector{Int64}
    perm::Vector{Int64}
    parent::Vector{Int64}
    nvalues::Vector{Int64}
end

# Cluster trees built on the device: the regular 2^3 tree (the TwoNTree the
# reference reads through halfsize/center/level, nearinteractions.jl:70-80) and
# the deterministic k-means tree (KMeansTree(pts; builder=KMeansTreeBuilder(...)),
# src/NestedCrossApproximation.jl:50-52). Both return a TreeC for the
# partition and Phase-1 bindings below.
function _treec(h::Ptr{Cvoid}, nn::Int64, npoints::Integer)
    parent, level, lo, hi = (Vector{Int64}(undef, nn) for _ in 1:4)
    childptr = Vector{Int64}(undef, nn + 1)
    children = Vector{Int64}(undef, max(nn - 1, 1))
    perm = Vector{Int64}(undef, max(npoints, 1))
    center = Vector{Float64}(undef, 3nn)
    radius = Vector{Float64}(undef, nn)
    try
        check(ccall((:nca_tree_export, libnca_hip), Int32,
            (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
             Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
            h, parent, level, lo, hi, childptr, children, perm, center, radius))
    finally
        ccall((:nca_tree_destroy, libnca_hip), Int32, (Ptr{Cvoid},), h)
    end
    return TreeC(nn, childptr, children[1:(nn - 1)], center, radius, lo, hi, perm[1:npoints], parent,
                 hi .- lo .+ 1)
end

function hiptwontree(ctx::Context, pts::Points; minvalues::Integer=64, minhalfsize::Real=0.0)
    nn, h = Ref{Int64}(0), Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:nca_tree_twon, libnca_hip), Int32,
        (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Float64, Ref{Int64}, Ref{Ptr{Cvoid}}),
        ctx.ptr, pts.ptr, minvalues, minhalfsize, nn, h))
    return _treec(h[], nn[], pts.n)
end

function hipkmeanstree(ctx::Context, pts::Points; numberofclusters::Integer=2, minvalues::Integer=32)
    nn, h = Ref{Int64}(0), Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:nca_tree_kmeans, libnca_hip), Int32,
        (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ref{Int64}, Ref{Ptr{Cvoid}}),
        ctx.ptr, pts.ptr, numberofclusters, minvalues, nn, h))
    return _treec(h[], nn[], pts.n)
end

# Admissible and near pairs: far iff max diameter <= eta * distance, or the
# wideband criterion isnearwideband (src/admissibility/nearinteractions.jl:83-103).
function hipinteractions(ctx::Context, a::TreeC, b::TreeC; eta=1.0, k=0.0, wideband::Bool=false,
                         etalf=1.0, etahf=5.0, gamma=1.0)
    nn, nf = Ref{Int64}(0), Ref{Int64}(0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    params = [eta, k, etalf, etahf, gamma]
    GC.@preserve a b params check(ccall((:nca_interactions, libnca_hip), Int32,
        (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64},
         Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Int32, Ptr{Float64},
         Ref{Int64}, Ref{Int64}, Ref{Ptr{Cvoid}}),
        ctx.ptr, a.nnodes, a.childptr, a.children, a.center, a.radius,
        b.nnodes, b.childptr, b.children, b.center, b.radius, Int32(wideband), params, nn, nf, h))
    nearptr = Vector{Int64}(undef, a.nnodes + 1)
    nearnodes = Vector{Int64}(undef, nn[])
    farptr = Vector{Int64}(undef, a.nnodes + 1)
    farnodes = Vector{Int64}(undef, nf[])
    check(ccall((:nca_partition_export, libnca_hip), Int32,
        (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}), h[], nearptr, nearnodes, farptr, farnodes))
    ccall((:nca_partition_destroy, libnca_hip), Int32, (Ptr{Cvoid},), h[])
    return (nearptr, nearnodes), (farptr, farnodes)      # 1-based CSR, as interactions returns
end

# Far pivots of every node of `self` among the points of `other` (far field =
# own far nodes, then the ancestors', farinteractions.jl:38-60), with the
# reference's count rule min(maxrank, |t|) (_selectfarpivots,
# batchedbottomup.jl:54-72) for oversample = 1; the selection is the geometric
# stand-in for TreeMimicryPivoting.
function hipfarpivots(ctx::Context, other::Points, self::TreeC, otree::TreeC, farptr, farnodes;
                      maxrank::Integer=50, weight=1.5, oversample=1.0)
    ptr = Vector{Int64}(undef, self.nnodes + 1)
    piv = Vector{Int64}(undef, self.nnodes * maxrank)
    fp, fn = Int64.(farptr), Int64.(farnodes)
    GC.@preserve self otree fp fn check(ccall((:nca_farpivots_tree, libnca_hip), Int32,
        (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
         Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Float64, Float64, Ptr{Int64}, Int64, Ptr{Int64}),
        ctx.ptr, other.ptr, self.nnodes, self.center, self.nvalues, self.parent, fp, fn,
        otree.nnodes, otree.lo, otree.hi, otree.perm, maxrank, weight, oversample, ptr, length(piv), piv))
    return [piv[(ptr[i] + 1):ptr[i + 1]] .+ 1 for i in 1:self.nnodes]   # farpivots[node], 1-based point ids
end
