/*
 * nca_hip.h — C-ABI of the MI355X (gfx950) hot path of NestedCrossApproximation.jl:
 * the batched bottom-up NCA compressor and the H^2 matrix-vector product.
 *
 * Every entry point replaces one seam of the reference (file:line cited at each
 * declaration) and is what the Julia package extension binds with `ccall`.
 *
 * Conventions
 *  - Status: every function returns int32_t, NCA_OK (0) on success; the message of
 *    the last failure of the calling thread is nca_last_error(). No C++ exception
 *    crosses the ABI.
 *  - Complex numbers are interleaved (re, im) doubles (Julia ComplexF64 layout);
 *    matrices are column-major.
 *  - Arrays that mirror Julia fields (CSR pointers, node ids, point ids, pivots)
 *    are 1-based, exactly as the Julia structs hold them; packed-block offsets are
 *    0-based element offsets (the binding creates them when it flattens a
 *    Vector{Matrix}).
 *  - Points are 3 x n column-major doubles (reinterpret of Vector{SVector{3,Float64}}).
 *  - Handles own device memory; host arrays stay owned by the caller (Julia pins
 *    them with GC.@preserve for the duration of the call).
 *  - A context and its stream serve one host task at a time.
 */
#ifndef NCA_HIP_H
#define NCA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nca_ctx nca_ctx;
typedef struct nca_points nca_points;
typedef struct nca_stack nca_stack;
typedef struct nca_coupling nca_coupling;
typedef struct nca_h2 nca_h2;
typedef struct nca_h2_graph nca_h2_graph;

enum {
    NCA_OK = 0,
    NCA_ERR_ARGUMENT = 1,  /* ArgumentError */
    NCA_ERR_DIMENSION = 2, /* DimensionMismatch */
    NCA_ERR_HIP = 3,       /* HIP runtime failure */
    NCA_ERR_SINGULAR = 4,  /* SingularCrossMatrixError */
    NCA_ERR_LIMIT = 5      /* a block exceeds a documented kernel limit */
};

enum { NCA_OP_N = 0, NCA_OP_T = 1, NCA_OP_C = 2 };
enum { NCA_PREC_FP64 = 0, NCA_PREC_MIXED = 1 };

/* Helmholtz point kernel G(x, y) = cis(sign * k * |x - y|) / (4 pi |x - y|),
 * zero at |x - y| == 0. sign = +1: README.md:104-108 (cis(k r)/(4 pi r));
 * sign = -1: test/nca/test_petrovgalerkinnca.jl:14-17 (exp(-i k r)). */
typedef struct {
    double wavenumber;
    int32_t sign;
    int32_t reserved;
} nca_helmholtz;

/* ---- runtime ------------------------------------------------------------ */
const char* nca_last_error(void);
int32_t nca_abi_version(void);
int32_t nca_ctx_create(int32_t device, nca_ctx** out);
int32_t nca_ctx_destroy(nca_ctx* ctx);
int32_t nca_ctx_synchronize(nca_ctx* ctx);
/* The hipStream_t all work of the context runs on (for device-pointer callers). */
int32_t nca_ctx_stream(nca_ctx* ctx, void** stream);
/* The device's architecture name (gcnArchName, e.g. "gfx950:sramecc+:xnack-"),
 * NUL-terminated in buf[len]; the key of tuning tables. */
int32_t nca_ctx_device_name(nca_ctx* ctx, char* buf, int64_t len);

/* Upload n points (3 x n, column-major). */
int32_t nca_points_create(nca_ctx* ctx, const double* xyz, int64_t n, nca_points** out);
int32_t nca_points_destroy(nca_points* pts);

/* ---- seam 1: Phase 2 level assembly ------------------------------------ */
/* Replaces assemble_level(farmatrix, rowidcs, colidcs; maxrank, ntasks,
 * transposed) (src/compressors/batchedbottomup.jl:101-116) for a HIP kernel
 * matrix; the device layout is that of the reference's GPU shard
 * (ext/NestedCrossApproximationCUDAExt/distributedblockstack.jl:7-27,80-86):
 * one (sum m_b) x maxrank column-major buffer, block b in rows
 * rowoffsets[b]+1 : rowoffsets[b+1], columns 1:n_b, zero padding elsewhere.
 * rowptr/colptr: 1-based CSR over the flattened rowidcs/colidcs (length nb+1).
 * transposed == 0: entry (a, c) = G(test[rowidx], trial[colidx]);
 * transposed != 0: entry (a, c) = G(test[colidx], trial[rowidx]), i.e.
 * transpose(A[colidcs[b], rowidcs[b]]) without conjugation (:98-99).
 * Errors: NCA_ERR_ARGUMENT when a block has more than maxrank columns
 * (distributedblockstack.jl:59-62). */
int32_t nca_level_assemble(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                           const nca_points* trial, int64_t nblocks, const int64_t* rowptr,
                           const int64_t* rowidx, const int64_t* colptr, const int64_t* colidx,
                           int64_t maxrank, int32_t transposed, nca_stack** out);
/* A stack from host data (same layout; rowoffsets 0-based cumulative, nb+1). */
int32_t nca_stack_upload(nca_ctx* ctx, const double* buffer, int64_t nblocks, const int64_t* rowoffsets,
                         const int64_t* ncols, int64_t maxrank, nca_stack** out);
int32_t nca_stack_shape(const nca_stack* st, int64_t* nblocks, int64_t* nrows, int64_t* maxrank);
int32_t nca_stack_download(nca_ctx* ctx, const nca_stack* st, double* buffer);
int32_t nca_stack_destroy(nca_stack* st);

/* ---- seam 2: Phase 3 compression --------------------------------------- */
/* Batched fully pivoted ACA with the sampled stop, on an internal working copy
 * (the stack stays pristine, as compresslevel.jl:31-38 copies it). Replaces the
 * ACA callable cfg(buffer, rowoffsets, ncols, maxrank; sampled=true)
 * (compresslevel.jl:34-36; CPU: _compressblock, batchedbottomup.jl:118-125).
 * Outputs (host): ranks[nb]; rowpivots/colpivots maxrank x nb column-major,
 * 1-based block-local, zero past the rank. */
int32_t nca_aca_fullpiv(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, int32_t* ranks,
                        int64_t* rowpivots, int64_t* colpivots);

/* Batched skeleton LU + solve: X_b = L_b * P_b^{-1}, L_b = A_b[:, colpivots],
 * P_b = L_b[rowpivots, :], m_b x r_b column-major at xoffsets[b] (0-based
 * complex offsets, nb+1, computed by the caller from ranks and block rows).
 * Replaces _gather_solve_tensors! + getrf/getrs_strided_batched!
 * (compresslevel.jl:52-137; Z = X^T) and, on the CPU, nestedtestbasis /
 * testtransfer with _crossfactorization (nestedbasis.jl:43-57,
 * transfermatrices.jl:183-211, utils.jl:57-80). independent[b] = -1 when the
 * cross matrix is regular, else the number of independent pivots of the
 * singularity rule min|U_ii|^2 <= (10 eps)^2 max|U_ii|^2 (utils.jl:72-79);
 * the call then returns NCA_ERR_SINGULAR. */
int32_t nca_skeleton_solve(nca_ctx* ctx, const nca_stack* st, const int32_t* ranks, const int64_t* rowpivots,
                           const int64_t* colpivots, const int64_t* xoffsets, double* X, int64_t* independent);

/* Both steps in one call, the pivots staying on the device in between
 * (compresslevel.jl:31-45). xoffsets is an output here (nb+1); X must hold
 * xcapacity complex numbers (sum_b m_b * min(m_b, n_b, maxrank) always suffices). */
int32_t nca_compress_level(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, int32_t* ranks,
                           int64_t* rowpivots, int64_t* colpivots, int64_t xcapacity, int64_t* xoffsets,
                           double* X, int64_t* independent);

/* Batched IACA: the factorization BottomUp runs per (cluster, direction)
 * block with IACA(MaximumValue(), TreeMimicryPivoting(...),
 * FNormExtrapolator(tol)) (compute_test_pivots!/compute_trial_pivots!
 * src/pivoting/pivots.jl:215-364 through the IACA _compute_raw_pivots! :348-364, called
 * from src/compressors/bottomup.jl:214-424). Each block's columns are its far
 * candidates in geometric order (the stand-in for tree mimicry; trial blocks
 * assembled transposed); step r takes column r, the row pivot is the first
 * maximum of its residual column, and the candidate stops the factorization
 * (uncounted) when r ||R_r(:, r)||^2 <= tol^2 sum_{l<r} ||A(:, l)||^2, its
 * residual is zero, or r reaches min(maxrank, m, n). Outputs as
 * nca_aca_fullpiv; the colpivots are 1..rank. The IACA and FNormExtrapolator of
 * AdaptiveCrossApproximation.jl are not vendored: this is the published method
 * restated, not that package's arithmetic. */
int32_t nca_iaca(nca_ctx* ctx, const nca_stack* st, double tol, int32_t* ranks, int64_t* rowpivots,
                 int64_t* colpivots);
/* nca_iaca + the skeleton solve, as nca_compress_level. */
int32_t nca_compress_level_iaca(nca_ctx* ctx, const nca_stack* st, double tol, int32_t* ranks, int64_t* rowpivots,
                                int64_t* colpivots, int64_t xcapacity, int64_t* xoffsets, double* X,
                                int64_t* independent);

/* ---- factors kept on the device ------------------------------------------ */
typedef struct nca_factors nca_factors;
/* A device pool of skeleton factors X for block ids 1..nids (the nodes of a
 * tree, or the direction indices of directional data). The reference's GPU
 * path downloads Z after every level (compresslevel.jl:131) and writes the
 * bases and transfers on the CPU (compresslevel.jl:242-316); with a pool the
 * factors stay on the device until nca_h2_create_resident /
 * nca_dh2_create_resident pack the stores from them. */
int32_t nca_factors_create(nca_ctx* ctx, int64_t nids, nca_factors** out);
int32_t nca_factors_destroy(nca_factors* f);
/* nca_compress_level (method 0, fully pivoted ACA with the sampled stop) or
 * nca_compress_level_iaca (method 1) with X kept in the pool: block b of the
 * stack becomes the factor of ids[b] (1-based; an id takes one factor).
 * ranks, pivots and independent as nca_compress_level. */
int32_t nca_compress_level_resident(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, int32_t method,
                                    nca_factors* f, const int64_t* ids, int32_t* ranks, int64_t* rowpivots,
                                    int64_t* colpivots, int64_t* independent);
/* The whole batched bottom-up pass of one side (the level loop of
 * testbases/trialbases, src/compressors/batchedbottomup.jl:273-332, with
 * assemble_level + compresslevel_test!/trial! per level,
 * batchedbottomup.jl:80-244) on the device: one launch per level, one
 * workgroup per block evaluates the block's entries, runs the ACA (method 0,
 * sampled stop) or IACA (method 1), solves the skeleton and keeps X in f; the
 * rows of an inner block are its children's pivots in child order, built on
 * the device from the level below, so nothing returns to the host between
 * levels. Same pivots, ranks and factors, bit for bit, as
 * nca_level_assemble + nca_compress_level[_iaca] level by level.
 * Tree over block ids 1..nids (nodes, or direction indices): childptr (nids+1)
 * / children CSR, lo / hi (values perm[lo:hi] of a leaf, 1-based inclusive),
 * perm (1-based ids of self's points); levptr (nlevels+1) / levids: the blocks
 * of each level, deepest level first; fpptr (nids+1) / fpidx: the far pivots
 * (columns, 1-based ids of other's points) of every block. transposed as
 * nca_level_assemble (the trial side). Outputs per block id: ranks, mrows
 * (rows of its block), independent (-1 regular; else the singularity rule's
 * count and the call returns NCA_ERR_SINGULAR); pivots: the row pivots (1-based
 * point ids) of the blocks in levids order, rank of them each (pivcap >= their
 * total; sum over blocks of min(maxrank, columns) always suffices). Limits (NCA_ERR_LIMIT):
 * maxrank <= 96, block rows <= 2048 (bounded up front from the children's
 * possible ranks). */
int32_t nca_bottomup_side(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* self,
                          const nca_points* other, int32_t transposed, int64_t nids, const int64_t* childptr,
                          const int64_t* children, const int64_t* lo, const int64_t* hi, const int64_t* perm,
                          int64_t nlevels, const int64_t* levptr, const int64_t* levids, const int64_t* fpptr,
                          const int64_t* fpidx, int64_t maxrank, double tol, int32_t sampled, int32_t method,
                          nca_factors* f, int32_t* ranks, int64_t* mrows, int64_t pivcap, int64_t* pivots,
                          int64_t* independent);
/* testbases / trialbases(::TopDown) (src/compressors/topdown.jl:51-206) with
 * the factorization ACA() on the device, the factors kept in f. Levels from
 * the root down (levptr/levids: the nodes of every level, root level first);
 * per level one launch, one workgroup per cluster: the partially pivoted ACA
 * (compute_test_pivots!/compute_trial_pivots!, src/pivoting/pivots.jl:215-300;
 * AdaptiveCrossApproximation.jl is not vendored: the published method
 * restated, row pivot 0 first, pivots by maximum value, the Frobenius-norm
 * stop |u|^2 |v|^2 <= tol^2 ||S||_F^2 with the cross kept) of the block of the
 * cluster's points against its far field, evaluated entrywise: the points of
 * its far nodes (farptr/farnodes, 1-based CSR into the other tree, whose
 * layout is olo/ohi/operm) and then the column pivots of its parent. Then the
 * factor of every cluster with pivots, X = A[rows, sigma] A[tau, sigma]^{-1},
 * rows = a leaf's points (nestedtestbasis, nestedbasis.jl:43-57) or the
 * children's pivots in child order (the stacked transfers of testtransfer,
 * transfermatrices.jl:126-130), so the stores come from f as for BottomUp.
 * transposed as nca_level_assemble (the trial side). Outputs: ranks (-1: no far
 * field), mrows, independent as nca_bottomup_side; pivots / cpivots: the row
 * (cluster) and column (far) pivots, 1-based point ids, of the clusters with
 * pivots in levids order (pivcap >= their total). Limits: maxrank <= 96, a
 * factor block of at most 2048 rows. The scratch of a launch is bounded by
 * 4 GiB (clusters of a level are split into launches). */
int32_t nca_topdown_side(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* self, const nca_points* other,
                         int32_t transposed, int64_t nn, const int64_t* parent, const int64_t* childptr,
                         const int64_t* children, const int64_t* lo, const int64_t* hi, const int64_t* perm,
                         int64_t nn_other, const int64_t* olo, const int64_t* ohi, const int64_t* operm,
                         const int64_t* farptr, const int64_t* farnodes, int64_t nlevels, const int64_t* levptr,
                         const int64_t* levids, int64_t maxrank, double tol, nca_factors* f, int32_t* ranks,
                         int64_t* mrows, int64_t pivcap, int64_t* pivots, int64_t* cpivots, int64_t* independent);
/* Total complex entries, and all factors downloaded: the X of every
 * nca_compress_level_resident or nca_bottomup_side call in call order (each
 * call's blocks in stack order, or level by level: the X of nca_compress_level). */
int32_t nca_factors_size(const nca_factors* f, int64_t* nentries);
int32_t nca_factors_download(nca_ctx* ctx, const nca_factors* f, double* X);

/* ---- cluster trees on the device ----------------------------------------- */
typedef struct nca_tree nca_tree;
/* Regular 2^3 tree over the points: the H2Trees.TwoNTree the reference reads
 * through halfsize/center/level (src/admissibility/nearinteractions.jl:70-80,
 * src/admissibility/directionalsubdivision/twondirections.jl:132-158; H2Trees
 * is not vendored). The bounding cube of the points (center = midpoint of the
 * bounding box, half size = largest half extent) is halved while a box holds
 * more than minvalues points and its half size exceeds minhalfsize; empty
 * octants are dropped; a box whose points all lie in one octant and whose half
 * size is at most 1e-12 stays a leaf. Octant o = (x >= c) | (y >= c) << 1 |
 * (z >= c) << 2, children in increasing o, node ids level by level in the
 * order of the parents, every node's values a contiguous range of the
 * permutation (stable inside a node). radius = half size * sqrt(3). Built by
 * level-synchronous kernels on the device. nnodes: the number of nodes. */
int32_t nca_tree_twon(nca_ctx* ctx, const nca_points* pts, int64_t minvalues, double minhalfsize,
                      int64_t* nnodes, nca_tree** out);
/* Deterministic k-means tree (the KMeansTree of src/NestedCrossApproximation.jl:50-52
 * and test/testutils.jl:32-42, KMeansTreeBuilder(numberofclusters, minvalues);
 * H2Trees and ParallelKMeans are not vendored): a node with more than minvalues
 * points is split into at most nclusters (2..8) clusters, seeded by
 * farthest-point traversal from the point farthest from its centroid, then
 * Lloyd iterations (at most 100, ties to the lower cluster); empty clusters are
 * dropped, fewer than two nonempty ones leave a leaf. center = centroid,
 * radius = largest distance to it. Sums over a node's points are fixed-order
 * (slot q mod 256 in order, then a halving tree). One workgroup per node of a
 * level. NCA_ERR_LIMIT above 8 clusters. */
int32_t nca_tree_kmeans(nca_ctx* ctx, const nca_points* pts, int64_t nclusters, int64_t minvalues,
                        int64_t* nnodes, nca_tree** out);
/* The tree in the 1-based layout of the Julia tree: parent (0 at the root),
 * level (1 at the root), lo/hi (values = perm[lo:hi]), childptr (nnodes+1) and
 * children (nnodes-1) as CSR, perm (npoints), center (3 x nnodes), radius. Any
 * output may be NULL. */
int32_t nca_tree_export(const nca_tree* tree, int64_t* parent, int64_t* level, int64_t* lo, int64_t* hi,
                        int64_t* childptr, int64_t* children, int64_t* perm, double* center, double* radius);
int32_t nca_tree_destroy(nca_tree* tree);

/* ---- admissibility partition on the device ------------------------------- */
typedef struct nca_partition nca_partition;
/* Near and far cluster pairs of a test and a trial tree, the result of
 * interactions(tree, isnear) (AdaptiveCrossApproximation.jl, not vendored; the
 * near CSR feeds assemblenears src/nca/abstractnca.jl:134-178, the far CSR
 * farinteractions src/admissibility/farinteractions.jl:62-92), computed by a
 * level-synchronous dual traversal on the device. Trees: 1-based child CSR
 * (childptr nn+1, children), centers 3 x nn, radii (half size * sqrt(3) for
 * boxes). wideband = 0: near iff max(diam) > params[0] * dist; wideband = 1:
 * isnearwideband with params = {-, k, eta_lf, eta_hf, gamma}
 * (src/admissibility/nearinteractions.jl:83-103); dist = |c_t - c_s| - (r_t + r_s).
 * A pair is split until it is far or both nodes are leaves. */
int32_t nca_interactions(nca_ctx* ctx, int64_t nn_test, const int64_t* childptr_test,
                         const int64_t* children_test, const double* center_test, const double* radius_test,
                         int64_t nn_trial, const int64_t* childptr_trial, const int64_t* children_trial,
                         const double* center_trial, const double* radius_trial, int32_t wideband,
                         const double* params, int64_t* nnear, int64_t* nfar, nca_partition** out);
/* 1-based CSR by test node, trial nodes of a row increasing: nearptr and
 * farptr nn_test+1, nearnodes nnear, farnodes nfar. */
int32_t nca_partition_export(const nca_partition* p, int64_t* nearptr, int64_t* nearnodes, int64_t* farptr,
                             int64_t* farnodes);
int32_t nca_partition_destroy(nca_partition* p);

/* ---- Phase 1 on the device ----------------------------------------------- */
/* Far pivots of every item (a cluster, or a (cluster, direction) pair) among
 * the points of its far field, computed on the device. Replaces the Phase-1
 * selection _selectfarpivots (src/compressors/batchedbottomup.jl:54-72,
 * farfield farinteractions.jl:38-60, count rule _clustersize pivots.jl:15-25;
 * the reference's TreeMimicryPivoting lives in AdaptiveCrossApproximation.jl
 * and is not vendored): n = min(maxrank, ceil(oversample |t|)) pivots, capped by
 * the far field's points; candidates an even subsample of every far node, the
 * first pivot nearest to the item's center, every next one maximizing
 * d_min^2 / d_c^(2 weight). Equal, for weights 1 and 1.5, to the host selection
 * ncat_farpivots. centers: 3 x nitems; nvals: |t| per item; ffptr/ffnodes:
 * 1-based CSR of far nodes of the other tree (other_lo/hi/perm: its 1-based
 * layout). Outputs pivptr (nitems+1, 0-based) and piv (0-based point ids);
 * pivcap must hold the total (else NCA_ERR_LIMIT with the total in the message). */
int32_t nca_farpivots(nca_ctx* ctx, const nca_points* other, int64_t nitems, const double* centers,
                      const int64_t* nvals, const int64_t* ffptr, const int64_t* ffnodes, int64_t nn_other,
                      const int64_t* other_lo, const int64_t* other_hi, const int64_t* other_perm,
                      int64_t maxrank, double weight, double oversample, int64_t* pivptr, int64_t pivcap,
                      int64_t* piv);
/* The same selection with the far fields built on the device, as farfield
 * (farinteractions.jl:38-60) builds them: the far nodes of a node, then those
 * of its parent and so on to the root. farptr/farnodes: 1-based CSR of the far
 * nodes of every node of the self tree (FarData, farinteractions.jl:62-92);
 * parent: 1-based parent of every node (0 at a root). Same outputs and limits
 * (a candidate pool above 4096 points is reported after the launch). */
int32_t nca_farpivots_tree(nca_ctx* ctx, const nca_points* other, int64_t nitems, const double* centers,
                           const int64_t* nvals, const int64_t* parent, const int64_t* farptr,
                           const int64_t* farnodes, int64_t nn_other, const int64_t* other_lo,
                           const int64_t* other_hi, const int64_t* other_perm, int64_t maxrank, double weight,
                           double oversample, int64_t* pivptr, int64_t pivcap, int64_t* piv);

/* ---- seam 3: coupling matrices ----------------------------------------- */
/* Replaces assemble_couplingstore(farmatrix, testpivots, trialpivots,
 * testfardata, trialfardata) (src/couplingmatrices.jl:11-53): block i of the
 * test CSR row of t is A[testpivots[t], trialpivots[farnodes[i]]] (r_t x r_s,
 * column-major), packed in CSR order. Kept on the device for nca_h2_create.
 * farptr/farnodes: CouplingTraversalPlan (1-based); pivptr/piv: 1-based CSR of
 * the pivots of every node (empty for clusters without pivots). */
int32_t nca_coupling_assemble(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                              const nca_points* trial, int64_t ntest, const int64_t* farptr,
                              const int64_t* farnodes, int64_t ntrial, const int64_t* testpivptr,
                              const int64_t* testpiv, const int64_t* trialpivptr, const int64_t* trialpiv,
                              nca_coupling** out);
int32_t nca_coupling_size(const nca_coupling* cp, int64_t* nblocks, int64_t* nentries);
/* offsets: nblocks+1 (0-based complex offsets); blocks: nentries complex. */
int32_t nca_coupling_download(nca_ctx* ctx, const nca_coupling* cp, int64_t* offsets, double* blocks);
int32_t nca_coupling_destroy(nca_coupling* cp);

/* ---- seam 4: the H^2 operator and its product --------------------------- */
/* Everything PetrovGalerkinNCACore holds for the far field
 * (src/nca/petrovgalerkinnca.jl:43-101, plan types src/nca/abstractnca.jl:1-108)
 * plus the near-field plan; the near blocks themselves are evaluated on the fly.
 * Tree layout: perm = values in tree order (1-based point ids), lo/hi = 1-based
 * inclusive positions of every node's values (hi = lo - 1 when empty). */
typedef struct {
    int64_t M, N, nn_test, nn_trial;
    const int64_t *test_perm, *test_lo, *test_hi;
    const int64_t *trial_perm, *trial_lo, *trial_hi;
    /* CoefficientPlan.globalidxtocoeffptr of both sides (nn+1, 1-based) */
    const int64_t *test_cptr, *trial_cptr;
    /* BasisStore: plan.nodes + blocks (test |t| x r_t, trial r_s x |s|) */
    int64_t nbasis_test, nbasis_trial;
    const int64_t *basis_test_nodes, *basis_test_off, *basis_trial_nodes, *basis_trial_off;
    const double *basis_test, *basis_trial;
    /* TransferStore: TransferTraversalPlan (leveltoidcsptr, levelidcs,
     * idxtochdsptr, chdsidcs) + blocks (test r_c x r_p, trial r_p x r_c) */
    int64_t nlev_test, nlev_trial;
    const int64_t *tt_levptr, *tt_levidcs, *tt_chdptr, *tt_chds, *tt_off;
    const int64_t *st_levptr, *st_levidcs, *st_chdptr, *st_chds, *st_off;
    const double *tt_blocks, *st_blocks;
    /* CouplingTraversalPlan (nodetofarnodesptr, farnodes); blocks from nca_coupling */
    const int64_t *cp_ptr, *cp_nodes;
    /* near interactions: test node -> near trial nodes (1-based CSR) */
    const int64_t *nearptr, *nearnodes;
} nca_h2_desc;

/* Build the device-resident operator. The coupling store is borrowed: it must
 * outlive the operator. precision: NCA_PREC_FP64 or NCA_PREC_MIXED. */
int32_t nca_h2_create(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                      const nca_points* trial, const nca_h2_desc* desc, const nca_coupling* coupling,
                      int32_t precision, nca_h2** out);
/* nca_h2_create with the basis and transfer stores packed on the device from
 * resident factors (indexed by node): basis of node t = X_t (test) or X_t^T
 * (trial); the transfer block of child c of p = the rows of X_p that hold c's
 * pivots (test) or their transpose (trial), the rows of X_p being the pivots
 * of the children in the order the transfer plan lists them. The desc's block
 * pointers (basis_*, tt_blocks, st_blocks) are not read; its offsets must fit
 * the factors (NCA_ERR_DIMENSION). The pools may be destroyed afterwards. */
int32_t nca_h2_create_resident(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                               const nca_points* trial, const nca_h2_desc* desc, const nca_coupling* coupling,
                               int32_t precision, const nca_factors* testf, const nca_factors* trialf,
                               nca_h2** out);
int32_t nca_h2_destroy(nca_h2* h2);
int32_t nca_h2_set_precision(nca_h2* h2, int32_t precision);

/* Directional (wideband) H^2: one basis per (cluster, direction). Everything
 * PetrovGalerkinNCACore holds for fardata::DirectionalData
 * (src/nca/petrovgalerkinnca.jl:43-101) with the directional plans of
 * src/nca/abstractnca.jl:5-9,41-48,78-82, all 1-based: coefficients, transfers
 * and coupling are addressed by direction index (diridx), trees, bases and the
 * near field by node. The coefficients of a node's directions must be
 * contiguous in diridx order, as plan_from_pivots gives them. */
typedef struct {
    int64_t M, N, nn_test, nn_trial, ne_test, ne_trial;
    const int64_t *test_perm, *test_lo, *test_hi;
    const int64_t *trial_perm, *trial_lo, *trial_hi;
    /* DirectionalData.nodetodirsptr of both trees (nn+1) */
    const int64_t *test_dirptr, *trial_dirptr;
    /* CoefficientPlan.globalidxtocoeffptr over direction indices (ne+1) */
    const int64_t *test_cptr, *trial_cptr;
    /* BasisStore with DirBasisTraversalPlan (nodes, idxtodiridxptr, diridx):
     * nbnodes nodes; block i (test |t| x r, trial r x |s|) of direction index
     * diridx[i] at offsets off[i] (0-based complex, nblocks+1) */
    int64_t nbnodes_test, nbnodes_trial;
    const int64_t *bt_nodes, *bt_dirptr, *bt_diridx, *bt_off;
    const int64_t *bs_nodes, *bs_dirptr, *bs_diridx, *bs_off;
    const double *bt_blocks, *bs_blocks;
    /* TransferStore with DirTransferTraversalPlan over direction indices
     * (leveltoidcsptr, levelidcs, idxtochdsptr, chdsidcs) + block offsets; the
     * group arrays of the plan are not needed: the downward pass sums per child */
    int64_t nlev_test, nlev_trial;
    const int64_t *tt_levptr, *tt_levidcs, *tt_chdptr, *tt_chds, *tt_off;
    const int64_t *st_levptr, *st_levidcs, *st_chdptr, *st_chds, *st_off;
    const double *tt_blocks, *st_blocks;
    /* DirCouplingTraversalPlan (nodetodiridxptr over test nodes, tidcs, sidcs);
     * the coupling store holds the blocks in this order, assembled by
     * nca_coupling_assemble over direction indices */
    const int64_t *cp_nodeptr, *cp_tidcs, *cp_sidcs;
    /* near interactions: test node -> near trial nodes */
    const int64_t *nearptr, *nearnodes;
} nca_dh2_desc;

/* Build the device-resident directional operator; the handle works with every
 * nca_h2_* call below. The coupling store must outlive it. */
int32_t nca_dh2_create(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                       const nca_points* trial, const nca_dh2_desc* desc, const nca_coupling* coupling,
                       int32_t precision, nca_h2** out);
/* nca_dh2_create with the stores packed on the device from resident factors
 * indexed by direction index (as nca_h2_create_resident; the bases of a
 * node's directions become the node's basis over all its coefficients). */
int32_t nca_dh2_create_resident(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                                const nca_points* trial, const nca_dh2_desc* desc, const nca_coupling* coupling,
                                int32_t precision, const nca_factors* testf, const nca_factors* trialf,
                                nca_h2** out);

/* Y = alpha * op(A) * X + beta * Y with op N/T/C, nrhs columns. Replaces
 * LinearMaps._unsafe_mul!(y, A::Union{PetrovGalerkinNCA, TransposeMap,
 * AdjointMap}, x) (src/nca/petrovgalerkinnca.jl:415-425, stages :431-589) and
 * the 5-argument mul!. Host pointers; the call returns when Y is written. */
int32_t nca_h2_matvec(nca_h2* h2, int32_t op, int64_t nrhs, const double* x, int64_t ldx, double* y,
                      int64_t ldy, const double* alpha, const double* beta);
/* Same with device pointers, enqueued on the context stream (no host sync). */
int32_t nca_h2_matvec_device(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                             int64_t ldy, const double* alpha, const double* beta);
/* Stored near field, the layout of the reference CPU path (assemblenears,
 * src/nca/abstractnca.jl:134-178: one dense block row per test cluster over
 * its near clusters' points, applied by BlockSparseMatrices at
 * src/nca/petrovgalerkinnca.jl:474). on = 1: forward products read the near
 * entries from HBM instead of evaluating them (entries made on the device
 * once per precision, 16 B (FP64) or 8 B (MIXED) each; transpose/adjoint
 * products still evaluate them); on = 0: drop them and evaluate on the fly
 * (the default). Changing it invalidates captured graphs of stored sweeps. */
int32_t nca_h2_store_near(nca_h2* h2, int32_t on);
/* No counterpart in the reference: nca_h2_matvec_device that allocates
 * nothing, for callers that cannot allocate device memory (code under hipGraph
 * capture, tight solver loops, harnesses that forbid hipMalloc). Its workspace
 * is a static 64 MiB device array per GPU, held until the process ends; a
 * product that needs more (about 16 B x nrhs per point plus 40 B x nrhs per
 * coefficient) returns NCA_ERR_LIMIT without launching anything. MIXED needs
 * the operator's c32 copies, which nca_h2_set_precision (or creation in MIXED)
 * makes; otherwise NCA_ERR_ARGUMENT. Every such call on one GPU shares the
 * array: calls whose device work overlaps on one GPU (other streams, other
 * operators) race; calls on one context stream run in order. */
int32_t nca_h2_matvec_device_noalloc(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx,
                                     void* y, int64_t ldy, const double* alpha, const double* beta);

/* hipGraph capture of one device sweep, no counterpart in the reference (its
 * CUDA extension launches each kernel from Julia): nca_h2_graph_create records
 * nca_h2_matvec_device(h2, op, nrhs, x, ldx, y, ldy, alpha, beta) once, bound to
 * those device buffers and to the current precision; nca_h2_graph_launch
 * replays it on the context stream (asynchronous, like the device matvec).
 * Launching after nca_h2_set_precision changed the precision is an
 * NCA_ERR_ARGUMENT; the graph must be destroyed before its operator. */
int32_t nca_h2_graph_create(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                            int64_t ldy, const double* alpha, const double* beta, nca_h2_graph** out);
int32_t nca_h2_graph_launch(nca_h2_graph* graph);
int32_t nca_h2_graph_destroy(nca_h2_graph* graph);

/* Autotuner, no counterpart in the reference: times the forward coupling and
 * the leaf expansion + near field over 16 right-hand sides in the current
 * precision for every launch choice (MFMA complex products with three or four
 * real products; tile orders; workgroup-to-XCD orders), `reps` timed sweeps
 * each (median), and keeps the fastest per stage. Choices are stored in the
 * operator and used by every later product; results are unchanged within the
 * precision's tolerance. out may be NULL. */
typedef struct {
    int32_t near_k3, leaf_order, near_xcd_order;
    int32_t couple_k3, couple_tile_order, couple_xcd_order;
    int32_t candidates;
    int32_t near_stagger;                  /* near-field walks start at a tile-dependent range */
    double near_ms, couple_ms;             /* stage times of the chosen configuration */
    double near_ms_start, couple_ms_start; /* stage times of the configuration before tuning */
} nca_h2_tuning;
int32_t nca_h2_autotune(nca_h2* h2, int32_t reps, nca_h2_tuning* out);
/* Applies launch choices that nca_h2_autotune returned earlier (a persisted
 * tuning table) to the operator's current precision, without timing; the
 * timing fields are not read. Captured graphs must be captured again
 * (NCA_ERR_ARGUMENT at launch). No counterpart in the reference. */
int32_t nca_h2_tuning_apply(nca_h2* h2, const nca_h2_tuning* t);

/* Measurement aid, no counterpart in the reference: one product as
 * nca_h2_matvec_device (nrhs <= 16) with HIP events recorded on the context
 * stream between the stages; stage_ms[0..5] = gather + clear, leaf projection,
 * upward transfer, coupling, downward transfer, leaf expansion + near field. */
int32_t nca_h2_matvec_profile(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                              int64_t ldy, const double* alpha, const double* beta, double* stage_ms);

typedef struct {
    int64_t coupling_blocks, coupling_entries;
    int64_t basis_entries_test, basis_entries_trial;
    int64_t transfer_entries_test, transfer_entries_trial;
    int64_t near_pairs, near_entries;
    int64_t coeffs_test, coeffs_trial;
    int64_t plan_bytes;
} nca_h2_stats;
int32_t nca_h2_get_stats(const nca_h2* h2, nca_h2_stats* out);

/* No counterpart in the reference: how the operator's forward products
 * evaluate the near field in the current precision, for measurement and
 * tests. form = 0: FP64 kernel values on the FP64 MFMA; 1: MIXED, FP32 values
 * on the f32 MFMA (or the stored near field); 2: MIXED, FP32 values split
 * into f16 parts on the f16 MFMA (the MIXED default when the near pairs'
 * distances span at most 2^17; error per product term below 2^-21 of it).
 * rmin, rmax (may be NULL): the smallest near-pair distance above the
 * zero-distance guard and the largest, measured once the MIXED copies exist
 * (before: 0 and the bound from the point boxes). */
int32_t nca_h2_near_form(nca_h2* h2, int32_t* form, double* rmin, double* rmax);
/* The same for the coupling (stage 3): form = 0 FP64 stored blocks; 1 MIXED,
 * the stored blocks' c32 copy on the f32 MFMA; 2 MIXED, the blocks evaluated
 * on the fly at the pivots (G(test pivots of t, trial pivots of s), the
 * store's entries) on the f16 MFMA with split operands, no c32 copy kept
 * (the MIXED default when the far pivot distances span at most 2^17; plain
 * and directional operators). rmin, rmax: those distances,
 * measured once the MIXED set-up ran (else 0). */
int32_t nca_h2_couple_form(nca_h2* h2, int32_t* form, double* rmin, double* rmax);

/* No counterpart in the reference: the distance form of the f16 near field
 * and coupling (forms 2 above) when every phase stays below one revolution:
 * FP32 differences of coordinates taken relative to each tile's first row,
 * instead of FP64 differences, for the tiles of the leaves (near field) and
 * blocks (coupling) whose every tile meets the condition measured at the
 * MIXED set-up: no pair strictly between 0 and the zero-distance guard, and
 * the tile's rows within 32 times its smallest pair distance of its first
 * row (a coordinate difference then errs by at most 33 * 2^-23 of the pair's
 * distance). mode = 1 uses it there (the default), 0 keeps FP64
 * differences everywhere, -1 only queries. near_on, couple_on (may be NULL):
 * whether it is in use for any tile; near_frac, couple_frac (may be NULL):
 * the fraction of tiles (both directions) that meet the condition (MIXED
 * only, else 0). Changing the mode invalidates captured graphs. */
int32_t nca_h2_rel32(nca_h2* h2, int32_t mode, int32_t* near_on, int32_t* couple_on, double* near_frac,
                     double* couple_frac);

#ifdef __cplusplus
}
#endif
#endif
