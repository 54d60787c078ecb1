"""Batched bottom-up compressor (src/compressors/batchedbottomup.jl) on the GPU.

The level driver keeps the structure of testbases/trialbases(::BatchedBottomUp)
(batchedbottomup.jl:246-447); Phase 2 and Phase 3 go through the two seams
assemble_level (:101-116) and compresslevel_test!/trial! (:145-244), which here
call the HIP C-ABI. Index sets are numpy CSR arrays instead of
Vector{Vector{Int}}; node ids are 0-based in Python.
"""
import ctypes
import os
import time

import numpy as np

from . import _abi
from ._native import hip_lib
from .runtime import call, _p
from . import trees as _trees


def concat_ranges(starts, lengths):
    """Indices of the concatenated ranges [s, s + l)."""
    lengths = np.asarray(lengths, np.int64)
    starts = np.asarray(starts, np.int64)
    total = int(lengths.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    ends = np.cumsum(lengths)
    out = np.ones(total, np.int64)
    nz = lengths > 0
    s_nz, l_nz, e_nz = starts[nz], lengths[nz], ends[nz]
    out[0] = s_nz[0]
    if len(s_nz) > 1:
        out[e_nz[:-1]] = s_nz[1:] - (s_nz[:-1] + l_nz[:-1] - 1)
    return np.cumsum(out)


class ACA:
    """Fully pivoted ACA with the Frobenius-norm stop, the dense compressor of
    BatchedBottomUp: ACA(FullPivoting(), FullPivoting(), FNormEstimator(tol))."""

    def __init__(self, tol=1e-4, rowpivoting="full", columnpivoting="full"):
        if rowpivoting != "full" or columnpivoting != "full":
            raise ValueError("the batched compressor takes fully pivoted ACA")
        self.tol = float(tol)


class TopDown:
    """TopDown(factorization=ACA(), representor=None) (topdown.jl:26-47): the
    nested bases from the root to the leaves, each cluster factorized against
    the points of its far nodes and the column pivots of its parent with the
    partially pivoted ACA (nca_topdown_side). Only the ACA factorization and no
    representor are taken on the device."""

    def __init__(self, factorization=None, representor=None):
        factorization = ACA() if factorization is None else factorization
        if not isinstance(factorization, ACA):
            raise ValueError("TopDown on the device takes ACA() (the reference's default factorization)")
        if representor is not None:
            raise ValueError("TopDown on the device takes no representor")
        self.factorization = factorization
        self.representor = representor


class IACA:
    """IACA(MaximumValue(), TreeMimicryPivoting(...), FNormExtrapolator(tol))
    as BottomUp takes it (bottomup.jl:214-424, pivots.jl:215-364): the column
    pivots follow the geometric far-pivot order of the Phase-1 selector (the
    stand-in for tree mimicry), the row pivots are the maxima of the residual
    columns, and the stop compares the candidate's residual column with the
    mean accepted column (see k_iaca). On the trial side the blocks are
    assembled transposed, so the same rule picks the cluster's columns by
    maximum value and the far rows geometrically, as IACA(trialpivoting,
    MaximumValue(), ...) does."""

    def __init__(self, tol=1e-4):
        self.tol = float(tol)


class GeometricPivoting:
    """Phase-1 far-pivot selector (stand-in for TreeMimicryPivoting; see
    nca_tree.cpp): up to min(maxrank, oversample |t|) far points per cluster,
    spread by maximal separation relative to (distance to the cluster)^weight.
    The default oversample = 1 is the reference's count rule
    min(maxrank, |t|) (batchedbottomup.jl:57, _clustersize pivots.jl:15-25);
    oversample > 1 draws more far pivots (an option, not the reference's rule).
    device: select on the GPU (nca_farpivots) when the caller passes the other
    side's device points, else on the host; both give the same pivots for
    weights 1 and 1.5."""

    def __init__(self, weight=1.5, oversample=1.0, device=True):
        self.weight = float(weight)
        self.oversample = float(oversample)
        self.device = bool(device)

    def select(self, selftree, othertree, fd, maxrank, otherpoints=None):
        if self.device and otherpoints is not None:
            return device_farpivots_tree(otherpoints, selftree, fd, othertree, maxrank, self.weight,
                                         self.oversample)
        return _trees.farpivots(selftree, othertree, fd, maxrank, self.weight, self.oversample)


def device_farpivots(otherpoints, centers, nvals, ffptr, ffnodes, othertree, maxrank, weight, oversample):
    """Phase 1 on the device (nca_farpivots) for items with explicit far fields
    (0-based CSR of nodes of othertree), centers (nitems x 3) and sizes nvals.
    Returns a 0-based CSR of point ids."""
    n = len(ffptr) - 1
    c = np.ascontiguousarray(centers, np.float64)
    nv = np.ascontiguousarray(nvals, np.int64)
    fp1 = np.ascontiguousarray(np.asarray(ffptr) + 1, np.int64)
    fn1 = np.ascontiguousarray(np.asarray(ffnodes) + 1, np.int64)
    lay = othertree.layout_1based()
    ptr = np.zeros(n + 1, np.int64)
    cap = max(n * int(maxrank), 1)
    piv = np.zeros(cap, np.int64)
    call(hip_lib().nca_farpivots(otherpoints.ctx.handle, otherpoints.handle, n, _p(c), _p(nv), _p(fp1),
                                 _p(fn1) if len(fn1) else None, othertree.nnodes, _p(lay["nodelo"]),
                                 _p(lay["nodehi"]), _p(lay["perm"]), int(maxrank), float(weight), float(oversample),
                                 _p(ptr), cap, _p(piv)))
    return ptr, piv[:ptr[-1]].copy()


def device_farpivots_tree(otherpoints, selftree, fd, othertree, maxrank, weight, oversample):
    """Phase 1 on the device (nca_farpivots_tree) for every node of selftree,
    its far field (own far nodes, then its ancestors') built on the device
    from the FarData CSR and the parent array. Returns a 0-based CSR of point
    ids, equal to device_farpivots on farfield_csr(selftree, fd)."""
    n = selftree.nnodes
    c = np.ascontiguousarray(selftree.center, np.float64)
    nv = np.ascontiguousarray(selftree.hi - selftree.lo, np.int64)
    par1 = np.ascontiguousarray(np.asarray(selftree.parent, np.int64) + 1)
    fp1 = np.ascontiguousarray(np.asarray(fd.nodetofarnodesptr, np.int64) + 1)
    fn1 = np.ascontiguousarray(np.asarray(fd.farnodes, np.int64) + 1)
    lay = othertree.layout_1based()
    ptr = np.zeros(n + 1, np.int64)
    cap = max(n * int(maxrank), 1)
    piv = np.zeros(cap, np.int64)
    call(hip_lib().nca_farpivots_tree(otherpoints.ctx.handle, otherpoints.handle, n, _p(c), _p(nv), _p(par1),
                                      _p(fp1), _p(fn1) if len(fn1) else None, othertree.nnodes,
                                      _p(lay["nodelo"]), _p(lay["nodehi"]), _p(lay["perm"]), int(maxrank),
                                      float(weight), float(oversample), _p(ptr), cap, _p(piv)))
    return ptr, piv[:ptr[-1]].copy()


class BatchedBottomUp:
    """BatchedBottomUp(; pivotselector, compressor) (batchedbottomup.jl:30-51)."""

    def __init__(self, pivotselector=None, compressor=None):
        if pivotselector is None or compressor is None:
            raise TypeError("BatchedBottomUp needs the keywords pivotselector and compressor")
        self.pivotselector = pivotselector
        self.compressor = compressor


def _timing():
    return os.environ.get("NCA_BATCHED_TIMING", "0") == "1"


# --------------------------------------------------------------------------
# Seam 1: Phase 2.
class LevelStack:
    """Device-resident level buffer returned by assemble_level for a HIP kernel
    matrix; its type routes compresslevel_*! to the HIP methods."""

    def __init__(self, farmatrix, handle, nb, rowoffsets, ncols, maxrank):
        self.farmatrix = farmatrix
        self.handle = handle
        self.nb = nb
        self.rowoffsets = rowoffsets
        self.ncols = ncols
        self.maxrank = maxrank

    def download(self):
        ld = int(self.rowoffsets[-1])
        buf = np.zeros((ld, self.maxrank), np.complex128, order="F")
        call(hip_lib().nca_stack_download(self.farmatrix.ctx.handle, self.handle, _p(buf)))
        return buf

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_stack_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def assemble_level(farmatrix, rowptr, rowidx, colptr, colidx, maxrank, transposed=False):
    """Phase 2 for one level: block b = A[rowidx[rowptr[b]:rowptr[b+1]], colidx[...]]
    (0-based CSR of 0-based point ids), or its transposed-assembly form."""
    nb = len(rowptr) - 1
    rp1 = np.ascontiguousarray(rowptr + 1, np.int64)
    cp1 = np.ascontiguousarray(colptr + 1, np.int64)
    ri1 = np.ascontiguousarray(rowidx + 1, np.int64)
    ci1 = np.ascontiguousarray(colidx + 1, np.int64)
    h = ctypes.c_void_p()
    kern = farmatrix.kernel.abi()
    call(hip_lib().nca_level_assemble(farmatrix.ctx.handle, ctypes.byref(kern), farmatrix.test.handle,
                                      farmatrix.trial.handle, nb, _p(rp1), _p(ri1) if len(ri1) else None, _p(cp1),
                                      _p(ci1) if len(ci1) else None, int(maxrank), int(bool(transposed)),
                                      ctypes.byref(h)))
    return LevelStack(farmatrix, h, nb, np.asarray(rowptr, np.int64).copy(), np.diff(colptr), maxrank)


def stack_from_host(farmatrix, buffer, rowoffsets, ncols, maxrank):
    """A LevelStack from host data in the level-buffer layout (for tests)."""
    buf = np.asfortranarray(np.asarray(buffer, dtype=np.complex128))
    ro = np.ascontiguousarray(rowoffsets, np.int64)
    nc = np.ascontiguousarray(ncols, np.int64)
    h = ctypes.c_void_p()
    call(hip_lib().nca_stack_upload(farmatrix.ctx.handle, _p(buf), len(nc), _p(ro), _p(nc), int(maxrank),
                                    ctypes.byref(h)))
    return LevelStack(farmatrix, h, len(nc), ro, nc, maxrank)


def aca_fullpiv(stack, tol, sampled=True):
    """Batched ACA on a stack: (ranks, rowpivots, colpivots) with maxrank x nb
    1-based pivots as the device method returns them."""
    nb, mr = stack.nb, stack.maxrank
    ranks = np.zeros(max(nb, 1), np.int32)
    rp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    cp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    call(hip_lib().nca_aca_fullpiv(stack.farmatrix.ctx.handle, stack.handle, float(tol), int(bool(sampled)),
                                   _p(ranks), _p(rp), _p(cp)))
    return ranks[:nb], rp[:, :nb], cp[:, :nb]


def skeleton_solve(stack, ranks, rowpivots, colpivots):
    """X_b = L_b / P_b per block; returns (xoffsets, X, independent)."""
    nb = stack.nb
    m = np.diff(stack.rowoffsets)
    xoff = np.zeros(nb + 1, np.int64)
    np.cumsum(m * np.asarray(ranks, np.int64), out=xoff[1:])
    X = np.zeros(max(int(xoff[-1]), 1), np.complex128)
    indep = np.zeros(max(nb, 1), np.int64)
    rk = np.ascontiguousarray(ranks, np.int32)
    rp = np.asfortranarray(rowpivots, np.int64)
    cp = np.asfortranarray(colpivots, np.int64)
    status = hip_lib().nca_skeleton_solve(stack.farmatrix.ctx.handle, stack.handle, _p(rk), _p(rp), _p(cp),
                                          _p(xoff), _p(X), _p(indep))
    return xoff, X[:xoff[-1]], indep[:nb], status


def iaca(stack, tol):
    """Batched IACA on a stack (columns = candidates in geometric order):
    (ranks, rowpivots, colpivots), maxrank x nb, 1-based."""
    nb, mr = stack.nb, stack.maxrank
    ranks = np.zeros(max(nb, 1), np.int32)
    rp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    cp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    call(hip_lib().nca_iaca(stack.farmatrix.ctx.handle, stack.handle, float(tol), _p(ranks), _p(rp), _p(cp)))
    return ranks[:nb], rp[:, :nb], cp[:, :nb]


def compress_level(stack, tol, sampled=True, method="aca"):
    """ACA (or IACA, method="iaca") + skeleton solve in one call (pivots stay
    on the device in between)."""
    nb, mr = stack.nb, stack.maxrank
    m = np.diff(stack.rowoffsets)
    cap = int(np.sum(m * np.minimum(np.minimum(m, stack.ncols), mr))) if nb else 0
    ranks = np.zeros(max(nb, 1), np.int32)
    rp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    cp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    xoff = np.zeros(nb + 1, np.int64)
    X = np.zeros(max(cap, 1), np.complex128)
    indep = np.zeros(max(nb, 1), np.int64)
    if method == "iaca":
        call(hip_lib().nca_compress_level_iaca(stack.farmatrix.ctx.handle, stack.handle, float(tol), _p(ranks),
                                               _p(rp), _p(cp), cap, _p(xoff), _p(X), _p(indep)))
    else:
        call(hip_lib().nca_compress_level(stack.farmatrix.ctx.handle, stack.handle, float(tol),
                                          int(bool(sampled)), _p(ranks), _p(rp), _p(cp), cap, _p(xoff), _p(X),
                                          _p(indep)))
    return ranks[:nb], rp[:, :nb], cp[:, :nb], xoff, X[:xoff[-1]]


class FactorPool:
    """Skeleton factors kept on the device (nca_factors) for block ids
    0..nids-1: each compress_level_resident call adds one level."""

    def __init__(self, ctx, nids):
        self.ctx = ctx
        self.nids = int(nids)
        h = ctypes.c_void_p()
        call(hip_lib().nca_factors_create(ctx.handle, self.nids, ctypes.byref(h)))
        self.handle = h

    def size(self):
        n = ctypes.c_int64()
        call(hip_lib().nca_factors_size(self.handle, ctypes.byref(n)))
        return n.value

    def download(self):
        """All factors, the levels in call order (the X of compress_level)."""
        X = np.zeros(max(self.size(), 1), np.complex128)
        call(hip_lib().nca_factors_download(self.ctx.handle, self.handle, _p(X)))
        return X[:self.size()]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_factors_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def compress_level_resident(stack, tol, pool, ids, sampled=True, method="aca"):
    """compress_level with X kept in ``pool``: block b becomes the factor of
    ids[b] (0-based). Returns (ranks, rowpivots, colpivots, xoff) with xoff the
    offsets X would have in compress_level's output."""
    nb, mr = stack.nb, stack.maxrank
    ranks = np.zeros(max(nb, 1), np.int32)
    rp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    cp = np.zeros((mr, max(nb, 1)), np.int64, order="F")
    indep = np.zeros(max(nb, 1), np.int64)
    ids1 = np.ascontiguousarray(np.asarray(ids, np.int64) + 1) if nb else np.zeros(1, np.int64)
    call(hip_lib().nca_compress_level_resident(stack.farmatrix.ctx.handle, stack.handle, float(tol),
                                               int(bool(sampled)), 1 if method == "iaca" else 0, pool.handle,
                                               _p(ids1), _p(ranks), _p(rp), _p(cp), _p(indep)))
    ranks = ranks[:nb]
    xoff = np.zeros(nb + 1, np.int64)
    np.cumsum(np.diff(stack.rowoffsets) * ranks, out=xoff[1:])
    return ranks, rp[:, :nb], cp[:, :nb], xoff


# --------------------------------------------------------------------------
class SideFactors:
    """Result of one side: pivots per node (0-based CSR of point ids; nodes
    without far field are unassigned), and per active node its block X (m x r,
    column-major) together with the child ranges of its rows."""

    def __init__(self, nn):
        self.nn = nn
        self.assigned = np.zeros(nn, bool)
        self.rank = np.zeros(nn, np.int64)
        self.pivots = [None] * nn
        self.X = [None] * nn
        self.m = np.zeros(nn, np.int64)
        self.level_transfer_nodes = {}

    def pivot_csr(self):
        ptr = np.zeros(self.nn + 1, np.int64)
        np.cumsum(self.rank, out=ptr[1:])
        idx = np.concatenate([p for p in self.pivots if p is not None and len(p)] or [np.zeros(0, np.int64)])
        return ptr, idx.astype(np.int64)


class LevelFactors:
    """SideFactors of the batched driver in flat arrays: per node its rank, rows
    m, pivots (a slice of one flat point-id array) and its block X (m x r,
    column-major) at xoff[t] of the packed factors of all levels. The per-node
    views pivots[t] and X[t] read like SideFactors."""

    def __init__(self, nn):
        self.nn = nn
        self.assigned = np.zeros(nn, bool)
        self.rank = np.zeros(nn, np.int64)
        self.m = np.zeros(nn, np.int64)
        self.pstart = np.zeros(nn, np.int64)
        self.xoff = np.zeros(nn, np.int64)
        self._piv = [np.zeros(0, np.int64)]
        self._npiv = 0
        self._X = [np.zeros(0, np.complex128)]
        self._nx = 0
        self.pool = None  # a FactorPool when the factors stay on the device
        self.level_transfer_nodes = {}
        self.pivots = _NodeView(self, "pivots")
        self.X = _NodeView(self, "X")

    @property
    def pivflat(self):
        if len(self._piv) > 1:
            self._piv = [np.concatenate(self._piv)]
        return self._piv[0]

    @property
    def Xflat(self):
        if self.pool is not None and self._nx and not len(self._X[0]):
            self._X = [self.pool.download()]  # resident factors: downloaded once, when read
        if len(self._X) > 1:
            self._X = [np.concatenate(self._X)]
        return self._X[0]

    def add_level(self, nodes, rows, ranks, pivots, xoff_level, X_level):
        """Record the blocks of one level (nodes in block order)."""
        k = len(nodes)
        ranks = np.asarray(ranks, np.int64)
        self.assigned[nodes] = True
        self.rank[nodes] = ranks
        self.m[nodes] = rows
        self.pstart[nodes] = self._npiv + np.cumsum(ranks) - ranks
        self._piv.append(np.asarray(pivots, np.int64))
        self._npiv += len(pivots)
        self.xoff[nodes] = self._nx + xoff_level[:k]
        if X_level is not None:
            self._X.append(np.asarray(X_level, np.complex128))
        self._nx += int(xoff_level[k])

    def pivot_csr(self):
        ptr = np.zeros(self.nn + 1, np.int64)
        r = np.where(self.assigned, self.rank, 0)
        np.cumsum(r, out=ptr[1:])
        return ptr, self.pivflat[concat_ranges(self.pstart, r)]


class _NodeView:
    def __init__(self, f, what):
        self.f, self.what = f, what

    def __getitem__(self, t):
        f = self.f
        if not f.assigned[t]:
            return None
        r, m = int(f.rank[t]), int(f.m[t])
        if self.what == "pivots":
            return f.pivflat[f.pstart[t]:f.pstart[t] + r]
        return f.Xflat[f.xoff[t]:f.xoff[t] + m * r].reshape((m, r), order="F")


def _level_rows(tree, f, active, leaf):
    """Rows of every active block: a leaf's points, else the concatenated
    pivots of its children in ChildIterator order (batchedbottomup.jl:282-303).
    Returns (rowptr, rowidx)."""
    isleaf = leaf[active]
    nch = np.diff(tree.childptr)[active]
    nseg = np.where(isleaf, 1, nch)
    segptr = np.zeros(len(active) + 1, np.int64)
    np.cumsum(nseg, out=segptr[1:])
    seg_start = np.zeros(segptr[-1], np.int64)
    seg_len = np.zeros(segptr[-1], np.int64)
    lf = active[isleaf]
    seg_start[segptr[:-1][isleaf]] = tree.lo[lf]
    seg_len[segptr[:-1][isleaf]] = tree.hi[lf] - tree.lo[lf]
    inner = ~isleaf
    if np.any(inner):
        ch = tree.children[concat_ranges(tree.childptr[active[inner]], nch[inner])]
        pos = concat_ranges(segptr[:-1][inner], nch[inner])
        npts = tree.perm.shape[0]
        seg_start[pos] = npts + f.pstart[ch]
        seg_len[pos] = np.where(f.assigned[ch], f.rank[ch], 0)
    src = np.concatenate([tree.perm.astype(np.int64), f.pivflat])
    counts = np.add.reduceat(seg_len, segptr[:-1]) if len(seg_len) else np.zeros(len(active), np.int64)
    counts = np.where(nseg > 0, counts, 0)
    rowptr = np.zeros(len(active) + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return rowptr, src[concat_ranges(seg_start, seg_len)]


def _compresslevel(cfg, storage, active, rowptr, rowidx, out):
    """compresslevel_test!/trial! for a HIP LevelStack (batchedbottomup.jl:145-244,
    compresslevel.jl:242-316): batched ACA + skeleton solve, then the pivots
    tpivots[t] = tvalues[t][rowpivots] and the factor X of every block (kept
    in out.pool on the device when the side has one)."""
    method = "iaca" if isinstance(cfg, IACA) else "aca"
    if out.pool is not None:
        ranks, rp, cp, xoff = compress_level_resident(storage, cfg.tol, out.pool, active, sampled=True,
                                                      method=method)
        X = None
    else:
        ranks, rp, cp, xoff, X = compress_level(storage, cfg.tol, sampled=True, method=method)
    ranks = ranks.astype(np.int64)
    m = np.diff(storage.rowoffsets)
    mr = rp.shape[0]
    take = np.arange(mr)[None, :] < ranks[:, None]
    pos = (rowptr[:-1, None] + rp.T - 1)[take]
    out.add_level(active, m, ranks, rowidx[pos], xoff, X)


def _bottomup_fused(compressor, farmatrix, selftree, side, maxrank, farpivots, out):
    """All levels of batched_side in one call (nca_bottomup_side): one launch
    per level assembles, compresses and solves every block, the rows of a
    level built on the device from the pivots of the level below, the factors
    kept in out.pool. Fills out as the level loop would."""
    fptr, fidx = farpivots
    nfar = np.diff(fptr)
    leaf = selftree.leafmask()
    levs = sorted(selftree.levels().tolist(), reverse=True)
    actives = []
    for lev in levs:
        nodes = selftree.levelnodes(lev)
        actives.append((lev, nodes, nodes[nfar[nodes] > 0]))
    levptr = np.zeros(len(actives) + 1, np.int64)
    np.cumsum([len(a) for _, _, a in actives], out=levptr[1:])
    levids = np.concatenate([a for _, _, a in actives] + [np.zeros(0, np.int64)]).astype(np.int64)
    nn = selftree.nnodes
    selfpts, otherpts = (farmatrix.test, farmatrix.trial) if side == 0 else (farmatrix.trial, farmatrix.test)
    i64 = lambda a: np.ascontiguousarray(a, np.int64) if len(a) else np.zeros(1, np.int64)
    arr = dict(cp=i64(selftree.childptr + 1), ch=i64(selftree.children + 1), lo=i64(selftree.lo + 1),
               hi=i64(selftree.hi), perm=i64(selftree.perm + 1), levptr=i64(levptr + 1), levids=i64(levids + 1),
               fptr=i64(np.asarray(fptr) + 1), fidx=i64(np.asarray(fidx) + 1))
    ranks = np.zeros(max(nn, 1), np.int32)
    mrows = np.zeros(max(nn, 1), np.int64)
    pivcap = int(np.minimum(nfar[levids], maxrank).sum()) if len(levids) else 0
    piv = np.empty(max(pivcap, 1), np.int64)
    indep = np.zeros(max(nn, 1), np.int64)
    cfg = compressor.compressor
    kern = farmatrix.kernel.abi()
    call(hip_lib().nca_bottomup_side(
        farmatrix.ctx.handle, ctypes.byref(kern), selfpts.handle, otherpts.handle, int(side == 1), nn,
        _p(arr["cp"]), _p(arr["ch"]), _p(arr["lo"]), _p(arr["hi"]), _p(arr["perm"]), len(actives),
        _p(arr["levptr"]), _p(arr["levids"]), _p(arr["fptr"]), _p(arr["fidx"]), int(maxrank), float(cfg.tol), 1,
        1 if isinstance(cfg, IACA) else 0, out.pool.handle, _p(ranks), _p(mrows), pivcap, _p(piv), _p(indep)))
    p0 = 0
    for lev, nodes, active in actives:
        if len(active):
            r = ranks[active].astype(np.int64)
            m = mrows[active]
            xoff = np.zeros(len(active) + 1, np.int64)
            np.cumsum(m * r, out=xoff[1:])
            p1 = p0 + int(r.sum())
            out.add_level(active, m, r, piv[p0:p1] - 1, xoff, None)
            p0 = p1
            out.level_transfer_nodes[lev] = nodes[~leaf[nodes] & out.assigned[nodes]].tolist()
    return out


def topdown_side(compressor, farmatrix, tree, fd, side, maxrank=50):
    """testbases (side 0) / trialbases (side 1) of TopDown: LevelFactors with
    the factors on the device (pool), the row pivots as pivots and the column
    (far) pivots in colpivots (0-based CSR). fd: the far data of this side."""
    selftree = tree.testtree if side == 0 else tree.trialtree
    othertree = tree.trialtree if side == 0 else tree.testtree
    nn = selftree.nnodes
    leaf = selftree.leafmask()
    levs = sorted(selftree.levels().tolist())
    levnodes = [selftree.levelnodes(lev) for lev in levs]
    levptr = np.zeros(len(levs) + 1, np.int64)
    np.cumsum([len(x) for x in levnodes], out=levptr[1:])
    levids = np.concatenate(levnodes + [np.zeros(0, np.int64)]).astype(np.int64)
    selfpts, otherpts = (farmatrix.test, farmatrix.trial) if side == 0 else (farmatrix.trial, farmatrix.test)
    i64 = lambda a: np.ascontiguousarray(a, np.int64) if len(a) else np.zeros(1, np.int64)
    arr = dict(pa=i64(selftree.parent + 1), cp=i64(selftree.childptr + 1), ch=i64(selftree.children + 1),
               lo=i64(selftree.lo + 1), hi=i64(selftree.hi), perm=i64(selftree.perm + 1),
               olo=i64(othertree.lo + 1), ohi=i64(othertree.hi), operm=i64(othertree.perm + 1),
               fptr=i64(fd.nodetofarnodesptr + 1), fnod=i64(fd.farnodes + 1), levptr=i64(levptr + 1),
               levids=i64(levids + 1))
    out = LevelFactors(nn)
    out.pool = FactorPool(farmatrix.ctx, nn)
    ranks = np.zeros(max(nn, 1), np.int32)
    mrows = np.zeros(max(nn, 1), np.int64)
    pivcap = int(np.minimum(selftree.hi - selftree.lo, maxrank).sum())
    piv = np.empty(max(pivcap, 1), np.int64)
    cpiv = np.empty(max(pivcap, 1), np.int64)
    indep = np.zeros(max(nn, 1), np.int64)
    kern = farmatrix.kernel.abi()
    call(hip_lib().nca_topdown_side(
        farmatrix.ctx.handle, ctypes.byref(kern), selfpts.handle, otherpts.handle, int(side == 1), nn,
        _p(arr["pa"]), _p(arr["cp"]), _p(arr["ch"]), _p(arr["lo"]), _p(arr["hi"]), _p(arr["perm"]),
        othertree.nnodes, _p(arr["olo"]), _p(arr["ohi"]), _p(arr["operm"]), _p(arr["fptr"]), _p(arr["fnod"]),
        len(levs), _p(arr["levptr"]), _p(arr["levids"]), int(maxrank), float(compressor.factorization.tol),
        out.pool.handle, _p(ranks), _p(mrows), pivcap, _p(piv), _p(cpiv), _p(indep)))
    p0 = 0
    cranks = np.zeros(nn, np.int64)
    cparts = []
    for lev, nodes in zip(levs, levnodes):
        active = nodes[ranks[nodes] >= 0]
        if len(active):
            r = ranks[active].astype(np.int64)
            m = mrows[active]
            xoff = np.zeros(len(active) + 1, np.int64)
            np.cumsum(m * r, out=xoff[1:])
            p1 = p0 + int(r.sum())
            out.add_level(active, m, r, piv[p0:p1] - 1, xoff, None)
            cparts.append((active, r, cpiv[p0:p1] - 1))
            cranks[active] = r
            p0 = p1
            out.level_transfer_nodes[lev] = nodes[~leaf[nodes] & out.assigned[nodes]].tolist()
    # column pivots by node (0-based CSR)
    cptr = np.zeros(nn + 1, np.int64)
    np.cumsum(cranks, out=cptr[1:])
    cidx = np.zeros(int(cptr[-1]), np.int64)
    for active, r, c in cparts:
        cidx[concat_ranges(cptr[active], r)] = c
    out.colpivots = (cptr, cidx)
    return out


def batched_side(compressor, farmatrix, tree, fd, side, maxrank=50, farpivots=None, resident=False, fused=None):
    """testbases (side 0) / trialbases (side 1) of BatchedBottomUp, returning
    LevelFactors. farpivots: optional Phase-1 result (0-based CSR). resident:
    the factors stay on the device (LevelFactors.pool) and are downloaded only
    when read. fused (default: resident): all levels in one call, one launch
    per level (nca_bottomup_side); else one assembly and one compression call
    per level."""
    selftree = tree.testtree if side == 0 else tree.trialtree
    othertree = tree.trialtree if side == 0 else tree.testtree
    if farpivots is None:
        farpivots = compressor.pivotselector.select(selftree, othertree, fd, maxrank,
                                                    farmatrix.trial if side == 0 else farmatrix.test)
    fptr, fidx = farpivots
    nfar = np.diff(fptr)
    out = LevelFactors(selftree.nnodes)
    out.farpivots = farpivots
    if resident:
        out.pool = FactorPool(farmatrix.ctx, selftree.nnodes)
    if fused is None:
        fused = resident
    if fused:
        if not resident:
            raise ValueError("batched_side: the fused driver keeps the factors on the device (resident=True)")
        t0 = time.perf_counter()
        _bottomup_fused(compressor, farmatrix, selftree, side, maxrank, farpivots, out)
        if _timing():
            print(f"[batchedbottomup {'test' if side == 0 else 'trial'}] fused levels: "
                  f"{time.perf_counter() - t0:.4f}s", flush=True)
        return out
    leaf = selftree.leafmask()
    for lev in sorted(selftree.levels().tolist(), reverse=True):
        nodes = selftree.levelnodes(lev)
        active = nodes[nfar[nodes] > 0]
        if len(active) == 0:
            continue
        t0 = time.perf_counter()
        rowptr, rowidx = _level_rows(selftree, out, active, leaf)
        colptr = np.zeros(len(active) + 1, np.int64)
        np.cumsum(nfar[active], out=colptr[1:])
        colidx = fidx[concat_ranges(fptr[active], nfar[active])]
        storage = assemble_level(farmatrix, rowptr, rowidx, colptr, colidx, maxrank, transposed=(side == 1))
        t1 = time.perf_counter()
        _compresslevel(compressor.compressor, storage, active, rowptr, rowidx, out)
        t2 = time.perf_counter()
        if _timing():
            print(f"[batchedbottomup {'test' if side == 0 else 'trial'}] level {lev + 1}: {len(active)} blocks | "
                  f"assemble={t1 - t0:.4f}s | compress(aca+basis/transfer)={t2 - t1:.4f}s", flush=True)
        out.level_transfer_nodes[lev] = nodes[~leaf[nodes] & out.assigned[nodes]].tolist()
    return out


class BasisStore:
    def __init__(self, nodes, offsets, blocks):
        self.nodes = nodes        # 0-based node ids
        self.offsets = offsets    # nblocks+1 complex offsets
        self.blocks = blocks      # packed column-major blocks


class TransferStore:
    def __init__(self, leveltoidcsptr, levelidcs, idxtochdsptr, chdsidcs, offsets, blocks):
        self.leveltoidcsptr = leveltoidcsptr  # 0-based
        self.levelidcs = levelidcs
        self.idxtochdsptr = idxtochdsptr
        self.chdsidcs = chdsidcs
        self.offsets = offsets
        self.blocks = blocks


def _offsets(sizes):
    off = np.zeros(len(sizes) + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    return off


def _gather_blocks(src, src_off, ld, rows, cols, transpose):
    """Pack complex blocks rows[b] x cols[b] (column-major) read from src at
    src_off[b] with leading dimension ld[b], as is or transposed (host lib)."""
    from ._native import tree_lib
    sz = rows * cols
    dst_off = np.zeros(len(sz) + 1, np.int64)
    np.cumsum(sz, out=dst_off[1:])
    out = np.empty(max(int(dst_off[-1]), 1), np.complex128)
    arrs = [np.ascontiguousarray(a, np.int64) for a in (src_off, ld, rows, cols, dst_off)]
    if len(sz):
        tree_lib().ncat_gather_blocks(len(sz), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), int(transpose),
                                      _p(src), _p(arrs[4]), _p(out))
    return dst_off, out[:dst_off[-1]]


def _stores_flat(f, tree, side, blocks=True):
    """stores_from_factors for LevelFactors: the same layout and order as the
    per-node path below, packed by the host library from the flat factors.
    blocks=False: plans and offsets only (blocks None), for stores that the
    device packs from resident factors."""
    leaf = tree.leafmask()
    Xf = np.ascontiguousarray(f.Xflat) if blocks else None
    # bases: leaves with a block, node order; test X (m x r), trial X^T (r x m)
    bn = np.nonzero(f.assigned & leaf)[0]
    m, r, xo = f.m[bn], f.rank[bn], f.xoff[bn]
    if not blocks:
        boff, bl = _offsets(m * r), None
    elif side == 0:
        boff, bl = _gather_blocks(Xf, xo, m, m, r, False)
    else:
        boff, bl = _gather_blocks(Xf, xo, m, r, m, True)
    basis = BasisStore(bn.astype(np.int64), boff, bl)
    # transfers: levels ascending, parents in level_transfer_nodes order, all children
    parents = [np.asarray(f.level_transfer_nodes.get(lev, []), np.int64) for lev in sorted(tree.levels().tolist())]
    levptr = np.zeros(len(parents) + 1, np.int64)
    np.cumsum([len(p_) for p_ in parents], out=levptr[1:])
    P = np.concatenate(parents) if len(parents) else np.zeros(0, np.int64)
    nch = np.diff(tree.childptr)[P]
    chdptr = np.zeros(len(P) + 1, np.int64)
    np.cumsum(nch, out=chdptr[1:])
    C = tree.children[concat_ranges(tree.childptr[P], nch)] if len(P) else np.zeros(0, np.int64)
    Pe = np.repeat(P, nch)                       # parent of every edge
    rc = np.where(f.assigned[C], f.rank[C], 0)
    # row offset of each child inside its parent's block: exclusive sum of rc per parent
    csum = np.cumsum(rc)
    first = np.repeat(chdptr[:-1], nch)
    oc = csum - rc - np.where(first > 0, csum[np.maximum(first - 1, 0)], 0) if len(C) else np.zeros(0, np.int64)
    rp, mp, xp = f.rank[Pe], f.m[Pe], f.xoff[Pe]
    if not blocks:
        toff, tl = _offsets(rc * rp), None
    elif side == 0:   # E = X_p[oc : oc + rc, :] (rc x rp)
        toff, tl = _gather_blocks(Xf, xp + oc, mp, rc, rp, False)
    else:           # E^T (rp x rc)
        toff, tl = _gather_blocks(Xf, xp + oc, mp, rp, rc, True)
    transfer = TransferStore(levptr, P, chdptr, C.astype(np.int64), toff, tl)
    return basis, transfer


def stores_from_factors(f, tree, side, blocks=True):
    """_build_basis_storage / _build_transfer_storage (nestedbasis.jl:1-4,
    transfermatrices.jl:58-95) from the block factors X: test basis X (m x r),
    trial basis X^T; test transfer X[childrows, :], trial transfer its transpose.
    blocks=False (LevelFactors only): plans and offsets, blocks None."""
    if isinstance(f, LevelFactors):
        return _stores_flat(f, tree, side, blocks)
    leaf = tree.leafmask()
    bnodes = [t for t in range(f.nn) if f.assigned[t] and leaf[t]]
    bl, boff = [], [0]
    for t in bnodes:
        X = f.X[t]
        B = X if side == 0 else X.T
        bl.append(np.asfortranarray(B).ravel(order="F"))
        boff.append(boff[-1] + X.size)
    basis = BasisStore(np.asarray(bnodes, np.int64), np.asarray(boff, np.int64),
                       np.concatenate(bl) if bl else np.zeros(0, np.complex128))
    levels = sorted(tree.levels().tolist())
    levptr, levidcs, chdptr, chds, tl, toff = [0], [], [0], [], [], [0]
    for lev in levels:
        for p in f.level_transfer_nodes.get(lev, []):
            levidcs.append(p)
            X = f.X[p]
            o = 0
            for c in tree.childrenof(p):
                rc = int(f.rank[c]) if f.assigned[c] else 0
                E = X[o:o + rc, :]
                o += rc
                B = E if side == 0 else E.T
                tl.append(np.asfortranarray(B).ravel(order="F"))
                toff.append(toff[-1] + E.size)
                chds.append(int(c))
            chdptr.append(len(chds))
        levptr.append(len(levidcs))
    transfer = TransferStore(np.asarray(levptr, np.int64), np.asarray(levidcs, np.int64),
                             np.asarray(chdptr, np.int64), np.asarray(chds, np.int64), np.asarray(toff, np.int64),
                             np.concatenate(tl) if tl else np.zeros(0, np.complex128))
    return basis, transfer


def testbases(compressor, farmatrix, tree, fardata, maxrank=50, farpivots=None, resident=False):
    f = batched_side(compressor, farmatrix, tree, fardata, 0, maxrank, farpivots, resident)
    b, t = stores_from_factors(f, tree.testtree, 0, blocks=not resident)
    return b, t, f


def trialbases(compressor, farmatrix, tree, fardata, maxrank=50, farpivots=None, resident=False):
    f = batched_side(compressor, farmatrix, tree, fardata, 1, maxrank, farpivots, resident)
    b, t = stores_from_factors(f, tree.trialtree, 1, blocks=not resident)
    return b, t, f


# --------------------------------------------------------------------------
class CouplingStore:
    """Device-resident coupling matrices in test-CSR order (couplingmatrices.jl:11-53)."""

    def __init__(self, farmatrix, handle, ptr, nodes):
        self.farmatrix = farmatrix
        self.handle = handle
        self.nodetofarnodesptr = ptr
        self.farnodes = nodes

    def size(self):
        nb = ctypes.c_int64()
        ne = ctypes.c_int64()
        call(hip_lib().nca_coupling_size(self.handle, ctypes.byref(nb), ctypes.byref(ne)))
        return nb.value, ne.value

    def download(self):
        nb, ne = self.size()
        off = np.zeros(nb + 1, np.int64)
        blocks = np.zeros(max(ne, 1), np.complex128)
        call(hip_lib().nca_coupling_download(self.farmatrix.ctx.handle, self.handle, _p(off), _p(blocks)))
        return off, blocks[:ne]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_coupling_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def assemble_couplingstore(farmatrix, testpivots, trialpivots, testfardata):
    """testpivots/trialpivots: 0-based CSR (ptr, point ids) of every node."""
    tptr, tidx = testpivots
    sptr, sidx = trialpivots
    fp = np.ascontiguousarray(testfardata.nodetofarnodesptr + 1, np.int64)
    fn = np.ascontiguousarray(testfardata.farnodes + 1, np.int64)
    a = [np.ascontiguousarray(x, np.int64) for x in (tptr + 1, tidx + 1, sptr + 1, sidx + 1)]
    h = ctypes.c_void_p()
    kern = farmatrix.kernel.abi()
    ntest = len(tptr) - 1
    ntrial = len(sptr) - 1
    call(hip_lib().nca_coupling_assemble(farmatrix.ctx.handle, ctypes.byref(kern), farmatrix.test.handle,
                                         farmatrix.trial.handle, ntest, _p(fp), _p(fn) if len(fn) else None, ntrial,
                                         _p(a[0]), _p(a[1]) if len(a[1]) else None, _p(a[2]),
                                         _p(a[3]) if len(a[3]) else None, ctypes.byref(h)))
    return CouplingStore(farmatrix, h, testfardata.nodetofarnodesptr, testfardata.farnodes)
