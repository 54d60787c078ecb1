"""Cluster trees, block trees, the admissibility partition and far data.

Host-side inputs of the compressor and the matvec. The reference takes them
from H2Trees.jl (trees) and AdaptiveCrossApproximation.jl (``interactions``);
neither is vendored, so ``libnca_tree.so`` provides deterministic stand-ins
with the same contracts (see nca_tree.cpp). Node ids and positions here are
0-based; ``*_1based`` helpers give the arrays in the form the Julia structs
hold them, which is what the C-ABI takes.
"""
import ctypes

import numpy as np

from ._native import tree_lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class ClusterTree:
    """A cluster tree whose every node owns a contiguous range of ``perm``.

    values(node) = perm[lo[node]:hi[node]] (0-based point ids). Node 0 is the
    root; ids are assigned level by level, so iterating the nodes of a level in
    increasing id order is the LevelIterator order.
    """

    def __init__(self, points, parent, level, lo, hi, childptr, children, perm, center, radius, kind):
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.parent = parent
        self.level = level
        self.lo = lo
        self.hi = hi
        self.childptr = childptr
        self.children = children
        self.perm = perm
        self.center = center
        self.radius = radius
        self.kind = kind

    @property
    def nnodes(self):
        return int(self.parent.shape[0])

    @property
    def npoints(self):
        return int(self.perm.shape[0])

    def values(self, node):
        return self.perm[self.lo[node]:self.hi[node]]

    def nvalues(self, node):
        return int(self.hi[node] - self.lo[node])

    def isleaf(self, node):
        return self.childptr[node + 1] == self.childptr[node]

    def leafmask(self):
        return self.childptr[1:] == self.childptr[:-1]

    def childrenof(self, node):
        return self.children[self.childptr[node]:self.childptr[node + 1]]

    def levels(self):
        return np.unique(self.level)

    def levelnodes(self, lev):
        return np.nonzero(self.level == lev)[0]

    # Arrays in the 1-based form of the C-ABI (Julia indexing).
    def layout_1based(self):
        return dict(
            perm=np.ascontiguousarray(self.perm + 1),
            nodelo=np.ascontiguousarray(self.lo + 1),
            nodehi=np.ascontiguousarray(self.hi.copy()),
            level=np.ascontiguousarray(self.level + 1),
            childptr=np.ascontiguousarray(self.childptr + 1),
            children=np.ascontiguousarray(self.children + 1),
        )


def _export_tree(handle, points, kind):
    L = tree_lib()
    nn = L.ncat_tree_nnodes(handle)
    n = points.shape[0]
    parent = np.empty(nn, np.int64)
    level = np.empty(nn, np.int64)
    lo = np.empty(nn, np.int64)
    hi = np.empty(nn, np.int64)
    childptr = np.empty(nn + 1, np.int64)
    children = np.empty(max(nn - 1, 0), np.int64)
    perm = np.empty(n, np.int64)
    center = np.empty((nn, 3), np.float64)
    radius = np.empty(nn, np.float64)
    L.ncat_tree_export(handle, _ptr(parent), _ptr(level), _ptr(lo), _ptr(hi), _ptr(childptr),
                       _ptr(children), _ptr(perm), _ptr(center), _ptr(radius))
    L.ncat_tree_free(handle)
    return ClusterTree(points, parent, level, lo, hi, childptr, children, perm, center, radius, kind)


def _device_tree(ctx, pts, kind, build):
    """A tree built on the device (nca_tree_twon / nca_tree_kmeans), exported
    from the 1-based layout of the C-ABI."""
    from ._native import hip_lib
    from .runtime import Points, call
    L = hip_lib()
    dpts = Points(ctx, pts)
    nn = ctypes.c_int64()
    h = ctypes.c_void_p()
    call(build(L, dpts.handle, ctypes.byref(nn), ctypes.byref(h)))
    try:
        nn, n = nn.value, pts.shape[0]
        parent, level, lo, hi = (np.empty(nn, np.int64) for _ in range(4))
        childptr = np.empty(nn + 1, np.int64)
        children = np.empty(max(nn - 1, 1), np.int64)
        perm = np.empty(max(n, 1), np.int64)
        center = np.empty((nn, 3), np.float64)
        radius = np.empty(nn, np.float64)
        call(L.nca_tree_export(h, _ptr(parent), _ptr(level), _ptr(lo), _ptr(hi), _ptr(childptr),
                               _ptr(children), _ptr(perm), _ptr(center), _ptr(radius)))
    finally:
        L.nca_tree_destroy(h)
    return ClusterTree(pts, parent - 1, level - 1, lo - 1, hi, childptr - 1, children[:nn - 1] - 1,
                       perm[:n] - 1, center, radius, kind)


def TwoNTree(points, minvalues=64, minhalfsize=0.0, ctx=None):
    """Regular 2^3 tree (octree) over ``points`` (n x 3); built on the device of
    ``ctx`` when one is given (the same tree)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if ctx is not None:
        return _device_tree(ctx, pts, "twon", lambda L, p, nn, h: L.nca_tree_twon(
            ctx.handle, p, int(minvalues), float(minhalfsize), nn, h))
    h = tree_lib().ncat_tree_twon(_ptr(pts), pts.shape[0], int(minvalues), float(minhalfsize))
    return _export_tree(h, pts, "twon")


def KMeansTree(points, numberofclusters=2, minvalues=32, ctx=None):
    """Deterministic k-means tree over ``points`` (n x 3); built on the device
    of ``ctx`` when one is given (the same tree)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if ctx is not None:
        return _device_tree(ctx, pts, "kmeans", lambda L, p, nn, h: L.nca_tree_kmeans(
            ctx.handle, p, int(numberofclusters), int(minvalues), nn, h))
    h = tree_lib().ncat_tree_kmeans(_ptr(pts), pts.shape[0], int(numberofclusters), int(minvalues))
    return _export_tree(h, pts, "kmeans")


class BlockTree:
    def __init__(self, testtree, trialtree):
        self.testtree = testtree
        self.trialtree = trialtree


class Partition:
    """Near and far interactions by test node (0-based CSR of trial nodes)."""

    def __init__(self, nearptr, nearnodes, farptr, farnodes):
        self.nearptr = nearptr
        self.nearnodes = nearnodes
        self.farptr = farptr
        self.farnodes = farnodes


def interactions(tree, eta=1.0):
    """Admissible and near cluster pairs: far iff max(diam) <= eta * dist."""
    a, b = tree.testtree, tree.trialtree
    L = tree_lib()
    h = L.ncat_interactions(a.nnodes, _ptr(a.childptr), _ptr(a.children), _ptr(a.center), _ptr(a.radius),
                            b.nnodes, _ptr(b.childptr), _ptr(b.children), _ptr(b.center), _ptr(b.radius),
                            float(eta))
    nnear = ctypes.c_int64()
    nfar = ctypes.c_int64()
    L.ncat_partition_sizes(h, ctypes.byref(nnear), ctypes.byref(nfar))
    nearptr = np.empty(a.nnodes + 1, np.int64)
    farptr = np.empty(a.nnodes + 1, np.int64)
    nearnodes = np.empty(nnear.value, np.int64)
    farnodes = np.empty(nfar.value, np.int64)
    L.ncat_partition_export(h, _ptr(nearptr), _ptr(nearnodes), _ptr(farptr), _ptr(farnodes))
    L.ncat_partition_free(h)
    return Partition(nearptr, nearnodes, farptr, farnodes)


def interactions_device(tree, ctx, eta=1.0, wideband=None):
    """The partition of ``interactions`` (or of the wideband criterion, with
    ``wideband`` = an IsNearWideband) computed on the device (nca_interactions);
    equal to the host traversal's."""
    from ._native import hip_lib
    from .runtime import call
    a, b = tree.testtree, tree.trialtree
    arr = []
    for t in (a, b):
        arr.append([np.ascontiguousarray(t.childptr + 1, np.int64),
                    np.ascontiguousarray(t.children + 1, np.int64) if len(t.children) else np.zeros(1, np.int64),
                    np.ascontiguousarray(t.center, np.float64), np.ascontiguousarray(t.radius, np.float64)])
    if wideband is None:
        params = np.array([float(eta), 0.0, 0.0, 0.0, 0.0])
    else:
        params = np.array([0.0, wideband.k, wideband.etalf, wideband.etahf, wideband.islf.gamma])
    nnear, nfar = ctypes.c_int64(), ctypes.c_int64()
    h = ctypes.c_void_p()
    call(hip_lib().nca_interactions(ctx.handle, a.nnodes, *[_ptr(x) for x in arr[0]], b.nnodes,
                                    *[_ptr(x) for x in arr[1]], int(wideband is not None), _ptr(params),
                                    ctypes.byref(nnear), ctypes.byref(nfar), ctypes.byref(h)))
    try:
        nearptr = np.empty(a.nnodes + 1, np.int64)
        farptr = np.empty(a.nnodes + 1, np.int64)
        nearnodes = np.empty(max(nnear.value, 1), np.int64)
        farnodes = np.empty(max(nfar.value, 1), np.int64)
        call(hip_lib().nca_partition_export(h, _ptr(nearptr), _ptr(nearnodes), _ptr(farptr), _ptr(farnodes)))
    finally:
        hip_lib().nca_partition_destroy(h)
    return Partition(nearptr - 1, nearnodes[:nnear.value] - 1, farptr - 1, farnodes[:nfar.value] - 1)


class FarData:
    """Far nodes of every node of one tree, CSR (farinteractions.jl:16-19), 0-based."""

    def __init__(self, ptr, nodes):
        self.nodetofarnodesptr = ptr
        self.farnodes = nodes

    def farnodesof(self, node):
        return self.farnodes[self.nodetofarnodesptr[node]:self.nodetofarnodesptr[node + 1]]


def transpose_csr(ptr, idx, ncols):
    """CSR transpose with rows kept in increasing order (farinteractions.jl:66-92)."""
    L = tree_lib()
    p = np.ascontiguousarray(ptr, np.int64)
    i = np.ascontiguousarray(idx, np.int64)
    h = L.ncat_csr_transpose(p.shape[0] - 1, _ptr(p), _ptr(i) if len(i) else None, int(ncols))
    tptr = np.empty(ncols + 1, np.int64)
    tidx = np.empty(L.ncat_csr_nnz(h), np.int64)
    L.ncat_csr_export(h, _ptr(tptr), _ptr(tidx))
    L.ncat_csr_free(h)
    return tptr, tidx


def fardata(tree, partition):
    """(test FarData, trial FarData): the trial side is the CSR transpose."""
    test = FarData(partition.farptr, partition.farnodes)
    tptr, tnodes = transpose_csr(partition.farptr, partition.farnodes, tree.trialtree.nnodes)
    return test, FarData(tptr, tnodes)


def farfield(tree, fd, node):
    """Own far nodes followed by those of every ancestor (farinteractions.jl:38-60)."""
    out = []
    n = node
    while n >= 0:
        out.extend(fd.farnodesof(n).tolist())
        n = tree.parent[n]
    return np.asarray(out, dtype=np.int64)


def farfield_csr(tree, fd):
    """farfield of every node (own far nodes, then the ancestors') as a 0-based CSR."""
    L = tree_lib()
    par = np.ascontiguousarray(tree.parent, np.int64)
    fp = np.ascontiguousarray(fd.nodetofarnodesptr, np.int64)
    fn = np.ascontiguousarray(fd.farnodes, np.int64)
    h = L.ncat_farfield_csr(tree.nnodes, _ptr(par), _ptr(fp), _ptr(fn) if len(fn) else None)
    ptr = np.empty(tree.nnodes + 1, np.int64)
    idx = np.empty(L.ncat_csr_nnz(h), np.int64)
    L.ncat_csr_export(h, _ptr(ptr), _ptr(idx) if len(idx) else None)
    L.ncat_csr_free(h)
    return ptr, idx


def farpivots(selftree, othertree, fd, maxrank, weight=1.5, oversample=1.0):
    """Phase 1 (batchedbottomup.jl:257-266): geometric far pivots of every node
    of ``selftree`` among the points of ``othertree``. Returns a 0-based CSR
    (ptr, point ids)."""
    L = tree_lib()
    pts = np.ascontiguousarray(othertree.points)
    h = L.ncat_farpivots(selftree.nnodes, _ptr(selftree.parent), _ptr(selftree.lo), _ptr(selftree.hi),
                         _ptr(selftree.center), _ptr(othertree.lo), _ptr(othertree.hi), _ptr(othertree.perm),
                         _ptr(pts), _ptr(fd.nodetofarnodesptr), _ptr(fd.farnodes), int(maxrank), float(weight), float(oversample))
    nnz = L.ncat_csr_nnz(h)
    ptr = np.empty(selftree.nnodes + 1, np.int64)
    idx = np.empty(nnz, np.int64)
    L.ncat_csr_export(h, _ptr(ptr), _ptr(idx))
    L.ncat_csr_free(h)
    return ptr, idx


def forest(trees):
    """One cluster tree over independent problems: a new root whose children are
    the roots of ``trees`` (each over its own point set). Points, positions and
    node ids of problem p are offset; ids stay level by level (level l + 1 of
    the forest holds level l of every tree, problem by problem), so one
    compressor launch per forest level serves all problems. Returns
    (tree, nodemaps, pointoffsets): nodemaps[p][i] = forest id of node i of
    tree p; problem p owns points (and positions) [pointoffsets[p],
    pointoffsets[p + 1])."""
    P = len(trees)
    poff = np.zeros(P + 1, np.int64)
    np.cumsum([t.npoints for t in trees], out=poff[1:])
    maxlev = max(int(t.level.max()) for t in trees)
    # forest ids: root, then for each tree level all trees' nodes of that level
    nodemaps = [np.empty(t.nnodes, np.int64) for t in trees]
    nid = 1
    for lev in range(maxlev + 1):
        for p, t in enumerate(trees):
            nodes = np.nonzero(t.level == lev)[0]
            nodemaps[p][nodes] = np.arange(nid, nid + len(nodes))
            nid += len(nodes)
    nn = nid
    parent = np.full(nn, -1, np.int64)
    level = np.zeros(nn, np.int64)
    lo = np.zeros(nn, np.int64)
    hi = np.zeros(nn, np.int64)
    center = np.zeros((nn, 3), np.float64)
    radius = np.zeros(nn, np.float64)
    nchild = np.zeros(nn, np.int64)
    lo[0], hi[0] = 0, poff[-1]
    nchild[0] = P
    pts = np.concatenate([t.points for t in trees], axis=0)
    perm = np.concatenate([t.perm + poff[p] for p, t in enumerate(trees)])
    center[0] = pts.mean(axis=0) if len(pts) else 0.0
    radius[0] = float(np.max(np.linalg.norm(pts - center[0], axis=1))) if len(pts) else 0.0
    for p, t in enumerate(trees):
        m = nodemaps[p]
        parent[m] = np.where(t.parent >= 0, m[np.maximum(t.parent, 0)], 0)
        level[m] = t.level + 1
        lo[m] = t.lo + poff[p]
        hi[m] = t.hi + poff[p]
        center[m] = t.center
        radius[m] = t.radius
        nchild[m] = np.diff(t.childptr)
    childptr = np.zeros(nn + 1, np.int64)
    np.cumsum(nchild, out=childptr[1:])
    children = np.empty(max(nn - 1, 0), np.int64)
    children[0:P] = [nodemaps[p][0] for p in range(P)]
    for p, t in enumerate(trees):
        m = nodemaps[p]
        for i in range(t.nnodes):
            c = t.children[t.childptr[i]:t.childptr[i + 1]]
            if len(c):
                children[childptr[m[i]]:childptr[m[i] + 1]] = m[c]
    tree = ClusterTree(pts, parent, level, lo, hi, childptr, children, perm, center, radius, trees[0].kind)
    return tree, nodemaps, poff


def forest_partition(testmaps, trialmaps, partitions, ntest):
    """The block-diagonal partition of a forest: problem p's near and far pairs
    with forest ids (no pairs across problems, none at the forest root); each
    node keeps its problem's order of near and far nodes."""

    def merged(ptrname, nodename):
        rows = np.concatenate([np.repeat(tm, np.diff(getattr(pt, ptrname)))
                               for tm, pt in zip(testmaps, partitions)])
        cols = np.concatenate([sm[getattr(pt, nodename)] for sm, pt in zip(trialmaps, partitions)])
        order = np.argsort(rows, kind="stable")
        ptr = np.zeros(ntest + 1, np.int64)
        np.cumsum(np.bincount(rows, minlength=ntest), out=ptr[1:])
        return ptr, cols[order].astype(np.int64)

    nearptr, nearnodes = merged("nearptr", "nearnodes")
    farptr, farnodes = merged("farptr", "farnodes")
    return Partition(nearptr, nearnodes, farptr, farnodes)
