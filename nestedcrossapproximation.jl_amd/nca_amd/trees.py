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


def TwoNTree(points, minvalues=64, minhalfsize=0.0):
    """Regular 2^3 tree (octree) over ``points`` (n x 3)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    h = tree_lib().ncat_tree_twon(_ptr(pts), pts.shape[0], int(minvalues), float(minhalfsize))
    return _export_tree(h, pts, "twon")


def KMeansTree(points, numberofclusters=2, minvalues=32):
    """Deterministic k-means tree over ``points`` (n x 3)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
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


class FarData:
    """Far nodes of every node of one tree, CSR (farinteractions.jl:16-19), 0-based."""

    def __init__(self, ptr, nodes):
        self.nodetofarnodesptr = ptr
        self.farnodes = nodes

    def farnodesof(self, node):
        return self.farnodes[self.nodetofarnodesptr[node]:self.nodetofarnodesptr[node + 1]]


def transpose_csr(ptr, idx, ncols):
    """CSR transpose with rows kept in increasing order (farinteractions.jl:66-92)."""
    counts = np.bincount(idx, minlength=ncols).astype(np.int64)
    tptr = np.zeros(ncols + 1, np.int64)
    np.cumsum(counts, out=tptr[1:])
    rows = np.repeat(np.arange(ptr.shape[0] - 1, dtype=np.int64), np.diff(ptr))
    order = np.argsort(idx, kind="stable")
    return tptr, rows[order]


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


def farpivots(selftree, othertree, fd, maxrank):
    """Phase 1 (batchedbottomup.jl:257-266): geometric far pivots of every node
    of ``selftree`` among the points of ``othertree``. Returns a 0-based CSR
    (ptr, point ids)."""
    L = tree_lib()
    pts = np.ascontiguousarray(othertree.points)
    h = L.ncat_farpivots(selftree.nnodes, _ptr(selftree.parent), _ptr(selftree.lo), _ptr(selftree.hi),
                         _ptr(selftree.center), _ptr(othertree.lo), _ptr(othertree.hi), _ptr(othertree.perm),
                         _ptr(pts), _ptr(fd.nodetofarnodesptr), _ptr(fd.farnodes), int(maxrank))
    nnz = L.ncat_csr_nnz(h)
    ptr = np.empty(selftree.nnodes + 1, np.int64)
    idx = np.empty(nnz, np.int64)
    L.ncat_csr_export(h, _ptr(ptr), _ptr(idx))
    L.ncat_csr_free(h)
    return ptr, idx
