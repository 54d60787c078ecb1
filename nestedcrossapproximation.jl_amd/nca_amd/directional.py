"""Directional (wideband) far data: the high-frequency subdivision of every
cluster's far interactions into directions.

Follows src/admissibility/nearinteractions.jl (isnearwideband, islf) and
src/admissibility/directionalsubdivision/{abstractdirections,
boundingballdirections,twondirections}.jl. Node ids, direction indices
(``diridx``) and local positions are 0-based here; direction *labels* keep the
reference's values: 0 is the low-frequency direction, a positive label is a
direction of the cluster (an index into its direction set, or a node of the
direction tree for TwoN trees). ``partochddirmap`` holds 0-based local
positions.
"""
import ctypes
import math

import numpy as np

from ._native import tree_lib
from . import trees as _trees


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Admissibility (nearinteractions.jl)
class IsLowFrequency:
    """IsLowFrequencyFunctor (nearinteractions.jl:105-158): a diameter d is low
    frequency iff k d <= gamma."""

    def __init__(self, gamma, k):
        self.gamma = float(gamma)
        self.k = float(k)

    def __call__(self, diam):
        return self.k * diam <= self.gamma

    def level(self, tree, level):
        """TwoN trees: the box diameter of a level, 2 sqrt(3) h0 / 2^level."""
        h0 = tree.radius[0] / math.sqrt(3.0)
        return self(2.0 * math.sqrt(3.0) * h0 / 2.0 ** level)

    def node(self, tree, node):
        if tree.kind == "twon":
            return self.level(tree, int(tree.level[node]))
        return self(2.0 * generalized_radius(tree)[node])


def generalized_radius(tree):
    """_bbgenradius (nearinteractions.jl:143-151) of every node: the root's
    radius, else min(parent's, mean radius of the siblings). Cached on the tree."""
    g = getattr(tree, "_genradius", None)
    if g is not None:
        return g
    g = np.empty(tree.nnodes)
    g[0] = tree.radius[0]
    for p in range(tree.nnodes):
        ch = tree.childrenof(p)
        if len(ch):
            g[ch] = min(g[p], float(np.sum(tree.radius[ch])) / len(ch))
    tree._genradius = g
    return g


class IsNearWideband:
    """isnearwideband(k; etalf=1, etahf=5, gamma=1) (nearinteractions.jl:33-81)."""

    def __init__(self, k, etalf=1.0, etahf=5.0, gamma=1.0, islf=None):
        self.k = float(k)
        self.etalf = float(etalf)
        self.etahf = float(etahf)
        self.islf = islf if islf is not None else IsLowFrequency(gamma, k)

    def __call__(self, ta, tb, a, b):
        """Near test of nodes a of ta and b of tb (radius = half size sqrt(3) for boxes)."""
        ra, rb = float(ta.radius[a]), float(tb.radius[b])
        d = float(np.linalg.norm(ta.center[a] - tb.center[b])) - (ra + rb)
        sep = max(d, 0.0)
        if self.islf(min(2 * ra, 2 * rb)):
            return 2 * max(ra, rb) > self.etalf * sep
        return 4 * self.k * max(ra * ra, rb * rb) > self.etahf * sep


def isnearwideband(k, etalf=1.0, etahf=5.0, gamma=1.0, islf=None):
    return IsNearWideband(k, etalf, etahf, gamma, islf)


def interactions(tree, isnear):
    """Near and far cluster pairs under the wideband admissibility (the
    traversal of ``trees.interactions`` with the wideband criterion)."""
    a, b = tree.testtree, tree.trialtree
    L = tree_lib()
    h = L.ncat_interactions_wideband(a.nnodes, _ptr(a.childptr), _ptr(a.children), _ptr(a.center),
                                     _ptr(a.radius), b.nnodes, _ptr(b.childptr), _ptr(b.children),
                                     _ptr(b.center), _ptr(b.radius), isnear.k, isnear.etalf, isnear.etahf,
                                     isnear.islf.gamma)
    nnear = ctypes.c_int64()
    nfar = ctypes.c_int64()
    L.ncat_partition_sizes(h, ctypes.byref(nnear), ctypes.byref(nfar))
    nearptr = np.empty(a.nnodes + 1, np.int64)
    farptr = np.empty(a.nnodes + 1, np.int64)
    nearnodes = np.empty(nnear.value, np.int64)
    farnodes = np.empty(nfar.value, np.int64)
    L.ncat_partition_export(h, _ptr(nearptr), _ptr(nearnodes), _ptr(farptr), _ptr(farnodes))
    L.ncat_partition_free(h)
    return _trees.Partition(nearptr, nearnodes, farptr, farnodes)


# --------------------------------------------------------------------------
# Direction sets (boundingballdirections.jl)
def sphericalfibonaccipoints(n):
    """n points of a spherical Fibonacci lattice (boundingballdirections.jl:19-28)."""
    i = np.arange(n, dtype=np.float64)
    ga = math.pi * (3.0 - math.sqrt(5.0))
    z = 1.0 - 2.0 * (i + 0.5) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(i * ga), r * np.sin(i * ga), z], axis=1)


def directions(dim, diam, k, gamma):
    """Direction set of a cluster of diameter diam (boundingballdirections.jl:5-17):
    2D n = max(4, ceil(pi / asin(min(1, gamma/kd)))) on the circle; 3D
    n = max(6, ceil(6 * 4^log2(acos(1/sqrt 3) / asin(min(1, gamma/kd))))) on a
    Fibonacci lattice; kd = max(k diam, eps)."""
    kd = max(k * diam, np.finfo(np.float64).eps)
    a = math.asin(min(1.0, gamma / kd))
    if dim == 2:
        n = max(4, int(math.ceil(math.pi / a)))
        j = np.arange(n)
        return np.stack([np.cos(2 * math.pi * j / n), np.sin(2 * math.pi * j / n)], axis=1)
    n = max(6, int(math.ceil(6.0 * 4.0 ** (math.log(math.acos(1.0 / math.sqrt(3.0)) / a) / math.log(2.0)))))
    return sphericalfibonaccipoints(n)


def _dots(evecs, vecs):
    """Dot products evecs[i] . vecs[p] as (len(vecs), len(evecs)), summed in
    coordinate order (no BLAS), so that argmax ties break the same everywhere."""
    out = evecs[None, :, 0] * vecs[:, None, 0]
    for d in range(1, evecs.shape[1]):
        out = out + evecs[None, :, d] * vecs[:, None, d]
    return out


def nearest_direction(evecs, vec):
    """_nearest_direction: 0-based index of the largest dot product, the first on ties."""
    return int(np.argmax(_dots(np.asarray(evecs), np.asarray(vec, np.float64)[None, :])[0]))


# --------------------------------------------------------------------------
# Direction tree of TwoN trees (twondirections.jl)
class DirectionTree:
    """DirectionTree{3}: 6 root directions (the face centres of the cube,
    +x -x +y -y +z -z), each refined into 4 nested patches per level. Node
    labels are 1-based in level order; ``parent``/``firstchild`` 0 = none.
    ``dirroot`` is the finest high-frequency cluster level (0-based)."""

    def __init__(self, maxlevel, dirroot, dim=3):
        if maxlevel < 1:
            raise ValueError("Directional tree maxlevel must be >= 1.")
        rootcount, branching = 2 * dim, 1 << (dim - 1)
        total = rootcount * ((branching ** maxlevel - 1) // (branching - 1))
        dirv = np.zeros((total + 1, dim))
        axis = np.zeros(total + 1, np.int64)
        parent = np.zeros(total + 1, np.int64)
        firstchild = np.zeros(total + 1, np.int64)
        nchildren = np.zeros(total + 1, np.int64)
        nxt = 1
        for ax in range(dim):
            for sgn in (1.0, -1.0):
                dirv[nxt, ax] = sgn
                axis[nxt] = ax
                nxt += 1
        # patch offsets of each axis: +-1 on the other axes, the lowest bit on the first of them
        offs = []
        for ax in range(dim):
            o = np.zeros((branching, dim))
            others = [d for d in range(dim) if d != ax]
            for mask in range(branching):
                for bit, d in enumerate(others):
                    o[mask, d] = 1.0 if (mask >> bit) & 1 else -1.0
            offs.append(o)
        start, count, scale = 1, rootcount, 0.5
        for _ in range(maxlevel - 1):
            for node in range(start, start + count):
                firstchild[node] = nxt
                nchildren[node] = branching
                for o in offs[axis[node]]:
                    dirv[nxt] = dirv[node] + scale * o
                    axis[nxt] = axis[node]
                    parent[nxt] = node
                    nxt += 1
            start += count
            count *= branching
            scale /= 2.0
        assert nxt == total + 1
        dirv[1:] /= np.linalg.norm(dirv[1:], axis=1)[:, None]
        self.level, self.dirroot, self.rootcount, self.branching = maxlevel, dirroot, rootcount, branching
        self.dir, self.parent, self.firstchild, self.nchildren = dirv, parent, firstchild, nchildren

    def children(self, node):
        if node == 0:
            return np.arange(1, self.rootcount + 1)
        return np.arange(self.firstchild[node], self.firstchild[node] + self.nchildren[node])

    def direction(self, vecs, clusterlevel):
        """direction(interaction, tree, clusterlevel) for rows of vecs: descend
        min(dirroot - clusterlevel + 1, level) levels by the largest dot product."""
        vecs = np.atleast_2d(np.asarray(vecs, np.float64))
        depth = min(self.dirroot - clusterlevel + 1, self.level)
        cur = np.zeros(len(vecs), np.int64)
        for lev in range(depth):
            if lev == 0:
                cand = np.broadcast_to(np.arange(1, self.rootcount + 1), (len(vecs), self.rootcount))
            else:
                cand = self.firstchild[cur][:, None] + np.arange(self.branching)[None, :]
            d = self.dir[cand]
            dots = d[:, :, 0] * vecs[:, None, 0]
            for c in range(1, d.shape[2]):
                dots = dots + d[:, :, c] * vecs[:, None, c]
            cur = cand[np.arange(len(vecs)), np.argmax(dots, axis=1)]
        return cur


def _dirroot_and_maxlevel(tree, ptr, islf):
    """_dirroot_and_maxlevel (twondirections.jl:136-155), 0-based levels."""
    lflevel = 0
    while not islf.level(tree, lflevel):
        lflevel += 1
    dirroot = lflevel - 1
    first = dirroot
    has = np.diff(ptr) > 0
    for lev in np.unique(tree.level[has]):
        if not islf.level(tree, int(lev)) and lev < first:
            first = int(lev)
    return dirroot, dirroot - first + 1


# --------------------------------------------------------------------------
class DirectionalData:
    """DirectionalData (abstractdirections.jl:3-50) in 0-based CSR form.

    fardata: the FarData whose far nodes are grouped by direction; per node
    its direction indices nodetodirsptr[t]:nodetodirsptr[t+1] with labels
    ``dirs``; per direction index its far nodes diridxtofarptr[e]:[e+1];
    per node, for each direction of its parent, the matching local direction
    (partochddirmap[nodetochdirptr[t]:nodetochdirptr[t+1]])."""

    def __init__(self, fardata, nodetodirsptr, dirs, diridxtofarptr, nodetochdirptr, partochddirmap, isnear):
        self.fardata = fardata
        self.nodetodirsptr = nodetodirsptr
        self.dirs = dirs
        self.diridxtofarptr = diridxtofarptr
        self.nodetochdirptr = nodetochdirptr
        self.partochddirmap = partochddirmap
        self.isnear = isnear

    @property
    def ndiridx(self):
        return int(self.nodetodirsptr[-1])

    def farptr(self):
        return self.fardata.nodetofarnodesptr

    def farnodes(self, node=None):
        if node is None:
            return self.fardata.farnodes
        return self.fardata.farnodesof(node)

    def nfarnodes(self, node):
        p = self.fardata.nodetofarnodesptr
        return int(p[node + 1] - p[node])

    def ndirections(self, node):
        return int(self.nodetodirsptr[node + 1] - self.nodetodirsptr[node])

    def dirrange(self, node):
        return range(int(self.nodetodirsptr[node]), int(self.nodetodirsptr[node + 1]))

    def dirsof(self, node):
        return self.dirs[self.nodetodirsptr[node]:self.nodetodirsptr[node + 1]]

    def _local(self, node, dir):
        hit = np.nonzero(self.dirsof(node) == dir)[0]
        if len(hit) == 0:
            raise KeyError(f"Direction {dir} is not present in node {node}.")
        return int(hit[0])

    def dirfarnodes(self, node, dir):
        e = int(self.nodetodirsptr[node]) + self._local(node, dir)
        return self.fardata.farnodes[self.diridxtofarptr[e]:self.diridxtofarptr[e + 1]]

    def diridxfromlocalfaridx(self, node, faridx):
        """Direction index of the faridx-th (0-based) far node of node."""
        if not 0 <= faridx < self.nfarnodes(node):
            raise IndexError(f"Local far index {faridx} exceeds the far nodes of node {node}.")
        pos = self.fardata.nodetofarnodesptr[node] + faridx
        r = self.dirrange(node)
        return int(r.start + np.searchsorted(self.diridxtofarptr[r.start + 1:r.stop + 1], pos, side="right"))

    def dirmap(self, node):
        return self.partochddirmap[self.nodetochdirptr[node]:self.nodetochdirptr[node + 1]]

    def node_of(self):
        return np.repeat(np.arange(len(self.nodetodirsptr) - 1), np.diff(self.nodetodirsptr))


def finalize_directionaldata(fardata, dirs, dircount, partochddirmap, partochddircount, isnear):
    """_finalize_directionaldata (abstractdirections.jl:52-103): the per-node
    lists into CSR; the counts of a node must add up to its far nodes."""
    ptr = fardata.nodetofarnodesptr
    nn = len(dirs)
    ndirs = np.array([len(d) for d in dirs], np.int64)
    nodetodirsptr = np.zeros(nn + 1, np.int64)
    np.cumsum(ndirs, out=nodetodirsptr[1:])
    nodetochdirptr = np.zeros(nn + 1, np.int64)
    np.cumsum(np.asarray(partochddircount, np.int64), out=nodetochdirptr[1:])
    ne = int(nodetodirsptr[-1])
    linedirs = np.zeros(ne, np.int64)
    diridxtofarptr = np.zeros(ne + 1, np.int64)
    linemap = np.zeros(int(nodetochdirptr[-1]), np.int64)
    for t in range(nn):
        a, b = nodetodirsptr[t], nodetodirsptr[t + 1]
        if b > a:
            linedirs[a:b] = dirs[t]
            c = np.asarray(dircount[t], np.int64)
            if len(c) != b - a or int(ptr[t]) + int(c.sum()) != int(ptr[t + 1]):
                raise AssertionError(f"Direction counts for node {t} do not sum to its far node count.")
            diridxtofarptr[a] = ptr[t]
            diridxtofarptr[a + 1:b] = ptr[t] + np.cumsum(c)[:-1]
        pa, pb = nodetochdirptr[t], nodetochdirptr[t + 1]
        if pb > pa:
            linemap[pa:pb] = partochddirmap[t]
    diridxtofarptr[-1] = ptr[-1]
    return DirectionalData(fardata, nodetodirsptr, linedirs, diridxtofarptr, nodetochdirptr, linemap, isnear)


def _run_counts(sorted_labels, unique_labels):
    """Far nodes per label for the first len(unique_labels) labels (sorted input)."""
    return np.array([int(np.count_nonzero(sorted_labels == u)) for u in unique_labels], np.int64)


def directionalfardata(fardata, tree, fartree, isnear):
    """directionalfardata (twondirections.jl:259-308 for TwoN trees,
    boundingballdirections.jl:44-152 for bounding-ball trees) in the host
    library (nca_tree.cpp ncat_directional_fardata). Reorders the far nodes of
    ``fardata`` by direction in place (a stable sort)."""
    if not isinstance(isnear.islf, IsLowFrequency):
        return directionalfardata_reference(fardata, tree, fartree, isnear)
    L = tree_lib()
    nn = tree.nnodes
    i64 = lambda a: np.ascontiguousarray(a, np.int64)
    par, lev, cp, ch = i64(tree.parent), i64(tree.level), i64(tree.childptr), i64(tree.children)
    ctr, rad = np.ascontiguousarray(tree.center, np.float64), np.ascontiguousarray(tree.radius, np.float64)
    fctr = np.ascontiguousarray(fartree.center, np.float64)
    fp = i64(fardata.nodetofarnodesptr)
    fn = i64(fardata.farnodes)
    h = L.ncat_directional_fardata(int(tree.kind == "twon"), nn, _ptr(par), _ptr(lev), _ptr(cp),
                                   _ptr(ch) if len(ch) else None, _ptr(ctr), _ptr(rad), _ptr(fp),
                                   _ptr(fn) if len(fn) else None, _ptr(fctr), isnear.k, isnear.islf.gamma,
                                   isnear.etahf)
    if not h:
        raise AssertionError("directional far data: direction counts do not sum to the far nodes")
    ne, nmap = ctypes.c_int64(), ctypes.c_int64()
    L.ncat_dirdata_sizes(h, ctypes.byref(ne), ctypes.byref(nmap))
    dirptr = np.empty(nn + 1, np.int64)
    dirs = np.empty(ne.value, np.int64)
    dfp = np.empty(ne.value + 1, np.int64)
    chptr = np.empty(nn + 1, np.int64)
    chmap = np.empty(nmap.value, np.int64)
    L.ncat_dirdata_export(h, _ptr(dirptr), _ptr(dirs) if ne.value else None, _ptr(dfp), _ptr(chptr),
                          _ptr(chmap) if nmap.value else None)
    L.ncat_dirdata_free(h)
    fardata.farnodes[:] = fn
    return DirectionalData(fardata, dirptr, dirs, dfp, chptr, chmap, isnear)


def directionalfardata_reference(fardata, tree, fartree, isnear):
    """The same subdivision written out in numpy (for any low-frequency
    classifier; the library form above is checked against it)."""
    ptr = fardata.nodetofarnodesptr
    nodes = fardata.farnodes
    nn = len(ptr) - 1
    k = isnear.k
    gamma = isnear.islf.gamma
    dirs = [None] * nn
    dircount = [np.zeros(0, np.int64) for _ in range(nn)]
    chmap = [np.zeros(0, np.int64) for _ in range(nn)]
    chcount = np.zeros(nn, np.int64)
    twon = tree.kind == "twon"
    if twon:
        dirtree = DirectionTree(*reversed(_dirroot_and_maxlevel(tree, ptr, isnear.islf)))
    else:
        dirsvec = [None] * nn
    order = np.argsort(tree.level, kind="stable")
    for node in order.tolist():
        level = int(tree.level[node])
        pnode = int(tree.parent[node])
        hasfars = ptr[node + 1] > ptr[node]
        hasparentdirs = pnode >= 0 and len(dirs[pnode]) > 0
        lf = isnear.islf.level(tree, level) if twon else isnear.islf.node(tree, node)
        if lf:
            if twon:
                inherit = hasparentdirs and level > 0 and isnear.islf.level(tree, level - 1)
            else:
                inherit = hasparentdirs and level > 0 and isnear.islf.node(tree, pnode)
            dirs[node] = np.array([0], np.int64) if (hasfars or inherit) else np.zeros(0, np.int64)
            chmap[node] = np.array([0], np.int64) if inherit else np.zeros(0, np.int64)
            continue
        hasparent = hasparentdirs if twon else (pnode >= 0 and dirsvec[pnode] is not None)
        if not hasfars and not hasparent:
            dirs[node] = np.zeros(0, np.int64)
            continue
        a, b = int(ptr[node]), int(ptr[node + 1])
        vec = tree.center[node][None, :] - fartree.center[nodes[a:b]]
        if twon:
            nodedirs = dirtree.direction(vec, level) if b > a else np.zeros(0, np.int64)
        else:
            total = directions(3, 2.0 * float(tree.radius[node]), k, gamma)
            nodedirs = (np.argmax(_dots(total, vec), axis=1) + 1) if b > a else np.zeros(0, np.int64)
        perm = np.argsort(nodedirs, kind="stable")
        nodes[a:b] = nodes[a:b][perm]
        sdirs = nodedirs[perm]
        uniq = list(np.unique(nodedirs).tolist())
        noriginal = len(uniq)
        partochddir = np.zeros(len(dirs[pnode]) if pnode >= 0 else 0, np.int64)
        if twon:
            if hasparentdirs:
                cd = [int(dirtree.parent[pd]) for pd in dirs[pnode].tolist()]
                for idx, c in enumerate(cd):
                    if c not in uniq:
                        uniq.append(c)
                    partochddir[idx] = uniq.index(c) + 1
        elif pnode >= 0:
            ctopc = tree.center[node] - tree.center[pnode]
            mult = k * (2 * tree.radius[pnode]) / isnear.etahf + 2 * tree.radius[pnode]
            if dirsvec[pnode] is not None and len(dirsvec[pnode]):
                target = ctopc[None, :] + mult * dirsvec[pnode]
                cd = (np.argmax(_dots(total, target), axis=1) + 1).tolist()
                for idx, c in enumerate(cd):
                    if c not in uniq:
                        uniq.append(c)
                    partochddir[idx] = uniq.index(c) + 1
        cnt = np.zeros(len(uniq), np.int64)
        cnt[:noriginal] = _run_counts(sdirs, uniq[:noriginal])
        if not twon:
            dirsvec[node] = total[np.asarray(uniq, np.int64) - 1]
        # 1-based positions (0 = none) -> 0-based
        chmap[node] = partochddir - 1
        chcount[node] = int(np.count_nonzero(partochddir))
        dirs[node] = np.asarray(uniq, np.int64)
        dircount[node] = cnt
    # all far nodes of a low-frequency node belong to the direction 0
    for node in range(nn):
        if len(dirs[node]) and len(dircount[node]) == 0:
            dircount[node] = np.array([max(0, int(ptr[node + 1] - ptr[node]))], np.int64)
            chcount[node] = len(chmap[node])
    return finalize_directionaldata(fardata, dirs, dircount, chmap, chcount, isnear)


def fardata(tree, isnear, partition=None):
    """fardata(tree, isnear::IsNearWidebandFunctor) (abstractdirections.jl:105-116):
    the far data of both trees, each subdivided into directions."""
    if partition is None:
        partition = interactions(tree, isnear)
    tf, sf = _trees.fardata(tree, partition)
    tf = _trees.FarData(tf.nodetofarnodesptr.copy(), tf.farnodes.copy())
    td = directionalfardata(tf, tree.testtree, tree.trialtree, isnear)
    sd = directionalfardata(sf, tree.trialtree, tree.testtree, isnear)
    return td, sd


def dirfarfield_all(tree, data):
    """dirfarfield (abstractdirections.jl:205-243) of every direction index:
    0-based CSR (ptr over direction indices, far nodes of the other tree)."""
    L = tree_lib()
    par = np.ascontiguousarray(tree.parent, np.int64)
    arrs = [np.ascontiguousarray(a, np.int64) for a in (data.nodetodirsptr, data.dirs, data.nodetochdirptr,
                                                         data.partochddirmap, data.diridxtofarptr,
                                                         data.fardata.farnodes)]
    h = L.ncat_dirfarfield(len(par), _ptr(par), *[_ptr(a) if len(a) else None for a in arrs])
    ne = data.ndiridx
    ptr = np.empty(ne + 1, np.int64)
    idx = np.empty(L.ncat_csr_nnz(h), np.int64)
    L.ncat_csr_export(h, _ptr(ptr), _ptr(idx))
    L.ncat_csr_free(h)
    return ptr, idx


def dirfarfield(tree, data, node, dir):
    """Far field of (node, dir): its own far nodes in the direction, then the
    inherited ones (abstractdirections.jl:205-243)."""
    loc = data._local(node, dir)
    ptr, idx = dirfarfield_all(tree, data)
    e = int(data.nodetodirsptr[node]) + loc
    return idx[ptr[e]:ptr[e + 1]]


def isbasisnode(tree, data, node, dir):
    """A leaf, or a high-frequency direction of a node whose first child has
    no direction or only the direction 0 (abstractdirections.jl:245-251)."""
    if tree.isleaf(node):
        return True
    if dir == 0:
        return False
    fc = int(tree.childrenof(node)[0])
    d = data.dirsof(fc)
    return len(d) == 0 or (len(d) == 1 and d[0] == 0)


def interaction(node, farnode, tree, fartree):
    """Vector from the centre of farnode to the centre of node."""
    return tree.center[node] - fartree.center[farnode]
