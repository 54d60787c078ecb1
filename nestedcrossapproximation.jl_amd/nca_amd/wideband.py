"""Directional H^2 (wideband) on the GPU: testbases/trialbases(::BottomUp,
..., ::DirectionalData) (src/compressors/bottomup.jl:214-424), the directional
coupling store (src/couplingmatrices.jl:55-111) and the directional plans
(src/nca/abstractnca.jl:5-9,41-48,78-82), behind the same device operator as
the plain H^2 (nca_dh2_create).

Every (cluster, direction) pair with a non-empty directional far field is one
block of the batched level compressor: its rows are the cluster's points at a
basis node (isbasisnode), else the pivots of the matching direction of every
child (dirmap) in ChildIterator order; its columns are Phase-1 far pivots
chosen from its dirfarfield. The pairs therefore form a tree-like structure
(``DirectionBlocks``) that the level driver of compressors.py walks like a
cluster tree. The reference factorizes each pair with IACA (tree-mimicry
pivots on the far side, maximum-value pivots on the cluster side) from the
un-vendored AdaptiveCrossApproximation.jl; here the far side is the geometric
Phase-1 selection and the cluster side is fully pivoted ACA with the sampled
stop, the batched compressor's factorization.
"""
import ctypes
import time

import numpy as np

from . import _abi
from ._native import hip_lib, tree_lib
from .runtime import call, _p, HIPKernelMatrix, default_context
from . import trees as _trees
from . import compressors as _cmp
from . import directional as _dir
from . import nca as _nca


class BottomUp:
    """BottomUp(; factorization, representor) (bottomup.jl:26-47) for the
    directional path: ``factorization`` factorizes every (cluster, direction)
    block, the fully pivoted ACA over its far pivots or the IACA (columns in
    the geometric order of the far pivots, rows by maximum value);
    ``pivotselector`` is the Phase-1 far-pivot selection."""

    def __init__(self, factorization=None, pivotselector=None, representor=None):
        if representor is not None:
            raise ValueError("representors are not supported by the directional GPU compressor")
        self.factorization = factorization if factorization is not None else _cmp.ACA(1e-4)
        self.pivotselector = pivotselector if pivotselector is not None else _cmp.GeometricPivoting()

    @property
    def compressor(self):
        return self.factorization


class DirectionBlocks:
    """The direction indices of one cluster tree as the nodes of the level
    driver: level and point range of their cluster, leaf = basis direction
    (isbasisnode), children = the matching direction of every child cluster
    (bottomup.jl:249-257). Only directions with a far field get children."""

    kind = "directions"

    def __init__(self, tree, data, ffptr):
        ne = data.ndiridx
        node = data.node_of()
        self.tree = tree
        self.data = data
        self.node = node
        self.points = tree.points
        self.perm = tree.perm
        self.level = tree.level[node]
        self.lo = tree.lo[node]
        self.hi = tree.hi[node]
        self.center = tree.center[node]
        self.radius = tree.radius[node]
        leafnode = tree.leafmask()
        nch = np.diff(tree.childptr)
        # first child without directions or with only the low-frequency one
        fc = np.where(nch > 0, tree.children[np.minimum(tree.childptr[:-1], max(len(tree.children) - 1, 0))]
                      if len(tree.children) else 0, -1)
        nd = np.diff(data.nodetodirsptr)
        fcd = np.zeros(tree.nnodes, bool)
        has = fc >= 0
        fnd = nd[fc[has]]
        first = data.dirs[np.minimum(data.nodetodirsptr[fc[has]], max(ne - 1, 0))] if ne else np.zeros(len(fnd), np.int64)
        fcd[has] = (fnd == 0) | ((fnd == 1) & (first == 0))
        dirs = data.dirs
        self.isbasis = leafnode[node] | ((dirs != 0) & fcd[node])
        active = np.diff(ffptr) > 0
        E = np.nonzero(~self.isbasis & active)[0]
        T = node[E]
        loc = E - data.nodetodirsptr[T]
        k = nch[T]
        C = tree.children[_cmp.concat_ranges(tree.childptr[T], k)] if len(E) else np.zeros(0, np.int64)
        locr = np.repeat(loc, k)
        nmap = data.nodetochdirptr[C + 1] - data.nodetochdirptr[C]
        if np.any(locr >= nmap):
            raise ValueError("DirectionalData: a child lacks the direction map of its parent's direction")
        cl = data.partochddirmap[data.nodetochdirptr[C] + locr]
        if np.any(cl < 0):
            raise ValueError("DirectionalData: a parent direction maps to no direction of a child")
        counts = np.zeros(ne, np.int64)
        counts[E] = k
        self.childptr = np.zeros(ne + 1, np.int64)
        np.cumsum(counts, out=self.childptr[1:])
        self.children = (data.nodetodirsptr[C] + cl).astype(np.int64)

    @property
    def nnodes(self):
        return int(len(self.node))

    @property
    def npoints(self):
        return int(self.perm.shape[0])

    def leafmask(self):
        return self.isbasis

    def isleaf(self, e):
        return bool(self.isbasis[e])

    def values(self, e):
        return self.perm[self.lo[e]:self.hi[e]]

    def nvalues(self, e):
        return int(self.hi[e] - self.lo[e])

    def childrenof(self, e):
        return self.children[self.childptr[e]:self.childptr[e + 1]]

    def levels(self):
        return np.unique(self.level)

    def levelnodes(self, lev):
        return np.nonzero(self.level == lev)[0]

    def layout_1based(self):
        return dict(
            perm=np.ascontiguousarray(self.perm + 1),
            nodelo=np.ascontiguousarray(self.lo + 1),
            nodehi=np.ascontiguousarray(self.hi.copy()),
            level=np.ascontiguousarray(self.level + 1),
            childptr=np.ascontiguousarray(self.childptr + 1),
            children=np.ascontiguousarray(self.children + 1) if len(self.children) else np.zeros(1, np.int64),
        )


def farpivots(blocks, othertree, ffptr, ffnodes, maxrank, selector, otherpoints=None):
    """Phase 1 of every direction index over its dirfarfield: on the device
    (nca_farpivots) when the selector allows it and the other side's device
    points are given, else on the host (nca_tree.cpp ncat_farpivots_lists).
    Returns a 0-based CSR of point ids."""
    if getattr(selector, "device", False) and otherpoints is not None:
        t = blocks.tree
        return _cmp.device_farpivots(otherpoints, t.center[blocks.node], (t.hi - t.lo)[blocks.node], ffptr, ffnodes,
                                     othertree, maxrank, selector.weight, selector.oversample)
    L = tree_lib()
    t = blocks.tree
    arrs = [np.ascontiguousarray(a, np.int64) for a in (blocks.node, t.lo, t.hi, othertree.lo, othertree.hi,
                                                         othertree.perm, ffptr, ffnodes)]
    ctr = np.ascontiguousarray(t.center, np.float64)
    pts = np.ascontiguousarray(othertree.points, np.float64)
    P = lambda a: _p(a) if len(a) else None
    h = L.ncat_farpivots_lists(blocks.nnodes, P(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(ctr), _p(arrs[3]),
                               _p(arrs[4]), _p(arrs[5]), _p(pts), _p(arrs[6]), P(arrs[7]), int(maxrank),
                               float(selector.weight), float(selector.oversample))
    nnz = L.ncat_csr_nnz(h)
    ptr = np.empty(blocks.nnodes + 1, np.int64)
    idx = np.empty(nnz, np.int64)
    L.ncat_csr_export(h, _p(ptr), _p(idx) if nnz else None)
    L.ncat_csr_free(h)
    return ptr, idx


class DirBasisStore:
    """BasisStore with a DirBasisTraversalPlan (abstractnca.jl:5-9): clusters
    ``nodes``, their direction indices diridx[idxtodiridxptr[i]:...[i+1]] and
    one block per direction index (offsets into blocks)."""

    def __init__(self, nodes, idxtodiridxptr, diridx, offsets, blocks):
        self.nodes, self.idxtodiridxptr, self.diridx = nodes, idxtodiridxptr, diridx
        self.offsets, self.blocks = offsets, blocks


class DirTransferStore:
    """TransferStore with a DirTransferTraversalPlan (abstractnca.jl:41-48):
    level -> parent direction indices -> child direction indices, plus the
    groups (the directions of one cluster) of each level."""

    def __init__(self, leveltoidcsptr, levelidcs, idxtochdsptr, chdsidcs, leveltogroupsptr, groupsptr, offsets,
                 blocks):
        self.leveltoidcsptr, self.levelidcs = leveltoidcsptr, levelidcs
        self.idxtochdsptr, self.chdsidcs = idxtochdsptr, chdsidcs
        self.leveltogroupsptr, self.groupsptr = leveltogroupsptr, groupsptr
        self.offsets, self.blocks = offsets, blocks


def dir_stores(f, blocks, side, host_blocks=True):
    """_build_dirbasis_storage / _build_dirtransfer_storage (nestedbasis.jl,
    transfermatrices.jl:40-99) from the factors of the direction indices
    (host_blocks=False: plans and offsets only, for resident factors)."""
    b, t = _cmp.stores_from_factors(f, blocks, side, blocks=host_blocks)
    node = blocks.node
    bn = node[b.nodes]
    nodes, counts = np.unique(bn, return_counts=True) if len(bn) else (np.zeros(0, np.int64), np.zeros(0, np.int64))
    dptr = np.zeros(len(nodes) + 1, np.int64)
    np.cumsum(counts, out=dptr[1:])
    basis = DirBasisStore(nodes.astype(np.int64), dptr, b.nodes.astype(np.int64), b.offsets, b.blocks)
    # groups: consecutive parents of one cluster within a level
    gl, gp = [0], [0]
    for lev in range(len(t.leveltoidcsptr) - 1):
        par = t.levelidcs[t.leveltoidcsptr[lev]:t.leveltoidcsptr[lev + 1]]
        if len(par):
            pn = node[par]
            starts = np.concatenate([[0], np.nonzero(np.diff(pn))[0] + 1, [len(pn)]])
            gp.extend((t.leveltoidcsptr[lev] + starts[1:]).tolist())
        gl.append(len(gp) - 1)
    transfer = DirTransferStore(t.leveltoidcsptr, t.levelidcs, t.idxtochdsptr, t.chdsidcs,
                                np.asarray(gl, np.int64), np.asarray(gp, np.int64), t.offsets, t.blocks)
    return basis, transfer


def _entry_of_position(data):
    """Direction index of every far position of a DirectionalData."""
    pos = np.arange(len(data.fardata.farnodes))
    return np.searchsorted(data.diridxtofarptr, pos, side="right") - 1


def coupling_plan(testdata, trialdata):
    """DirCouplingTraversalPlan (couplingmatrices.jl:55-111): per far pair of
    the test CSR (in its direction order) the test and the trial direction
    index. Returns (nodetodiridxptr, tidcs, sidcs), 0-based."""
    tptr, tnodes = testdata.fardata.nodetofarnodesptr, testdata.fardata.farnodes
    sptr, snodes = trialdata.fardata.nodetofarnodesptr, trialdata.fardata.farnodes
    ntest = len(tptr) - 1
    tidcs = _entry_of_position(testdata)
    sowner = np.repeat(np.arange(len(sptr) - 1), np.diff(sptr))
    skey = sowner * ntest + snodes
    order = np.argsort(skey, kind="stable")
    sentry = _entry_of_position(trialdata)
    towner = np.repeat(np.arange(ntest), np.diff(tptr))
    qkey = tnodes * ntest + towner
    at = np.searchsorted(skey[order], qkey)
    if len(qkey) and (np.any(at >= len(order)) or np.any(skey[order][np.minimum(at, len(order) - 1)] != qkey)):
        raise ValueError("coupling_plan: the trial far data is not the transpose of the test far data")
    sidcs = sentry[order[at]] if len(qkey) else np.zeros(0, np.int64)
    return tptr.copy(), tidcs.astype(np.int64), sidcs.astype(np.int64)


def dh2desc(tree, partition, testdata, trialdata, testf, trialf, tb, tt, sb, st, cplan):
    """The nca_dh2_desc arrays, 1-based."""
    ta, sa = tree.testtree, tree.trialtree
    i64 = lambda a: np.ascontiguousarray(a, np.int64)
    c128 = lambda a: np.ascontiguousarray(a, np.complex128) if a is not None and len(a) else np.zeros(1, np.complex128)
    nodeptr, tidcs, sidcs = cplan
    d = dict(
        M=ta.npoints, N=sa.npoints, nn_test=ta.nnodes, nn_trial=sa.nnodes,
        ne_test=testdata.ndiridx, ne_trial=trialdata.ndiridx,
        test_perm=i64(ta.perm + 1), test_lo=i64(ta.lo + 1), test_hi=i64(ta.hi),
        trial_perm=i64(sa.perm + 1), trial_lo=i64(sa.lo + 1), trial_hi=i64(sa.hi),
        test_dirptr=i64(testdata.nodetodirsptr + 1), trial_dirptr=i64(trialdata.nodetodirsptr + 1),
        test_cptr=i64(_nca.coefficient_ptr(testf) + 1), trial_cptr=i64(_nca.coefficient_ptr(trialf) + 1),
        nbnodes_test=len(tb.nodes), nbnodes_trial=len(sb.nodes),
        bt_nodes=i64(tb.nodes + 1), bt_dirptr=i64(tb.idxtodiridxptr + 1), bt_diridx=i64(tb.diridx + 1),
        bt_off=i64(tb.offsets),
        bs_nodes=i64(sb.nodes + 1), bs_dirptr=i64(sb.idxtodiridxptr + 1), bs_diridx=i64(sb.diridx + 1),
        bs_off=i64(sb.offsets),
        bt_blocks=c128(tb.blocks), bs_blocks=c128(sb.blocks),
        nlev_test=len(tt.leveltoidcsptr) - 1, nlev_trial=len(st.leveltoidcsptr) - 1,
        tt_levptr=i64(tt.leveltoidcsptr + 1), tt_levidcs=i64(tt.levelidcs + 1),
        tt_chdptr=i64(tt.idxtochdsptr + 1), tt_chds=i64(tt.chdsidcs + 1), tt_off=i64(tt.offsets),
        st_levptr=i64(st.leveltoidcsptr + 1), st_levidcs=i64(st.levelidcs + 1),
        st_chdptr=i64(st.idxtochdsptr + 1), st_chds=i64(st.chdsidcs + 1), st_off=i64(st.offsets),
        tt_blocks=c128(tt.blocks), st_blocks=c128(st.blocks),
        cp_nodeptr=i64(nodeptr + 1), cp_tidcs=i64(tidcs + 1), cp_sidcs=i64(sidcs + 1),
        nearptr=i64(partition.nearptr + 1), nearnodes=i64(partition.nearnodes + 1),
    )
    for key in list(d):
        if isinstance(d[key], np.ndarray) and d[key].size == 0:
            d[key] = np.zeros(1, d[key].dtype)
    return d


def _abi_dh2desc(d):
    desc = _abi.DH2Desc()
    for name, _ in _abi.DH2Desc._fields_:
        v = d[name]
        setattr(desc, name, _p(v) if isinstance(v, np.ndarray) else int(v))
    return desc


class DirectionalNCACore(_nca.PetrovGalerkinNCACore):
    """PetrovGalerkinNCACore for fardata::DirectionalData: the directional
    bases of both trees (BottomUp), the directional coupling store and the
    device operator of nca_dh2_create. Products, graphs, autotuning and
    statistics are those of the plain core."""

    def __init__(self, kernel, testpoints, trialpoints, tree, testcompressor, trialcompressor, isnear,
                 maxrank=50, precision=_abi.NCA_PREC_FP64, ctx=None, verbose=False, partition=None,
                 reuse_symmetric=True, resident=True):
        self.ctx = ctx or default_context()
        self.kernel = kernel
        self.tree = tree
        self.maxrank = maxrank
        self.isnear = isnear
        self.resident = res = bool(resident)
        timings = {}
        t0 = time.perf_counter()
        if partition is None:
            partition = _trees.interactions_device(tree, self.ctx, wideband=isnear)
        self.partition = partition
        self.testdata, self.trialdata = _dir.fardata(tree, isnear, self.partition)
        timings["partition"] = time.perf_counter() - t0
        self.farmatrix = HIPKernelMatrix(kernel, testpoints, trialpoints, self.ctx)
        t0 = time.perf_counter()
        td, sd = self.testdata, self.trialdata
        ts, ss = testcompressor.pivotselector, trialcompressor.pivotselector
        # one tree, one point set, the same directional data and compressors:
        # the trial side repeats the test side bit for bit (symmetric kernel)
        self.symmetric = bool(
            reuse_symmetric and tree.testtree is tree.trialtree and _nca._same_points(testpoints, trialpoints)
            and type(ts) is type(ss) and vars(ts) == vars(ss)
            and vars(testcompressor.compressor) == vars(trialcompressor.compressor)
            and all(np.array_equal(getattr(td, n), getattr(sd, n)) for n in
                    ("nodetodirsptr", "dirs", "diridxtofarptr", "nodetochdirptr", "partochddirmap"))
            and np.array_equal(td.fardata.farnodes, sd.fardata.farnodes))
        tff = _dir.dirfarfield_all(tree.testtree, td)
        self.testblocks = DirectionBlocks(tree.testtree, td, tff[0])
        fp_test = farpivots(self.testblocks, tree.trialtree, tff[0], tff[1], maxrank, ts, self.farmatrix.trial)
        if self.symmetric:
            self.trialblocks, fp_trial = self.testblocks, fp_test
        else:
            sff = _dir.dirfarfield_all(tree.trialtree, sd)
            self.trialblocks = DirectionBlocks(tree.trialtree, sd, sff[0])
            fp_trial = farpivots(self.trialblocks, tree.testtree, sff[0], sff[1], maxrank, ss, self.farmatrix.test)
        timings["phase1"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        bt = _trees.BlockTree(self.testblocks, self.trialblocks)
        self.testfactors = _cmp.batched_side(testcompressor, self.farmatrix, bt, None, 0, maxrank, fp_test, res)
        if self.symmetric:
            self.trialfactors = self.testfactors
        else:
            self.trialfactors = _cmp.batched_side(trialcompressor, self.farmatrix, bt, None, 1, maxrank, fp_trial,
                                                  res)
        self.testbasis, self.testtransfer = dir_stores(self.testfactors, self.testblocks, 0, not res)
        self.trialbasis, self.trialtransfer = dir_stores(self.trialfactors, self.trialblocks, 1, not res)
        self.testpivots = self.testfactors.pivot_csr()
        self.trialpivots = self.trialfactors.pivot_csr()
        self.couplingplan = coupling_plan(self.testdata, self.trialdata)
        self.couplingstore = _cmp.assemble_couplingstore(
            self.farmatrix, self.testpivots, self.trialpivots,
            _trees.FarData(self.testdata.diridxtofarptr, self.couplingplan[2]))
        timings["assembly"] = time.perf_counter() - t0
        self._desc = dh2desc(tree, self.partition, self.testdata, self.trialdata, self.testfactors,
                             self.trialfactors, self.testbasis, self.testtransfer, self.trialbasis,
                             self.trialtransfer, self.couplingplan)
        self._desc_blocks = not res
        t0 = time.perf_counter()
        h = ctypes.c_void_p()
        kern = kernel.abi()
        abidesc = _abi_dh2desc(self._desc)
        if res:
            call(hip_lib().nca_dh2_create_resident(
                self.ctx.handle, ctypes.byref(kern), self.farmatrix.test.handle, self.farmatrix.trial.handle,
                ctypes.byref(abidesc), self.couplingstore.handle, int(precision), self.testfactors.pool.handle,
                self.trialfactors.pool.handle, ctypes.byref(h)))
        else:
            call(hip_lib().nca_dh2_create(self.ctx.handle, ctypes.byref(kern), self.farmatrix.test.handle,
                                          self.farmatrix.trial.handle, ctypes.byref(abidesc),
                                          self.couplingstore.handle, int(precision), ctypes.byref(h)))
        self.handle = h
        timings["h2_create"] = time.perf_counter() - t0
        self.timings = timings
        self.dim = (tree.testtree.npoints, tree.trialtree.npoints)
        if verbose:
            print("directional assembly timings:", {k: round(v, 4) for k, v in timings.items()}, flush=True)

    def host_stores(self):
        """The directional stores with their blocks on the host (packed from the
        downloaded factors when the factors are resident)."""
        if not getattr(self, "_desc_blocks", True):
            self.testbasis, self.testtransfer = dir_stores(self.testfactors, self.testblocks, 0)
            self.trialbasis, self.trialtransfer = dir_stores(self.trialfactors, self.trialblocks, 1)
            self._desc = dh2desc(self.tree, self.partition, self.testdata, self.trialdata, self.testfactors,
                                 self.trialfactors, self.testbasis, self.testtransfer, self.trialbasis,
                                 self.trialtransfer, self.couplingplan)
            self._desc_blocks = True
        return self.testbasis, self.testtransfer, self.trialbasis, self.trialtransfer
