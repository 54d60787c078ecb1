"""Shared set-up of the parity tests: the same tree, partition and Phase-1 far
pivots feed both the HIP path and the oracle."""
import numpy as np

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from nca_amd import nca as NC
from nca_amd import trees as T
from oracle import oracle as O


class Problem:
    def __init__(self, testpts, trialpts, tree, k, sign=1, tol=1e-4, maxrank=50, eta=1.0):
        self.testpts, self.trialpts, self.tree = testpts, trialpts, tree
        self.k, self.sign, self.tol, self.maxrank = k, sign, tol, maxrank
        self.partition = nca_amd.interactions(tree, eta)
        self.tf, self.sf = nca_amd.fardata(tree, self.partition)
        self.fpt = nca_amd.farpivots(tree.testtree, tree.trialtree, self.tf, maxrank)
        self.fps = nca_amd.farpivots(tree.trialtree, tree.testtree, self.sf, maxrank)


def sphere_problem(n, k=1.0, tol=1e-4, minvalues=64, maxrank=50):
    pts = geometry.fibonacci_sphere(n)
    tr = nca_amd.TwoNTree(pts, minvalues=minvalues)
    return Problem(pts, pts, nca_amd.BlockTree(tr, tr), k, 1, tol, maxrank)


def grid_problem(tol=1e-4):
    """32 x 32 grid of [0,1]^2, trial = test + (1.1, 0, 0.2), k = 2 pi / 10, 2-means trees, minvalues 32."""
    tp = geometry.grid_square(32)
    sp = tp + np.array([1.1, 0.0, 0.2])
    tree = nca_amd.BlockTree(nca_amd.KMeansTree(tp, 2, 32), nca_amd.KMeansTree(sp, 2, 32))
    return Problem(tp, sp, tree, 2 * np.pi / 10, 1, tol, 50)


def oracle_factors(res, tree, fp):
    pivptr, piv, xoff, X, st = res
    f = C.SideFactors(tree.nnodes)
    nfar = np.diff(fp[0])
    leaf = tree.leafmask()
    for lev in sorted(tree.levels().tolist(), reverse=True):
        for t in tree.levelnodes(lev):
            if nfar[t] == 0:
                continue
            f.assigned[t] = True
            r = pivptr[t + 1] - pivptr[t]
            f.rank[t] = r
            f.pivots[t] = piv[pivptr[t]:pivptr[t + 1]]
            m = tree.nvalues(t) if leaf[t] else int(sum(f.rank[c] for c in tree.childrenof(t) if f.assigned[c]))
            f.m[t] = m
            f.X[t] = X[xoff[t]:xoff[t + 1]].reshape((m, r), order="F")
    for lev in tree.levels():
        f.level_transfer_nodes[lev] = [int(t) for t in tree.levelnodes(lev) if not leaf[t] and f.assigned[t]]
    return f


def oracle_topdown(p, side, ranks_from=None):
    """oracle.topdown for one side of problem p; returns SideFactors with the
    rank of every node from ``ranks_from`` (assigned flags of the device
    result) or from the oracle's pivots (nodes with pivots)."""
    tree = p.tree.testtree if side == 0 else p.tree.trialtree
    other = p.tree.trialtree if side == 0 else p.tree.testtree
    fd = p.tf if side == 0 else p.sf
    res = O.topdown(side, p.k, p.sign, p.testpts, p.trialpts, tree, other, fd.nodetofarnodesptr, fd.farnodes,
                    p.maxrank, p.tol)
    assert res[6] == 0
    pivptr, piv, cpivptr, cpiv, xoff, X, st = res
    f = C.SideFactors(tree.nnodes)
    leaf = tree.leafmask()
    assigned = np.diff(pivptr) > 0 if ranks_from is None else ranks_from
    for lev in sorted(tree.levels().tolist(), reverse=True):
        for t in tree.levelnodes(lev):
            if not assigned[t]:
                continue
            f.assigned[t] = True
            r = pivptr[t + 1] - pivptr[t]
            f.rank[t] = r
            f.pivots[t] = piv[pivptr[t]:pivptr[t + 1]]
            m = tree.nvalues(t) if leaf[t] else int(sum(f.rank[c] for c in tree.childrenof(t) if f.assigned[c]))
            f.m[t] = m
            f.X[t] = X[xoff[t]:xoff[t + 1]].reshape((m, r), order="F")
    for lev in tree.levels():
        f.level_transfer_nodes[lev] = [int(t) for t in tree.levelnodes(lev) if not leaf[t] and f.assigned[t]]
    f.colpivots = (cpivptr, cpiv)
    return f


def oracle_assembly(p):
    tres = O.bases(0, p.k, p.sign, p.testpts, p.trialpts, p.tree.testtree, p.fpt[0], p.fpt[1], p.maxrank, p.tol)
    sres = O.bases(1, p.k, p.sign, p.testpts, p.trialpts, p.tree.trialtree, p.fps[0], p.fps[1], p.maxrank, p.tol)
    assert tres[4] == 0 and sres[4] == 0
    ft = oracle_factors(tres, p.tree.testtree, p.fpt)
    fs = oracle_factors(sres, p.tree.trialtree, p.fps)
    return ft, fs


def oracle_operator(p, ft, fs):
    tb, tt = C.stores_from_factors(ft, p.tree.testtree, 0)
    sb, st = C.stores_from_factors(fs, p.tree.trialtree, 1)
    d = NC.h2desc(p.tree, p.partition, p.tf, ft, fs, tb, tt, sb, st)
    tp, sp = ft.pivot_csr(), fs.pivot_csr()
    off, blocks = O.coupling(p.k, p.sign, p.testpts, p.trialpts, p.tf.nodetofarnodesptr, p.tf.farnodes,
                             tp[0], tp[1], sp[0], sp[1])
    return O.OracleH2(d, p.k, p.sign, p.testpts, p.trialpts, off, blocks), (off, blocks)


def oracle_operator_rows(p, ft, fs, rows):
    """The oracle's operator for the rows ``rows`` only: coupling blocks are
    assembled (and applied) for the test clusters that hold one of the rows,
    which are all the clusters whose coupling reaches those rows through the
    downward pass. matvec(..., rows=rows) is then exact on those rows."""
    tt = p.tree.testtree
    need = O.nodes_holding(tt.lo, tt.hi, tt.perm, rows)
    fptr, fnodes = O.restrict_rows(p.tf.nodetofarnodesptr, p.tf.farnodes, need)
    tb, tT = C.stores_from_factors(ft, tt, 0)
    sb, sT = C.stores_from_factors(fs, p.tree.trialtree, 1)
    far = T.FarData(fptr, fnodes)
    d = NC.h2desc(p.tree, p.partition, far, ft, fs, tb, tT, sb, sT)
    tp, sp = ft.pivot_csr(), fs.pivot_csr()
    off, blocks = O.coupling(p.k, p.sign, p.testpts, p.trialpts, fptr, fnodes, tp[0], tp[1], sp[0], sp[1])
    return O.OracleH2(d, p.k, p.sign, p.testpts, p.trialpts, off, blocks)


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def bitwise_equal(a, b):
    """Same shape and the same IEEE bit patterns in every entry."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


class DeviceBuffers:
    """Device memory through the HIP runtime (test plumbing). hipMemcpy runs on
    the null stream: synchronize the context stream before reading results."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def upload(self, arr):
        arr = np.asfortranarray(arr)
        p = self.ct.c_void_p()
        assert self.lib.hipMalloc(self.ct.byref(p), self.ct.c_size_t(max(arr.nbytes, 16))) == 0
        self.held.append(p.value)
        self.copy_in(p.value, arr)
        return p.value

    def copy_in(self, dst, arr):
        arr = np.asfortranarray(arr)
        assert self.lib.hipMemcpy(self.ct.c_void_p(dst), arr.ctypes.data_as(self.ct.c_void_p), arr.nbytes, 1) == 0

    def download(self, src, shape):
        out = np.zeros(shape, np.complex128, order="F")
        assert self.lib.hipMemcpy(out.ctypes.data_as(self.ct.c_void_p), self.ct.c_void_p(src), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(self.ct.c_void_p(p))
        self.held = []
