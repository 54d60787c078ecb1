"""ORACLE — test infrastructure only.

ctypes front end of liboracle.so (nca_oracle.c), the FP64 CPU restatement of the
batched bottom-up compressor and the H^2 matvec. Imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
D = ctypes.c_double


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


class H2(ctypes.Structure):
    _fields_ = [
        ("M", I64), ("N", I64), ("nn_test", I64), ("nn_trial", I64),
        ("k", D), ("sign", I32),
        ("testpts", P), ("trialpts", P),
        ("test_perm", P), ("test_lo", P), ("test_hi", P), ("trial_perm", P), ("trial_lo", P), ("trial_hi", P),
        ("test_cptr", P), ("trial_cptr", P),
        ("nbasis_test", I64), ("nbasis_trial", I64),
        ("basis_test_nodes", P), ("basis_test_off", P), ("basis_trial_nodes", P), ("basis_trial_off", P),
        ("basis_test", P), ("basis_trial", P),
        ("nlev_test", I64), ("nlev_trial", I64),
        ("tt_levptr", P), ("tt_levidcs", P), ("tt_chdptr", P), ("tt_chds", P), ("tt_off", P),
        ("st_levptr", P), ("st_levidcs", P), ("st_chdptr", P), ("st_chds", P), ("st_off", P),
        ("tt_blocks", P), ("st_blocks", P),
        ("cp_ptr", P), ("cp_nodes", P), ("cp_off", P), ("cp_blocks", P),
        ("nearptr", P), ("nearnodes", P),
        ("near_store", P),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_sincos.argtypes = [D, P, P]
        L.orc_near_assemble.restype = P
        L.orc_near_assemble.argtypes = [P]
        L.orc_near_entries.restype = I64
        L.orc_near_entries.argtypes = [P]
        L.orc_near_free.argtypes = [P]
        L.orc_entry.argtypes = [D, I32, P, P, P]
        L.orc_assemble_level.restype = I32
        L.orc_assemble_level.argtypes = [D, I32, P, P, I64, P, P, P, P, I64, I32, P]
        L.orc_aca_fullpiv.restype = I64
        L.orc_aca_fullpiv.argtypes = [P, I64, I64, I64, I64, D, I32, P, P]
        L.orc_iaca.restype = I64
        L.orc_iaca.argtypes = [P, I64, I64, I64, I64, D, P, P]
        L.orc_skeleton_solve.restype = I64
        L.orc_skeleton_solve.argtypes = [P, I64, I64, P, P]
        L.orc_bases.restype = I64
        L.orc_bases.argtypes = [I32, D, I32, P, P, I64, P, P, P, P, P, P, P, P, I64, D, I32, P, P, P, P, I64]
        L.orc_coupling.argtypes = [D, I32, P, P, I64, P, P, P, P, P, P, P, P]
        L.orc_paca.restype = I64
        L.orc_paca.argtypes = [D, I32, I32, P, P, I64, P, I64, P, I64, D, P, P, P, P]
        L.orc_topdown.restype = I64
        L.orc_topdown.argtypes = [I32, D, I32, P, P, I64, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, P, P, P, P, P,
                                  P, I64]
        L.orc_h2_matvec.argtypes = [P, I32, I64, P, I64, P, I64, P, P, I32, I64, P]
        L.orc_dense_rows.argtypes = [D, I32, P, P, I64, I64, P, I64, P, I64, P]
        L.orc_num_threads.restype = I32
        L.orc_set_num_threads.argtypes = [I32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(P)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def sincos(x):
    s = ctypes.c_double()
    c = ctypes.c_double()
    lib().orc_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def entry(k, sign, x, y):
    g = np.zeros(2)
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ya = np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_entry(float(k), int(sign), _p(xa), _p(ya), _p(g))
    return complex(g[0], g[1])


def assemble_level(k, sign, testpts, trialpts, rowidcs, colidcs, maxrank, transposed=False):
    """rowidcs/colidcs: lists of 1-based index arrays. Returns (buffer, rowoffsets)."""
    nb = len(rowidcs)
    rowptr = np.ones(nb + 1, np.int64)
    rowptr[1:] = 1 + np.cumsum([len(r) for r in rowidcs])
    colptr = np.ones(nb + 1, np.int64)
    colptr[1:] = 1 + np.cumsum([len(c) for c in colidcs])
    rowidx = _i64(np.concatenate([np.asarray(r, np.int64) for r in rowidcs]) if nb else np.zeros(0))
    colidx = _i64(np.concatenate([np.asarray(c, np.int64) for c in colidcs]) if nb else np.zeros(0))
    ld = int(rowptr[-1] - 1)
    buf = np.zeros((maxrank, max(ld, 0)), np.complex128).T.copy(order="F")
    tp = np.ascontiguousarray(testpts, dtype=np.float64)
    sp = np.ascontiguousarray(trialpts, dtype=np.float64)
    st = lib().orc_assemble_level(float(k), int(sign), _p(tp), _p(sp), nb, _p(rowptr), _p(rowidx), _p(colptr),
                                  _p(colidx), int(maxrank), int(bool(transposed)), _p(buf))
    if st != 0:
        raise ValueError("a block has more columns than maxrank")
    return buf, rowptr - 1


def aca_fullpiv(block, maxrank, tol, sampled=True):
    """Fully pivoted ACA on a copy of ``block``; returns (rank, rowpiv, colpiv) 1-based."""
    A = np.asfortranarray(np.array(block, dtype=np.complex128, copy=True))
    m, n = A.shape
    rp = np.zeros(max(maxrank, 1), np.int64)
    cp = np.zeros(max(maxrank, 1), np.int64)
    r = lib().orc_aca_fullpiv(_p(A), m, n, max(m, 1), int(maxrank), float(tol), int(bool(sampled)), _p(rp), _p(cp))
    return int(r), rp[:r].copy(), cp[:r].copy()


def iaca(block, maxrank, tol):
    """IACA with geometric column order on a copy of ``block`` (its columns
    are the candidates in order); returns (rank, rowpiv, colpiv) 1-based."""
    A = np.asfortranarray(np.array(block, dtype=np.complex128, copy=True))
    m, n = A.shape
    rp = np.zeros(max(maxrank, 1), np.int64)
    cp = np.zeros(max(maxrank, 1), np.int64)
    r = lib().orc_iaca(_p(A), m, n, max(m, 1), int(maxrank), float(tol), _p(rp), _p(cp))
    return int(r), rp[:r].copy(), cp[:r].copy()


def skeleton_solve(L, rowpiv):
    """X = L / L[rowpiv, :]; returns (X, independent) with independent = -1 when regular."""
    Lm = np.asfortranarray(np.array(L, dtype=np.complex128))
    m, r = Lm.shape
    X = np.zeros((m, r), np.complex128, order="F")
    rp = _i64(rowpiv)
    st = lib().orc_skeleton_solve(_p(Lm), m, r, _p(rp), _p(X))
    return X, int(st)


def bases(side, k, sign, testpts, trialpts, tree, farpivptr, farpiv, maxrank, tol, sampled=True, iaca=False):
    """testbases/trialbases(::BatchedBottomUp). tree: ClusterTree of this side;
    farpivptr/farpiv: 0-based CSR of 0-based point ids (Phase 1). iaca: the
    incomplete ACA over the far pivots in their order instead of the fully
    pivoted ACA. Returns (pivptr 0-based, pivots 0-based, xoffsets, X, status)."""
    if iaca:
        sampled = 2
    lay = tree.layout_1based()
    nn = tree.nnodes
    fptr = _i64(np.asarray(farpivptr) + 1)
    fpiv = _i64(np.asarray(farpiv) + 1)
    pivptr = np.zeros(nn + 1, np.int64)
    pivots = np.zeros(max(nn * maxrank, 1), np.int64)
    xoff = np.zeros(nn + 1, np.int64)
    sizes = tree.hi - tree.lo
    cap = int(2 * np.sum(np.where(tree.leafmask(), sizes * np.minimum(maxrank, sizes), 0)))
    tp = np.ascontiguousarray(testpts, dtype=np.float64)
    sp = np.ascontiguousarray(trialpts, dtype=np.float64)
    for _ in range(2):
        X = np.zeros(max(cap, 1), np.complex128)
        st = lib().orc_bases(int(side), float(k), int(sign), _p(tp), _p(sp), nn, _p(lay["perm"]), _p(lay["nodelo"]),
                             _p(lay["nodehi"]), _p(lay["level"]), _p(lay["childptr"]), _p(lay["children"]),
                             _p(fptr), _p(fpiv), int(maxrank), float(tol), int(sampled), _p(pivptr),
                             _p(pivots), _p(xoff), _p(X), cap)
        if st != 1:
            break
        cap = int(xoff[-1])
    pivptr0 = pivptr - 1
    return pivptr0, pivots[:pivptr0[-1]] - 1, xoff, X[:xoff[-1]].copy(), int(st)


def paca(k, sign, transposed, rowpts, colpts, rows, cols, maxrank, tol):
    """Partially pivoted ACA of the block G(rows, cols) (0-based point ids;
    transposed: G(cols, rows)); returns (rank, rowpiv, colpiv) 1-based
    positions and the factors U (m x rank), V (n x rank)."""
    rp_ = np.ascontiguousarray(rowpts, dtype=np.float64)
    cp_ = np.ascontiguousarray(colpts, dtype=np.float64)
    r1, c1 = _i64(np.asarray(rows) + 1), _i64(np.asarray(cols) + 1)
    m, n = len(rows), len(cols)
    km = max(min(maxrank, m, n), 1)
    rpv, cpv = np.zeros(km, np.int64), np.zeros(km, np.int64)
    U = np.zeros((max(m, 1), km), np.complex128, order="F")
    V = np.zeros((max(n, 1), km), np.complex128, order="F")
    r = lib().orc_paca(float(k), int(sign), int(bool(transposed)), _p(rp_), _p(cp_), m, _p(r1), n, _p(c1),
                       int(maxrank), float(tol), _p(rpv), _p(cpv), _p(U), _p(V))
    U = np.asfortranarray(U.ravel(order="F")[:m * r].reshape((m, r), order="F"))
    V = np.asfortranarray(V.ravel(order="F")[:n * r].reshape((n, r), order="F"))
    return int(r), rpv[:r].copy(), cpv[:r].copy(), U, V


def topdown(side, k, sign, testpts, trialpts, tree, othertree, farptr, farnodes, maxrank, tol):
    """testbases/trialbases(::TopDown) of one side. tree / othertree:
    ClusterTrees of this and of the other side; farptr / farnodes: 0-based CSR
    of the far nodes of every node. Returns (pivptr, pivots, cpivptr, cpivots,
    xoffsets, X, status), pivots 0-based point ids."""
    lay, olay = tree.layout_1based(), othertree.layout_1based()
    nn = tree.nnodes
    parent = _i64(np.asarray(tree.parent) + 1)
    fptr, fnod = _i64(np.asarray(farptr) + 1), _i64(np.asarray(farnodes) + 1) if len(farnodes) else np.zeros(1, np.int64)
    pivptr, cpivptr = np.zeros(nn + 1, np.int64), np.zeros(nn + 1, np.int64)
    pivots = np.zeros(max(nn * maxrank, 1), np.int64)
    cpivots = np.zeros(max(nn * maxrank, 1), np.int64)
    xoff = np.zeros(nn + 1, np.int64)
    tp = np.ascontiguousarray(testpts, dtype=np.float64)
    sp = np.ascontiguousarray(trialpts, dtype=np.float64)
    sizes = tree.hi - tree.lo
    cap = int(np.sum(sizes * maxrank))
    for _ in range(2):
        X = np.zeros(max(cap, 1), np.complex128)
        st = lib().orc_topdown(int(side), float(k), int(sign), _p(tp), _p(sp), nn, _p(lay["perm"]),
                               _p(lay["nodelo"]), _p(lay["nodehi"]), _p(lay["level"]), _p(parent),
                               _p(lay["childptr"]), _p(lay["children"]), _p(olay["perm"]), _p(olay["nodelo"]),
                               _p(olay["nodehi"]), _p(fptr), _p(fnod), int(maxrank), float(tol), _p(pivptr),
                               _p(pivots), _p(cpivptr), _p(cpivots), _p(xoff), _p(X), cap)
        if st != 1:
            break
        cap = int(xoff[-1])
    pp, cpp = pivptr - 1, cpivptr - 1
    return pp, pivots[:pp[-1]] - 1, cpp, cpivots[:cpp[-1]] - 1, xoff, X[:xoff[-1]].copy(), int(st)


def coupling(k, sign, testpts, trialpts, farptr, farnodes, testpivptr, testpiv, trialpivptr, trialpiv):
    """All inputs 0-based CSR; returns (offsets, blocks)."""
    ntest = len(farptr) - 1
    rt = np.diff(testpivptr)
    rs = np.diff(trialpivptr)
    rows = np.repeat(np.arange(ntest), np.diff(farptr))
    sizes = rt[rows] * rs[np.asarray(farnodes, np.int64)]
    offs = np.zeros(len(farnodes) + 1, np.int64)
    np.cumsum(sizes, out=offs[1:])
    blocks = np.zeros(max(int(offs[-1]), 1), np.complex128)
    tp = np.ascontiguousarray(testpts, dtype=np.float64)
    sp = np.ascontiguousarray(trialpts, dtype=np.float64)
    a = [_i64(np.asarray(farptr) + 1), _i64(np.asarray(farnodes) + 1), _i64(np.asarray(testpivptr) + 1),
         _i64(np.asarray(testpiv) + 1), _i64(np.asarray(trialpivptr) + 1), _i64(np.asarray(trialpiv) + 1)]
    lib().orc_coupling(float(k), int(sign), _p(tp), _p(sp), ntest, *[_p(x) for x in a], _p(offs), _p(blocks))
    return offs, blocks[:offs[-1]]


def nodes_holding(lo, hi, perm, rows):
    """Mask of the nodes whose value range [lo, hi) (tree positions, 0-based)
    holds a point of ``rows`` (original indices); perm[position] = index."""
    iperm = np.empty(len(perm), np.int64)
    iperm[np.asarray(perm, np.int64)] = np.arange(len(perm))
    pos = np.sort(iperm[np.asarray(rows, np.int64)])
    return np.searchsorted(pos, hi) > np.searchsorted(pos, lo)


def restrict_rows(ptr, idx, keep):
    """The CSR (ptr, idx) with the rows where ``keep`` is false emptied."""
    cnt = np.where(keep, np.diff(ptr), 0)
    nptr = np.zeros(len(ptr), np.int64)
    np.cumsum(cnt, out=nptr[1:])
    sel = np.repeat(keep, np.diff(ptr))
    return nptr, np.asarray(idx)[sel]


class OracleH2:
    """Holds the 1-based arrays of an H^2 description (see nca_amd.h2desc) for orc_h2_matvec."""

    def __init__(self, desc, k, sign, testpts, trialpts, cp_off, cp_blocks):
        self._keep = []
        h = H2()
        h.M, h.N, h.nn_test, h.nn_trial = desc["M"], desc["N"], desc["nn_test"], desc["nn_trial"]
        h.k = float(k)
        h.sign = int(sign)
        for name in ("testpts", "trialpts"):
            a = np.ascontiguousarray(testpts if name == "testpts" else trialpts, dtype=np.float64)
            self._keep.append(a)
            setattr(h, name, _p(a))
        for name, _ in H2._fields_:
            if name in ("M", "N", "nn_test", "nn_trial", "k", "sign", "testpts", "trialpts", "near_store"):
                continue
            if name in ("nbasis_test", "nbasis_trial", "nlev_test", "nlev_trial"):
                setattr(h, name, int(desc[name]))
                continue
            if name == "cp_off":
                a = _i64(cp_off)
            elif name == "cp_blocks":
                a = _c128(cp_blocks) if len(cp_blocks) else np.zeros(1, np.complex128)
            else:
                a = desc[name]
            self._keep.append(a)
            setattr(h, name, _p(a))
        self.h = h
        self.M, self.N = desc["M"], desc["N"]
        self._near = None

    def store_near(self):
        """Assemble and keep the near field as dense block rows (the reference's
        CPU layout); later forward products over all rows use it. Returns the
        number of stored entries."""
        self.drop_near()
        self._near = lib().orc_near_assemble(ctypes.byref(self.h))
        self.h.near_store = self._near
        return int(lib().orc_near_entries(self._near))

    def drop_near(self):
        if self._near:
            self.h.near_store = None
            lib().orc_near_free(self._near)
            self._near = None

    def __del__(self):
        try:
            self.drop_near()
        except Exception:
            pass

    def matvec(self, x, op=0, alpha=1.0, beta=0.0, y=None, near=True, rows=None):
        x = np.asfortranarray(np.asarray(x, dtype=np.complex128).reshape(x.shape[0], -1))
        R = x.shape[1]
        ny = self.M if op == 0 else self.N
        yout = np.zeros((ny, R), np.complex128, order="F") if y is None else np.asfortranarray(
            np.array(y, dtype=np.complex128).reshape(ny, -1))
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        if rows is None:
            nrows, rp = -1, np.zeros(1, np.int64)
        else:
            rp = _i64(rows)
            nrows = len(rp)
        lib().orc_h2_matvec(ctypes.byref(self.h), int(op), R, _p(x), x.shape[0], _p(yout), ny, _p(al), _p(be),
                            int(bool(near)), nrows, _p(rp))
        return yout


def dense_rows(k, sign, testpts, trialpts, rows, x):
    x = np.asfortranarray(np.asarray(x, dtype=np.complex128).reshape(np.asarray(x).shape[0], -1))
    rows = _i64(rows)
    R = x.shape[1]
    y = np.zeros((len(rows), R), np.complex128, order="F")
    tp = np.ascontiguousarray(testpts, dtype=np.float64)
    sp = np.ascontiguousarray(trialpts, dtype=np.float64)
    lib().orc_dense_rows(float(k), int(sign), _p(tp), _p(sp), sp.shape[0], len(rows), _p(rows), R, _p(x),
                         x.shape[0], _p(y))
    return y


def set_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return int(lib().orc_num_threads())


def _near_block(k, sign, tp, sp):
    """Helmholtz block exp(i sign k r) / (4 pi r), 0 at r = 0 (the oracle's entry)."""
    d = tp[:, None, :] - sp[None, :, :]
    r2 = (d[:, :, 0] * d[:, :, 0] + d[:, :, 1] * d[:, :, 1]) + d[:, :, 2] * d[:, :, 2]
    r = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = (np.cos(k * r) + 1j * sign * np.sin(k * r)) / (4 * np.pi * r)
    g[r2 == 0.0] = 0.0
    return g


class OracleDH2:
    """The directional H^2 product restated stage by stage from the
    reference's directional methods, in numpy, for small problems:
    _project_to_coefficients! (src/nca/aggregate.jl:21-42),
    _aggregate_coefficients! (aggregate.jl:44-67), _couple_forward! with a
    DirCouplingTraversalPlan (src/nca/couple.jl:26-50),
    _disaggregate_coefficients! by groups (src/nca/disaggregate.jl:27-60),
    _project_to_output! (disaggregate.jl:82-103), then y += Near x
    (src/nca/petrovgalerkinnca.jl:474). Takes the 1-based nca_dh2_desc arrays
    (as a dict) and the coupling blocks (0-based complex offsets, blocks)."""

    def __init__(self, desc, k, sign, testpts, trialpts, cp_off, cp_blocks):
        self.d, self.k, self.sign = desc, float(k), int(sign)
        self.tp = np.asarray(testpts, np.float64)
        self.sp = np.asarray(trialpts, np.float64)
        self.cp_off, self.cp_blocks = np.asarray(cp_off), np.asarray(cp_blocks)

    def _values(self, side, node):
        d = self.d
        perm, lo, hi = (d["test_perm"], d["test_lo"], d["test_hi"]) if side == 0 else \
            (d["trial_perm"], d["trial_lo"], d["trial_hi"])
        return perm[lo[node] - 1:hi[node]] - 1

    def matvec(self, X, near=True, rows=None):
        """Y = A X; with ``rows`` (original test indices) only those rows of Y
        are formed (the others are left zero): the coupling, the leaf
        expansion and the near field then run only for the test clusters
        that hold one of the rows, which is exact for those rows."""
        d = self.d
        X = np.asarray(X, np.complex128)
        vec = X.ndim == 1
        X = X.reshape(X.shape[0], -1)
        R = X.shape[1]
        if rows is None:
            need = np.ones(d["nn_test"], bool)
            want = None
        else:
            need = nodes_holding(d["test_lo"] - 1, d["test_hi"], d["test_perm"] - 1, rows)
            want = np.zeros(d["M"], bool)
            want[np.asarray(rows, np.int64)] = True
        tc, sc = d["test_cptr"] - 1, d["trial_cptr"] - 1
        xh = np.zeros((sc[d["ne_trial"]], R), np.complex128)
        yh = np.zeros((tc[d["ne_test"]], R), np.complex128)
        Y = np.zeros((d["M"], R), np.complex128)
        # stage 1: every direction of a basis cluster projects the cluster's values
        for i in range(d["nbnodes_trial"]):
            node = d["bs_nodes"][i] - 1
            xn = X[self._values(1, node)]
            m = len(xn)
            for l in range(d["bs_dirptr"][i] - 1, d["bs_dirptr"][i + 1] - 1):
                e = d["bs_diridx"][l] - 1
                r = sc[e + 1] - sc[e]
                V = d["bs_blocks"][d["bs_off"][l]:d["bs_off"][l + 1]].reshape((r, m), order="F")
                xh[sc[e]:sc[e + 1]] = V @ xn
        # stage 2: levels from the deepest, parent direction += E child direction
        lp, li, cp, ch, off = (d["st_levptr"] - 1, d["st_levidcs"] - 1, d["st_chdptr"] - 1, d["st_chds"] - 1,
                               d["st_off"])
        for lev in range(d["nlev_trial"] - 1, -1, -1):
            for idx in range(lp[lev], lp[lev + 1]):
                p = li[idx]
                rp = sc[p + 1] - sc[p]
                for c_ in range(cp[idx], cp[idx + 1]):
                    c = ch[c_]
                    rc = sc[c + 1] - sc[c]
                    E = d["st_blocks"][off[c_]:off[c_ + 1]].reshape((rp, rc), order="F")
                    xh[sc[p]:sc[p + 1]] += E @ xh[sc[c]:sc[c + 1]]
        # stage 3: per test cluster its (test direction, trial direction) blocks
        nptr, tid, sid = d["cp_nodeptr"] - 1, d["cp_tidcs"] - 1, d["cp_sidcs"] - 1
        for t in np.nonzero(need)[0]:
            for i in range(nptr[t], nptr[t + 1]):
                a, b = tid[i], sid[i]
                rt, rs = tc[a + 1] - tc[a], sc[b + 1] - sc[b]
                S = self.cp_blocks[self.cp_off[i]:self.cp_off[i + 1]].reshape((rt, rs), order="F")
                yh[tc[a]:tc[a + 1]] += S @ xh[sc[b]:sc[b + 1]]
        # stage 4: levels from the root, child direction += E parent direction
        lp, li, cp, ch, off = (d["tt_levptr"] - 1, d["tt_levidcs"] - 1, d["tt_chdptr"] - 1, d["tt_chds"] - 1,
                               d["tt_off"])
        for lev in range(d["nlev_test"]):
            for idx in range(lp[lev], lp[lev + 1]):
                p = li[idx]
                rp = tc[p + 1] - tc[p]
                for c_ in range(cp[idx], cp[idx + 1]):
                    c = ch[c_]
                    rc = tc[c + 1] - tc[c]
                    E = d["tt_blocks"][off[c_]:off[c_ + 1]].reshape((rc, rp), order="F")
                    yh[tc[c]:tc[c + 1]] += E @ yh[tc[p]:tc[p + 1]]
        # stage 5: every direction of a basis cluster adds to the cluster's values
        for i in range(d["nbnodes_test"]):
            node = d["bt_nodes"][i] - 1
            if not need[node]:
                continue
            rows_ = self._values(0, node)
            m = len(rows_)
            for l in range(d["bt_dirptr"][i] - 1, d["bt_dirptr"][i + 1] - 1):
                e = d["bt_diridx"][l] - 1
                r = tc[e + 1] - tc[e]
                U = d["bt_blocks"][d["bt_off"][l]:d["bt_off"][l + 1]].reshape((m, r), order="F")
                Y[rows_] += U @ yh[tc[e]:tc[e + 1]]
        if near:
            np_, nn_ = d["nearptr"] - 1, d["nearnodes"] - 1
            for t in np.nonzero(need)[0]:
                if np_[t + 1] == np_[t]:
                    continue
                rows_ = self._values(0, t)
                if want is not None:
                    rows_ = rows_[want[rows_]]
                cols = np.concatenate([self._values(1, s) for s in nn_[np_[t]:np_[t + 1]]])
                Y[rows_] += _near_block(self.k, self.sign, self.tp[rows_], self.sp[cols]) @ X[cols]
        if want is not None:
            Y[~want] = 0.0
        return Y[:, 0] if vec else Y

    def dense(self, near=True):
        return self.matvec(np.eye(self.d["N"], dtype=np.complex128), near=near)
