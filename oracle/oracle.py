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
        L.orc_skeleton_solve.restype = I64
        L.orc_skeleton_solve.argtypes = [P, I64, I64, P, P]
        L.orc_bases.restype = I64
        L.orc_bases.argtypes = [I32, D, I32, P, P, I64, P, P, P, P, P, P, P, P, I64, D, I32, P, P, P, P, I64]
        L.orc_coupling.argtypes = [D, I32, P, P, I64, P, P, P, P, P, P, P, P]
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


def skeleton_solve(L, rowpiv):
    """X = L / L[rowpiv, :]; returns (X, independent) with independent = -1 when regular."""
    Lm = np.asfortranarray(np.array(L, dtype=np.complex128))
    m, r = Lm.shape
    X = np.zeros((m, r), np.complex128, order="F")
    rp = _i64(rowpiv)
    st = lib().orc_skeleton_solve(_p(Lm), m, r, _p(rp), _p(X))
    return X, int(st)


def bases(side, k, sign, testpts, trialpts, tree, farpivptr, farpiv, maxrank, tol, sampled=True):
    """testbases/trialbases(::BatchedBottomUp). tree: ClusterTree of this side;
    farpivptr/farpiv: 0-based CSR of 0-based point ids (Phase 1).
    Returns (pivptr 0-based, pivots 0-based, xoffsets, X, status)."""
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
                             _p(fptr), _p(fpiv), int(maxrank), float(tol), int(bool(sampled)), _p(pivptr),
                             _p(pivots), _p(xoff), _p(X), cap)
        if st != 1:
            break
        cap = int(xoff[-1])
    pivptr0 = pivptr - 1
    return pivptr0, pivots[:pivptr0[-1]] - 1, xoff, X[:xoff[-1]].copy(), int(st)


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
