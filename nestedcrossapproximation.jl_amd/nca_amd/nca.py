"""PetrovGalerkinNCA on the GPU: assembly through the batched bottom-up
compressor and the H^2 product as a linear map (src/nca/petrovgalerkinnca.jl)."""
import ctypes
import threading
import time

import numpy as np

from . import _abi
from ._native import hip_lib
from .runtime import call, _p, HIPKernelMatrix, default_context
from . import trees as _trees
from . import compressors as _cmp


def coefficient_ptr(f):
    """plan_from_pivots (abstractnca.jl:110-119), 0-based."""
    ptr = np.zeros(f.nn + 1, np.int64)
    np.cumsum(np.where(f.assigned, f.rank, 0), out=ptr[1:])
    return ptr


def _same_points(a, b):
    if a is b:
        return True
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def h2desc(tree, partition, testfardata, testf, trialf, testbasis, testtransfer, trialbasis, trialtransfer):
    """The nca_h2_desc arrays, 1-based as the Julia structs hold them."""
    tt, st = tree.testtree, tree.trialtree
    i64 = lambda a: np.ascontiguousarray(a, np.int64)
    c128 = lambda a: np.ascontiguousarray(a, np.complex128) if a is not None and len(a) else np.zeros(1, np.complex128)
    d = dict(
        M=tt.npoints, N=st.npoints, nn_test=tt.nnodes, nn_trial=st.nnodes,
        test_perm=i64(tt.perm + 1), test_lo=i64(tt.lo + 1), test_hi=i64(tt.hi),
        trial_perm=i64(st.perm + 1), trial_lo=i64(st.lo + 1), trial_hi=i64(st.hi),
        test_cptr=i64(coefficient_ptr(testf) + 1), trial_cptr=i64(coefficient_ptr(trialf) + 1),
        nbasis_test=len(testbasis.nodes), nbasis_trial=len(trialbasis.nodes),
        basis_test_nodes=i64(testbasis.nodes + 1), basis_test_off=i64(testbasis.offsets),
        basis_trial_nodes=i64(trialbasis.nodes + 1), basis_trial_off=i64(trialbasis.offsets),
        basis_test=c128(testbasis.blocks), basis_trial=c128(trialbasis.blocks),
        nlev_test=len(testtransfer.leveltoidcsptr) - 1, nlev_trial=len(trialtransfer.leveltoidcsptr) - 1,
        tt_levptr=i64(testtransfer.leveltoidcsptr + 1), tt_levidcs=i64(testtransfer.levelidcs + 1),
        tt_chdptr=i64(testtransfer.idxtochdsptr + 1), tt_chds=i64(testtransfer.chdsidcs + 1),
        tt_off=i64(testtransfer.offsets),
        st_levptr=i64(trialtransfer.leveltoidcsptr + 1), st_levidcs=i64(trialtransfer.levelidcs + 1),
        st_chdptr=i64(trialtransfer.idxtochdsptr + 1), st_chds=i64(trialtransfer.chdsidcs + 1),
        st_off=i64(trialtransfer.offsets),
        tt_blocks=c128(testtransfer.blocks), st_blocks=c128(trialtransfer.blocks),
        cp_ptr=i64(testfardata.nodetofarnodesptr + 1), cp_nodes=i64(testfardata.farnodes + 1),
        nearptr=i64(partition.nearptr + 1), nearnodes=i64(partition.nearnodes + 1),
    )
    for key in list(d):
        if isinstance(d[key], np.ndarray) and d[key].size == 0:
            d[key] = np.zeros(1, d[key].dtype)
    return d


def _abi_desc(d):
    desc = _abi.H2Desc()
    for name, _ in _abi.H2Desc._fields_:
        v = d[name]
        if isinstance(v, np.ndarray):
            setattr(desc, name, _p(v))
        else:
            setattr(desc, name, int(v))
    return desc


class ConcurrentWorkspaceError(RuntimeError):
    def __init__(self):
        super().__init__("ConcurrentWorkspaceError: the workspace is in use by another matrix-vector product. "
                         "Create one H2Workspace per task; they can share the core.")


class PetrovGalerkinNCACore:
    """Read-only part: the device operator plus the host-side plans and factors."""

    def __init__(self, kernel, testpoints, trialpoints, tree, testcompressor, trialcompressor, eta=1.0,
                 maxrank=50, precision=_abi.NCA_PREC_FP64, ctx=None, verbose=False, partition=None,
                 reuse_symmetric=True, resident=True):
        self.ctx = ctx or default_context()
        self.kernel = kernel
        self.tree = tree
        self.maxrank = maxrank
        # resident: the factors stay on the device and the stores are packed
        # there (nca_h2_create_resident); the host blocks are made on request
        self.resident = bool(resident)
        timings = {}
        t0 = time.perf_counter()
        if partition is None:
            partition = _trees.interactions_device(tree, self.ctx, eta)
        self.partition = partition
        self.testfardata, self.trialfardata = _trees.fardata(tree, self.partition)
        timings["partition"] = time.perf_counter() - t0
        self.farmatrix = HIPKernelMatrix(kernel, testpoints, trialpoints, self.ctx)
        t0 = time.perf_counter()
        tf, sf = self.testfardata, self.trialfardata
        one_side = (tree.testtree is tree.trialtree and np.array_equal(tf.nodetofarnodesptr, sf.nodetofarnodesptr)
                    and np.array_equal(tf.farnodes, sf.farnodes))
        td_test, td_trial = isinstance(testcompressor, _cmp.TopDown), isinstance(trialcompressor, _cmp.TopDown)
        # Phase 1 (far pivots) for the BatchedBottomUp sides
        fp_test = fp_trial = None
        if not td_test:
            fp_test = testcompressor.pivotselector.select(tree.testtree, tree.trialtree, tf, maxrank,
                                                          self.farmatrix.trial)
        if not td_trial:
            ts, ss = trialcompressor.pivotselector, getattr(testcompressor, "pivotselector", None)
            if fp_test is not None and one_side and type(ts) is type(ss) and vars(ts) == vars(ss):
                fp_trial = fp_test  # one tree, symmetric partition: the same selection
            else:
                fp_trial = ts.select(tree.trialtree, tree.testtree, sf, maxrank, self.farmatrix.test)
        timings["phase1"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        res = self.resident
        if td_test:
            self.testfactors = _cmp.topdown_side(testcompressor, self.farmatrix, tree, tf, 0, maxrank)
            self.testbasis, self.testtransfer = _cmp.stores_from_factors(self.testfactors, tree.testtree, 0,
                                                                         blocks=not res)
        else:
            self.testbasis, self.testtransfer, self.testfactors = _cmp.testbases(
                testcompressor, self.farmatrix, tree, tf, maxrank, fp_test, res)
        # One tree, one point set, one partition and the same compressor: the
        # trial blocks transpose(A[cols, rows]) hold the test blocks' numbers
        # (the kernel is symmetric), so the trial pass would repeat the test
        # pass bit for bit; its factors are reused and packed transposed.
        if td_test and td_trial:
            same_cfg = vars(testcompressor.factorization) == vars(trialcompressor.factorization)
            same_inputs = one_side
        elif not td_test and not td_trial:
            same_cfg = vars(testcompressor.compressor) == vars(trialcompressor.compressor)
            same_inputs = fp_trial is fp_test
        else:
            same_cfg = same_inputs = False
        self.symmetric = bool(reuse_symmetric and same_inputs and same_cfg and _same_points(testpoints, trialpoints))
        if self.symmetric:
            self.trialfactors = self.testfactors
            self.trialbasis, self.trialtransfer = _cmp.stores_from_factors(self.testfactors, tree.trialtree, 1,
                                                                           blocks=not res)
        elif td_trial:
            self.trialfactors = _cmp.topdown_side(trialcompressor, self.farmatrix, tree, sf, 1, maxrank)
            self.trialbasis, self.trialtransfer = _cmp.stores_from_factors(self.trialfactors, tree.trialtree, 1,
                                                                           blocks=not res)
        else:
            self.trialbasis, self.trialtransfer, self.trialfactors = _cmp.trialbases(
                trialcompressor, self.farmatrix, tree, sf, maxrank, fp_trial, res)
        self.testpivots = self.testfactors.pivot_csr()
        self.trialpivots = self.trialfactors.pivot_csr()
        self.couplingstore = _cmp.assemble_couplingstore(self.farmatrix, self.testpivots, self.trialpivots,
                                                         self.testfardata)
        timings["assembly"] = time.perf_counter() - t0
        self._desc = h2desc(tree, self.partition, self.testfardata, self.testfactors, self.trialfactors,
                            self.testbasis, self.testtransfer, self.trialbasis, self.trialtransfer)
        self._desc_blocks = not res
        t0 = time.perf_counter()
        h = ctypes.c_void_p()
        kern = kernel.abi()
        abidesc = _abi_desc(self._desc)
        if res:
            call(hip_lib().nca_h2_create_resident(
                self.ctx.handle, ctypes.byref(kern), self.farmatrix.test.handle, self.farmatrix.trial.handle,
                ctypes.byref(abidesc), self.couplingstore.handle, int(precision), self.testfactors.pool.handle,
                self.trialfactors.pool.handle, ctypes.byref(h)))
        else:
            call(hip_lib().nca_h2_create(self.ctx.handle, ctypes.byref(kern), self.farmatrix.test.handle,
                                         self.farmatrix.trial.handle, ctypes.byref(abidesc),
                                         self.couplingstore.handle, int(precision), ctypes.byref(h)))
        self.handle = h
        timings["h2_create"] = time.perf_counter() - t0
        self.timings = timings
        self.dim = (tree.testtree.npoints, tree.trialtree.npoints)
        if verbose:
            print("assembly timings:", {k: round(v, 4) for k, v in timings.items()}, flush=True)

    def host_stores(self):
        """The basis and transfer stores with their blocks on the host (packed
        from the downloaded factors when the factors are resident)."""
        if not getattr(self, "_desc_blocks", True):
            t, s = self.tree.testtree, self.tree.trialtree
            self.testbasis, self.testtransfer = _cmp.stores_from_factors(self.testfactors, t, 0)
            self.trialbasis, self.trialtransfer = _cmp.stores_from_factors(self.trialfactors, s, 1)
            self._desc = h2desc(self.tree, self.partition, self.testfardata, self.testfactors, self.trialfactors,
                                self.testbasis, self.testtransfer, self.trialbasis, self.trialtransfer)
            self._desc_blocks = True
        return self.testbasis, self.testtransfer, self.trialbasis, self.trialtransfer

    @property
    def desc(self):
        """The nca_h2_desc arrays with the stores' blocks (made on request for
        resident factors)."""
        if not getattr(self, "_desc_blocks", True):
            self.host_stores()
        return self._desc

    @desc.setter
    def desc(self, d):
        self._desc = d
        self._desc_blocks = True

    def stats(self):
        s = _abi.H2Stats()
        call(hip_lib().nca_h2_get_stats(self.handle, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in _abi.H2Stats._fields_}

    def near_form(self):
        """(form, rmin, rmax) of nca_h2_near_form: 0 FP64, 1 MIXED on the f32
        MFMA (or stored), 2 MIXED on the f16 MFMA with split operands."""
        f = ctypes.c_int32()
        a = ctypes.c_double()
        b = ctypes.c_double()
        call(hip_lib().nca_h2_near_form(self.handle, ctypes.byref(f), ctypes.byref(a), ctypes.byref(b)))
        return f.value, a.value, b.value

    def couple_form(self):
        """(form, rmin, rmax) of nca_h2_couple_form: 0 FP64 stored, 1 MIXED
        stored (c32 copy), 2 MIXED evaluated on the fly on the f16 MFMA."""
        f = ctypes.c_int32()
        a = ctypes.c_double()
        b = ctypes.c_double()
        call(hip_lib().nca_h2_couple_form(self.handle, ctypes.byref(f), ctypes.byref(a), ctypes.byref(b)))
        return f.value, a.value, b.value

    def rel32(self, mode=-1, fractions=False):
        """(near_on, couple_on) of nca_h2_rel32: whether the f16 near field and
        coupling take FP32 distances from tile-relative coordinates on the
        tiles that allow it; mode 1 allows it (default), 0 keeps FP64
        differences, -1 only queries. fractions: also the fraction of tiles
        that meet the condition (near, coupling)."""
        a = ctypes.c_int32()
        b = ctypes.c_int32()
        fa = ctypes.c_double()
        fb = ctypes.c_double()
        call(hip_lib().nca_h2_rel32(self.handle, int(mode), ctypes.byref(a), ctypes.byref(b), ctypes.byref(fa),
                                    ctypes.byref(fb)))
        if fractions:
            return bool(a.value), bool(b.value), fa.value, fb.value
        return bool(a.value), bool(b.value)

    def set_precision(self, precision):
        call(hip_lib().nca_h2_set_precision(self.handle, int(precision)))

    def matvec_device(self, op, nrhs, xptr, ldx, yptr, ldy, alpha=1.0, beta=0.0):
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        call(hip_lib().nca_h2_matvec_device(self.handle, int(op), int(nrhs), ctypes.c_void_p(xptr), int(ldx),
                                            ctypes.c_void_p(yptr), int(ldy), _p(al), _p(be)))

    def store_near(self, on=True):
        """Forward products read the near field from a store made on the device
        (the reference CPU path's assemblenears layout) instead of evaluating
        it; store_near(False) drops it (nca_h2_store_near)."""
        call(hip_lib().nca_h2_store_near(self.handle, int(bool(on))))

    def matvec_device_noalloc(self, op, nrhs, xptr, ldx, yptr, ldy, alpha=1.0, beta=0.0):
        """matvec_device on the library's static device scratch: allocates
        nothing (nca_h2_matvec_device_noalloc)."""
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        call(hip_lib().nca_h2_matvec_device_noalloc(self.handle, int(op), int(nrhs), ctypes.c_void_p(xptr),
                                                    int(ldx), ctypes.c_void_p(yptr), int(ldy), _p(al), _p(be)))

    def autotune(self, reps=3):
        """Time the launch choices of the coupling and near-field kernels in the
        current precision and keep the fastest (nca_h2_autotune)."""
        t = _abi.H2Tuning()
        call(hip_lib().nca_h2_autotune(self.handle, int(reps), ctypes.byref(t)))
        return {n: getattr(t, n) for n, _ in _abi.H2Tuning._fields_}

    def graph(self, op, nrhs, xptr, ldx, yptr, ldy, alpha=1.0, beta=0.0):
        """Capture matvec_device(op, nrhs, ...) on these device buffers as a
        hipGraph; ``launch()`` replays it on the context stream."""
        return H2Graph(self, op, nrhs, xptr, ldx, yptr, ldy, alpha, beta)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_h2_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class H2Graph:
    """One captured device sweep of a core (bound to its device buffers and to
    the precision at capture time)."""

    def __init__(self, core, op, nrhs, xptr, ldx, yptr, ldy, alpha=1.0, beta=0.0):
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        g = ctypes.c_void_p()
        call(hip_lib().nca_h2_graph_create(core.handle, int(op), int(nrhs), ctypes.c_void_p(xptr), int(ldx),
                                           ctypes.c_void_p(yptr), int(ldy), _p(al), _p(be), ctypes.byref(g)))
        self.core = core  # the graph must not outlive its operator
        self.handle = g

    def launch(self):
        call(hip_lib().nca_h2_graph_launch(self.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_h2_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class H2Workspace:
    """The H^2 matrix as a linear map over a shared core; one product at a time
    per workspace (petrovgalerkinnca.jl:253-300). Workspaces of several threads
    may share one core: the library holds the core's lock for the whole of each
    product, so their products run one after the other on the device."""

    def __init__(self, core, op=_abi.NCA_OP_N):
        self.core = core
        self.op = op
        self._busy = threading.Lock()

    @property
    def shape(self):
        M, N = self.core.dim
        return (M, N) if self.op == _abi.NCA_OP_N else (N, M)

    @property
    def T(self):
        return transpose(self)

    def _mul(self, y, x, alpha, beta):
        if not self._busy.acquire(blocking=False):
            raise ConcurrentWorkspaceError()
        try:
            M, N = self.shape
            x = np.asarray(x)
            vec = x.ndim == 1
            X = np.asfortranarray(x.reshape(x.shape[0], -1).astype(np.complex128))
            if X.shape[0] != N or y.shape[0] != M or (not vec and y.reshape(M, -1).shape[1] != X.shape[1]):
                raise ValueError(f"DimensionMismatch: A is {M} x {N}, x has {X.shape[0]} rows, y has {y.shape[0]}")
            Y = np.asfortranarray(y.reshape(M, -1).astype(np.complex128))
            al = np.array([complex(alpha).real, complex(alpha).imag])
            be = np.array([complex(beta).real, complex(beta).imag])
            call(hip_lib().nca_h2_matvec(self.core.handle, int(self.op), X.shape[1], _p(X), X.shape[0], _p(Y),
                                         Y.shape[0], _p(al), _p(be)))
            y[...] = Y.reshape(y.shape)
            return y
        finally:
            self._busy.release()

    def __matmul__(self, x):
        x = np.asarray(x)
        y = np.zeros((self.shape[0],) + x.shape[1:], np.complex128)
        return self._mul(y, x, 1.0, 0.0)


def mul(y, A, x, alpha=1.0, beta=0.0):
    """mul!(y, A, x, alpha, beta)."""
    return A._mul(y, x, alpha, beta)


def transpose(A):
    ops = {_abi.NCA_OP_N: _abi.NCA_OP_T, _abi.NCA_OP_T: _abi.NCA_OP_N}
    if A.op not in ops:
        raise ValueError("transpose of an adjoint map")
    w = H2Workspace(A.core, ops[A.op])
    w._busy = A._busy
    return w


def adjoint(A):
    ops = {_abi.NCA_OP_N: _abi.NCA_OP_C, _abi.NCA_OP_C: _abi.NCA_OP_N}
    if A.op not in ops:
        raise ValueError("adjoint of a transposed map")
    w = H2Workspace(A.core, ops[A.op])
    w._busy = A._busy
    return w


def PetrovGalerkinNCA(kernel, testpoints, trialpoints, tree, testcompressor, trialcompressor, **kw):
    """Assemble the H^2 matrix and return it as an H2Workspace (petrovgalerkinnca.jl:327-331)."""
    return H2Workspace(PetrovGalerkinNCACore(kernel, testpoints, trialpoints, tree, testcompressor,
                                             trialcompressor, **kw))


class BatchedNCA:
    """Independent problems as one block-diagonal H^2 operator.

    Problem p has its own test and trial points and cluster trees; the trees
    become one forest (trees.forest) whose partition holds each problem's near
    and far pairs only. One compressor launch per forest level then serves all
    problems, and one matvec sweep (optionally one captured hipGraph) applies
    all of them. Vectors are stacked problem by problem: rows
    [testoffsets[p], testoffsets[p + 1]) of y and [trialoffsets[p],
    trialoffsets[p + 1]) of x belong to problem p.
    """

    def __init__(self, kernel, testpoints, trialpoints, testtrees, trialtrees, testcompressor, trialcompressor,
                 eta=1.0, **kw):
        if not (len(testpoints) == len(trialpoints) == len(testtrees) == len(trialtrees)) or not testtrees:
            raise ValueError("BatchedNCA: one test/trial point set and tree per problem")
        self.nproblems = len(testtrees)
        ttest, self.testmaps, self.testoffsets = _trees.forest(testtrees)
        if all(a is b for a, b in zip(testtrees, trialtrees)):
            # one tree per problem on both sides: one forest, so a symmetric
            # batch (same points, same trees) assembles one side and reuses it
            ttrial, self.trialmaps, self.trialoffsets = ttest, self.testmaps, self.testoffsets
        else:
            ttrial, self.trialmaps, self.trialoffsets = _trees.forest(trialtrees)
        parts = [_trees.interactions(_trees.BlockTree(a, b), eta) for a, b in zip(testtrees, trialtrees)]
        self.partitions = parts
        part = _trees.forest_partition(self.testmaps, self.trialmaps, parts, ttest.nnodes)
        self.tree = _trees.BlockTree(ttest, ttrial)
        tp = np.concatenate([np.asarray(x, np.float64) for x in testpoints], axis=0)
        sp = np.concatenate([np.asarray(x, np.float64) for x in trialpoints], axis=0)
        self.core = PetrovGalerkinNCACore(kernel, tp, sp, self.tree, testcompressor, trialcompressor, eta=eta,
                                          partition=part, **kw)
        self.workspace = H2Workspace(self.core)

    def split(self, y, side="test"):
        off = self.testoffsets if side == "test" else self.trialoffsets
        return [y[off[p]:off[p + 1]] for p in range(self.nproblems)]

    def stack(self, xs):
        return np.concatenate([np.asarray(x) for x in xs], axis=0)

    def __matmul__(self, x):
        return self.workspace @ x
