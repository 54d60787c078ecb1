"""Parity of the HIP path (through the C-ABI) with the oracle on the same inputs.

Bar: entries, ACA ranks and pivots, pivots of the whole bottom-up pass, skeleton
factors and coupling blocks bit for bit (both sides evaluate the same
operations in the same order); the H^2 product within 1e-12 relative (the
matvec sums in a different order)."""
import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from nca_amd import nca as NC
from oracle import oracle as O
from helpers import grid_problem, sphere_problem, oracle_assembly, oracle_operator, relerr, bitwise_equal

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    return nca_amd.default_context()


def farmatrix(ctx, k, tp, sp, sign=1):
    return nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(k, sign), tp, sp, ctx)


def random_level(rng, nt, ns, nb, maxrank, mmax=70):
    rows, cols = [], []
    for _ in range(nb):
        m = int(rng.integers(0, mmax))
        n = int(rng.integers(0, maxrank + 1))
        rows.append(rng.integers(1, nt + 1, m))
        cols.append(rng.integers(1, ns + 1, n))
    return rows, cols


def csr0(lists):
    ptr = np.zeros(len(lists) + 1, np.int64)
    np.cumsum([len(v) for v in lists], out=ptr[1:])
    idx = np.concatenate([np.asarray(v, np.int64) - 1 for v in lists]) if ptr[-1] else np.zeros(0, np.int64)
    return ptr, idx


@pytest.mark.parametrize("transposed", [False, True])
@pytest.mark.parametrize("sign", [1, -1])
def test_level_entries_bitwise(ctx, transposed, sign):
    say(f"level entries transposed={transposed} sign={sign}")
    rng = np.random.default_rng(7)
    tp = rng.standard_normal((500, 3))
    sp = rng.standard_normal((300, 3)) * 0.5 + 2.0
    sp[:5] = tp[:5]  # coincident points: zero entries
    k = 7.3
    rows, cols = random_level(rng, 300 if transposed else 500, 500 if transposed else 300, 40, 12)
    ref, ro = O.assemble_level(k, sign, tp, sp, rows, cols, 12, transposed)
    A = farmatrix(ctx, k, tp, sp, sign)
    rp, ri = csr0(rows)
    cp, ci = csr0(cols)
    st = C.assemble_level(A, rp, ri, cp, ci, 12, transposed)
    got = st.download()
    assert got.shape == ref.shape
    assert bitwise_equal(got, ref)


def test_level_assemble_rejects_wide_block(ctx):
    A = farmatrix(ctx, 1.0, np.zeros((3, 3)), np.ones((3, 3)))
    with pytest.raises(nca_amd.ArgumentError):
        C.assemble_level(A, np.array([0, 1]), np.array([0]), np.array([0, 3]), np.array([0, 1, 2]), 2)


def _aca_compare(ctx, buf, ro, ncols, maxrank, tol):
    A = farmatrix(ctx, 1.0, np.zeros((1, 3)), np.zeros((1, 3)))
    st = C.stack_from_host(A, buf, ro, ncols, maxrank)
    ranks, rp, cp = C.aca_fullpiv(st, tol, sampled=True)
    for b in range(len(ncols)):
        blk = buf[ro[b]:ro[b + 1], :ncols[b]]
        r, orp, ocp = O.aca_fullpiv(blk, maxrank, tol, sampled=True)
        assert ranks[b] == r, (b, ranks[b], r)
        assert np.array_equal(rp[:r, b], orp) and np.array_equal(cp[:r, b], ocp)
        assert np.all(rp[r:, b] == 0)
    return st, ranks, rp, cp


@pytest.mark.parametrize("tol", [1e-2, 1e-4, 1e-6])
def test_aca_bitwise_pivots_on_kernel_level(ctx, tol):
    say(f"aca kernel level tol={tol}")
    rng = np.random.default_rng(9)
    tp = rng.standard_normal((2000, 3))
    sp = rng.standard_normal((2000, 3)) + np.array([6.0, 0.0, 0.0])
    rows, cols = random_level(rng, 2000, 2000, 120, 50, mmax=90)
    # a few tall blocks take the global-memory path
    rows += [rng.integers(1, 2001, 400), rng.integers(1, 2001, 300)]
    cols += [rng.integers(1, 2001, 50), rng.integers(1, 2001, 33)]
    buf, ro = O.assemble_level(1.5, 1, tp, sp, rows, cols, 50)
    _aca_compare(ctx, buf, ro, np.array([len(c) for c in cols]), 50, tol)


@pytest.mark.parametrize("tol", [1e-2, 1e-4, 1e-8])
def test_iaca_bitwise_pivots_on_kernel_level(ctx, tol):
    """nca_iaca (geometric column order) against the oracle's orc_iaca, bit for
    bit, on kernel blocks of the LDS and of the global-memory path."""
    say(f"iaca kernel level tol={tol}")
    rng = np.random.default_rng(19)
    tp = rng.standard_normal((2000, 3))
    sp = rng.standard_normal((2000, 3)) + np.array([6.0, 0.0, 0.0])
    rows, cols = random_level(rng, 2000, 2000, 120, 50, mmax=90)
    rows += [rng.integers(1, 2001, 400), rng.integers(1, 2001, 300)]
    cols += [rng.integers(1, 2001, 50), rng.integers(1, 2001, 33)]
    buf, ro = O.assemble_level(1.5, 1, tp, sp, rows, cols, 50)
    ncols = np.array([len(c) for c in cols])
    A = farmatrix(ctx, 1.5, tp, sp)
    st = C.stack_from_host(A, buf, ro, ncols, 50)
    ranks, rp, cp = C.iaca(st, tol)
    for b in range(len(ncols)):
        r, orp, ocp = O.iaca(buf[ro[b]:ro[b + 1], :ncols[b]], 50, tol)
        assert ranks[b] == r, (b, ranks[b], r)
        assert np.array_equal(rp[:r, b], orp) and np.array_equal(cp[:r, b], ocp)
        assert np.all(rp[r:, b] == 0)
    # the zero blocks of test/gpu/test_compresslevel.jl:191-211
    z = np.zeros((13, 30), np.complex128, order="F")
    st0 = C.stack_from_host(A, z, [0, 0, 4, 13], [3, 3, 3], 30)
    assert list(C.iaca(st0, tol)[0]) == [0, 0, 0]


def test_aca_zero_blocks(ctx):
    """test/gpu/test_compresslevel.jl:191-211: rowoffsets [0, 0, 4, 13] of zeros."""
    buf = np.zeros((13, 30), np.complex128, order="F")
    A = farmatrix(ctx, 1.0, np.zeros((1, 3)), np.zeros((1, 3)))
    st = C.stack_from_host(A, buf, [0, 0, 4, 13], [3, 3, 3], 30)
    ranks, rp, cp = C.aca_fullpiv(st, 1e-3)
    assert list(ranks) == [0, 0, 0] and rp.shape == (30, 3) and not rp.any()


def test_aca_decaying_blocks_invariant(ctx):
    """test/gpu/test_compresslevel.jl:127-189 (complex, zero block included)."""
    rng = np.random.default_rng(41)
    cg = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    mr, specs = 30, [(40, 25), (90, 30), (65, 18)]
    ro = np.concatenate([[0], np.cumsum([m for m, _ in specs])])
    buf = np.zeros((ro[-1], mr), np.complex128, order="F")
    for i, (m, n) in enumerate(specs):
        if i > 0:
            U = np.linalg.qr(cg(m, n))[0][:, :n]
            V = np.linalg.qr(cg(n, n))[0]
            buf[ro[i]:ro[i + 1], :n] = U @ np.diag(10.0 ** np.linspace(0, -7, n)) @ V.conj().T
    st, ranks, rp, cp = _aca_compare(ctx, buf, ro, [n for _, n in specs], mr, 1e-3)
    xoff, X, indep, status = C.skeleton_solve(st, ranks, rp, cp)
    assert status == 0 and np.all(indep == -1)
    for i, (m, n) in enumerate(specs):
        r = ranks[i]
        blk = buf[ro[i]:ro[i + 1], :n]
        if r == 0:
            assert not blk.any()
            continue
        Xi = X[xoff[i]:xoff[i + 1]].reshape((m, r), order="F")
        recon = Xi @ blk[rp[:r, i] - 1, :]
        assert r < min(m, n)
        assert relerr(recon, blk) * np.sqrt(n / (n - r)) <= 1.01e-3


def test_skeleton_solve_closed_form_and_bitwise(ctx):
    """Ranks 0, 8, 35, 80 on blocks of 40, 90, 100, 160 rows
    (test/gpu/test_compresslevel.jl:65-125), against L / L[rows, :] and the oracle."""
    say("skeleton solve")
    rng = np.random.default_rng(17)
    mr, lengths, ranks = 80, [40, 90, 100, 160], [0, 8, 35, 80]
    ro = np.concatenate([[0], np.cumsum(lengths)])
    buf = np.zeros((ro[-1], mr), np.complex128, order="F")
    rp = np.zeros((mr, 4), np.int64, order="F")
    cp = np.zeros((mr, 4), np.int64, order="F")
    expected = []
    for i, (m, r) in enumerate(zip(lengths, ranks)):
        if r == 0:
            expected.append(None)
            continue
        rows = rng.permutation(m)[:r]
        cols = rng.permutation(mr)[:r]
        L = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
        L[rows, :] += r * np.eye(r)
        buf[ro[i] + np.arange(m)[:, None], cols[None, :]] = L
        rp[:r, i] = rows + 1
        cp[:r, i] = cols + 1
        expected.append((L, rows))
    A = farmatrix(ctx, 1.0, np.zeros((1, 3)), np.zeros((1, 3)))
    st = C.stack_from_host(A, buf, ro, [mr] * 4, mr)
    xoff, X, indep, status = C.skeleton_solve(st, np.array(ranks, np.int32), rp, cp)
    assert status == 0
    for i, e in enumerate(expected):
        if e is None:
            assert xoff[i + 1] == xoff[i]
            continue
        L, rows = e
        m, r = L.shape
        Xi = X[xoff[i]:xoff[i + 1]].reshape((m, r), order="F")
        ref = np.linalg.solve(L[rows, :].T, L.T).T
        np.testing.assert_allclose(Xi, ref, rtol=1e-11, atol=1e-11)
        Xo, sto = O.skeleton_solve(L, rows + 1)
        assert sto == -1
        assert bitwise_equal(Xi, Xo)


def test_skeleton_solve_singular(ctx):
    """A dependent pivot set raises SingularCrossMatrixError with the rank."""
    buf = np.zeros((2, 2), np.complex128, order="F")
    buf[:, :] = [[1, 2], [1, 2]]
    A = farmatrix(ctx, 1.0, np.zeros((1, 3)), np.zeros((1, 3)))
    st = C.stack_from_host(A, buf, [0, 2], [2], 2)
    rp = np.array([[1], [2]], np.int64)
    cp = np.array([[1], [2]], np.int64)
    xoff, X, indep, status = C.skeleton_solve(st, np.array([2], np.int32), rp, cp)
    assert status == nca_amd._abi.NCA_ERR_SINGULAR and indep[0] == 1


# ------------------------------------------------------------------ pipeline
def gpu_assembly(ctx, p):
    A = farmatrix(ctx, p.k, p.testpts, p.trialpts, p.sign)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(p.tol))
    ft = C.batched_side(cmp, A, p.tree, p.tf, 0, p.maxrank, p.fpt)
    fs = C.batched_side(cmp, A, p.tree, p.sf, 1, p.maxrank, p.fps)
    return A, ft, fs


def compare_sides(fg, fo, nn):
    for t in range(nn):
        assert fg.assigned[t] == fo.assigned[t]
        if not fg.assigned[t]:
            continue
        assert fg.rank[t] == fo.rank[t], t
        assert np.array_equal(fg.pivots[t], fo.pivots[t]), t
        assert bitwise_equal(fg.X[t], fo.X[t]), t


@pytest.fixture(scope="module")
def grid_case(ctx):
    p = grid_problem(1e-4)
    A, ft, fs = gpu_assembly(ctx, p)
    oft, ofs = oracle_assembly(p)
    return p, A, ft, fs, oft, ofs


@pytest.fixture(scope="module")
def sphere_case(ctx):
    p = sphere_problem(8192, k=1.0, tol=1e-6)
    A, ft, fs = gpu_assembly(ctx, p)
    oft, ofs = oracle_assembly(p)
    return p, A, ft, fs, oft, ofs


@pytest.mark.parametrize("case", ["grid_case", "sphere_case"])
def test_bottom_up_pivots_bitwise(case, request):
    say(f"bottom-up pivots {case}")
    p, A, ft, fs, oft, ofs = request.getfixturevalue(case)
    compare_sides(ft, oft, p.tree.testtree.nnodes)
    compare_sides(fs, ofs, p.tree.trialtree.nnodes)


@pytest.mark.parametrize("case", ["grid_case", "sphere_case"])
def test_coupling_bitwise(case, request):
    say(f"coupling {case}")
    p, A, ft, fs, oft, ofs = request.getfixturevalue(case)
    cs = C.assemble_couplingstore(A, ft.pivot_csr(), fs.pivot_csr(), p.tf)
    off, blocks = cs.download()
    tp, sp = oft.pivot_csr(), ofs.pivot_csr()
    ooff, oblocks = O.coupling(p.k, p.sign, p.testpts, p.trialpts, p.tf.nodetofarnodesptr, p.tf.farnodes,
                               tp[0], tp[1], sp[0], sp[1])
    assert np.array_equal(off, ooff)
    assert bitwise_equal(blocks, oblocks)


def build_operator(ctx, p, A, ft, fs):
    tb, tt = C.stores_from_factors(ft, p.tree.testtree, 0)
    sb, st = C.stores_from_factors(fs, p.tree.trialtree, 1)
    cs = C.assemble_couplingstore(A, ft.pivot_csr(), fs.pivot_csr(), p.tf)
    core = object.__new__(NC.PetrovGalerkinNCACore)
    core.ctx, core.dim, core.couplingstore, core.farmatrix = ctx, (p.testpts.shape[0], p.trialpts.shape[0]), cs, A
    core.desc = NC.h2desc(p.tree, p.partition, p.tf, ft, fs, tb, tt, sb, st)
    import ctypes
    h = ctypes.c_void_p()
    kern = A.kernel.abi()
    d = NC._abi_desc(core.desc)
    from nca_amd.runtime import call
    call(nca_amd.hip_lib().nca_h2_create(ctx.handle, ctypes.byref(kern), A.test.handle, A.trial.handle,
                                         ctypes.byref(d), cs.handle, 0, ctypes.byref(h)))
    core.handle = h
    return NC.H2Workspace(core)


@pytest.mark.parametrize("case", ["grid_case", "sphere_case"])
def test_matvec_parity(ctx, case, request):
    say(f"matvec {case}")
    p, A, ft, fs, oft, ofs = request.getfixturevalue(case)
    W = build_operator(ctx, p, A, ft, fs)
    H, _ = oracle_operator(p, oft, ofs)
    rng = np.random.default_rng(2)
    N, M = p.trialpts.shape[0], p.testpts.shape[0]
    # 5, 8 and 24 (= 16 + 8) run a last chunk padded to 16 columns on the MFMA
    # kernels; 1, 3 and 19 (= 16 + 3) a last chunk on the VALU kernels
    for R in (1, 3, 5, 8, 16, 19, 24):
        x = geometry.complex_gaussian(rng, (N, R))
        u = geometry.complex_gaussian(rng, (M, R))
        assert relerr(W @ x, H.matvec(x)) <= 1e-12
        assert relerr(NC.transpose(W) @ u, H.matvec(u, op=1)) <= 1e-12
        assert relerr(NC.adjoint(W) @ u, H.matvec(u, op=2)) <= 1e-12
    al, be = 0.3 + 0.2j, -1.1
    for shape in ((N,), (N, 16)):
        x = geometry.complex_gaussian(rng, shape)
        y0 = geometry.complex_gaussian(rng, (M,) + shape[1:])
        y = y0.copy()
        NC.mul(y, W, x, al, be)
        ref = H.matvec(x, alpha=al, beta=be, y=y0)
        assert relerr(y, ref.reshape(y.shape)) <= 1e-12
    x = geometry.complex_gaussian(rng, (N,))
    # repeated products are bitwise stable
    assert np.array_equal(W @ x, W @ x)
    with pytest.raises(ValueError):
        W @ x[:-1]


def test_matvec_vs_direct_summation_sphere(ctx, sphere_case):
    p, A, ft, fs, oft, ofs = sphere_case
    W = build_operator(ctx, p, A, ft, fs)
    rng = np.random.default_rng(5)
    x = geometry.complex_gaussian(rng, (p.trialpts.shape[0],))
    rows = rng.choice(p.testpts.shape[0], 200, replace=False)
    ref = O.dense_rows(p.k, p.sign, p.testpts, p.trialpts, rows, x)[:, 0]
    assert relerr((W @ x)[rows], ref) < 5 * p.tol


def test_petrovgalerkinnca_end_to_end(ctx):
    """The public constructor on the 32 x 32 grid problem (x = ones)."""
    say("PetrovGalerkinNCA end to end")
    p = grid_problem(1e-4)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    W = NC.PetrovGalerkinNCA(nca_amd.HelmholtzKernel(p.k), p.testpts, p.trialpts, p.tree, cmp, cmp, ctx=ctx)
    x = np.ones(1024, np.complex128)
    y = W @ x
    ref = O.dense_rows(p.k, 1, p.testpts, p.trialpts, np.arange(1024), x)[:, 0]
    assert relerr(y, ref) < 5e-4


@pytest.mark.parametrize("case", ["grid_case", "sphere_case"])
def test_matvec_mixed_precision(ctx, case, request):
    """MIXED (complex FP32 operators, source coefficients and near field; FP64
    destination coefficients, distance and phase) within 1e-5 relative of the
    FP64 oracle for op N, T and C at 16, 3, 8, 19 and 40 right-hand sides (the
    last two in several column chunks); FP64 again after switching back."""
    say(f"mixed precision {case}")
    p, A, ft, fs, oft, ofs = request.getfixturevalue(case)
    W = build_operator(ctx, p, A, ft, fs)
    H, _ = oracle_operator(p, oft, ofs)
    rng = np.random.default_rng(12)
    N, M = p.trialpts.shape[0], p.testpts.shape[0]
    W.core_set = W.core.set_precision if hasattr(W.core, "set_precision") else None
    NC.PetrovGalerkinNCACore.set_precision(W.core, nca_amd.NCA_PREC_MIXED)
    for R in (16, 3, 8, 19, 40):
        x = geometry.complex_gaussian(rng, (N, R))
        u = geometry.complex_gaussian(rng, (M, R))
        assert relerr(W @ x, H.matvec(x)) <= 1e-5
        assert relerr(NC.transpose(W) @ u, H.matvec(u, op=1)) <= 1e-5
        assert relerr(NC.adjoint(W) @ u, H.matvec(u, op=2)) <= 1e-5
    NC.PetrovGalerkinNCACore.set_precision(W.core, nca_amd.NCA_PREC_FP64)
    x = geometry.complex_gaussian(rng, (N, 16))
    assert relerr(W @ x, H.matvec(x)) <= 1e-12


def test_matvec_negative_phase_sign(ctx):
    """exp(-i k r)/(4 pi r): the device products (N, T, C; FP64 and MIXED,
    16 and 3 right-hand sides) against the oracle's product on the operator's
    own factors."""
    say("matvec sign -1")
    pts = geometry.fibonacci_sphere(3000)
    tree = nca_amd.TwoNTree(pts, minvalues=32)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    core = nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(1.5, -1), pts, pts,
                                         nca_amd.BlockTree(tree, tree), cmp, cmp, ctx=ctx)
    off, blocks = core.couplingstore.download()
    H = O.OracleH2(core.desc, 1.5, -1, pts, pts, off, blocks)
    W = NC.H2Workspace(core)
    rng = np.random.default_rng(31)
    for prec, bound in ((nca_amd.NCA_PREC_FP64, 1e-12), (nca_amd.NCA_PREC_MIXED, 1e-5)):
        core.set_precision(prec)
        for R in (16, 3):
            x = geometry.complex_gaussian(rng, (pts.shape[0], R))
            assert relerr(W @ x, H.matvec(x)) <= bound
            assert relerr(NC.transpose(W) @ x, H.matvec(x, op=1)) <= bound
            assert relerr(NC.adjoint(W) @ x, H.matvec(x, op=2)) <= bound


def near_turns(k, pts, tree, part):
    """The largest phase k r / (2 pi) a near pair can reach: the farthest
    corners of the two point boxes (the bound the operator uses to form the
    MIXED phase in FP32 up to one revolution)."""
    box = [(pts[tree.values(t)].min(0), pts[tree.values(t)].max(0)) if len(tree.values(t)) else None
           for t in range(tree.nnodes)]
    d2 = 0.0
    for t in range(tree.nnodes):
        for s in part.nearnodes[part.nearptr[t]:part.nearptr[t + 1]]:
            if box[t] is None or box[s] is None:
                continue
            g = np.maximum(box[t][1] - box[s][0], box[s][1] - box[t][0])
            d2 = max(d2, float(g @ g))
    return abs(k) / (2 * np.pi) * np.sqrt(d2)


@pytest.mark.parametrize("k, short", [(1.5, True), (40.0, False)])
def test_matvec_mixed_phase_forms(ctx, k, short):
    """MIXED forms the near-field distance and phase in FP32 when every near
    pair stays within one revolution, else in FP64: both forms within 1e-5 of
    the FP64 oracle (N, T, C at 16, 3 and 1 right-hand sides)."""
    pts = geometry.fibonacci_sphere(3000)
    tree = nca_amd.TwoNTree(pts, minvalues=32)
    bt = nca_amd.BlockTree(tree, tree)
    assert (near_turns(k, pts, tree, nca_amd.interactions(bt, 1.0)) <= 1.0) == short
    say(f"mixed phase forms k={k} (FP32 phase: {short})")
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    core = nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), pts, pts, bt, cmp, cmp, ctx=ctx)
    off, blocks = core.couplingstore.download()
    H = O.OracleH2(core.desc, k, 1, pts, pts, off, blocks)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    rng = np.random.default_rng(32)
    for R in (16, 3, 1):
        x = geometry.complex_gaussian(rng, (pts.shape[0], R))
        assert relerr(W @ x, H.matvec(x)) <= 1e-5
        assert relerr(NC.transpose(W) @ x, H.matvec(x, op=1)) <= 1e-5
        assert relerr(NC.adjoint(W) @ x, H.matvec(x, op=2)) <= 1e-5


def test_config_sphere65k_fp64(ctx):
    """The FP64 one-GPU configuration at its full size: N = 65,536 Fibonacci
    sphere, k = 1, octree minvalues 64, tol 1e-6, one seeded complex Gaussian
    right-hand side. Pivots, ranks, factors and coupling bit for bit against the
    oracle, the product within 1e-12 of the oracle's, and within 5 tol of dense
    direct summation on 200 sampled rows."""
    say("config sphere 65536 fp64")
    p = sphere_problem(65536, k=1.0, tol=1e-6)
    A, ft, fs = gpu_assembly(ctx, p)
    oft, ofs = oracle_assembly(p)
    compare_sides(ft, oft, p.tree.testtree.nnodes)
    compare_sides(fs, ofs, p.tree.trialtree.nnodes)
    cs = C.assemble_couplingstore(A, ft.pivot_csr(), fs.pivot_csr(), p.tf)
    off, blocks = cs.download()
    H, (ooff, oblocks) = oracle_operator(p, oft, ofs)
    assert np.array_equal(off, ooff) and bitwise_equal(blocks, oblocks)
    W = build_operator(ctx, p, A, ft, fs)
    x = geometry.complex_gaussian(np.random.default_rng(2), (p.trialpts.shape[0],))
    y = W @ x
    assert relerr(y, H.matvec(x)[:, 0]) <= 1e-12
    rows = np.random.default_rng(3).choice(p.testpts.shape[0], 200, replace=False)
    ref = O.dense_rows(p.k, p.sign, p.testpts, p.trialpts, rows, x)[:, 0]
    assert relerr(y[rows], ref) < 5 * p.tol
