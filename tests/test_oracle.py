"""Pins the oracle (oracle/nca_oracle.c) against the reference's known answers,
closed forms and brute force, before it is trusted as the parity checker."""
import mpmath
import numpy as np
import pytest

from oracle import oracle as O
from helpers import (grid_problem, sphere_problem, oracle_assembly, oracle_operator, oracle_operator_rows,
                     oracle_topdown, relerr, Problem)
from nca_amd import geometry


def G(k, x, y, sign=1):
    r = np.linalg.norm(np.asarray(x) - np.asarray(y))
    return 0j if r == 0 else np.exp(1j * sign * k * r) / (4 * np.pi * r)


# ---------------------------------------------------------------- sin / cos
def test_sincos_against_mpmath():
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(-120, 120, 3000), rng.uniform(-1, 1, 500),
                         np.arange(-80, 81) * np.pi / 2 + rng.uniform(-1e-9, 1e-9, 161), [0.0, 1e-300, 5e-8]])
    worst = 0.0
    for x in xs:
        s, c = O.sincos(x)
        ms, mc = mpmath.sin(mpmath.mpf(x)), mpmath.cos(mpmath.mpf(x))
        # absolute error in units of 2^-53 (the values lie in [-1, 1])
        err = max(abs(float(mpmath.mpf(s) - ms)), abs(float(mpmath.mpf(c) - mc))) / 2.0**-53
        worst = max(worst, err)
    assert worst <= 2.0, worst


def test_sincos_quadrants_exact_at_zero():
    assert O.sincos(0.0) == (0.0, 1.0)
    s, c = O.sincos(np.pi / 2)
    assert abs(s - 1.0) < 2e-16 and abs(c) < 2e-16


# ---------------------------------------------------------------- entries
def G_exact(k, x, y, sign):
    mpmath.mp.prec = 200
    r = mpmath.sqrt(sum((mpmath.mpf(float(a)) - mpmath.mpf(float(b))) ** 2 for a, b in zip(x, y)))
    v = mpmath.expjpi(sign * mpmath.mpf(k) * r / mpmath.pi) / (4 * mpmath.pi * r)
    return complex(v)


@pytest.mark.parametrize("sign", [1, -1])
def test_entry_closed_form(sign):
    """cis(sign k r)/(4 pi r) against a 200-bit evaluation at the same double
    inputs; the bound is the rounding of r, k r (phase error ~ k r ulp) and the
    final division."""
    rng = np.random.default_rng(3)
    for _ in range(400):
        x, y = rng.standard_normal(3), rng.standard_normal(3)
        k = rng.uniform(0.1, 60)
        g = O.entry(k, sign, x, y)
        ref = G_exact(k, x, y, sign)
        kr = k * np.linalg.norm(x - y)
        assert abs(g - ref) <= abs(ref) * 2.0**-52 * (3.0 + 2.0 * kr), (g, ref, kr)
        assert abs(g - G(k, x, y, sign)) <= abs(ref) * 2.0**-52 * (6.0 + 4.0 * kr)
    assert O.entry(2.0, 1, [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]) == 0j


def test_assemble_level_transposed_layout():
    """blocks[i] == transpose(dense[fartest[i], owntrial[i]]) (test/compressors/test_trialassembly.jl:23-30)."""
    rng = np.random.default_rng(17)
    for nt, ns in ((160, 160), (96, 160), (160, 96)):
        tp = rng.random((nt, 3))
        sp = rng.random((ns, 3)) + np.array([10.0, 0.0, 0.0])
        owntrial = [[ns - 4, ns - 3, ns - 2, ns - 1, ns], [2, 4, 6, 8]]
        fartest = [[1, 3, 5], [nt - 1, nt]]
        buf, ro = O.assemble_level(0.7, 1, tp, sp, owntrial, fartest, 10, transposed=True)
        for i in range(2):
            blk = buf[ro[i]:ro[i + 1], :]
            ref = np.array([[G(0.7, tp[c - 1], sp[r - 1]) for c in fartest[i]] for r in owntrial[i]])
            np.testing.assert_allclose(blk[:, :len(fartest[i])], ref, rtol=1e-14)
            assert np.all(blk[:, len(fartest[i]):] == 0)
    with pytest.raises(ValueError):
        O.assemble_level(0.7, 1, tp, sp, [[1]], [list(range(1, 12))], 10)


# ---------------------------------------------------------------- ACA
def test_aca_sampled_stop_known_answer():
    """test/compressors/test_batchedbottomup.jl:43-52."""
    A = np.zeros((5, 3))
    A[0, 0], A[1, 1], A[2, 2] = 1, 0.009, 0.001
    assert O.aca_fullpiv(A, 3, 1e-2, sampled=False)[0] == 1
    assert O.aca_fullpiv(A, 3, 1e-2, sampled=True)[0] == 2


def test_aca_exact_rank_blocks():
    """Exact-rank blocks give their rank and CUR relerr < 1e-9
    (examples/batchedbottomup_compression_validation.jl:40-73)."""
    rng = np.random.default_rng(1)
    cg = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    for m, n, r in ((120, 25, 8), (64, 30, 17), (200, 12, 5), (90, 20, 20)):
        A = cg(m, r) @ cg(r, n)
        rank, rp, cp = O.aca_fullpiv(A, 30, 1e-12, sampled=True)
        assert rank == r
        rpz, cpz = rp - 1, cp - 1
        recon = A[:, cpz] @ np.linalg.solve(A[np.ix_(rpz, cpz)], A[rpz, :])
        assert relerr(recon, A) < 1e-9
        assert len(set(rp)) == rank and len(set(cp)) == rank


def test_aca_zero_block():
    assert O.aca_fullpiv(np.zeros((4, 3)), 30, 1e-3)[0] == 0
    assert O.aca_fullpiv(np.zeros((0, 3)), 30, 1e-3)[0] == 0


def test_aca_residual_invariant():
    """||X A[rp,:] - A|| / ||A|| * sqrt(n / (n - r)) <= 1.01 tol and r < min(m, n)
    on decaying blocks (test/gpu/test_compresslevel.jl:127-189)."""
    rng = np.random.default_rng(41)
    cg = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    tol = 1e-3
    for m, n in ((90, 30), (65, 18)):
        U = np.linalg.qr(cg(m, n))[0][:, :n]
        V = np.linalg.qr(cg(n, n))[0]
        A = U @ np.diag(10.0 ** np.linspace(0, -7, n)) @ V.conj().T
        rank, rp, cp = O.aca_fullpiv(A, 30, tol, sampled=True)
        assert 0 < rank < min(m, n)
        X, st = O.skeleton_solve(A[:, cp - 1], rp)
        assert st == -1
        recon = X @ A[rp - 1, :]
        assert relerr(recon, A) * np.sqrt(n / (n - rank)) <= 1.01 * tol


def brute_force_aca(A, maxrank, tol, sampled):
    """Greedy full pivoting written with numpy (argmax over |R|^2 in column-major
    order, exact residual Frobenius norms)."""
    R = np.array(A, dtype=np.complex128)
    m, n = R.shape
    nA2 = np.sum(np.abs(R) ** 2)
    rows, cols = [], []
    while len(rows) < min(maxrank, m, n):
        r = len(rows)
        n2 = np.sum(np.abs(R) ** 2)
        if n2 * (n / (n - r) if sampled else 1.0) <= tol * tol * nA2:
            break
        e = int(np.argmax((np.abs(R) ** 2).ravel(order="F")))
        i, j = e % m, e // m
        if R[i, j] == 0:
            break
        R = R - np.outer(R[:, j], R[i, :] / R[i, j])
        R[i, :] = 0
        R[:, j] = 0
        rows.append(i + 1)
        cols.append(j + 1)
    return len(rows), rows, cols


def test_aca_matches_brute_force_on_kernel_blocks():
    rng = np.random.default_rng(5)
    for trial in range(20):
        m, n = rng.integers(2, 40), rng.integers(2, 30)
        x = rng.random((m, 3))
        y = rng.random((n, 3)) + np.array([2.0 + rng.random(), 0, 0])
        A = np.array([[G(3.0, a, b) for b in y] for a in x])
        for tol in (1e-2, 1e-4, 1e-8):
            got = O.aca_fullpiv(A, 50, tol, sampled=True)
            ref = brute_force_aca(A, 50, tol, True)
            assert got[0] == ref[0] and list(got[1]) == ref[1] and list(got[2]) == ref[2]


# ---------------------------------------------------------------- IACA
def brute_force_iaca(A, maxrank, tol):
    """The IACA with geometric column order written with numpy: step r takes
    column r, the row pivot is the first maximum of its residual column, and
    the candidate stops the factorization when r ||R(:, r)||^2 <= tol^2
    sum_{l<r} ||A(:, l)||^2."""
    R = np.array(A, dtype=np.complex128)
    m, n = R.shape
    rows = []
    acc = 0.0
    while len(rows) < min(maxrank, m, n):
        r = len(rows)
        col = R[:, r]
        c2 = np.sum(np.abs(col) ** 2)
        if r > 0 and c2 * r <= tol * tol * acc:
            break
        i = int(np.argmax(np.abs(col) ** 2))
        if col[i] == 0:
            break
        R = R - np.outer(col, R[i, :] / col[i])
        R[i, :] = 0
        R[:, r] = 0
        rows.append(i + 1)
        acc += np.sum(np.abs(A[:, r]) ** 2)
    return len(rows), rows, list(range(1, len(rows) + 1))


def test_iaca_contracts():
    """Rank at most min(maxrank, m, n) (_ranklimit, test/pivoting/test_pivots.jl:77-82),
    unique row pivots, column pivots 1..rank in order, zero blocks rank 0."""
    rng = np.random.default_rng(3)
    assert O.iaca(np.zeros((4, 3)), 30, 1e-3)[0] == 0
    assert O.iaca(np.zeros((0, 3)), 30, 1e-3)[0] == 0
    for m, n, mr in ((30, 20, 50), (7, 40, 50), (40, 40, 5), (12, 9, 12)):
        A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        r, rp, cp = O.iaca(A, mr, 1e-14)
        assert r == min(m, n, mr)
        assert len(set(rp.tolist())) == r and list(cp) == list(range(1, r + 1))


def test_iaca_stop_rule_scaling():
    """A known answer fixed by the stop rule: column 0 = e_0 (norm 1), column
    1 = delta e_1 (its residual after the first cross). With tol 1e-2 the
    second candidate stops the factorization when delta^2 <= tol^2 * 1."""
    for delta, want in ((0.5e-2, 1), (2e-2, 2)):
        A = np.zeros((3, 2), np.complex128)
        A[0, 0] = 1.0
        A[1, 1] = delta
        assert O.iaca(A, 50, 1e-2)[0] == want
    # the mean over accepted columns: two unit columns, a third of 0.015
    A = np.zeros((4, 3), np.complex128)
    A[0, 0], A[1, 1], A[2, 2] = 1.0, 1.0, 0.015
    assert O.iaca(A, 50, 1e-2)[0] == 3        # 2 * 0.015^2 > 1e-4 * 2
    A[2, 2] = 0.009
    assert O.iaca(A, 50, 1e-2)[0] == 2        # 2 * 0.009^2 <= 1e-4 * 2


def test_iaca_exact_rank_blocks():
    """Exact-rank blocks: the rank and CUR error < 1e-9 once the first r
    columns span the column space (generic blocks)."""
    rng = np.random.default_rng(11)
    cg = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    for m, n, r in ((120, 25, 8), (64, 30, 17), (40, 40, 12)):
        A = cg(m, r) @ cg(r, n)
        rank, rp, cp = O.iaca(A, 50, 1e-10)
        assert rank == r
        rpz, cpz = rp - 1, cp - 1
        recon = A[:, cpz] @ np.linalg.solve(A[np.ix_(rpz, cpz)], A[rpz, :])
        assert relerr(recon, A) < 1e-9


def test_iaca_matches_brute_force_on_kernel_blocks():
    rng = np.random.default_rng(6)
    for trial in range(20):
        m, n = rng.integers(2, 40), rng.integers(2, 30)
        x = rng.random((m, 3))
        y = rng.random((n, 3)) + np.array([2.0 + rng.random(), 0, 0])
        A = np.array([[G(3.0, a, b) for b in y] for a in x])
        for tol in (1e-2, 1e-4, 1e-8):
            got = O.iaca(A, 50, tol)
            ref = brute_force_iaca(A, 50, tol)
            assert got[0] == ref[0] and list(got[1]) == ref[1] and list(got[2]) == ref[2], (trial, tol)


# ---------------------------------------------------------------- skeleton solve
def test_skeleton_solve_closed_form():
    """Z = transpose(L / L[rows, :]) at rtol 1e-11, ranks up to 80
    (test/gpu/test_compresslevel.jl:65-125)."""
    rng = np.random.default_rng(17)
    for m, r in ((90, 8), (100, 35), (160, 80)):
        rows = rng.permutation(m)[:r]
        L = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
        L[rows, :] += r * np.eye(r)
        X, st = O.skeleton_solve(L, rows + 1)
        assert st == -1
        ref = np.linalg.solve(L[rows, :].T, L.T).T
        np.testing.assert_allclose(X, ref, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(X[rows, :], np.eye(r), atol=1e-12)


def test_skeleton_solve_singular():
    """[1 2; 1 2]: 2 pivots, rank 1 (test/utils/test_utils.jl:54-71); duplicated
    rows are dependent (test/test_nestedbasis.jl:55-65)."""
    L = np.array([[1, 2], [1, 2]], dtype=np.complex128)
    _, st = O.skeleton_solve(L, [1, 2])
    assert st == 1
    rng = np.random.default_rng(3)
    B = rng.random((8, 2)) + 0j
    B[6, :] = B[2, :]
    _, st = O.skeleton_solve(B, [3, 7])
    assert st >= 0
    _, st = O.skeleton_solve(np.array([[2, 1], [1, 3]], dtype=np.complex128), [1, 2])
    assert st == -1


# ---------------------------------------------------------------- coupling
def test_coupling_known_layout():
    """Blocks in test-CSR order, empty blocks kept (test/test_couplingmatrices.jl:12-39)."""
    rng = np.random.default_rng(1)
    tp, sp = rng.random((9, 3)), rng.random((7, 3)) + np.array([3.0, 0, 0])
    farptr = np.array([1, 3, 3, 5]) - 1
    farnodes = np.array([2, 4, 1, 4]) - 1
    testpiv = [[1, 4], [2], [5, 9, 3]]
    trialpiv = [[7, 1], [3], [2, 6], []]
    csr = lambda L: (np.concatenate([[0], np.cumsum([len(v) for v in L])]),
                     np.array([i - 1 for v in L for i in v], np.int64))
    tpp, tpi = csr(testpiv)
    spp, spi = csr(trialpiv)
    off, blocks = O.coupling(1.3, 1, tp, sp, farptr, farnodes, tpp, tpi, spp, spi)
    expect = [[[1, 4], [3]], [[1, 4], []], [[5, 9, 3], [7, 1]], [[5, 9, 3], []]]
    assert list(np.diff(off)) == [2, 0, 6, 0]
    for i, (rows, cols) in enumerate(expect):
        blk = blocks[off[i]:off[i + 1]].reshape((len(rows), len(cols)), order="F")
        ref = np.array([[G(1.3, tp[a - 1], sp[b - 1]) for b in cols] for a in rows]).reshape(len(rows), len(cols))
        np.testing.assert_allclose(blk, ref, rtol=1e-14)


# ---------------------------------------------------------------- matvec
def dense_h2(p, ft, fs, H):
    """The compressed operator as a dense matrix, assembled block by block from
    the factors: near blocks exactly, every far pair as U_t S_ts V_s with U_t and
    V_s expanded through the transfers."""
    tt, st = p.tree.testtree, p.tree.trialtree

    def expand(f, tree, side):
        # basis of every active node in its own point order: leaf X (or X^T);
        # inner: stack child bases times transfer blocks
        B = {}
        for lev in sorted(tree.levels().tolist(), reverse=True):
            for t in tree.levelnodes(lev):
                if not f.assigned[t]:
                    continue
                if tree.isleaf(t):
                    B[t] = f.X[t]
                else:
                    parts, o = [], 0
                    for c in tree.childrenof(c_ := t):
                        rc = int(f.rank[c]) if f.assigned[c] else 0
                        E = f.X[t][o:o + rc, :]
                        o += rc
                        parts.append(B[c] @ E if c in B else np.zeros((tree.nvalues(c), f.rank[t])))
                    B[t] = np.vstack(parts)
        return B

    UB = expand(ft, tt, 0)
    VB = expand(fs, st, 1)
    A = np.zeros((tt.npoints, st.npoints), np.complex128)
    off, blocks = H
    tr, sr = ft.rank, fs.rank
    for t in range(tt.nnodes):
        for i in range(p.tf.nodetofarnodesptr[t], p.tf.nodetofarnodesptr[t + 1]):
            s = p.tf.farnodes[i]
            S = blocks[off[i]:off[i + 1]].reshape((tr[t], sr[s]), order="F")
            A[np.ix_(tt.values(t), st.values(s))] += UB[t] @ S @ VB[s].T
        for i in range(p.partition.nearptr[t], p.partition.nearptr[t + 1]):
            s = p.partition.nearnodes[i]
            for a in tt.values(t):
                for b in st.values(s):
                    A[a, b] += O.entry(p.k, p.sign, p.testpts[a], p.trialpts[b])
    return A


@pytest.fixture(scope="module")
def grid():
    p = grid_problem(1e-4)
    ft, fs = oracle_assembly(p)
    H, cp = oracle_operator(p, ft, fs)
    return p, ft, fs, H, cp


def test_partition_covers_every_entry_once(grid):
    """test/admissibility/test_farinteractions.jl:71-80."""
    p = grid[0]
    tt, st = p.tree.testtree, p.tree.trialtree
    cover = np.zeros((tt.npoints, st.npoints), np.int64)
    for t in range(tt.nnodes):
        for i in range(p.tf.nodetofarnodesptr[t], p.tf.nodetofarnodesptr[t + 1]):
            cover[np.ix_(tt.values(t), st.values(p.tf.farnodes[i]))] += 1
        for i in range(p.partition.nearptr[t], p.partition.nearptr[t + 1]):
            cover[np.ix_(tt.values(t), st.values(p.partition.nearnodes[i]))] += 1
    assert np.all(cover == 1)


def test_matvec_is_the_compressed_operator(grid):
    """The 5-stage product equals the dense assembly of U S V + near (all three ops)."""
    p, ft, fs, H, cp = grid
    A = dense_h2(p, ft, fs, cp)
    rng = np.random.default_rng(2)
    x = geometry.complex_gaussian(rng, (p.trialpts.shape[0], 2))
    u = geometry.complex_gaussian(rng, (p.testpts.shape[0], 2))
    assert relerr(H.matvec(x), A @ x) < 1e-13
    assert relerr(H.matvec(u, op=1), A.T @ u) < 1e-13
    assert relerr(H.matvec(u, op=2), A.conj().T @ u) < 1e-13
    y0 = geometry.complex_gaussian(rng, (p.testpts.shape[0], 2))
    al, be = 0.3 + 0.2j, -1.1
    assert relerr(H.matvec(x, alpha=al, beta=be, y=y0), al * (A @ x) + be * y0) < 1e-13
    rows = np.array([0, 5, 77, 1023])
    yr = H.matvec(x, rows=rows)
    assert relerr(yr[rows], (A @ x)[rows]) < 1e-13
    # coupling assembled only for the clusters that hold the rows: same rows
    Hr = oracle_operator_rows(p, ft, fs, rows)
    assert np.array_equal(Hr.matvec(x, rows=rows)[rows], yr[rows])
    need = O.nodes_holding(p.tree.testtree.lo, p.tree.testtree.hi, p.tree.testtree.perm, rows)
    assert 0 < need.sum() < p.tree.testtree.nnodes


def test_matvec_far_plus_near(grid):
    p, ft, fs, H, cp = grid
    rng = np.random.default_rng(4)
    x = geometry.complex_gaussian(rng, (p.trialpts.shape[0],))
    far = H.matvec(x, near=False)
    full = H.matvec(x)
    rows = np.arange(p.testpts.shape[0])
    dense = O.dense_rows(p.k, p.sign, p.testpts, p.trialpts, rows, x)
    assert relerr(full, dense) < 5 * p.tol
    assert relerr(far, full) > 0


@pytest.mark.parametrize("tol", [1e-2, 1e-4, 1e-6])
def test_matvec_vs_direct_summation(tol):
    """norm(nca * x - y) / norm(y) < 5 tol (test/nca/test_petrovgalerkinnca.jl:102)."""
    p = sphere_problem(4096, k=1.0, tol=tol)
    ft, fs = oracle_assembly(p)
    H, _ = oracle_operator(p, ft, fs)
    rng = np.random.default_rng(2)
    x = geometry.complex_gaussian(rng, (4096,))
    rows = rng.choice(4096, 300, replace=False)
    ref = O.dense_rows(p.k, p.sign, p.testpts, p.trialpts, rows, x)[:, 0]
    assert relerr(H.matvec(x)[rows, 0], ref) < 5 * tol
    ref_t = O.dense_rows(p.k, p.sign, p.trialpts, p.testpts, rows, x)[:, 0]  # symmetric kernel
    assert relerr(H.matvec(x, op=1)[rows, 0], ref_t) < 5 * tol


def test_stored_near_field_matches_on_the_fly():
    """The stored near field (dense block rows, the reference's CPU layout)
    gives the on-the-fly product: same entries, other summation order."""
    p = sphere_problem(2048, tol=1e-4)
    ft, fs = oracle_assembly(p)
    H, _ = oracle_operator(p, ft, fs)
    rng = np.random.default_rng(9)
    x = geometry.complex_gaussian(rng, (p.trialpts.shape[0], 5))
    y0 = H.matvec(x)
    n = H.store_near()
    nn = p.partition.nearptr
    sz_t = p.tree.testtree.hi - p.tree.testtree.lo
    sz_s = p.tree.trialtree.hi - p.tree.trialtree.lo
    rows = np.repeat(np.arange(len(nn) - 1), np.diff(nn))
    assert n == int(np.sum(sz_t[rows] * sz_s[p.partition.nearnodes]))
    y1 = H.matvec(x)
    assert relerr(y1, y0) <= 1e-13
    # restricted rows and transposed products still evaluate on the fly
    u = geometry.complex_gaussian(rng, (p.testpts.shape[0], 2))
    H2 = oracle_operator(p, ft, fs)[0]
    assert np.array_equal(H.matvec(u, op=1), H2.matvec(u, op=1))
    H.drop_near()
    assert np.array_equal(H.matvec(x), y0)


@pytest.mark.parametrize("kind,n,lam", [("kmeans", 3000, 1.5), ("twon", 3000, 1.5)])
def test_oracle_directional_vs_direct_summation(kind, n, lam):
    """The oracle's directional product (OracleDH2: the reference's directional
    stages restated in numpy) over the oracle's own level driver on the
    (cluster, direction) blocks, with the fully pivoted ACA and with the IACA:
    within 10 tol of dense direct summation for tol 1e-2 and 1e-4
    (test/nca/test_petrovgalerkinnca.jl:104-153, wideband with etalf = etahf =
    gamma = 1)."""
    import math
    import nca_amd
    from nca_amd import directional as D, wideband as W, compressors as C, trees as T, geometry
    from helpers import oracle_factors
    pts = geometry.fibonacci_sphere(n)
    tr = nca_amd.TwoNTree(pts, minvalues=16) if kind == "twon" else nca_amd.KMeansTree(pts, 2, 16)
    tree = T.BlockTree(tr, tr)
    k = 2 * math.pi / lam
    isn = D.isnearwideband(k, etalf=1.0, etahf=1.0, gamma=1.0)
    part = D.interactions(tree, isn)
    td, sd = D.fardata(tree, isn, part)
    tff, sff = D.dirfarfield_all(tr, td), D.dirfarfield_all(tr, sd)
    tb, sb = W.DirectionBlocks(tr, td, tff[0]), W.DirectionBlocks(tr, sd, sff[0])
    sel = C.GeometricPivoting(device=False)
    fpt = W.farpivots(tb, tr, tff[0], tff[1], 50, sel)
    fps = W.farpivots(sb, tr, sff[0], sff[1], 50, sel)
    x = np.random.default_rng(4).standard_normal(n) + 1j * np.random.default_rng(5).standard_normal(n)
    rows = np.arange(0, n, 7)
    yd = O.dense_rows(k, 1, pts, pts, rows, x)[:, 0]
    for tol, iaca in ((1e-2, False), (1e-4, False), (1e-2, True), (1e-4, True)):
        tres = O.bases(0, k, 1, pts, pts, tb, fpt[0], fpt[1], 50, tol, iaca=iaca)
        sres = O.bases(1, k, 1, pts, pts, sb, fps[0], fps[1], 50, tol, iaca=iaca)
        assert tres[4] == 0 and sres[4] == 0
        ft, fs = oracle_factors(tres, tb, fpt), oracle_factors(sres, sb, fps)
        tB, tT = W.dir_stores(ft, tb, 0)
        sB, sT = W.dir_stores(fs, sb, 1)
        cplan = W.coupling_plan(td, sd)
        tp_, sp_ = ft.pivot_csr(), fs.pivot_csr()
        off, blocks = O.coupling(k, 1, pts, pts, td.diridxtofarptr, cplan[2], tp_[0], tp_[1], sp_[0], sp_[1])
        A = O.OracleDH2(W.dh2desc(tree, part, td, sd, ft, fs, tB, tT, sB, sT, cplan), k, 1, pts, pts, off, blocks)
        y = A.matvec(x)
        err = np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd)
        assert err < 10 * tol, (tol, iaca, err)
        # the row-restricted product (only clusters holding the rows) is exact there
        yr = A.matvec(x, rows=rows[::5])
        assert relerr(yr[rows[::5]], y[rows[::5]]) < 1e-14
        assert not np.any(np.delete(yr, rows[::5]))


# ---------------------------------------------------------------- TopDown: partially pivoted ACA
def brute_force_paca(A, maxrank, tol):
    """The partially pivoted ACA written with numpy from its description: row
    pivot 0 first, column pivot = first maximum of the residual row over unused
    columns, next row pivot = first maximum of the new column over unused rows,
    stop (keeping the cross) when |u| |v| <= tol ||U V^T||_F, exact norms."""
    m, n = A.shape
    U = np.zeros((m, 0), np.complex128)
    V = np.zeros((n, 0), np.complex128)
    usedr, usedc = np.zeros(m, bool), np.zeros(n, bool)
    rows, cols = [], []
    i = 0
    while len(rows) < min(maxrank, m, n):
        row = A[i, :] - U[i, :] @ V.T
        w = np.where(usedc, -1.0, np.abs(row) ** 2)
        j = int(np.argmax(w))
        if not w[j] > 0:
            break
        v = row / row[j]
        u = A[:, j] - U @ V[j, :]
        U = np.column_stack([U, u])
        V = np.column_stack([V, v])
        rows.append(i + 1)
        cols.append(j + 1)
        usedr[i] = usedc[j] = True
        if np.sum(np.abs(u) ** 2) * np.sum(np.abs(v) ** 2) <= tol * tol * np.sum(np.abs(U @ V.T) ** 2):
            break
        wr = np.where(usedr, -1.0, np.abs(u) ** 2)
        if not np.any(~usedr):
            break
        i = int(np.argmax(wr))
    return len(rows), rows, cols, U, V


def kernel_block(k, x, y, sign=1):
    return np.array([[O.entry(k, sign, a, b) for b in y] for a in x])


def test_paca_matches_brute_force_on_kernel_blocks():
    """Ranks and pivots of the oracle's partially pivoted ACA equal a numpy
    restatement's on separated random clusters (both block orientations), and
    U V^T reproduces the block to a small multiple of tol."""
    rng = np.random.default_rng(12)
    for trial in range(12):
        m, n = int(rng.integers(3, 60)), int(rng.integers(3, 300))
        x = rng.random((m, 3))
        y = rng.random((n, 3)) * 1.5 + np.array([2.5 + rng.random(), 0.3, 0])
        for tol in (1e-2, 1e-4, 1e-8):
            for transposed in (0, 1):
                A = kernel_block(2.0, x, y) if not transposed else kernel_block(2.0, y, x).T
                r, rp, cp, U, V = O.paca(2.0, 1, transposed, x, y, np.arange(m), np.arange(n), 40, tol)
                rr, rrp, rcp, RU, RV = brute_force_paca(A, 40, tol)
                assert r == rr and list(rp) == rrp and list(cp) == rcp, (trial, tol, transposed)
                assert np.allclose(U, RU, rtol=1e-10, atol=1e-14 * np.abs(A).max())
                assert np.linalg.norm(A - U @ V.T) <= 20 * tol * np.linalg.norm(A) or r == min(40, m, n)


def test_paca_exact_rank_one_and_zero_blocks():
    """Columns of one repeated point: the block has rank 1 and the ACA stops
    after one cross with U V^T = A; coincident points give a zero block and
    rank 0."""
    x = np.random.default_rng(3).random((17, 3))
    y = np.repeat(np.array([[3.0, 0.5, 0.2]]), 9, axis=0)
    A = kernel_block(1.5, x, y)
    r, rp, cp, U, V = O.paca(1.5, 1, 0, x, y, np.arange(17), np.arange(9), 20, 1e-12)
    assert r == 1 and rp[0] == 1
    assert np.allclose(U @ V.T, A, rtol=0, atol=1e-15 * np.abs(A).max())
    z = np.zeros((5, 3))
    r, *_ = O.paca(1.5, 1, 0, z, z, np.arange(5), np.arange(5), 5, 1e-4)
    assert r == 0


def separated_squares(tol, sign=-1):
    """The reference's "single far interaction" geometry
    (test/nca/test_petrovgalerkinnca.jl:155-229): two 7 x 7 grids of the unit
    square, the second shifted by (3, 0, 0); k = 2 pi / (100 h), h = 0.16,
    exp(-i k r); 2-means trees with minvalues 60."""
    tp = geometry.grid_square(7)
    sp = tp + np.array([3.0, 0.0, 0.0])
    tree = nca_amd_tree(tp, sp, 60)
    return Problem(tp, sp, tree, 2 * np.pi / (100 * 0.16), sign, tol, 50)


def nca_amd_tree(tp, sp, minvalues):
    import nca_amd
    return nca_amd.BlockTree(nca_amd.KMeansTree(tp, 2, minvalues), nca_amd.KMeansTree(sp, 2, minvalues))


@pytest.mark.parametrize("tol", [1e-2, 1e-4, 1e-6])
def test_topdown_oracle_single_far_interaction(tol):
    """TopDown with ACA on the reference's single-far-interaction case: the
    compressed operator within 5 tol of the dense matrix (the reference's
    bound, test_petrovgalerkinnca.jl:229)."""
    p = separated_squares(tol)
    ft, fs = oracle_topdown(p, 0), oracle_topdown(p, 1)
    _, H = oracle_operator(p, ft, fs)
    A = dense_h2(p, ft, fs, H)
    D = kernel_block(p.k, p.testpts, p.trialpts, p.sign)
    x = geometry.complex_gaussian(np.random.default_rng(1), (p.trialpts.shape[0], 1))[:, 0]
    assert relerr(A @ x, D @ x) < 5 * tol


def test_topdown_oracle_readme_case():
    """The README's example (README.md:95-137: two 32 x 32 grids, k = 2 pi / 10,
    2-means trees with minvalues 32, TopDown with ACA(tol=1e-4)): a leaf's
    factor interpolates at its pivots (X[tau, :] = I), the stacked transfers
    of a parent are A[tau_c, sigma] A[tau, sigma]^{-1}, and the operator is
    within 20 tol of direct summation (the README states no bound)."""
    p = grid_problem(1e-4)
    ft, fs = oracle_topdown(p, 0), oracle_topdown(p, 1)
    tt = p.tree.testtree
    checked = 0
    for t in range(tt.nnodes):
        if not ft.assigned[t] or ft.rank[t] == 0:
            continue
        piv = ft.pivots[t]
        sigma = ft.colpivots[1][ft.colpivots[0][t]:ft.colpivots[0][t + 1]]
        P = kernel_block(p.k, p.testpts[piv], p.trialpts[sigma])
        if tt.isleaf(t):
            pos = [int(np.nonzero(tt.values(t) == q)[0][0]) for q in piv]
            assert np.allclose(ft.X[t][pos, :], np.eye(len(piv)), atol=1e-9)
        else:
            rows = np.concatenate([ft.pivots[c] for c in tt.childrenof(t) if ft.assigned[c]])
            L = kernel_block(p.k, p.testpts[rows], p.trialpts[sigma])
            assert np.allclose(ft.X[t] @ P, L, rtol=1e-8, atol=1e-10 * np.abs(L).max())
        checked += 1
    assert checked > 10
    _, H = oracle_operator(p, ft, fs)
    A = dense_h2(p, ft, fs, H)
    D = kernel_block(p.k, p.testpts, p.trialpts)
    x = geometry.complex_gaussian(np.random.default_rng(2), (p.trialpts.shape[0], 1))[:, 0]
    assert relerr(A @ x, D @ x) < 20 * p.tol
