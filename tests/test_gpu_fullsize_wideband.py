"""The directional wideband workload at its full size on the GPU path: N =
262,144 Fibonacci points of the unit sphere (test = trial), k = 16 pi (16
wavelengths across), 2-means tree with minvalues 64, isnearwideband(k) with
etalf 1, etahf 5, gamma 1, BottomUp with the IACA per (cluster, direction)
block at tol 1e-4, maxrank 50, 8 seeded complex Gaussian right-hand sides.

Checks: ranks, pivots and skeleton factors of all 775k test (cluster,
direction) blocks bit for bit against the oracle's level driver on the same
blocks and Phase-1 pivots; the coupling blocks of the clusters that hold 1,000
seeded rows bit for bit against the oracle's; on those rows the FP64 product
within 1e-12 and MIXED within 1e-5 of the oracle's directional product
(OracleDH2) over all 8 right-hand sides; MIXED within 1e-5 of FP64; both within
10 tol of dense direct summation on sampled rows (the reference's wideband test
bound, test/nca/test_petrovgalerkinnca.jl:104-153); and <y, A x> = <A^H y, x>,
y^T (A x) = (A^T y)^T x and linearity across right-hand sides."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import nca as NC
from nca_amd import wideband as W
from oracle import oracle as O
from helpers import bitwise_equal, oracle_factors, relerr

pytestmark = pytest.mark.gpu

N, K, TOL, MAXRANK, NRHS = 262144, 16 * math.pi, 1e-4, 50, 8


def say(msg):
    print(f"[gpu-fullsize-wideband] {msg}", flush=True)


@pytest.fixture(scope="module")
def full():
    say(f"assemble N={N} k=16 pi tol={TOL}")
    pts = geometry.fibonacci_sphere(N)
    tr = nca_amd.KMeansTree(pts, 2, 64)
    core = W.DirectionalNCACore(nca_amd.HelmholtzKernel(K), pts, pts, nca_amd.BlockTree(tr, tr),
                                W.BottomUp(C.IACA(TOL)), W.BottomUp(C.IACA(TOL)), D.isnearwideband(K),
                                maxrank=MAXRANK, ctx=nca_amd.default_context())
    X = np.asfortranarray(geometry.complex_gaussian(np.random.default_rng(8), (N, NRHS)))
    return pts, core, X


def product(core, precision, X, op=nca_amd.NCA_OP_N):
    core.set_precision(precision)
    Y = np.zeros((N, X.shape[1]), np.complex128)
    return NC.mul(Y, NC.H2Workspace(core, op), X)


def test_fullsize_wideband_pivots_equal_oracle(full):
    pts, core, _ = full
    blocks, f = core.testblocks, core.testfactors
    say(f"oracle bases over {blocks.nnodes} direction indices (test side)")
    fp = f.farpivots
    res = O.bases(0, K, 1, pts, pts, blocks, fp[0], fp[1], MAXRANK, TOL, iaca=True)
    assert res[4] == 0
    ref = oracle_factors(res, blocks, fp)
    assert np.array_equal(ref.assigned, f.assigned)
    assert np.array_equal(ref.rank, f.rank)
    for a, b in zip(ref.pivot_csr(), f.pivot_csr()):
        assert np.array_equal(a, b)
    for e in np.nonzero(f.assigned)[0]:
        assert bitwise_equal(ref.X[e], f.X[e]), e


def test_fullsize_wideband_product_equals_oracle_on_rows(full):
    """On 1,000 seeded rows: the coupling blocks of every test cluster that
    holds one of them bit for bit against the oracle's, then the FP64 product
    within 1e-12 and MIXED within 1e-5 of the oracle's directional product
    (its coupling blocks; factors bit for bit equal, checked above)."""
    pts, core, X = full
    rows = np.sort(np.random.default_rng(1000).choice(N, 1000, replace=False))
    tt = core.tree.testtree
    need = O.nodes_holding(tt.lo, tt.hi, tt.perm, rows)
    td = core.testdata
    fptr, snodes = td.diridxtofarptr, core.couplingplan[2]
    keep = need[core.testblocks.node]
    rptr, rnodes = O.restrict_rows(fptr, snodes, keep)
    tp, sp = core.testpivots, core.trialpivots
    say(f"oracle coupling for {int(keep.sum())} test direction indices")
    roff, rblocks = O.coupling(K, 1, pts, pts, rptr, rnodes, tp[0], tp[1], sp[0], sp[1])
    # back to the store's numbering: blocks of the other clusters have size 0
    kept = np.nonzero(np.repeat(keep, np.diff(fptr)))[0]
    sizes = np.zeros(len(snodes), np.int64)
    sizes[kept] = np.diff(roff)
    off = np.zeros(len(snodes) + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    goff, gblocks = core.couplingstore.download()
    ks = sizes[kept]
    gidx = np.repeat(goff[kept] - np.cumsum(ks) + ks, ks) + np.arange(int(ks.sum()))
    assert bitwise_equal(gblocks[gidx], rblocks)
    del gblocks
    A = O.OracleDH2(core.desc, K, 1, pts, pts, off, rblocks)
    say("oracle directional product on 1,000 rows")
    ref = A.matvec(X, rows=rows)[rows]
    Y64 = product(core, nca_amd.NCA_PREC_FP64, X)[rows]
    Ymix = product(core, nca_amd.NCA_PREC_MIXED, X)[rows]
    e64, emix = relerr(Y64, ref), relerr(Ymix, ref)
    say(f"vs oracle on 1,000 rows x 8 RHS: FP64 {e64:.3e}, MIXED {emix:.3e}")
    assert e64 <= 1e-12
    assert emix <= 1e-5


def test_fullsize_wideband_mixed_and_direct_summation(full):
    pts, core, X = full
    say("FP64 and MIXED products, 8 RHS")
    Y64 = product(core, nca_amd.NCA_PREC_FP64, X)
    Ymix = product(core, nca_amd.NCA_PREC_MIXED, X)
    err = relerr(Ymix, Y64)
    say(f"MIXED vs FP64 {err:.3e}")
    assert err <= 1e-5
    rows = np.sort(np.random.default_rng(9).choice(N, 256, replace=False))
    ref = O.dense_rows(K, 1, pts, pts, rows, np.asfortranarray(X[:, :2]))
    e64, emix = relerr(Y64[rows, :2], ref), relerr(Ymix[rows, :2], ref)
    say(f"vs direct summation: FP64 {e64:.3e}, MIXED {emix:.3e}")
    assert e64 < 10 * TOL and emix < 10 * TOL


@pytest.mark.parametrize("precision,bound", [(nca_amd.NCA_PREC_FP64, 1e-12), (nca_amd.NCA_PREC_MIXED, 1e-5)])
def test_fullsize_wideband_adjoint_transpose_linearity(full, precision, bound):
    _, core, X = full
    say(f"identities, precision {precision}")
    x, y = X[:, :2], X[:, 2:4]
    Ax = product(core, precision, x)
    AHy = product(core, precision, y, nca_amd.NCA_OP_C)
    ATy = product(core, precision, y, nca_amd.NCA_OP_T)
    for q in range(2):
        scale = np.linalg.norm(y[:, q]) * np.linalg.norm(Ax[:, q])
        assert abs(np.vdot(y[:, q], Ax[:, q]) - np.vdot(AHy[:, q], x[:, q])) <= bound * scale
        assert abs(y[:, q] @ Ax[:, q] - ATy[:, q] @ x[:, q]) <= bound * scale
    comb = product(core, precision, np.asfortranarray(x[:, :1] + 2j * x[:, 1:2]))[:, 0]
    assert relerr(comb, Ax[:, 0] + 2j * Ax[:, 1]) <= bound
