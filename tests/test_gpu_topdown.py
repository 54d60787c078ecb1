"""TopDown (the README's compressor, README.md:95-137) on the GPU
(nca_topdown_side): the partially pivoted ACA of every cluster against its
far field and its parent's column pivots, level by level from the root, then
the factors. Ranks, row and column pivots and every factor X bit for bit
against the oracle's restatement (oracle.topdown) on both sides; the product
against the oracle's operator (1e-12) and against dense direct summation
within the reference's bound of 5 tol on its single-far-interaction case
(test/nca/test_petrovgalerkinnca.jl:155-229) and within 20 tol on the README
case, whose text states no bound;
mixed compressors and the reuse of the test side's factors on one tree."""
import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from nca_amd import nca as NC
from oracle import oracle as O
from helpers import Problem, bitwise_equal, grid_problem, oracle_operator, oracle_topdown, relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-topdown] {msg}", flush=True)


def separated_squares(tol):
    tp = geometry.grid_square(7)
    sp = tp + np.array([3.0, 0.0, 0.0])
    tree = nca_amd.BlockTree(nca_amd.KMeansTree(tp, 2, 60), nca_amd.KMeansTree(sp, 2, 60))
    return Problem(tp, sp, tree, 2 * np.pi / (100 * 0.16), -1, tol, 50)


def core_of(p, testcomp=None, trialcomp=None, **kw):
    td = C.TopDown(factorization=C.ACA(p.tol))
    return nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(p.k, p.sign), p.testpts, p.trialpts, p.tree,
                                         testcomp or td, trialcomp or td, maxrank=p.maxrank, partition=p.partition,
                                         **kw)


def same_as_oracle(f, ref, tree):
    assert np.array_equal(f.assigned, ref.assigned)
    assert np.array_equal(f.rank[f.assigned], ref.rank[ref.assigned])
    for t in np.nonzero(f.assigned)[0]:
        assert np.array_equal(f.pivots[t], ref.pivots[t]), t
        assert np.array_equal(f.colpivots[1][f.colpivots[0][t]:f.colpivots[0][t + 1]],
                              ref.colpivots[1][ref.colpivots[0][t]:ref.colpivots[0][t + 1]]), t
        assert bitwise_equal(f.X[t], ref.X[t]), t


def dense(p):
    return np.array([[O.entry(p.k, p.sign, a, b) for b in p.trialpts] for a in p.testpts])


@pytest.mark.parametrize("tol", [1e-2, 1e-4, 1e-6])
def test_topdown_single_far_interaction(tol):
    p = separated_squares(tol)
    say(f"single far interaction tol={tol}")
    core = core_of(p)
    for side, f in ((0, core.testfactors), (1, core.trialfactors)):
        ref = oracle_topdown(p, side, ranks_from=f.assigned)
        same_as_oracle(f, ref, p.tree.testtree if side == 0 else p.tree.trialtree)
    x = geometry.complex_gaussian(np.random.default_rng(1), (p.trialpts.shape[0], 2))
    y = NC.H2Workspace(core) @ x
    assert relerr(y, dense(p) @ x) < 5 * tol


def test_topdown_readme_case():
    p = grid_problem(1e-4)
    say("README case")
    core = core_of(p)
    assert not core.symmetric
    ft = oracle_topdown(p, 0, ranks_from=core.testfactors.assigned)
    fs = oracle_topdown(p, 1, ranks_from=core.trialfactors.assigned)
    same_as_oracle(core.testfactors, ft, p.tree.testtree)
    same_as_oracle(core.trialfactors, fs, p.tree.trialtree)
    H, _ = oracle_operator(p, ft, fs)
    x = geometry.complex_gaussian(np.random.default_rng(3), (p.trialpts.shape[0], 4))
    W = NC.H2Workspace(core)
    y = W @ x
    assert relerr(y, H.matvec(x)) <= 1e-12
    # the README states no bound for its example; TopDown's local errors add up
    # over the 2-means tree's levels (measured: 1.05e-3 at tol 1e-4, the same
    # for the oracle's operator, which equals this product to 1e-12)
    assert relerr(y, dense(p) @ x) < 20 * p.tol
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    assert relerr(NC.H2Workspace(core) @ x, y) <= 1e-5


def test_topdown_one_tree_reuse_and_identities():
    """One tree and one point set: the trial side reuses the test side's
    factors (the same products bit for bit as two passes), and
    <u, A x> = <A^H u, x>, u^T A x = (A^T u)^T x."""
    pts = geometry.fibonacci_sphere(6000)
    tr = nca_amd.TwoNTree(pts, minvalues=32)
    td = C.TopDown(factorization=C.ACA(1e-4))
    say("one tree")
    cores = [nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(2.0), pts, pts, nca_amd.BlockTree(tr, tr), td, td,
                                           reuse_symmetric=r) for r in (True, False)]
    assert cores[0].symmetric and not cores[1].symmetric
    rng = np.random.default_rng(4)
    x = geometry.complex_gaussian(rng, (pts.shape[0], 3))
    u = geometry.complex_gaussian(rng, (pts.shape[0], 3))
    a, b = (NC.H2Workspace(c) for c in cores)
    assert np.array_equal(a @ x, b @ x)
    Ax, AHu, ATu = a @ x, NC.adjoint(a) @ u, NC.transpose(a) @ u
    for q in range(3):
        s = np.linalg.norm(u[:, q]) * np.linalg.norm(Ax[:, q])
        assert abs(np.vdot(u[:, q], Ax[:, q]) - np.vdot(AHu[:, q], x[:, q])) <= 1e-12 * s
        assert abs(u[:, q] @ Ax[:, q] - ATu[:, q] @ x[:, q]) <= 1e-12 * s
    rows = np.sort(rng.choice(pts.shape[0], 200, replace=False))
    ref = O.dense_rows(2.0, 1, pts, pts, rows, np.asfortranarray(x))
    assert relerr(Ax[rows], ref) < 10e-4


def test_topdown_mixed_with_bottomup():
    """TopDown on the test side, BatchedBottomUp on the trial side."""
    p = grid_problem(1e-4)
    say("TopDown + BatchedBottomUp")
    core = core_of(p, trialcomp=C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(p.tol)))
    x = geometry.complex_gaussian(np.random.default_rng(5), (p.trialpts.shape[0], 2))
    assert relerr(NC.H2Workspace(core) @ x, dense(p) @ x) < 20 * p.tol


def test_topdown_rejects_unsupported_configurations():
    with pytest.raises(ValueError, match="ACA"):
        C.TopDown(factorization=C.IACA(1e-4))
    with pytest.raises(ValueError, match="representor"):
        C.TopDown(representor=object())
