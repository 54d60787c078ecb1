"""Degenerate and ragged inputs through the C-ABI: a single point, problems
with no admissible pair (near field only), test and trial sets of different
sizes far apart (far field only at the root), zero right-hand sides, a
wideband problem without any direction, and coincident points under different
test and trial trees. Each product is checked against the
oracle's (dense direct summation where the operator is exact)."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import nca as NC
from nca_amd import trees as T
from nca_amd import wideband as W
from oracle import oracle as O
from helpers import relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-edge] {msg}", flush=True)


def core_of(tp, sp, k=1.0, tol=1e-4, minvalues=16):
    tt = nca_amd.KMeansTree(tp, 2, minvalues)
    st = tt if sp is tp else nca_amd.KMeansTree(sp, 2, minvalues)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(tol))
    return nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), tp, sp, nca_amd.BlockTree(tt, st), cmp, cmp)


@pytest.mark.parametrize("n", [1, 2, 7, 33])
def test_near_field_only(n):
    """Too few points for any admissible pair: the product is the near field,
    i.e. dense direct summation (zero on the diagonal)."""
    say(f"near only n={n}")
    pts = geometry.fibonacci_sphere(n) if n > 1 else np.zeros((1, 3))
    core = core_of(pts, pts, minvalues=64)
    # one point: a cluster of radius 0 is admissible with itself (diameter 0
    # is not > eta * distance 0), a zero block of rank 0
    assert core.stats()["coupling_entries"] == 0
    A = NC.H2Workspace(core)
    rng = np.random.default_rng(n)
    for R in (1, 5, 16):
        x = geometry.complex_gaussian(rng, (n, R))
        ref = O.dense_rows(1.0, 1, pts, pts, np.arange(n), x)
        assert relerr(A @ x, ref) <= 1e-12
        assert relerr(NC.transpose(A) @ x, ref) <= 1e-12  # symmetric kernel
    if n == 1:
        assert np.all((A @ np.ones(1, np.complex128)) == 0)


def test_far_apart_sets_of_different_sizes():
    """M = 37 test and N = 300 trial points far apart: one admissible pair at
    the roots; the product equals the oracle's on the same factors."""
    say("far apart 37 x 300")
    tp = geometry.fibonacci_sphere(37) * 0.3
    sp = geometry.fibonacci_sphere(300) * 0.5 + np.array([6.0, 0.0, 0.0])
    core = core_of(tp, sp, k=0.7)
    st = core.stats()
    assert st["near_pairs"] == 0 and st["coupling_blocks"] >= 1
    off, blocks = core.couplingstore.download()
    H = O.OracleH2(core.desc, 0.7, 1, tp, sp, off, blocks)
    A = NC.H2Workspace(core)
    rng = np.random.default_rng(3)
    x = geometry.complex_gaussian(rng, (300, 3))
    u = geometry.complex_gaussian(rng, (37, 3))
    assert relerr(A @ x, H.matvec(x)) <= 1e-12
    assert relerr(NC.adjoint(A) @ u, H.matvec(u, op=2)) <= 1e-12
    assert relerr(A @ x, O.dense_rows(0.7, 1, tp, sp, np.arange(37), x)) < 5e-4


def test_zero_right_hand_sides_and_shapes():
    pts = geometry.fibonacci_sphere(500)
    core = core_of(pts, pts)
    A = NC.H2Workspace(core)
    y = np.ones((500, 0), np.complex128)
    out = A._mul(y, np.zeros((500, 0), np.complex128), 1.0, 0.0)
    assert out.shape == (500, 0)
    with pytest.raises(ValueError):
        A @ np.zeros(499, np.complex128)


def test_wideband_without_directions():
    """A wideband problem whose clusters are all near (no direction index):
    the directional operator is its near field."""
    say("wideband, no directions")
    pts = geometry.fibonacci_sphere(200)
    tr = nca_amd.KMeansTree(pts, 2, 16)
    k = 2 * math.pi / 0.1
    bu = W.BottomUp(C.ACA(1e-4))
    core = W.DirectionalNCACore(nca_amd.HelmholtzKernel(k), pts, pts, T.BlockTree(tr, tr), bu, bu,
                                D.isnearwideband(k))
    assert core.testdata.ndiridx == 0
    x = geometry.complex_gaussian(np.random.default_rng(2), (200, 4))
    assert relerr(NC.H2Workspace(core) @ x, O.dense_rows(k, 1, pts, pts, np.arange(200), x)) <= 1e-12


def test_coincident_points_in_different_trees():
    """The same 2,000 points on both sides under different trees (2^n tree for
    the test side, 2-means for the trial side): coincident points meet in leaf
    pairs other than a leaf with itself, where the near field must still give
    G = 0 at distance 0. FP64 within 1e-12 and MIXED within 1e-5 of the
    oracle's product on the same factors, op N and C."""
    say("coincident points, different trees")
    pts = geometry.fibonacci_sphere(2000)
    tt = nca_amd.TwoNTree(pts, minvalues=32)
    st = nca_amd.KMeansTree(pts, 2, 24)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    core = nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(2.0), pts, pts, nca_amd.BlockTree(tt, st), cmp,
                                         cmp)
    assert core.stats()["near_pairs"] > 0
    off, blocks = core.couplingstore.download()
    H = O.OracleH2(core.desc, 2.0, 1, pts, pts, off, blocks)
    A = NC.H2Workspace(core)
    rng = np.random.default_rng(17)
    for prec, bound in ((nca_amd.NCA_PREC_FP64, 1e-12), (nca_amd.NCA_PREC_MIXED, 1e-5)):
        core.set_precision(prec)
        for R in (16, 5):
            x = geometry.complex_gaussian(rng, (2000, R))
            assert relerr(A @ x, H.matvec(x)) <= bound, (prec, R)
            assert relerr(NC.adjoint(A) @ x, H.matvec(x, op=2)) <= bound, (prec, R)
    core.set_precision(nca_amd.NCA_PREC_FP64)
