"""The fused bottom-up driver (nca_bottomup_side: one launch per level that
assembles, compresses and solves every block, the rows built on the device
from the level below) against the level-by-level calls (nca_level_assemble +
nca_compress_level[_iaca]): ranks, pivots and every factor X bit for bit, for
the fully pivoted ACA and the IACA, on the test side and on the transposed
trial side, for a plain tree pair and for directional blocks; and the
argument checks of the driver."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import trees as T
from nca_amd import wideband as W
from helpers import bitwise_equal

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-fused] {msg}", flush=True)


def same_factors(a, b):
    assert np.array_equal(a.assigned, b.assigned)
    assert np.array_equal(a.rank, b.rank)
    assert np.array_equal(a.m, b.m)
    for x, y in zip(a.pivot_csr(), b.pivot_csr()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.xoff[a.assigned], b.xoff[b.assigned])
    assert bitwise_equal(a.Xflat, b.Xflat)
    for lev in a.level_transfer_nodes:
        assert a.level_transfer_nodes[lev] == b.level_transfer_nodes[lev]


@pytest.mark.parametrize("method", ["aca", "iaca"])
def test_fused_equals_level_calls_plain(method):
    ctx = nca_amd.default_context()
    tp = geometry.ellipsoid(9000, (1.2, 0.8, 1.0))
    sp = geometry.ellipsoid(7000, (0.9, 1.1, 0.7)) + np.array([0.05, -0.1, 0.0])
    tree = T.BlockTree(nca_amd.TwoNTree(tp, 32), nca_amd.KMeansTree(sp, 2, 24))
    part = T.interactions(tree, 1.0)
    tf, sf = T.fardata(tree, part)
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(3.0), tp, sp, ctx)
    cfg = C.IACA(1e-5) if method == "iaca" else C.ACA(1e-5)
    comp = C.BatchedBottomUp(C.GeometricPivoting(), cfg)
    for side, fd, (st, ot) in ((0, tf, (tree.testtree, tree.trialtree)), (1, sf, (tree.trialtree, tree.testtree))):
        say(f"plain {method} side {side}")
        fp = comp.pivotselector.select(st, ot, fd, 40, A.trial if side == 0 else A.test)
        fused = C.batched_side(comp, A, tree, fd, side, 40, fp, resident=True, fused=True)
        level = C.batched_side(comp, A, tree, fd, side, 40, fp, resident=True, fused=False)
        host = C.batched_side(comp, A, tree, fd, side, 40, fp, resident=False)
        same_factors(fused, level)
        same_factors(fused, host)


@pytest.mark.parametrize("method", ["aca", "iaca"])
def test_fused_equals_level_calls_directional(method):
    pts = geometry.fibonacci_sphere(4000)
    tr = nca_amd.TwoNTree(pts, minvalues=16)
    k = 2 * math.pi
    isn = D.isnearwideband(k, etalf=1.0, etahf=1.0, gamma=1.0)
    cfg = C.IACA(1e-4) if method == "iaca" else C.ACA(1e-4)
    say(f"directional {method}")
    cores = [W.DirectionalNCACore(nca_amd.HelmholtzKernel(k), pts, pts, T.BlockTree(tr, tr), W.BottomUp(cfg),
                                  W.BottomUp(cfg), isn, maxrank=50, resident=r) for r in (True, False)]
    same_factors(cores[0].testfactors, cores[1].testfactors)
    x = geometry.complex_gaussian(np.random.default_rng(5), (pts.shape[0], 4))
    from nca_amd import nca as NC
    assert bitwise_equal(NC.H2Workspace(cores[0]) @ x, NC.H2Workspace(cores[1]) @ x)


def test_fused_driver_rejects_bad_arguments():
    ctx = nca_amd.default_context()
    pts = geometry.fibonacci_sphere(3000)
    tr = nca_amd.TwoNTree(pts, 32)
    tree = T.BlockTree(tr, tr)
    tf, _ = T.fardata(tree, T.interactions(tree, 1.0))
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(1.0), pts, pts, ctx)
    comp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    fp = comp.pivotselector.select(tr, tr, tf, 40, A.trial)
    with pytest.raises(nca_amd.NCAError, match="maxrank"):
        C.batched_side(comp, A, tree, tf, 0, 97, fp, resident=True, fused=True)
    with pytest.raises(nca_amd.ArgumentError, match="more columns than maxrank"):
        C.batched_side(comp, A, tree, tf, 0, 10, fp, resident=True, fused=True)
    # a pool already holding the factors of a block
    f = C.batched_side(comp, A, tree, tf, 0, 40, fp, resident=True, fused=True)
    with pytest.raises(nca_amd.ArgumentError, match="factor already"):
        C._bottomup_fused(comp, A, tr, 0, 40, fp, f)
