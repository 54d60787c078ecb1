"""The batching workload at its full size on the GPU path: 64 independent
ellipsoid surfaces (semi-axes uniform in [0.5, 2], sizes uniform in [4,096,
65,536] Fibonacci-mapped points, seeded), Helmholtz k = 1, ACA tol 1e-4, all
compressed as one forest (one launch per forest level) and applied to 32
seeded complex Gaussian right-hand sides per problem in one captured sweep.

Checks: ranks, pivots and skeleton factors of every test cluster of the
forest bit for bit against the oracle (same forest, partition and Phase-1 far
pivots); for FP64 and MIXED, three hipGraph replays equal to the eager sweep;
MIXED within 1e-5 of FP64; FP64 within 5 tol of dense direct summation on
sampled rows of four problems."""
import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from oracle import oracle as O
from helpers import DeviceBuffers, oracle_factors, relerr, bitwise_equal

pytestmark = pytest.mark.gpu

P, R, K, TOL, MAXRANK = 64, 32, 1.0, 1e-4, 50


def say(msg):
    print(f"[gpu-fullsize-batch] {msg}", flush=True)


@pytest.fixture(scope="module")
def full():
    rng = np.random.default_rng(64)
    axes = rng.uniform(0.5, 2.0, (P, 3))
    sizes = rng.integers(4096, 65537, P)
    sets = [geometry.ellipsoid(int(n), a) for n, a in zip(sizes, axes)]
    say(f"assemble {P} problems, {int(sizes.sum())} points")
    trees = [nca_amd.TwoNTree(p, minvalues=64) for p in sets]
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(TOL))
    B = nca_amd.BatchedNCA(nca_amd.HelmholtzKernel(K), sets, sets, trees, trees, cmp, cmp, maxrank=MAXRANK,
                           ctx=nca_amd.default_context())
    X = np.asfortranarray(geometry.complex_gaussian(rng, (int(B.testoffsets[-1]), R)))
    return sets, B, X


def test_fullsize_batch_pivots_equal_oracle(full):
    sets, B, _ = full
    pts = np.concatenate(sets, axis=0)
    tree = B.tree
    fpt = nca_amd.farpivots(tree.testtree, tree.trialtree, B.core.testfardata, MAXRANK)
    say(f"oracle bases over the forest ({tree.testtree.nnodes} clusters, test side)")
    res = O.bases(0, K, 1, pts, pts, tree.testtree, fpt[0], fpt[1], MAXRANK, TOL)
    assert res[4] == 0
    fo = oracle_factors(res, tree.testtree, fpt)
    fg = B.core.testfactors
    nn = tree.testtree.nnodes
    assert np.array_equal(fg.assigned[:nn], fo.assigned[:nn])
    for t in np.nonzero(fo.assigned[:nn])[0]:
        assert fg.rank[t] == fo.rank[t], t
        assert np.array_equal(fg.pivots[t], fo.pivots[t]), t
        assert bitwise_equal(fg.X[t], fo.X[t]), t


def test_fullsize_batch_graph_mixed_direct_summation(full):
    sets, B, X = full
    N = X.shape[0]
    ctx = B.core.ctx
    out = {}
    dev = DeviceBuffers()
    try:
        dx, dy = dev.upload(X), dev.upload(np.zeros_like(X))
        for label, prec in (("fp64", nca_amd.NCA_PREC_FP64), ("mixed", nca_amd.NCA_PREC_MIXED)):
            say(f"{label}: eager sweep and three graph replays, {R} RHS")
            B.core.set_precision(prec)
            B.core.matvec_device(0, R, dx, N, dy, N)
            ctx.synchronize()
            eager = dev.download(dy, X.shape)
            g = B.core.graph(0, R, dx, N, dy, N)
            for rep in range(3):
                dev.copy_in(dy, np.zeros_like(X))
                g.launch()
                ctx.synchronize()
                assert bitwise_equal(dev.download(dy, X.shape), eager), (label, rep)
            del g
            out[label] = eager
    finally:
        dev.free()
    err = relerr(out["mixed"], out["fp64"])
    say(f"MIXED vs FP64 {err:.3e}")
    assert err <= 1e-5
    rng = np.random.default_rng(65)
    for p in (0, 21, 42, 63):
        rows = np.sort(rng.choice(sets[p].shape[0], 64, replace=False))
        xp = np.asfortranarray(X[B.trialoffsets[p]:B.trialoffsets[p + 1], :2])
        ref = O.dense_rows(K, 1, sets[p], sets[p], rows, xp)
        got = out["fp64"][B.testoffsets[p] + rows, :2]
        e = relerr(got, ref)
        say(f"problem {p}: {sets[p].shape[0]} points, vs direct summation {e:.3e}")
        assert e < 5 * TOL
