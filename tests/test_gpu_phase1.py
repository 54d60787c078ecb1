"""Phase 1 on the device (nca_farpivots, and nca_farpivots_tree, which builds
the far fields itself) against the host selection (nca_tree.cpp): the same far
pivots, bit for bit, for weight 1 — plain clusters (own + ancestors' far field)
of one and of two trees, and (cluster, direction) items over their
dirfarfield."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import trees as T
from nca_amd import wideband as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-phase1] {msg}", flush=True)


@pytest.mark.parametrize("case", ["sphere", "grid"])
def test_device_phase1_equals_host(case):
    say(f"plain {case}")
    if case == "sphere":
        tp = sp = geometry.fibonacci_sphere(20000)
        tt = st = nca_amd.TwoNTree(tp, minvalues=32)
    else:
        tp = geometry.grid_square(64)
        sp = tp + np.array([1.1, 0.0, 0.2])
        tt, st = nca_amd.KMeansTree(tp, 2, 32), nca_amd.KMeansTree(sp, 2, 32)
    tree = nca_amd.BlockTree(tt, st)
    part = nca_amd.interactions(tree)
    tf, sf = nca_amd.fardata(tree, part)
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(1.0), tp, sp)
    sel = C.GeometricPivoting()
    for selftree, othertree, fd, pts in ((tt, st, tf, A.trial), (st, tt, sf, A.test)):
        for maxrank, over in ((50, 2.0), (30, 1.0)):
            host = T.farpivots(selftree, othertree, fd, maxrank, 1.0, over)
            ff = T.farfield_csr(selftree, fd)
            for t in range(0, selftree.nnodes, 97):
                assert np.array_equal(ff[1][ff[0][t]:ff[0][t + 1]], T.farfield(selftree, fd, t))
            dev = C.device_farpivots(pts, selftree.center, selftree.hi - selftree.lo, ff[0], ff[1], othertree,
                                     maxrank, 1.0, over)
            assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
            # far fields built on the device from the FarData CSR and the parents
            devt = C.device_farpivots_tree(pts, selftree, fd, othertree, maxrank, 1.0, over)
            assert np.array_equal(host[0], devt[0]) and np.array_equal(host[1], devt[1])
    # the selector routes to the device when it gets the other side's points
    got = sel.select(tt, st, tf, 50, A.trial)
    want = T.farpivots(tt, st, tf, 50)
    assert np.array_equal(got[1], want[1])


def test_device_phase1_directional():
    say("directional")
    pts = geometry.fibonacci_sphere(3000)
    tr = nca_amd.KMeansTree(pts, 2, 16)
    isn = D.isnearwideband(2 * math.pi / 1.5, etalf=1.0, etahf=1.0, gamma=1.0)
    td, _ = D.fardata(T.BlockTree(tr, tr), isn)
    ff = D.dirfarfield_all(tr, td)
    blocks = W.DirectionBlocks(tr, td, ff[0])
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(1.0), pts, pts)
    host = W.farpivots(blocks, tr, ff[0], ff[1], 50, C.GeometricPivoting(device=False))
    dev = W.farpivots(blocks, tr, ff[0], ff[1], 50, C.GeometricPivoting(), A.trial)
    assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])


def test_device_phase1_default_weight_bitwise():
    """The default selector (weight 1.5, the reference's count rule): the
    device computes d^1.5 as d sqrt(d), as the host does, so the pivots agree
    bit for bit, with the far fields built on the device or given."""
    pts = geometry.fibonacci_sphere(6000)
    tr = nca_amd.TwoNTree(pts, minvalues=32)
    tree = nca_amd.BlockTree(tr, tr)
    tf, _ = nca_amd.fardata(tree, nca_amd.interactions(tree))
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(1.0), pts, pts)
    sel = C.GeometricPivoting()
    assert sel.weight == 1.5 and sel.oversample == 1.0
    host = T.farpivots(tr, tr, tf, 50)
    dev = sel.select(tr, tr, tf, 50, A.trial)
    assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
    ff = T.farfield_csr(tr, tf)
    devc = C.device_farpivots(A.trial, tr.center, tr.hi - tr.lo, ff[0], ff[1], tr, 50, 1.5, 1.0)
    assert np.array_equal(devc[0], host[0]) and np.array_equal(devc[1], host[1])
    # the count rule: min(maxrank, |t|) capped by the far field's points
    nfar = np.array([sum(tr.nvalues(f) for f in ff[1][ff[0][t]:ff[0][t + 1]]) for t in range(tr.nnodes)])
    assert np.array_equal(np.diff(host[0]), np.minimum(np.minimum(50, tr.hi - tr.lo), nfar))


def test_device_phase1_weight():
    """weight != 1 goes through pow on both sides: same counts, pivots from the far field."""
    pts = geometry.fibonacci_sphere(6000)
    tr = nca_amd.TwoNTree(pts, minvalues=32)
    tree = nca_amd.BlockTree(tr, tr)
    tf, _ = nca_amd.fardata(tree, nca_amd.interactions(tree))
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(1.0), pts, pts)
    host = T.farpivots(tr, tr, tf, 50, 1.3, 2.0)
    ff = T.farfield_csr(tr, tf)
    dev = C.device_farpivots(A.trial, tr.center, tr.hi - tr.lo, ff[0], ff[1], tr, 50, 1.3, 2.0)
    assert np.array_equal(host[0], dev[0])
    same = np.mean(host[1] == dev[1])
    say(f"  weight 1.3: {100 * same:.2f} % of the pivots identical")
    devt = C.device_farpivots_tree(A.trial, tr, tf, tr, 50, 1.3, 2.0)
    assert np.array_equal(dev[0], devt[0]) and np.array_equal(dev[1], devt[1])
    for t in range(0, tr.nnodes, 37):
        farpts = set(np.concatenate([tr.values(f) for f in ff[1][ff[0][t]:ff[0][t + 1]]] or [[]]).tolist())
        assert set(dev[1][dev[0][t]:dev[0][t + 1]].tolist()) <= farpts


def test_device_phase1_wave_and_block_items():
    """Explicit far fields of very different sizes in one call: pools of at
    most 1,024 candidates take the one-wave kernel, larger ones (more than
    1,024 far nodes, one candidate each) the workgroup kernel; every item's
    pivots equal the host selection bit for bit (weights 1.5 and 1)."""
    from types import SimpleNamespace
    pts = geometry.fibonacci_sphere(60000)
    tr = nca_amd.TwoNTree(pts, minvalues=16)
    leaves = np.array([t for t in range(tr.nnodes) if tr.isleaf(t)])
    assert len(leaves) > 4100
    rng = np.random.default_rng(7)
    big = np.array([t for t in range(1, tr.nnodes) if 60 <= len(tr.values(t)) <= 400])
    items = rng.choice(big, 8, replace=False)
    sizes = [3, 100, 900, 1024, 1100, 2000, 4096, 40]
    lists = [np.sort(rng.choice(leaves, nf, replace=False)) for nf in sizes]
    ffptr = np.zeros(len(items) + 1, np.int64)
    ffptr[1:] = np.cumsum([len(v) for v in lists])
    ffnodes = np.concatenate(lists).astype(np.int64)
    blocks = SimpleNamespace(tree=tr, node=np.asarray(items, np.int64), nnodes=len(items))
    A = nca_amd.HIPKernelMatrix(nca_amd.HelmholtzKernel(1.0), pts, pts)
    say("wave and block items")
    for weight in (1.5, 1.0):
        for maxrank in (50, 12):
            host = W.farpivots(blocks, tr, ffptr, ffnodes, maxrank, C.GeometricPivoting(weight, device=False))
            dev = W.farpivots(blocks, tr, ffptr, ffnodes, maxrank, C.GeometricPivoting(weight), A.trial)
            assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1]), (weight, maxrank)
