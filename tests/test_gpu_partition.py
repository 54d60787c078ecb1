"""The admissibility partition on the device (nca_interactions) equals the
host traversal's: same near and far CSR, rows sorted — standard criterion on
one and on two trees, and the wideband criterion on both tree kinds."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import trees as T

pytestmark = pytest.mark.gpu


def same(p, q):
    for n in ("nearptr", "nearnodes", "farptr", "farnodes"):
        assert np.array_equal(getattr(p, n), getattr(q, n)), n


@pytest.mark.parametrize("case", ["sphere_twon", "grid_kmeans", "ellipsoid_kmeans"])
def test_device_partition_standard(case):
    ctx = nca_amd.default_context()
    if case == "sphere_twon":
        pts = geometry.fibonacci_sphere(50000)
        tt = st = nca_amd.TwoNTree(pts, minvalues=32)
    elif case == "grid_kmeans":
        tp = geometry.grid_square(64)
        tt, st = nca_amd.KMeansTree(tp, 2, 32), nca_amd.KMeansTree(tp + np.array([1.1, 0.0, 0.2]), 2, 32)
    else:
        tt = nca_amd.KMeansTree(geometry.ellipsoid(20000, (1.5, 0.7, 1.1)), 2, 40)
        st = nca_amd.KMeansTree(geometry.ellipsoid(7000, (0.5, 0.9, 1.3)) + 0.2, 2, 40)
    tree = T.BlockTree(tt, st)
    for eta in (1.0, 0.5):
        same(T.interactions_device(tree, ctx, eta), T.interactions(tree, eta))


@pytest.mark.parametrize("kind,n,lam", [("kmeans", 20000, 0.3), ("twon", 20000, 0.3), ("kmeans", 5000, 3.0)])
def test_device_partition_wideband(kind, n, lam):
    ctx = nca_amd.default_context()
    pts = geometry.fibonacci_sphere(n)
    tr = nca_amd.TwoNTree(pts, minvalues=32) if kind == "twon" else nca_amd.KMeansTree(pts, 2, 32)
    tree = T.BlockTree(tr, tr)
    isn = D.isnearwideband(2 * math.pi / lam)
    same(T.interactions_device(tree, ctx, wideband=isn), D.interactions(tree, isn))
