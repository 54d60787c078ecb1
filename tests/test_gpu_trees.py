"""Cluster trees built on the device (nca_tree_twon, nca_tree_kmeans) equal the
host builders' trees in every field: parent, level, value ranges, child CSR,
permutation, centers and radii (bit for bit) — on spheres up to the 1M-point
headline sphere, a regular grid (points on box midplanes), random and
clustered points with coincident duplicates (the 1e-12 half-size rule), tiny
sets, minhalfsize > 0 and 2, 3 and 8 clusters per k-means node."""
import numpy as np
import pytest

import nca_amd
from nca_amd import geometry

pytestmark = pytest.mark.gpu

FIELDS = ("parent", "level", "lo", "hi", "childptr", "children", "perm", "center", "radius")


def say(msg):
    print(f"[gpu-trees] {msg}", flush=True)


def same(a, b):
    assert a.nnodes == b.nnodes
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape, f
        assert np.array_equal(x, y), f


def pointsets():
    rng = np.random.default_rng(31)
    dup = rng.random((300, 3))
    return {
        "sphere20k": geometry.fibonacci_sphere(20000),
        "grid64": geometry.grid_square(64),
        "cube5k": rng.random((5000, 3)),
        "duplicates": np.concatenate([dup, np.repeat(dup[:3], 150, axis=0), np.zeros((90, 3))]),
        "one": np.array([[0.3, -0.2, 0.1]]),
        "two": np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]),
    }


@pytest.mark.parametrize("name", list(pointsets()))
@pytest.mark.parametrize("minvalues,minhalfsize", [(32, 0.0), (1, 0.0), (64, 0.05)])
def test_device_twon_equals_host(name, minvalues, minhalfsize):
    ctx = nca_amd.default_context()
    pts = pointsets()[name]
    say(f"twon {name} n={len(pts)} minvalues={minvalues} minhalfsize={minhalfsize}")
    same(nca_amd.TwoNTree(pts, minvalues, minhalfsize, ctx=ctx), nca_amd.TwoNTree(pts, minvalues, minhalfsize))


@pytest.mark.parametrize("name", list(pointsets()))
@pytest.mark.parametrize("k,minvalues", [(2, 32), (3, 20), (8, 64), (2, 1)])
def test_device_kmeans_equals_host(name, k, minvalues):
    ctx = nca_amd.default_context()
    pts = pointsets()[name]
    say(f"kmeans {name} n={len(pts)} k={k} minvalues={minvalues}")
    same(nca_amd.KMeansTree(pts, k, minvalues, ctx=ctx), nca_amd.KMeansTree(pts, k, minvalues))


def test_device_trees_headline_sizes():
    """The 1M-point sphere's octree (minvalues 64) and the 262k-point sphere's
    2-means tree (minvalues 64), the trees of the two full-size workloads."""
    ctx = nca_amd.default_context()
    pts = geometry.fibonacci_sphere(1048576)
    say("twon sphere 1M")
    same(nca_amd.TwoNTree(pts, 64, ctx=ctx), nca_amd.TwoNTree(pts, 64))
    pts = geometry.fibonacci_sphere(262144)
    say("kmeans sphere 262k")
    same(nca_amd.KMeansTree(pts, 2, 64, ctx=ctx), nca_amd.KMeansTree(pts, 2, 64))


def test_device_tree_rejects_too_many_clusters():
    ctx = nca_amd.default_context()
    with pytest.raises(Exception, match="clusters"):
        nca_amd.KMeansTree(geometry.fibonacci_sphere(100), 9, 10, ctx=ctx)
