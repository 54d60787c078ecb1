"""Directional (wideband) far data: admissibility, direction sets, direction
tree, DirectionalData and dirfarfield. CPU only.

Restates the known answers and properties of
test/admissibility/test_nearinteractions.jl:54-111 and
test/admissibility/directionalsubdivision/{test_abstractdirections,
test_boundingballdirections,test_twondirections}.jl on this package's trees
(node ids, direction indices and local positions 0-based; direction labels as
in the reference, 0 = low frequency)."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import trees as T


def _unitvectors(dim, n=10000, seed=5):
    v = np.random.default_rng(seed).standard_normal((n, dim))
    return v / np.linalg.norm(v, axis=1)[:, None]


class _Ball:
    """One-node 'tree' of radius r centred at (d, 0, 0)."""

    def __init__(self, r, d):
        self.radius = np.array([float(r)])
        self.center = np.array([[float(d), 0.0, 0.0]])


def test_isnearwideband_known_answers():
    """test_nearinteractions.jl:74-111 (d = distance of the centres)."""
    def wb(k, etalf):
        return D.isnearwideband(k, etalf=etalf, etahf=5.0, gamma=1.0)

    def isnear(f, ra, rb, d):
        return f(_Ball(ra, 0.0), _Ball(rb, d), 0, 0)

    lf, hf = wb(0.1, 1), wb(40, 1)
    assert isnear(lf, 1, 1, 0) and isnear(hf, 1, 1, 0)
    assert isnear(lf, 1, 1, 3.9) and not isnear(lf, 1, 1, 4.1)
    assert not isnear(wb(0.1, 4), 1, 1, 3.9)
    assert not isnear(hf, 1, 1, 40) and isnear(wb(60, 1), 1, 1, 40)
    assert not isnear(hf, 1, 0.01, 4) and isnear(hf, 1, 1, 4)
    assert isnear(hf, 1, 0.01, 4) == isnear(hf, 0.01, 1, 4)
    g = D.isnearwideband(2.0)
    assert (g.etalf, g.etahf, g.islf.gamma, g.islf.k) == (1.0, 5.0, 1.0, 2.0)


@pytest.mark.parametrize("kind", ["twon", "kmeans"])
def test_wideband_partition(kind):
    """The host traversal with the wideband criterion: near and far blocks
    cover every pair once, every far pair is admissible, and the near pairs
    are leaves that are not."""
    pts = np.random.default_rng(3).random((300, 3))
    tr = nca_amd.TwoNTree(pts, minvalues=8) if kind == "twon" else nca_amd.KMeansTree(pts, 2, 8)
    isn = D.isnearwideband(2 * math.pi / 0.3)
    part = D.interactions(T.BlockTree(tr, tr), isn)
    cover = np.zeros((300, 300), np.int64)
    for t in range(tr.nnodes):
        for s in part.farnodes[part.farptr[t]:part.farptr[t + 1]]:
            assert not isn(tr, tr, t, s)
            cover[np.ix_(tr.values(t), tr.values(s))] += 1
        for s in part.nearnodes[part.nearptr[t]:part.nearptr[t + 1]]:
            assert tr.isleaf(t) and tr.isleaf(s) and isn(tr, tr, t, s)
            cover[np.ix_(tr.values(t), tr.values(s))] += 1
    assert np.all(cover == 1)


def test_finalize_directionaldata():
    """test_abstractdirections.jl:12-55: ragged per-node input, CSR out."""
    far = T.FarData(np.array([0, 3, 3, 5]), np.array([11, 12, 13, 21, 22]))
    data = D.finalize_directionaldata(far, [np.array([5, 7]), np.zeros(0, np.int64), np.array([9])],
                                      [np.array([2, 1]), np.zeros(0, np.int64), np.array([2])],
                                      [np.zeros(0, np.int64), np.zeros(0, np.int64), np.array([0, 1])],
                                      [0, 0, 2], "sentinel")
    assert data.nodetodirsptr.tolist() == [0, 2, 2, 3]
    assert data.dirs.tolist() == [5, 7, 9]
    assert data.diridxtofarptr.tolist() == [0, 2, 3, 5]
    assert data.nodetochdirptr.tolist() == [0, 0, 0, 2]
    assert data.partochddirmap.tolist() == [0, 1]
    assert data.isnear == "sentinel"
    assert [data.ndirections(n) for n in range(3)] == [2, 0, 1]
    assert [list(data.dirrange(n)) for n in range(3)] == [[0, 1], [], [2]]
    assert [data.dirsof(n).tolist() for n in range(3)] == [[5, 7], [], [9]]
    assert [data.nfarnodes(n) for n in range(3)] == [3, 0, 2]
    assert [data.farnodes(n).tolist() for n in range(3)] == [[11, 12, 13], [], [21, 22]]
    assert [data.dirmap(n).tolist() for n in range(3)] == [[], [], [0, 1]]
    assert data.dirfarnodes(0, 7).tolist() == [13]
    with pytest.raises(AssertionError):
        D.finalize_directionaldata(far, [np.array([5, 7]), np.zeros(0, np.int64), np.array([9])],
                                   [np.array([2, 2]), np.zeros(0, np.int64), np.array([2])],
                                   [np.zeros(0, np.int64), np.zeros(0, np.int64), np.array([0, 1])],
                                   [0, 0, 2], "sentinel")


@pytest.mark.parametrize("n", [6, 100, 1600])
def test_sphericalfibonaccipoints(n):
    """test_boundingballdirections.jl:17-31."""
    pts = D.sphericalfibonaccipoints(n)
    assert pts.shape == (n, 3)
    assert np.allclose(np.linalg.norm(pts, axis=1), 1.0)
    v = _unitvectors(3)
    cover = np.max(np.arccos(np.clip(np.max(v @ pts.T, axis=1), -1, 1)))
    assert cover < 3 / math.sqrt(n)
    d2 = np.sum((pts[:, None, :] - pts[None, :, :]) ** 2, axis=2)
    d2[np.diag_indices(n)] = np.inf
    assert math.sqrt(d2.min()) > 1.3 / math.sqrt(n)


@pytest.mark.parametrize("dim", [2, 3])
def test_directions(dim):
    """test_boundingballdirections.jl:33-52."""
    assert len(D.directions(dim, 1.0, 0.1, 1.0)) == 2 * dim
    v = _unitvectors(dim)
    counts = []
    for kd in (5.0, 10.0, 20.0):
        e = D.directions(dim, kd, 1.0, 1.0)
        counts.append(len(e))
        assert np.allclose(np.linalg.norm(e, axis=1), 1.0)
        worst = np.max(np.arccos(np.clip(np.max(v @ e.T, axis=1), -1, 1)))
        assert worst < 1.2 * math.asin(1.0 / kd)
    for i in range(2):
        assert counts[i + 1] / counts[i] == pytest.approx(2 ** (dim - 1), rel=0.05)


def test_nearest_direction():
    """test_boundingballdirections.jl:54-67."""
    e = np.eye(3)
    assert D.nearest_direction(e, [0.9, 0.1, 0.1]) == 0
    assert D.nearest_direction(e, [0.1, 0.9, 0.1]) == 1
    assert D.nearest_direction(e, [0.1, 0.1, 0.9]) == 2
    pool = D.sphericalfibonaccipoints(50)
    for v in _unitvectors(3, 500):
        assert D.nearest_direction(pool, v) == int(np.argmax([pool[i] @ v for i in range(50)]))


@pytest.mark.parametrize("dim", [2, 3])
def test_direction_tree(dim):
    """test_twondirections.jl:16-51."""
    maxlevel, branching = 3, 2 ** (dim - 1)
    dt = D.DirectionTree(maxlevel, 5, dim)
    nn = len(dt.dir) - 1
    assert nn == 2 * dim * (branching ** maxlevel - 1) // (branching - 1)
    assert dt.children(0).tolist() == list(range(1, 2 * dim + 1))
    roots = [np.eye(dim)[a] * s for a in range(dim) for s in (1.0, -1.0)]
    assert np.array_equal(dt.dir[1:2 * dim + 1], np.array(roots))
    assert np.allclose(np.linalg.norm(dt.dir[1:], axis=1), 1.0)
    assert all(dt.parent[r] == 0 for r in range(1, 2 * dim + 1))
    for n in range(1, nn + 1):
        for c in dt.children(n):
            assert dt.parent[c] == n
    leaves = [n for n in range(1, nn + 1) if dt.nchildren[n] == 0]
    assert len(leaves) == 2 * dim * branching ** (maxlevel - 1)
    with pytest.raises(ValueError):
        D.DirectionTree(0, 1)


@pytest.mark.parametrize("dim,cones", [(2, (0.79, 0.47, 0.25, 0.13)), (3, (0.96, 0.62, 0.34, 0.18))])
def test_direction_picks(dim, cones):
    """test_twondirections.jl:60-99 (levels 0-based: the reference's level - 1)."""
    dirroot, maxlevel = 5, 4
    branching = 2 ** (dim - 1)
    dt = D.DirectionTree(maxlevel, dirroot, dim)
    v = _unitvectors(dim)
    roots = np.arange(1, 2 * dim + 1)
    assert np.array_equal(dt.direction(dt.dir[roots], dirroot), roots)
    assert np.all(dt.direction(v, dirroot + 1) == 0)
    assert np.array_equal(dt.direction(v, dirroot - maxlevel - 3), dt.direction(v, dirroot - maxlevel + 1))
    depth_of = np.zeros(len(dt.dir), np.int64)
    for n in range(1, len(dt.dir)):
        depth_of[n] = 1 if dt.parent[n] == 0 else depth_of[dt.parent[n]] + 1
    prev = None
    for depth in range(1, maxlevel + 1):
        level = dirroot - depth + 1
        picks = dt.direction(v, level)
        assert np.all(depth_of[picks] == depth)
        assert len(np.unique(picks)) == 2 * dim * branching ** (depth - 1)
        if prev is not None:
            assert np.array_equal(dt.parent[picks], prev)
        ang = np.arccos(np.clip(np.sum(v * dt.dir[picks], axis=1), -1, 1))
        assert ang.max() <= cones[depth - 1]
        prev = picks


def _twon200():
    pts = np.random.default_rng(3).random((200, 3))
    return nca_amd.TwoNTree(pts, minvalues=2, minhalfsize=1 / 64)


def test_dirroot_and_maxlevel():
    """test_twondirections.jl:108-141: a far pointer with one far node on each
    of the given levels (0-based here)."""
    tr = _twon200()
    assert (tr.level.min(), tr.level.max()) == (0, 4)

    def ptron(levels):
        nodes = [int(np.nonzero(tr.level == lev)[0][0]) for lev in levels]
        has = np.array([1 if n in nodes else 0 for n in range(tr.nnodes)])
        return np.concatenate([[0], np.cumsum(has)])

    def dm(levels, lam):
        return D._dirroot_and_maxlevel(tr, ptron(levels), D.IsLowFrequency(1.0, 2 * math.pi / lam))

    islf = D.IsLowFrequency(1.0, 2 * math.pi)
    assert not islf.level(tr, 3) and islf.level(tr, 4)
    assert dm([], 1.0) == (3, 1)
    assert dm([3], 1.0) == (3, 1)
    assert dm([2], 1.0) == (3, 2)
    assert dm([1, 3], 1.0) == (3, 3)
    assert dm([4], 1.0) == (3, 1)
    assert dm([2, 4], 1.0) == (3, 2)
    assert dm([2], 50.0) == (-1, 1)
    assert dm([2], 0.2) == (5, 4)


def _cases():
    pts = np.random.default_rng(3).random((200, 3))
    isn = D.isnearwideband(2 * math.pi / 1.5, gamma=1.0)
    out = []
    for name, tr in (("TwoNTree", nca_amd.TwoNTree(pts, minvalues=2, minhalfsize=1 / 64)),
                     ("BoundingBallTree", nca_amd.KMeansTree(pts, 2, 4))):
        td, sd = D.fardata(T.BlockTree(tr, tr), isn)
        out.append((name, tr, td, sd, isn))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fardata_directional(case):
    """test_abstractdirections.jl:87-103, 105-124, 166-183."""
    _, tr, td, sd, isn = case
    nn = tr.nnodes
    tp = {(t, int(s)) for t in range(nn) for s in td.farnodes(t)}
    sp = {(int(t), s) for s in range(nn) for t in sd.farnodes(s)}
    assert tp and tp == sp
    alldirs = [td.dirsof(n).tolist() for n in range(nn)]
    assert any(d == [0] for d in alldirs) and any(d and d != [0] for d in alldirs)
    for n in range(nn):
        split = []
        for d in td.dirsof(n):
            split.extend(td.dirfarnodes(n, int(d)).tolist())
        assert split == td.farnodes(n).tolist()
        first = int(td.farptr()[n])
        for i in range(td.nfarnodes(n)):
            e = td.diridxfromlocalfaridx(n, i)
            assert e in td.dirrange(n) and td.diridxtofarptr[e] <= first + i < td.diridxtofarptr[e + 1]
    node = next(n for n in range(nn) if td.ndirections(n) > 1)
    with pytest.raises(KeyError):
        td.dirfarnodes(node, int(td.dirsof(node).max()) + 1)
    with pytest.raises(IndexError):
        td.diridxfromlocalfaridx(0, td.nfarnodes(0))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_dirfarfield(case):
    """test_abstractdirections.jl:126-164."""
    _, tr, td, _, _ = case
    ptr, idx = D.dirfarfield_all(tr, td)
    for n in range(tr.nnodes):
        chain = [n]
        while tr.parent[chain[-1]] >= 0:
            chain.append(int(tr.parent[chain[-1]]))
        chainfar = set()
        for a in chain:
            chainfar |= set(td.farnodes(a).tolist())
        for loc, d in enumerate(td.dirsof(n).tolist()):
            e = int(td.nodetodirsptr[n]) + loc
            own = td.dirfarnodes(n, d).tolist()
            field = idx[ptr[e]:ptr[e + 1]].tolist()
            assert field == D.dirfarfield(tr, td, n, d).tolist()
            assert field[:len(own)] == own
            assert len(set(field)) == len(field)
            assert set(field) <= chainfar
            if d != 0:
                continue
            want = []
            for a in chain:
                if 0 not in td.dirsof(a).tolist():
                    break
                want.extend(td.dirfarnodes(a, 0).tolist())
            assert field == want


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_isbasisnode(case):
    """test_abstractdirections.jl:185-219."""
    _, tr, td, _, _ = case
    nn = tr.nnodes
    leaf = next(n for n in range(nn) if tr.isleaf(n) and td.ndirections(n) > 0)
    inner = next(n for n in range(nn) if not tr.isleaf(n) and td.dirsof(n).tolist() == [0])
    assert all(D.isbasisnode(tr, td, leaf, int(d)) for d in td.dirsof(leaf))
    assert not D.isbasisnode(tr, td, inner, 0)
    ptr, idx = D.dirfarfield_all(tr, td)
    for t in range(nn):
        for loc, e in enumerate(td.dirrange(t)):
            d = int(td.dirs[e])
            if ptr[e + 1] == ptr[e] or D.isbasisnode(tr, td, t, d):
                continue
            for c in tr.childrenof(t):
                assert 0 <= td.dirmap(c)[loc] < td.ndirections(c)


@pytest.mark.parametrize("lam,haslf,hashf,mixed", [(50.0, True, False, False), (0.02, False, True, False),
                                                   (1.5, True, True, True)])
def test_bounding_ball_regimes(lam, haslf, hashf, mixed):
    """test_boundingballdirections.jl:69-109."""
    pts = np.random.default_rng(3).random((200, 3))
    tr = nca_amd.KMeansTree(pts, 2, 4)
    isn = D.isnearwideband(2 * math.pi / lam, gamma=1.0)
    data, _ = D.fardata(T.BlockTree(tr, tr), isn)
    nn = tr.nnodes
    alldirs = [data.dirsof(n).tolist() for n in range(nn)]
    lf = [isn.islf.node(tr, n) for n in range(nn)]
    par = tr.parent
    assert any(d == [0] for d in alldirs) == haslf
    assert any(d and d != [0] for d in alldirs) == hashf
    assert all((alldirs[n] == [0]) == lf[n] for n in range(nn) if alldirs[n])
    assert all((not lf[par[n]]) or lf[n] for n in range(nn) if par[n] >= 0)
    boundary = [n for n in range(nn) if par[n] >= 0 and lf[n] and not lf[par[n]]]
    assert bool(boundary) == mixed
    assert all((alldirs[n] == [0]) == (data.nfarnodes(n) > 0) for n in boundary)
    # Below it, an LF node under an LF parent with [0] inherits [0] whether or
    # not it has far nodes of its own (on this stand-in tree all of them have).
    below = [n for n in range(nn) if par[n] >= 0 and lf[n] and lf[par[n]] and alldirs[par[n]] == [0]]
    assert bool(below) or not mixed
    assert all(alldirs[n] == [0] and data.dirmap(n).tolist() == [0] for n in below)


@pytest.mark.parametrize("gamma,ownbound,inheritedbound", [(1.0, 1.18, 2.18), (0.5, 0.6, 1.14)])
def test_directional_admissibility(gamma, ownbound, inheritedbound):
    """test_boundingballdirections.jl:111-151: sin of the angle between the
    interaction vector and the direction, relative to the aperture 1/(k diam).
    The reference measures an icosphere mesh; here a 2,562-point Fibonacci
    sphere (the vertex count of a 4-times refined icosahedron)."""
    pts = geometry.fibonacci_sphere(2562)
    tr = nca_amd.KMeansTree(pts, 2, 50)
    k = 2 * math.pi / 0.25
    isn = D.isnearwideband(k, gamma=gamma)
    data, _ = D.fardata(T.BlockTree(tr, tr), isn)
    assert not any(isn.islf.node(tr, n) for n in range(tr.nnodes))
    ptr, idx = D.dirfarfield_all(tr, data)

    def sinangle(v, c):
        if v @ c <= 0:
            return 1.0
        return np.linalg.norm(np.cross(v, c)) / (np.linalg.norm(v) * np.linalg.norm(c))

    worst_own = worst_inh = 0.0
    for n in range(tr.nnodes):
        diam = 2 * tr.radius[n]
        ev = D.directions(3, diam, k, gamma)
        for loc, e in enumerate(data.dirrange(n)):
            d = int(data.dirs[e])
            own = int(data.diridxtofarptr[e + 1] - data.diridxtofarptr[e])
            for i, f in enumerate(idx[ptr[e]:ptr[e + 1]]):
                out = sinangle(D.interaction(n, f, tr, tr), ev[d - 1]) * k * diam
                if i < own:
                    worst_own = max(worst_own, out)
                else:
                    worst_inh = max(worst_inh, out)
    assert worst_own <= ownbound
    assert worst_inh <= inheritedbound


@pytest.mark.parametrize("kind,lam", [("twon", 1.5), ("twon", 0.2), ("kmeans", 1.5), ("kmeans", 0.02),
                                      ("kmeans", 50.0)])
def test_library_directional_data_equals_numpy(kind, lam):
    """The host library's subdivision equals the numpy statement of the same
    algorithm: direction labels, counts, maps and the far-node order."""
    pts = np.random.default_rng(3).random((400, 3))
    tr = nca_amd.TwoNTree(pts, minvalues=4) if kind == "twon" else nca_amd.KMeansTree(pts, 2, 4)
    tree = T.BlockTree(tr, tr)
    isn = D.isnearwideband(2 * math.pi / lam, gamma=1.0)
    part = D.interactions(tree, isn)
    tf, _ = T.fardata(tree, part)
    a = T.FarData(tf.nodetofarnodesptr.copy(), tf.farnodes.copy())
    b = T.FarData(tf.nodetofarnodesptr.copy(), tf.farnodes.copy())
    da = D.directionalfardata(a, tr, tr, isn)
    db = D.directionalfardata_reference(b, tr, tr, isn)
    for name in ("nodetodirsptr", "dirs", "diridxtofarptr", "nodetochdirptr", "partochddirmap"):
        assert np.array_equal(getattr(da, name), getattr(db, name)), name
    assert np.array_equal(a.farnodes, b.farnodes)
