"""Host-side logic (trees, partition, far data, Phase 1, plans) and the C-ABI
library's exports. No GPU compute."""
import ctypes
import os
import re

import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from nca_amd import nca as NC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "nca_hip.h")).read()
    names = set(re.findall(r"\b(nca_[a-z0-9_]+)\s*\(", header))
    assert len(names) >= 25
    lib = ctypes.CDLL(nca_amd._native.HIP_LIB)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(nca_amd._abi.SIGNATURES) == names


def test_missing_library_fails_loudly(monkeypatch):
    from nca_amd import _native
    monkeypatch.setattr(_native, "_hip", None)
    monkeypatch.setattr(_native, "HIP_LIB", "/nonexistent/libnca_hip.so")
    with pytest.raises(_native.NativeLibraryMissing):
        _native.hip_lib()


@pytest.mark.parametrize("kind", ["twon", "kmeans"])
def test_tree_structure(kind):
    pts = geometry.fibonacci_sphere(3000)
    t = nca_amd.TwoNTree(pts, minvalues=40) if kind == "twon" else nca_amd.KMeansTree(pts, 2, 40)
    assert sorted(t.perm.tolist()) == list(range(3000))
    assert t.lo[0] == 0 and t.hi[0] == 3000 and t.parent[0] == -1
    cover = np.zeros(3000, np.int64)
    for n in range(t.nnodes):
        ch = t.childrenof(n)
        if len(ch):
            # children partition the parent, in order
            assert t.lo[ch[0]] == t.lo[n] and t.hi[ch[-1]] == t.hi[n]
            assert np.all(t.lo[ch[1:]] == t.hi[ch[:-1]])
            assert np.all(t.level[ch] == t.level[n] + 1)
        else:
            assert t.nvalues(n) <= 40 or kind == "twon"
            cover[t.lo[n]:t.hi[n]] += 1
    assert np.all(cover == 1)
    # level order: ids increase with the level
    assert np.all(np.diff(t.level) >= 0)


def test_partition_and_fardata():
    tp = geometry.grid_square(32)
    sp = tp + np.array([1.1, 0.0, 0.2])
    tree = nca_amd.BlockTree(nca_amd.KMeansTree(tp, 2, 32), nca_amd.KMeansTree(sp, 2, 32))
    part = nca_amd.interactions(tree)
    tt, st = tree.testtree, tree.trialtree
    cover = np.zeros((1024, 1024), np.int64)
    for t in range(tt.nnodes):
        for s in part.farnodes[part.farptr[t]:part.farptr[t + 1]]:
            cover[np.ix_(tt.values(t), st.values(s))] += 1
        for s in part.nearnodes[part.nearptr[t]:part.nearptr[t + 1]]:
            assert tt.isleaf(t) and st.isleaf(s)
            cover[np.ix_(tt.values(t), st.values(s))] += 1
    assert np.all(cover == 1)
    tf, sf = nca_amd.fardata(tree, part)
    pairs = {(t, s) for t in range(tt.nnodes) for s in tf.farnodesof(t)}
    tpairs = {(t, s) for s in range(st.nnodes) for t in sf.farnodesof(s)}
    assert pairs == tpairs
    for s in range(st.nnodes):
        assert list(sf.farnodesof(s)) == sorted(sf.farnodesof(s))


def test_farfield_and_phase1_pivots():
    pts = geometry.fibonacci_sphere(4096)
    tr = nca_amd.TwoNTree(pts, minvalues=64)
    tree = nca_amd.BlockTree(tr, tr)
    part = nca_amd.interactions(tree)
    tf, _ = nca_amd.fardata(tree, part)
    # the default (oversample 1) is the reference's count rule _selectfarpivots/_clustersize
    p1, i1 = nca_amd.farpivots(tr, tr, tf, 50)
    assert np.array_equal(nca_amd.farpivots(tr, tr, tf, 50)[1], i1)  # deterministic
    # oversample 2: the opt-in that draws twice as many far pivots
    ptr, idx = nca_amd.farpivots(tr, tr, tf, 50, oversample=2.0)
    for t in range(tr.nnodes):
        F = nca_amd.farfield(tr, tf, t)
        own = tf.farnodesof(t)
        assert list(F[:len(own)]) == list(own)
        piv = idx[ptr[t]:ptr[t + 1]]
        piv1 = i1[p1[t]:p1[t + 1]]
        if len(F) == 0:
            assert len(piv) == 0 and len(piv1) == 0
            continue
        farpts = np.concatenate([tr.values(f) for f in F])
        # _selectfarpivots: min(maxrank, |t|, far points), distinct, from the far field
        assert len(piv1) == min(50, tr.nvalues(t), len(farpts))
        assert len(piv) == min(50, 2 * tr.nvalues(t), len(farpts))
        for pv in (piv, piv1):
            assert len(set(pv.tolist())) == len(pv)
            assert set(pv.tolist()) <= set(farpts.tolist())


def test_plan_from_pivots():
    """test/nca/test_abstractnca.jl:9-24."""
    f = C.SideFactors(4)
    f.assigned[:] = True
    f.rank[:] = [2, 3, 0, 1]
    assert list(NC.coefficient_ptr(f) + 1) == [1, 3, 6, 6, 7]
    g = C.SideFactors(4)
    g.assigned[1] = g.assigned[3] = True
    g.rank[1], g.rank[3] = 2, 0
    assert list(NC.coefficient_ptr(g) + 1) == [1, 1, 3, 3, 3]
    assert list(NC.coefficient_ptr(C.SideFactors(0)) + 1) == [1]


def test_concat_ranges():
    """ptrfromcounts-style CSR helpers (test/utils/test_utils.jl:6-20)."""
    assert list(C.concat_ranges([5, 0, 10, 3], [2, 0, 3, 1])) == [5, 6, 10, 11, 12, 3]
    assert len(C.concat_ranges([], [])) == 0
    assert list(C.concat_ranges([7], [0])) == []


def test_batched_bottom_up_constructor():
    with pytest.raises(TypeError):
        C.BatchedBottomUp(pivotselector=C.GeometricPivoting())
    with pytest.raises(TypeError):
        C.BatchedBottomUp(compressor=C.ACA(1e-3))
    b = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-3))
    assert b.compressor.tol == 1e-3


def test_forest_is_block_diagonal():
    """A forest of independent trees: nested ranges, level-by-level ids, and
    near + far blocks (with their ancestors) cover every point pair of each
    problem exactly once and no pair across problems."""
    from nca_amd import trees as T
    sets = [geometry.fibonacci_sphere(700), geometry.ellipsoid(500, (1.5, 0.8, 1.1)),
            geometry.ellipsoid(300, (0.6, 1.3, 0.9))]
    trees = [nca_amd.TwoNTree(p, minvalues=32) for p in sets]
    f, maps, off = T.forest(trees)
    assert f.nnodes == 1 + sum(t.nnodes for t in trees)
    assert np.all(np.diff(f.level) >= 0)
    for n in range(f.nnodes):
        ch = f.childrenof(n)
        if len(ch):
            assert f.lo[ch].min() == f.lo[n] and f.hi[ch].max() == f.hi[n]
            assert np.all(f.parent[ch] == n)
    parts = [T.interactions(T.BlockTree(t, t)) for t in trees]
    P = T.forest_partition(maps, maps, parts, f.nnodes)
    n = int(off[-1])
    cover = np.zeros((n, n), np.int8)
    for t in range(f.nnodes):
        for ptr, nodes in ((P.nearptr, P.nearnodes), (P.farptr, P.farnodes)):
            for s in nodes[ptr[t]:ptr[t + 1]]:
                cover[np.ix_(f.perm[f.lo[t]:f.hi[t]], f.perm[f.lo[s]:f.hi[s]])] += 1
    blk = np.zeros((n, n), np.int8)
    for p in range(len(sets)):
        blk[off[p]:off[p + 1], off[p]:off[p + 1]] = 1
    assert np.array_equal(cover, blk)


def test_flat_factors_match_per_node_path():
    """The batched driver's flat factors (LevelFactors) give the same rows per
    level, pivots and packed bases/transfers as the per-node construction."""
    from helpers import sphere_problem, oracle_assembly
    p = sphere_problem(4096, tol=1e-4)
    ft, fs = oracle_assembly(p)
    for f, tree, side in ((ft, p.tree.testtree, 0), (fs, p.tree.trialtree, 1)):
        g = C.LevelFactors(f.nn)
        leaf = tree.leafmask()
        for lev in sorted(tree.levels().tolist(), reverse=True):
            nodes = np.array([t for t in tree.levelnodes(lev) if f.assigned[t]], np.int64)
            if len(nodes) == 0:
                continue
            # rows of the level from the flat factors of the deeper levels
            rowptr, rowidx = C._level_rows(tree, g, nodes, leaf)
            for i, t in enumerate(nodes):
                want = tree.perm[tree.lo[t]:tree.hi[t]] if leaf[t] else np.concatenate(
                    [f.pivots[c] for c in tree.childrenof(t) if f.assigned[c]] or [np.zeros(0, np.int64)])
                assert np.array_equal(rowidx[rowptr[i]:rowptr[i + 1]], want)
            ranks = f.rank[nodes]
            X = [f.X[t].ravel(order="F") for t in nodes]
            xoff = np.zeros(len(nodes) + 1, np.int64)
            np.cumsum([len(x) for x in X], out=xoff[1:])
            g.add_level(nodes, f.m[nodes], ranks, np.concatenate([f.pivots[t] for t in nodes]), xoff,
                        np.concatenate(X))
            g.level_transfer_nodes[lev] = [int(t) for t in tree.levelnodes(lev) if not leaf[t] and f.assigned[t]]
        for t in range(f.nn):
            if f.assigned[t]:
                assert np.array_equal(g.pivots[t], f.pivots[t]) and np.array_equal(g.X[t], f.X[t])
        for a, b in zip(g.pivot_csr(), f.pivot_csr()):
            assert np.array_equal(a, b)
        f.level_transfer_nodes = {k: v for k, v in g.level_transfer_nodes.items()}
        b1, t1 = C.stores_from_factors(f, tree, side)
        b2, t2 = C.stores_from_factors(g, tree, side)
        assert np.array_equal(b1.nodes, b2.nodes) and np.array_equal(b1.offsets, b2.offsets)
        assert np.array_equal(b1.blocks, b2.blocks)
        for n in ("leveltoidcsptr", "levelidcs", "idxtochdsptr", "chdsidcs", "offsets", "blocks"):
            assert np.array_equal(getattr(t1, n), getattr(t2, n)), n


def test_julia_extension_binds_declared_symbols():
    """The Julia package extension (nestedcrossapproximation.jl_amd/julia/ext)
    calls only entry points that include/nca_hip.h declares and the library
    exports, and its descriptor structs list the C structs' fields in order."""
    import glob
    import re
    root = os.path.join(ROOT, "nestedcrossapproximation.jl_amd", "julia", "ext", "NestedCrossApproximationHIPExt")
    files = sorted(glob.glob(os.path.join(root, "*.jl")))
    assert len(files) >= 6
    header = open(os.path.join(ROOT, "include", "nca_hip.h")).read()
    declared = set(re.findall(r"\b(nca_[a-z0-9_]+)\(", header))
    used = set()
    for f in files:
        used |= set(re.findall(r"ccall\(\(:(nca_[a-z0-9_]+),", open(f).read()))
    assert used and used <= declared, sorted(used - declared)
    lib = nca_amd.hip_lib()
    for name in used:
        assert hasattr(lib, name), name
    main = open(os.path.join(root, "NestedCrossApproximationHIPExt.jl")).read()
    for f in files:
        if not f.endswith("NestedCrossApproximationHIPExt.jl"):
            assert f'include("{os.path.basename(f)}")' in main

    def c_fields(name):
        end = header.index("} " + name + ";")
        body = header[header.rindex("typedef struct {", 0, end) + len("typedef struct {"):end]
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        out = []
        for decl in body.split(";"):
            if decl.strip():  # "const int64_t *a, *b": the last identifier of every comma piece
                out += [re.findall(r"[A-Za-z_]\w*", piece)[-1] for piece in decl.split(",")]
        return out

    def jl_fields(name, text):
        body = re.search(r"struct " + name + r"\b[^\n]*\n(.*?)\nend", text, re.S).group(1)
        body = re.sub(r"#.*", "", body)
        return re.findall(r"(\w+)::", body)

    op = open(os.path.join(root, "operator.jl")).read()
    dr = open(os.path.join(root, "directional.jl")).read()
    assert jl_fields("H2DescC", op) == c_fields("nca_h2_desc")
    assert jl_fields("DH2DescC", dr) == c_fields("nca_dh2_desc")


def test_tuning_table_file(tmp_path):
    """The tuning table keeps the launch choices (not the timing fields) per
    key and reads back what it wrote; the shipped table parses and holds only
    valid choices."""
    from nca_amd import tuning as TU
    t = TU.TuningTable(str(tmp_path / "t.json"))
    choice = dict(near_k3=1, leaf_order=2, near_xcd_order=0, near_stagger=1, couple_k3=0, couple_tile_order=3,
                  couple_xcd_order=1, near_ms=1.23456, couple_ms=0.5, candidates=40, near_ms_start=2.0)
    t.put("gfx950|mixed|M1|N1", choice, workload="w")
    t.save()
    u = TU.TuningTable(t.path)
    e = u.get("gfx950|mixed|M1|N1")
    assert {n: e[n] for n in TU.CHOICES} == {n: choice[n] for n in TU.CHOICES}
    assert e["workload"] == "w" and "candidates" not in e
    shipped = TU.TuningTable()
    ranges = dict(near_k3=2, leaf_order=3, near_xcd_order=2, near_stagger=2, couple_k3=2, couple_tile_order=4,
                  couple_xcd_order=2)
    for key, e in shipped.entries.items():
        assert key.startswith("gfx950|")
        # near_k3 / couple_k3 = 2 (the f16 forms) are MIXED choices only
        k3 = 3 if "|mixed|" in key else 2
        for n, hi in dict(ranges, near_k3=k3, couple_k3=k3).items():
            assert 0 <= e[n] < hi, (key, n)
