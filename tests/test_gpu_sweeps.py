"""hipGraph-captured sweeps, block-diagonal batching of independent problems and
the autotuner.

Graph replays are checked against the eager sweep on every replay, with eager
work (uploads, other products) between the replays. The batched operator (one
forest, one launch per level and per matvec stage for all problems) is checked
against the same problems built one by one: pivots bitwise, products to 1e-12,
and each problem against dense direct summation."""
import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C, geometry, nca as NC
from oracle import oracle as O
from helpers import DeviceBuffers, bitwise_equal, relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-batch-graph] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    return nca_amd.default_context()


def _core(ctx, pts, tol=1e-4, k=1.0, minvalues=32):
    tree = nca_amd.TwoNTree(pts, minvalues=minvalues)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(tol))
    return nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), pts, pts, nca_amd.BlockTree(tree, tree),
                                         cmp, cmp, ctx=ctx)


@pytest.mark.parametrize("precision", [nca_amd.NCA_PREC_FP64, nca_amd.NCA_PREC_MIXED])
def test_graph_replays_match_eager(ctx, precision):
    say(f"graph replay precision={precision}")
    pts = geometry.fibonacci_sphere(3000)
    core = _core(ctx, pts)
    core.set_precision(precision)
    N = pts.shape[0]
    rng = np.random.default_rng(21)
    R = 19
    X = geometry.complex_gaussian(rng, (N, R))
    Y0 = geometry.complex_gaussian(rng, (N, R))
    dev = DeviceBuffers()
    try:
        dx, dy = dev.upload(X), dev.upload(Y0)
        al, be = 0.7 - 0.1j, 0.5 + 0.25j
        core.matvec_device(0, R, dx, N, dy, N, al, be)
        ctx.synchronize()
        eager = dev.download(dy, (N, R))
        g = core.graph(0, R, dx, N, dy, N, al, be)
        other = dev.upload(np.zeros((N, 3), np.complex128))
        for rep in range(4):
            dev.copy_in(dy, Y0)
            g.launch()
            ctx.synchronize()
            got = dev.download(dy, (N, R))
            assert np.array_equal(got, eager), f"replay {rep}"
            # eager work between the replays
            core.matvec_device(1, 3, dx, N, other, N)
            ctx.synchronize()
        del g
    finally:
        dev.free()


def test_graph_rejects_precision_change(ctx):
    pts = geometry.fibonacci_sphere(1500)
    core = _core(ctx, pts)
    N = pts.shape[0]
    dev = DeviceBuffers()
    try:
        dx = dev.upload(np.ones((N, 2), np.complex128))
        dy = dev.upload(np.zeros((N, 2), np.complex128))
        g = core.graph(0, 2, dx, N, dy, N)
        core.set_precision(nca_amd.NCA_PREC_MIXED)
        with pytest.raises(nca_amd.ArgumentError):
            g.launch()
        core.set_precision(nca_amd.NCA_PREC_FP64)
        g.launch()
        ctx.synchronize()
        del g
    finally:
        dev.free()


def test_batched_forest_matches_separate_problems(ctx):
    say("batched forest of 3 problems")
    sets = [geometry.fibonacci_sphere(1200),
            geometry.ellipsoid(1600, (1.5, 0.8, 1.1)),
            geometry.ellipsoid(900, (0.6, 1.3, 0.9))]
    tol, k = 1e-4, 1.0
    trees = [nca_amd.TwoNTree(p, minvalues=32) for p in sets]
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(tol))
    B = nca_amd.BatchedNCA(nca_amd.HelmholtzKernel(k), sets, sets, trees, trees, cmp, cmp, ctx=ctx)
    singles = [nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), p, p, nca_amd.BlockTree(t, t), cmp, cmp,
                                             ctx=ctx) for p, t in zip(sets, trees)]
    # pivots: the forest's are the problems' own, shifted by the point offsets
    for side in ("testpivots", "trialpivots"):
        fptr, fidx = getattr(B.core, side)
        for p, s in enumerate(singles):
            ptr, idx = getattr(s, side)
            m = B.testmaps[p] if side == "testpivots" else B.trialmaps[p]
            off = B.testoffsets[p] if side == "testpivots" else B.trialoffsets[p]
            for i in range(len(m)):
                a = fidx[fptr[m[i]]:fptr[m[i] + 1]]
                b = idx[ptr[i]:ptr[i + 1]] + off
                assert np.array_equal(a, b), (side, p, i)
    # products: stacked = per problem
    rng = np.random.default_rng(8)
    xs = [geometry.complex_gaussian(rng, (p.shape[0], 16)) for p in sets]
    y = B @ B.stack(xs)
    for p, (s, yp) in enumerate(zip(singles, B.split(y))):
        ref = NC.H2Workspace(s) @ xs[p]
        assert relerr(yp, ref) <= 1e-12, p
        rows = rng.choice(sets[p].shape[0], 100, replace=False)
        d = O.dense_rows(k, 1, sets[p], sets[p], rows, xs[p][:, :1])[:, 0]
        assert relerr(yp[rows, 0], d) < 5 * tol, p
    # one captured sweep applies all problems, 32 right-hand sides
    X = B.stack([geometry.complex_gaussian(rng, (p.shape[0], 32)) for p in sets])
    N = X.shape[0]
    eager = B @ X
    dev = DeviceBuffers()
    try:
        dx, dy = dev.upload(X), dev.upload(np.zeros_like(X))
        g = B.core.graph(0, 32, dx, N, dy, N)
        for rep in range(3):
            g.launch()
            ctx.synchronize()
            assert relerr(dev.download(dy, X.shape), eager) <= 1e-14, rep
        del g
    finally:
        dev.free()


@pytest.mark.parametrize("precision", [nca_amd.NCA_PREC_FP64, nca_amd.NCA_PREC_MIXED])
def test_autotune_keeps_results(ctx, precision):
    """The autotuner only changes launch choices: products stay within the
    precision's tolerance of the oracle, the choice is a timed candidate and
    the tuned stage times are not above the starting ones."""
    say(f"autotune precision={precision}")
    from helpers import sphere_problem
    p = sphere_problem(4096)
    core = _core(ctx, p.testpts, tol=p.tol, k=p.k, minvalues=64)
    core.set_precision(precision)
    t = core.autotune(reps=2)
    # MIXED also times the f16 near field (near_k3 = 2) and the coupling
    # evaluated on the fly (couple_k3 = 2): 3 x 3 x 4 near + 3 x 4 x 2 coupling candidates
    mixed = precision == nca_amd.NCA_PREC_MIXED
    assert t["candidates"] == (60 if mixed else 40) and t["near_stagger"] in (0, 1)
    assert t["near_k3"] in ((0, 1, 2) if mixed else (0, 1)) and t["leaf_order"] in (0, 1, 2)
    assert t["near_xcd_order"] in (0, 1)
    assert t["couple_k3"] in ((0, 1, 2) if mixed else (0, 1)) and t["couple_tile_order"] in (0, 1, 2, 3)
    assert t["couple_xcd_order"] in (0, 1)
    assert t["near_ms"] <= t["near_ms_start"] and t["couple_ms"] <= t["couple_ms_start"]
    rng = np.random.default_rng(4)
    x = geometry.complex_gaussian(rng, (p.testpts.shape[0], 16))
    W = NC.H2Workspace(core)
    W2 = _core(ctx, p.testpts, tol=p.tol, k=p.k, minvalues=64)
    W2.set_precision(precision)
    ref = NC.H2Workspace(W2) @ x
    bound = 1e-12 if precision == nca_amd.NCA_PREC_FP64 else 1e-5
    assert relerr(W @ x, ref) <= bound
    u = geometry.complex_gaussian(rng, (p.testpts.shape[0], 16))
    assert relerr(NC.adjoint(W) @ u, NC.adjoint(NC.H2Workspace(W2)) @ u) <= bound


@pytest.mark.parametrize("precision", [nca_amd.NCA_PREC_FP64, nca_amd.NCA_PREC_MIXED])
def test_tuning_table_round_trip(ctx, precision, tmp_path):
    """A tuning table: the first operator is measured and stored, a second one
    of the same shape takes the stored choices without timing (the same
    choices, so the same products bit for bit), the table survives a write and
    a read, and choices out of range are refused."""
    from nca_amd import tuning as TU
    say(f"tuning table precision={precision}")
    pts = geometry.fibonacci_sphere(5000)
    a, b = _core(ctx, pts), _core(ctx, pts)
    for c in (a, b):
        c.set_precision(precision)
    path = str(tmp_path / "table.json")
    table = TU.TuningTable(path)
    ra = TU.tune(a, precision, table, reps=1, workload="sphere 5000")
    assert ra["source"] == "measured" and ra["key"] in table.entries
    table.save()
    rb = TU.tune(b, precision, TU.TuningTable(path))
    assert rb["source"] == "table" and rb["key"] == ra["key"]
    assert all(rb[n] == ra[n] for n in TU.CHOICES)
    x = geometry.complex_gaussian(np.random.default_rng(6), (pts.shape[0], 16))
    assert np.array_equal(NC.H2Workspace(a) @ x, NC.H2Workspace(b) @ x)
    bad = dict(ra, couple_tile_order=7)
    with pytest.raises(nca_amd.ArgumentError, match="out of range"):
        TU.apply(b, bad)


def test_symmetric_reuse_equals_trial_pass():
    """One tree and one point set: the reused (transposed) test factors equal
    the ones a separate trial pass computes, bit for bit, and so do the products."""
    say("symmetric reuse")
    pts = geometry.fibonacci_sphere(6000)
    tr = nca_amd.TwoNTree(pts, minvalues=32)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    cores = [nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(1.3), pts, pts, nca_amd.BlockTree(tr, tr),
                                           cmp, cmp, reuse_symmetric=r) for r in (True, False)]
    assert cores[0].symmetric and not cores[1].symmetric
    a, b = cores
    a.host_stores()
    b.host_stores()
    for x, y in zip(a.trialfactors.pivot_csr(), b.trialfactors.pivot_csr()):
        assert np.array_equal(x, y)
    for name in ("nodes", "offsets", "blocks"):
        assert np.array_equal(getattr(a.trialbasis, name), getattr(b.trialbasis, name))
    for name in ("levelidcs", "chdsidcs", "offsets", "blocks"):
        assert np.array_equal(getattr(a.trialtransfer, name), getattr(b.trialtransfer, name))
    x = geometry.complex_gaussian(np.random.default_rng(1), (pts.shape[0], 16))
    assert np.array_equal(NC.H2Workspace(a) @ x, NC.H2Workspace(b) @ x)


def test_threads_share_one_core(ctx):
    """Two host threads, each with its own H2Workspace over one core, with
    different right-hand side counts (3 and 19, so the second grows the device
    workspace while the first may be running): every product equals the same
    product run serially, bit for bit."""
    import threading
    say("two threads over one core")
    pts = geometry.fibonacci_sphere(4000)
    core = _core(ctx, pts)
    rng = np.random.default_rng(33)
    xa = geometry.complex_gaussian(rng, (pts.shape[0], 3))
    xb = geometry.complex_gaussian(rng, (pts.shape[0], 19))
    # serial references from a fresh core (its workspace starts empty as well)
    ref_core = _core(ctx, pts)
    ra, rb = NC.H2Workspace(ref_core) @ xa, NC.adjoint(NC.H2Workspace(ref_core)) @ xb
    wa, wb = NC.H2Workspace(core), NC.adjoint(NC.H2Workspace(core))
    out, errors = {"a": [], "b": []}, []

    def run(name, w, x, n):
        try:
            for _ in range(n):
                out[name].append(w @ x)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    ta = threading.Thread(target=run, args=("a", wa, xa, 6))
    tb = threading.Thread(target=run, args=("b", wb, xb, 6))
    ta.start()
    tb.start()
    ta.join()
    tb.join()
    assert not errors, errors
    assert len(out["a"]) == len(out["b"]) == 6
    for y in out["a"]:
        assert bitwise_equal(y, ra)
    for y in out["b"]:
        assert bitwise_equal(y, rb)


def test_graph_rejects_reallocated_workspace(ctx):
    """A captured sweep points to the operator's workspace and launch plans. A
    product with more right-hand sides (which reallocates the workspace) or the
    autotuner (which rewrites the plans) makes every later launch of that graph
    fail on the host, without launching it; a graph captured afterwards runs and
    matches the eager product."""
    say("graph generation check")
    pts = geometry.fibonacci_sphere(2000)
    core = _core(ctx, pts)
    N = pts.shape[0]
    rng = np.random.default_rng(5)
    X2 = geometry.complex_gaussian(rng, (N, 2))
    dev = DeviceBuffers()
    try:
        dx = dev.upload(X2)
        dy = dev.upload(np.zeros((N, 2), np.complex128))
        g = core.graph(0, 2, dx, N, dy, N)
        g.launch()
        ctx.synchronize()
        # an eager product with 32 right-hand sides grows the workspace
        NC.H2Workspace(core) @ geometry.complex_gaussian(rng, (N, 32))
        with pytest.raises(nca_amd.ArgumentError):
            g.launch()
        del g
        g = core.graph(0, 2, dx, N, dy, N)
        g.launch()
        ctx.synchronize()
        assert relerr(dev.download(dy, (N, 2)), NC.H2Workspace(core) @ X2) <= 1e-14
        core.autotune(reps=1)
        with pytest.raises(nca_amd.ArgumentError):
            g.launch()
        del g
        g = core.graph(0, 2, dx, N, dy, N)
        g.launch()
        ctx.synchronize()
        assert relerr(dev.download(dy, (N, 2)), NC.H2Workspace(core) @ X2) <= 1e-12
        del g
    finally:
        dev.free()
