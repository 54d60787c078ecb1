"""The MIXED forms on f16 MFMA with split operands (every FP32 kernel value
and right-hand side or source coefficient as two f16 parts; the MIXED
defaults), through the C-ABI. The near field: the form MIXED products take
when the near pairs allow it (nca_h2_near_form), within the MIXED bar of the
oracle's compressed FP64 operator for N, T and C on 16-, 8- and 4-row tiles
and one or two 16-column chunks, no less accurate than the f32 MFMA form;
exact under power-of-two scaling of each right-hand side (the per-column
scales), exact zeros for zero columns, and near pairs spanning more than 2^17
in distance fall back to the f32 form. The coupling evaluated on the fly at
the pivots (nca_h2_couple_form): within the bar for N, T and C at 16, 3 and
40 right-hand sides, no less accurate than the stored c32 blocks. Distances
in FP32 from tile-relative coordinates where the tile statistics allow them
(nca_h2_rel32), within the bar, and FP64 differences kept on the tiles whose
rows lie more than 32 times their smallest pair distance apart. A point at
the origin (where the sources past a range end read their coordinates)
leaves every form finite."""
import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from nca_amd import nca as NC
from nca_amd import tuning as TU
from oracle import oracle as O
from helpers import relerr

pytestmark = pytest.mark.gpu

# the launch choices of the defaults, with the near-field form varied
BASE = dict(leaf_order=2, near_xcd_order=0, near_stagger=0, couple_k3=2, couple_tile_order=3, couple_xcd_order=0)


def say(msg):
    print(f"[gpu-near-f16] {msg}", flush=True)


def core_of(pts, k, minvalues=32, tol=1e-4):
    tree = nca_amd.TwoNTree(pts, minvalues=minvalues)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(tol))
    core = nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), pts, pts, nca_amd.BlockTree(tree, tree), cmp,
                                         cmp, ctx=nca_amd.default_context())
    off, blocks = core.couplingstore.download()
    return core, O.OracleH2(core.desc, k, 1, pts, pts, off, blocks)


def tile_shapes(core_pts, minvalues):
    tree = nca_amd.TwoNTree(core_pts, minvalues=minvalues)
    sizes = np.array([tree.nvalues(t) for t in np.where(tree.leafmask())[0]])
    rem = sizes % 16
    return {"16": bool(np.any(sizes >= 16)), "8": bool(np.any((rem > 4) & (rem <= 8))),
            "4": bool(np.any((rem > 0) & (rem <= 4)))}


def test_f16_form_default_and_within_bar():
    """8,192 sphere points (k = 2, octree leaves of 1..32 points: 16-, 8- and
    4-row tiles): MIXED takes the f16 form, N/T/C products at 16, 8 and 40
    right-hand sides are within 1e-5 of the FP64 oracle, not above twice the
    f32 form's error, and differ from it (another arithmetic)."""
    pts = geometry.fibonacci_sphere(8192)
    assert all(tile_shapes(pts, 32).values())
    core, H = core_of(pts, 2.0)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    form, rmin, rmax = core.near_form()
    say(f"form {form}, near distances {rmin:.3e} .. {rmax:.3e}")
    assert form == 2 and 0 < rmin < rmax <= 2 ** 17 * rmin
    rng = np.random.default_rng(41)
    for R in (16, 8, 40):
        x = geometry.complex_gaussian(rng, (pts.shape[0], R))
        refs = [H.matvec(x, op=op) for op in (0, 1, 2)]
        res = {}
        for k3 in (2, 1):
            TU.apply(core, dict(BASE, near_k3=k3))
            assert core.near_form()[0] == (2 if k3 == 2 else 1)
            res[k3] = [W @ x, NC.transpose(W) @ x, NC.adjoint(W) @ x]
        for op in range(3):
            e16, e32 = relerr(res[2][op], refs[op]), relerr(res[1][op], refs[op])
            say(f"R={R} op={op}: f16 {e16:.2e} f32 {e32:.2e}")
            assert e16 <= 1e-5 and e32 <= 1e-5
            assert e16 <= 2 * e32
            assert not np.array_equal(res[2][op], res[1][op])


def test_f16_power_of_two_scaling_and_zero_columns():
    """Each right-hand side scaled by its own power of two (2^-60 .. 2^60):
    every output column is the unscaled one times the same power, bit for
    bit (the per-column scales absorb the factor); a zero column gives exact
    zeros and leaves the others unchanged."""
    pts = geometry.fibonacci_sphere(4096)
    core, _ = core_of(pts, 1.0)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    assert core.near_form()[0] == 2
    rng = np.random.default_rng(42)
    x = geometry.complex_gaussian(rng, (pts.shape[0], 16))
    y0 = W @ x
    e = np.array([-60, -33, -7, 0, 1, 5, 17, 60, -1, 2, -2, 9, -20, 30, 44, -50], dtype=float)
    ys = W @ (x * 2.0 ** e)
    for q in range(16):
        assert np.array_equal(ys[:, q], y0[:, q] * 2.0 ** e[q]), q
    xz = x.copy()
    xz[:, 5] = 0
    yz = W @ xz
    assert np.all(yz[:, 5] == 0)
    others = [q for q in range(16) if q != 5]
    assert np.array_equal(yz[:, others], y0[:, others])


@pytest.mark.parametrize("gap, form", [(1e-4, 2), (1e-7, 1)])
def test_f16_distance_span(gap, form):
    """A 4,096-point sphere plus one point at distance `gap` from the first
    (its near pairs reach 1.5 apart, through a sparse coarse leaf): near
    pairs spanning up to 2^17 in distance keep the f16 form (1e-4: the point
    scale moves), beyond it (1e-7) MIXED falls back to the f32 form; both
    within 1e-5 of the FP64 oracle."""
    base = geometry.fibonacci_sphere(4096)
    d = np.cross(base[0], [0.0, 0.0, 1.0])
    d /= np.linalg.norm(d)
    pts = np.vstack([base, base[0] + gap * d])
    core, H = core_of(pts, 1.0)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    f, rmin, rmax = core.near_form()
    say(f"gap {gap}: form {f}, near distances {rmin:.3e} .. {rmax:.3e}")
    assert f == form and abs(rmin - gap) <= 1e-3 * gap
    # the close pair's tiles hold rows more than 32 times its distance apart:
    # they keep FP64 differences, the others take the relative form
    nr, _, fn, _ = core.rel32(fractions=True)
    say(f"gap {gap}: tiles on the relative form {fn:.4f}")
    assert nr == (form == 2) and 0.5 < fn < 1
    rng = np.random.default_rng(43)
    x = geometry.complex_gaussian(rng, (pts.shape[0], 16))
    assert relerr(W @ x, H.matvec(x)) <= 1e-5
    assert relerr(NC.adjoint(W) @ x, H.matvec(x, op=2)) <= 1e-5


def test_couple_on_the_fly_within_bar():
    """8,192 sphere points, k = 2: MIXED evaluates the coupling on the fly
    (form 2, far pivot distances measured); N, T and C at 16, 3 and 40
    right-hand sides within 1e-5 of the FP64 oracle, not above twice the
    stored c32 form's error, and different from it."""
    pts = geometry.fibonacci_sphere(8192)
    core, H = core_of(pts, 2.0)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    form, rmin, rmax = core.couple_form()
    say(f"coupling form {form}, far pivot distances {rmin:.3e} .. {rmax:.3e}")
    assert form == 2 and 0 < rmin < rmax <= 2 ** 17 * rmin
    rng = np.random.default_rng(44)
    for R in (16, 3, 40):
        x = geometry.complex_gaussian(rng, (pts.shape[0], R))
        refs = [H.matvec(x, op=op) for op in (0, 1, 2)]
        res = {}
        for ck in (2, 1):
            TU.apply(core, dict(BASE, near_k3=2, couple_k3=ck))
            assert core.couple_form()[0] == ck
            res[ck] = [W @ x, NC.transpose(W) @ x, NC.adjoint(W) @ x]
        for op in range(3):
            e2, e1 = relerr(res[2][op], refs[op]), relerr(res[1][op], refs[op])
            say(f"R={R} op={op}: on the fly {e2:.2e} stored {e1:.2e}")
            assert e2 <= 1e-5 and e1 <= 1e-5
            assert e2 <= 2 * e1
            assert not np.array_equal(res[2][op], res[1][op])


def test_rel32_distances_within_bar():
    """8,192 sphere points, k = 2: the f16 near field and coupling take their
    distances in FP32 from tile-relative coordinates (nca_h2_rel32, allowed
    by the tile statistics); N, T and C at 16 and 3 right-hand sides within
    1e-5 of the FP64 oracle with it and without it (FP64 differences), the
    relative form not above 1.5 times the other's error, and different from
    it; switching back restores the first result bit for bit."""
    pts = geometry.fibonacci_sphere(8192)
    core, H = core_of(pts, 2.0)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    TU.apply(core, dict(BASE, near_k3=2, couple_k3=2))
    assert core.near_form()[0] == 2 and core.couple_form()[0] == 2
    nr, cr, fn, fc = core.rel32(fractions=True)
    say(f"tiles on the relative form: near {fn:.4f}, coupling {fc:.4f}")
    assert (nr, cr) == (True, True) and fn >= 0.9 and fc >= 0.9
    rng = np.random.default_rng(45)
    for R in (16, 3):
        x = geometry.complex_gaussian(rng, (pts.shape[0], R))
        refs = [H.matvec(x, op=op) for op in (0, 1, 2)]
        res = {}
        for mode in (1, 0):
            assert core.rel32(mode) == (bool(mode), bool(mode))
            res[mode] = [W @ x, NC.transpose(W) @ x, NC.adjoint(W) @ x]
        core.rel32(1)
        assert np.array_equal(W @ x, res[1][0])
        for op in range(3):
            er, ed = relerr(res[1][op], refs[op]), relerr(res[0][op], refs[op])
            say(f"R={R} op={op}: relative FP32 {er:.2e} FP64 differences {ed:.2e}")
            assert er <= 1e-5 and ed <= 1e-5
            assert er <= 1.5 * ed
            assert not np.array_equal(res[1][op], res[0][op])


def test_point_at_origin_every_form():
    """The 32 x 32 grid of [0, 1]^2 (a point at the origin, the coordinates
    read for sources past a range end) with itself, k = 2 pi / 10: every MIXED
    form (near field f16, f32 four- or three-product; coupling on the fly or
    stored) gives finite products within 1e-5 of the FP64 oracle, N and C."""
    pts = geometry.grid_square(32)
    assert np.any(np.all(pts == 0, axis=1))
    core, H = core_of(pts, 2 * np.pi / 10)
    W = NC.H2Workspace(core)
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    rng = np.random.default_rng(45)
    x = geometry.complex_gaussian(rng, (pts.shape[0], 16))
    ref, refc = H.matvec(x), H.matvec(x, op=2)
    for nk in (2, 1, 0):
        for ck in (2, 1):
            TU.apply(core, dict(BASE, near_k3=nk, couple_k3=ck))
            y, yc = W @ x, NC.adjoint(W) @ x
            say(f"near_k3={nk} couple_k3={ck}: {relerr(y, ref):.2e} {relerr(yc, refc):.2e}")
            assert np.all(np.isfinite(y)) and np.all(np.isfinite(yc))
            assert relerr(y, ref) <= 1e-5 and relerr(yc, refc) <= 1e-5
