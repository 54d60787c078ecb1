"""The benchmark's per-stage byte model (bench.stage_bytes) on a small worked
example, computed by hand below, and the properties the model must keep."""
import bench

ST = dict(coeffs_trial=10, coeffs_test=12, coeffs_trial_leaves=6, coeffs_test_leaves=8,
          basis_entries_trial=100, basis_entries_test=120, transfer_entries_trial=30, transfer_entries_test=40,
          coupling_entries=500, coupling_columns=50, coupling_blocks=9, near_pairs=7)


def test_mixed_worked_example():
    # R = 2, M = N = 20, MIXED: operators and gathered/source vectors 8 B,
    # destination coefficients 16 B, the coupling result 8 B
    b = bench.stage_bytes(ST, 2, 20, 20, True)
    assert b["gather"] == 16 * 2 * 20 + 8 * 2 * 20 + 4 * 20 == 1040
    assert b["leaf_projection"] == 8 * 100 + 8 * 2 * 20 + 8 * 2 * 6 == 1216
    assert b["upward"] == 8 * 30 + 8 * 2 * (10 + 4) == 464
    assert b["coupling"] == 8 * 500 + 4 * 50 + 8 * 2 * 10 + 8 * 2 * 12 == 4552
    assert b["downward"] == 8 * 40 + 16 * 2 * 4 + (8 + 16) * 2 * 12 == 1024
    assert b["leaf_expansion_near"] == 8 * 120 + 16 * 2 * 8 + 8 * 2 * 20 + 24 * 40 + 12 * 7 + 16 * 2 * 20 == 3220


def test_mixed_f16_near_field():
    # the f16 near field adds to the gather a read of the c32 copy (8 B per
    # point and column), the write of the paired split copy (16 B) and the
    # column maxima (4 B per 64 points and column, written and read); the near
    # field reads the split copy (16 B) instead of the c32 one (8 B)
    b, h = bench.stage_bytes(ST, 2, 20, 20, True), bench.stage_bytes(ST, 2, 20, 20, True, h16=True)
    assert h["gather"] - b["gather"] == (8 + 16) * 2 * 20 + 2 * 4 * 2 * 1
    assert h["leaf_expansion_near"] - b["leaf_expansion_near"] == (16 - 8) * 2 * 20
    assert all(h[k] == b[k] for k in b if k not in ("gather", "leaf_expansion_near"))


def test_mixed_coupling_on_the_fly():
    # the coupling evaluated on the fly reads no coupling entries or column
    # maps: the source coefficients read (8 B), their paired split copy
    # written and read (16 B each), the pivot coordinates of both sides and
    # the ranges; it writes the same c32 result
    b, o = bench.stage_bytes(ST, 2, 20, 20, True), bench.stage_bytes(ST, 2, 20, 20, True, otf=True)
    assert o["coupling"] == (8 + 16 + 16) * 2 * 10 + 24 * (10 + 12) + 8 * 9 + 8 * 2 * 12
    assert all(o[k] == b[k] for k in b if k != "coupling")


def test_fp64_worked_example():
    b = bench.stage_bytes(ST, 2, 20, 20, False)
    assert b["gather"] == 16 * 2 * 20 * 2 + 4 * 20
    assert b["coupling"] == 16 * 500 + 4 * 50 + 16 * 2 * 10 + 16 * 2 * 12
    assert b["downward"] == 16 * 40 + 16 * 2 * 4 + 32 * 2 * 12


def test_model_properties():
    m, f = bench.stage_bytes(ST, 16, 20, 20, True), bench.stage_bytes(ST, 16, 20, 20, False)
    assert all(m[k] < f[k] for k in m)
    # y is read only when beta is nonzero
    mb = bench.stage_bytes(ST, 16, 20, 20, True, beta=True)
    assert mb["leaf_expansion_near"] - m["leaf_expansion_near"] == 16 * 16 * 20
    assert all(mb[k] == m[k] for k in m if k != "leaf_expansion_near")
    # operator bytes do not grow with the right-hand sides
    one, two = bench.stage_bytes(ST, 1, 20, 20, True), bench.stage_bytes(ST, 2, 20, 20, True)
    assert two["coupling"] - one["coupling"] == one["coupling"] - 8 * 500 - 4 * 50
