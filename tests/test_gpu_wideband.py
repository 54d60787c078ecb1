"""Directional (wideband) H^2 on the GPU, through the C-ABI (nca_dh2_create).

Parity bar as for the plain operator: ranks, pivots and factors of every
(cluster, direction) block bit for bit against the oracle's level driver on
the same blocks and Phase-1 pivots, with the fully pivoted ACA and with the
IACA (BottomUp's factorization); the directional product within 1e-12 relative of
the oracle's stage-by-stage restatement (oracle.OracleDH2) with the same
factors, for op N, T and C; the error against dense direct summation below
10 tol (test/nca/test_petrovgalerkinnca.jl:104-153, wideband with
etalf = etahf = gamma = 1); MIXED within 1e-5 of FP64 at 16 and 19
right-hand sides."""
import math

import numpy as np
import pytest

import nca_amd
from nca_amd import _abi
from nca_amd import compressors as C
from nca_amd import directional as D
from nca_amd import geometry
from nca_amd import trees as T
from nca_amd import wideband as W
from oracle import oracle as O
from helpers import bitwise_equal, oracle_factors, relerr

pytestmark = pytest.mark.gpu

CASES = [  # tree kind, points, wavelength: together they hold low- and high-frequency
    ("kmeans", 3000, 1.5),  # clusters, bases above the leaves and child directions with several parents
    ("twon", 3000, 1.5),
    ("twon", 4000, 1.0),
]


def say(msg):
    print(f"[gpu-wideband] {msg}", flush=True)


_cores = {}


def build(case, tol=1e-4, iaca=False):
    key = (case, tol, iaca)
    if key not in _cores:
        kind, n, lam = case
        pts = geometry.fibonacci_sphere(n)
        tr = nca_amd.TwoNTree(pts, minvalues=16) if kind == "twon" else nca_amd.KMeansTree(pts, 2, 16)
        k = 2 * math.pi / lam
        isn = D.isnearwideband(k, etalf=1.0, etahf=1.0, gamma=1.0)
        comp = W.BottomUp(C.IACA(tol) if iaca else C.ACA(tol))
        core = W.DirectionalNCACore(nca_amd.HelmholtzKernel(k), pts, pts, T.BlockTree(tr, tr), comp, comp, isn,
                                    maxrank=50)
        _cores[key] = (core, pts, k)
    return _cores[key]


def oracle_of(core, pts, k):
    off, blocks = core.couplingstore.download()
    return O.OracleDH2(core.desc, k, 1, pts, pts, off, blocks)


@pytest.mark.parametrize("iaca", [False, True], ids=["aca", "iaca"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}" for c in CASES])
def test_directional_pivots_bitwise(case, iaca):
    say(f"pivots {case} iaca={iaca}")
    core, pts, k = build(case, iaca=iaca)
    for side, blocks, f in ((0, core.testblocks, core.testfactors), (1, core.trialblocks, core.trialfactors)):
        fp = f.farpivots
        res = O.bases(side, k, 1, pts, pts, blocks, fp[0], fp[1], 50, 1e-4, iaca=iaca)
        assert res[4] == 0
        ref = oracle_factors(res, blocks, fp)
        assert np.array_equal(ref.assigned, f.assigned)
        assert np.array_equal(ref.rank, f.rank)
        for a, b in zip(ref.pivot_csr(), f.pivot_csr()):
            assert np.array_equal(a, b)
        for e in np.nonzero(f.assigned)[0]:
            assert bitwise_equal(ref.X[e], f.X[e]), (side, e)
    ib = core.testblocks.isbasis & ~core.tree.testtree.leafmask()[core.testblocks.node] & core.testfactors.assigned
    say(f"  direction indices {core.testblocks.nnodes}, bases above the leaves {int(ib.sum())}")


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}" for c in CASES])
def test_directional_matvec_parity(case):
    say(f"matvec {case}")
    core, pts, k = build(case)
    ref = oracle_of(core, pts, k)
    n = len(pts)
    rng = np.random.default_rng(11)
    A = nca_amd.H2Workspace(core)
    for R in (1, 16, 19):
        X = rng.standard_normal((n, R)) + 1j * rng.standard_normal((n, R))
        Y = A @ X
        assert relerr(Y, ref.matvec(X)) < 1e-12, R
    # T and C through the bilinear identities Z^T (A^T X) = (A Z)^T X and
    # Z^H (A^H X) = (A Z)^H X: 16 x 16 independent sums each
    x = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    Z = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    AZ = ref.matvec(Z)
    assert relerr(Z.T @ (nca_amd.nca.transpose(A) @ x), AZ.T @ x) < 1e-12
    assert relerr(Z.conj().T @ (nca_amd.nca.adjoint(A) @ x), AZ.conj().T @ x) < 1e-12
    y0 = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    y = y0.copy()
    nca_amd.nca.mul(y, A, x, 0.5 - 0.25j, 2.0)
    assert relerr(y, (0.5 - 0.25j) * ref.matvec(x) + 2.0 * y0) < 1e-12


@pytest.mark.parametrize("iaca", [False, True], ids=["aca", "iaca"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}" for c in CASES])
def test_directional_direct_sum_and_mixed(case, iaca):
    say(f"direct sum + mixed {case} iaca={iaca}")
    core, pts, k = build(case, iaca=iaca)
    n = len(pts)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    A = nca_amd.H2Workspace(core)
    y64 = A @ x
    rows = np.sort(rng.choice(n, 400, replace=False))
    yd = O.dense_rows(k, 1, pts, pts, rows, x)
    err = relerr(y64[rows], yd)
    say(f"  direct-sum error {err:.2e}")
    assert err < 10 * 1e-4
    x19 = rng.standard_normal((n, 19)) + 1j * rng.standard_normal((n, 19))
    y19 = A @ x19
    core.set_precision(_abi.NCA_PREC_MIXED)
    try:
        ym = A @ x
        ym19 = A @ x19  # two column chunks (16 + 3)
    finally:
        core.set_precision(_abi.NCA_PREC_FP64)
    assert relerr(ym, y64) < 1e-5
    assert relerr(ym19, y19) < 1e-5


def test_directional_graph_replays():
    """A captured directional sweep replays bit for bit like the eager one."""
    import torch
    core, pts, k = build(CASES[0])
    n = len(pts)
    rng = np.random.default_rng(3)
    x = torch.tensor(rng.standard_normal((16, n)) + 1j * rng.standard_normal((16, n)), dtype=torch.complex128,
                     device="cuda")
    y = torch.zeros((16, n), dtype=torch.complex128, device="cuda")
    core.matvec_device(_abi.NCA_OP_N, 16, x.data_ptr(), n, y.data_ptr(), n)
    torch.cuda.synchronize()
    want = y.clone()
    g = core.graph(_abi.NCA_OP_N, 16, x.data_ptr(), n, y.data_ptr(), n)
    for _ in range(3):
        y.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(y, want)
    del g
