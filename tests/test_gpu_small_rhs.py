"""One to three right-hand sides on both routes of the product, through the
C-ABI: the VALU kernels (the FP64 default) and the MFMA chunk kernels padded to
16 columns (the MIXED default, with the f16 near field and coupling), each
forced by NCA_MFMA_SMALL at creation, for N, T and C against the oracle's
compressed FP64 operator (FP64 within 1e-12, MIXED within 1e-5), and the two
routes within the same bars of each other."""
import os

import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C
from nca_amd import geometry
from nca_amd import nca as NC
from oracle import oracle as O
from helpers import relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-small-rhs] {msg}", flush=True)


def core_with(pts, k, route):
    """A core created with NCA_MFMA_SMALL = route (the variable is read at
    creation only)."""
    old = os.environ.get("NCA_MFMA_SMALL")
    os.environ["NCA_MFMA_SMALL"] = route
    try:
        tree = nca_amd.TwoNTree(pts, minvalues=32)
        cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
        return nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), pts, pts, nca_amd.BlockTree(tree, tree),
                                             cmp, cmp, ctx=nca_amd.default_context())
    finally:
        if old is None:
            del os.environ["NCA_MFMA_SMALL"]
        else:
            os.environ["NCA_MFMA_SMALL"] = old


def test_small_rhs_both_routes():
    """8,192 sphere points, k = 2: R = 1, 2, 3 on the VALU and MFMA routes, in
    FP64 and MIXED, N / T / C."""
    pts = geometry.fibonacci_sphere(8192)
    cores = {r: core_with(pts, 2.0, r) for r in ("0", "1")}
    off, blocks = cores["0"].couplingstore.download()
    H = O.OracleH2(cores["0"].desc, 2.0, 1, pts, pts, off, blocks)
    rng = np.random.default_rng(51)
    for prec, bar in ((nca_amd.NCA_PREC_FP64, 1e-12), (nca_amd.NCA_PREC_MIXED, 1e-5)):
        for c in cores.values():
            c.set_precision(prec)
        ws = {r: NC.H2Workspace(c) for r, c in cores.items()}
        for R in (1, 2, 3):
            x = geometry.complex_gaussian(rng, (pts.shape[0], R))
            for op, f in ((0, lambda w: w), (1, NC.transpose), (2, NC.adjoint)):
                ref = H.matvec(x, op=op)
                y = {r: f(w) @ x for r, w in ws.items()}
                e = {r: relerr(y[r], ref) for r in y}
                say(f"prec {prec} R={R} op={op}: VALU {e['0']:.2e}, MFMA {e['1']:.2e}")
                assert e["0"] <= bar and e["1"] <= bar
                assert relerr(y["1"], y["0"]) <= 2 * bar
