"""Stored near field (nca_h2_store_near): forward products read the near
entries, made once on the device in the layout of the reference CPU path
(assemblenears, src/nca/abstractnca.jl:134-178), instead of evaluating them.

Checks: the FP64 product with the stored near field within 1e-12 of the
oracle (and of the on-the-fly product) for 1, 3, 16, 19 and 32 right-hand
sides with alpha/beta; MIXED within 1e-5; transpose/adjoint products (which
evaluate on the fly) unchanged; a captured sweep replays bit for bit; the
allocation-free entry gives the same bits; dropping the store returns the
on-the-fly products bit for bit."""
import numpy as np
import pytest

import nca_amd
from nca_amd import compressors as C, geometry
from nca_amd import nca as NC
from helpers import DeviceBuffers, bitwise_equal, grid_problem, oracle_assembly, oracle_operator, relerr, \
    sphere_problem

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-stored-near] {msg}", flush=True)


def _core(p):
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(p.tol))
    return nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(p.k), p.testpts, p.trialpts, p.tree, cmp, cmp,
                                         partition=p.partition)


@pytest.mark.parametrize("which", ["grid", "sphere"])
def test_stored_near_matches_oracle(which):
    p = grid_problem(1e-4) if which == "grid" else sphere_problem(6000, k=2.0)
    say(f"{which}: {p.testpts.shape[0]} x {p.trialpts.shape[0]}")
    core = _core(p)
    ft, fs = oracle_assembly(p)
    H, _ = oracle_operator(p, ft, fs)
    M, N = core.dim
    rng = np.random.default_rng(12)
    W = NC.H2Workspace(core)
    fly = {}
    for R in (1, 3, 16, 19, 32):
        fly[R] = W @ geometry.complex_gaussian(np.random.default_rng(R), (N, R))
    core.store_near(True)
    for R in (1, 3, 16, 19, 32):
        X = geometry.complex_gaussian(np.random.default_rng(R), (N, R))
        Y0 = geometry.complex_gaussian(rng, (M, R))
        al, be = 0.5 - 0.75j, 1.25 + 0.5j
        got = NC.mul(Y0.copy(), W, X, al, be)
        ref = H.matvec(X, alpha=al, beta=be, y=Y0)
        assert relerr(got, ref) <= 1e-12, R
        assert relerr(W @ X, fly[R]) <= 1e-12, R
    # transpose and adjoint evaluate on the fly
    U = geometry.complex_gaussian(rng, (M, 16))
    assert relerr(NC.transpose(W) @ U, H.matvec(U, op=1)) <= 1e-12
    assert relerr(NC.adjoint(W) @ U, H.matvec(U, op=2)) <= 1e-12
    # MIXED: its own c32 entries
    core.set_precision(nca_amd.NCA_PREC_MIXED)
    X = geometry.complex_gaussian(rng, (N, 16))
    assert relerr(W @ X, H.matvec(X)) <= 1e-5
    core.set_precision(nca_amd.NCA_PREC_FP64)
    # dropping the store: the on-the-fly products again, bit for bit
    core.store_near(False)
    assert bitwise_equal(W @ geometry.complex_gaussian(np.random.default_rng(16), (N, 16)), fly[16])


@pytest.mark.parametrize("precision", [nca_amd.NCA_PREC_FP64, nca_amd.NCA_PREC_MIXED])
def test_stored_near_graph_and_noalloc(precision):
    """A captured stored sweep replays bit for bit (eager work in between), and
    the allocation-free entry gives the eager bits."""
    ctx = nca_amd.default_context()
    p = sphere_problem(4000, k=1.5)
    core = _core(p)
    core.set_precision(precision)
    core.store_near(True)
    N = p.testpts.shape[0]
    rng = np.random.default_rng(31)
    X = geometry.complex_gaussian(rng, (N, 16))
    dev = DeviceBuffers()
    try:
        dx, dy, dz = dev.upload(X), dev.upload(np.zeros((N, 16), np.complex128)), dev.upload(np.zeros((N, 16), np.complex128))
        core.matvec_device(0, 16, dx, N, dy, N)
        ctx.synchronize()
        eager = dev.download(dy, (N, 16))
        g = core.graph(0, 16, dx, N, dy, N)
        for rep in range(3):
            dev.copy_in(dy, np.zeros((N, 16), np.complex128))
            g.launch()
            ctx.synchronize()
            assert bitwise_equal(dev.download(dy, (N, 16)), eager), rep
            core.matvec_device(1, 3, dx, N, dz, N)  # eager work between the replays
            ctx.synchronize()
        core.matvec_device_noalloc(0, 16, dx, N, dz, N)
        ctx.synchronize()
        assert bitwise_equal(dev.download(dz, (N, 16)), eager)
        # dropping the store invalidates the graph (it reads the stored entries)
        core.store_near(False)
        with pytest.raises(nca_amd.ArgumentError):
            g.launch()
        del g
    finally:
        dev.free()
