"""nca_h2_matvec_device_noalloc: the device product on the library's static
scratch array instead of the operator's workspace. Its results equal
nca_h2_matvec_device's bit for bit (the same kernels), on two calls in a row,
the second with a smaller operator and fewer right-hand sides (so it runs on
the stale contents of the first); a product that needs more than the scratch
holds is refused on the host before anything is launched."""
import numpy as np
import pytest

import nca_amd
from nca_amd import _abi, compressors as C, geometry
from helpers import DeviceBuffers, bitwise_equal

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


def _core(n, k=1.0):
    pts = geometry.fibonacci_sphere(n)
    tree = nca_amd.TwoNTree(pts, minvalues=32)
    cmp = C.BatchedBottomUp(C.GeometricPivoting(), C.ACA(1e-4))
    return nca_amd.PetrovGalerkinNCACore(nca_amd.HelmholtzKernel(k), pts, pts, nca_amd.BlockTree(tree, tree),
                                         cmp, cmp)


@pytest.mark.parametrize("precision", [nca_amd.NCA_PREC_FP64, nca_amd.NCA_PREC_MIXED])
def test_noalloc_equals_device_product(precision):
    ctx = nca_amd.default_context()
    big, small = _core(4000), _core(1500, k=1.7)
    rng = np.random.default_rng(77)
    dev = DeviceBuffers()
    try:
        # first call: 4000 points, 16 right-hand sides, alpha/beta; second: 1500
        # points, 3 right-hand sides (the VALU chunk that clears its vectors), op T
        for core, R, op, al, be in ((big, 16, 0, 0.5 - 0.25j, 1.5), (small, 3, 1, 1.0, 0.0)):
            core.set_precision(precision)
            n = core.dim[0]
            say(f"precision {precision}: n={n} R={R} op={op}")
            X = geometry.complex_gaussian(rng, (n, R))
            Y0 = geometry.complex_gaussian(rng, (n, R))
            dx, dy1, dy2 = dev.upload(X), dev.upload(Y0), dev.upload(Y0)
            core.matvec_device(op, R, dx, n, dy1, n, al, be)
            ctx.synchronize()
            ref = dev.download(dy1, (n, R))
            core.matvec_device_noalloc(op, R, dx, n, dy2, n, al, be)
            ctx.synchronize()
            got = dev.download(dy2, (n, R))
            assert bitwise_equal(got, ref), (n, R, op)
    finally:
        dev.free()


def test_noalloc_refuses_oversized_products():
    """4000 points x 2048 right-hand sides need about 130 MB of workspace:
    NCA_ERR_LIMIT from the host, nothing launched, y untouched."""
    ctx = nca_amd.default_context()
    core = _core(4000)
    n, R = core.dim[0], 2048
    dev = DeviceBuffers()
    try:
        X = np.ones((n, R), np.complex128)
        dx, dy = dev.upload(X), dev.upload(X)
        with pytest.raises(nca_amd.NCAError) as e:
            core.matvec_device_noalloc(0, R, dx, n, dy, n)
        assert e.value.code == _abi.NCA_ERR_LIMIT
        ctx.synchronize()
        assert np.array_equal(dev.download(dy, (n, R)), X)
    finally:
        dev.free()
