"""Contexts, point sets and the HIP kernel matrix (the `farmatrix` of the reference)."""
import ctypes

import numpy as np

from . import _abi
from ._native import hip_lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def call(status):
    _abi.check(hip_lib(), status)


class Context:
    """One device and its stream; serves one host task at a time."""

    def __init__(self, device=0):
        L = hip_lib()
        h = ctypes.c_void_p()
        call(L.nca_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        call(hip_lib().nca_ctx_synchronize(self.handle))

    def stream(self):
        s = ctypes.c_void_p()
        call(hip_lib().nca_ctx_stream(self.handle, ctypes.byref(s)))
        return s.value

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class Points:
    def __init__(self, ctx, pts):
        self.ctx = ctx
        self.host = np.ascontiguousarray(pts, dtype=np.float64)
        h = ctypes.c_void_p()
        call(hip_lib().nca_points_create(ctx.handle, _p(self.host), self.host.shape[0], ctypes.byref(h)))
        self.handle = h
        self.n = self.host.shape[0]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                hip_lib().nca_points_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class HelmholtzKernel:
    """G(x, y) = cis(sign k |x - y|) / (4 pi |x - y|), zero at coincident points."""

    def __init__(self, k, sign=1):
        self.k = float(k)
        self.sign = 1 if sign >= 0 else -1

    def abi(self):
        return _abi.Helmholtz(self.k, self.sign, 0)

    def __call__(self, x, y):
        r = float(np.linalg.norm(np.asarray(x) - np.asarray(y)))
        if r == 0.0:
            return 0j
        return np.exp(1j * self.sign * self.k * r) / (4 * np.pi * r)


class HIPKernelMatrix:
    """The kernel matrix A[i, j] = G(test_i, trial_j) with both point sets on the
    device; plays the role of the AbstractKernelMatrix `farmatrix`, and its type
    selects the HIP methods of assemble_level / compresslevel_*! /
    assemble_couplingstore."""

    def __init__(self, kernel, testpoints, trialpoints, ctx=None):
        self.ctx = ctx or default_context()
        self.kernel = kernel
        self.test = Points(self.ctx, testpoints)
        self.trial = Points(self.ctx, trialpoints) if trialpoints is not testpoints else self.test
        self.shape = (self.test.n, self.trial.n)
