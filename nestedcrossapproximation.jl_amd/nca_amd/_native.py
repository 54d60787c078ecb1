"""Loading of the in-tree native libraries.

``libnca_hip.so`` holds the HIP kernels for gfx950 and the C-ABI of the hot
path; ``libnca_tree.so`` holds the host-side tree, partition and Phase-1 pivot
code. Both are built in-tree by ``build.py``. There is no fallback: a missing
library is an error.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_LIB = os.path.join(PKG_DIR, "lib", "libnca_hip.so")
TREE_LIB = os.path.join(PKG_DIR, "lib", "libnca_tree.so")

_hip = None
_tree = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path, what):
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{what} not built: {path} is missing (run build.py or __graft_entry__.build())")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def tree_lib():
    global _tree
    if _tree is None:
        _tree = _load(TREE_LIB, "host tree library")
        _declare_tree(_tree)
    return _tree


def hip_lib():
    global _hip
    if _hip is None:
        _hip = _load(HIP_LIB, "HIP extension")
        from . import _abi
        _abi.declare(_hip)
    return _hip


def _declare_tree(L):
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    D = ctypes.c_double
    L.ncat_tree_twon.restype = P
    L.ncat_tree_twon.argtypes = [P, I64, I64, D]
    L.ncat_tree_kmeans.restype = P
    L.ncat_tree_kmeans.argtypes = [P, I64, I64, I64]
    L.ncat_tree_nnodes.restype = I64
    L.ncat_tree_nnodes.argtypes = [P]
    L.ncat_tree_export.restype = None
    L.ncat_tree_export.argtypes = [P] * 10
    L.ncat_tree_free.restype = None
    L.ncat_tree_free.argtypes = [P]
    L.ncat_interactions.restype = P
    L.ncat_interactions.argtypes = [I64, P, P, P, P, I64, P, P, P, P, D]
    L.ncat_interactions_wideband.restype = P
    L.ncat_interactions_wideband.argtypes = [I64, P, P, P, P, I64, P, P, P, P, D, D, D, D]
    L.ncat_partition_sizes.restype = None
    L.ncat_partition_sizes.argtypes = [P, P, P]
    L.ncat_partition_export.restype = None
    L.ncat_partition_export.argtypes = [P, P, P, P, P]
    L.ncat_partition_free.restype = None
    L.ncat_partition_free.argtypes = [P]
    L.ncat_gather_blocks.restype = None
    L.ncat_gather_blocks.argtypes = [I64, P, P, P, P, ctypes.c_int32, P, P, P]
    L.ncat_csr_transpose.restype = P
    L.ncat_csr_transpose.argtypes = [I64, P, P, I64]
    L.ncat_farpivots.restype = P
    L.ncat_farpivots.argtypes = [I64, P, P, P, P, P, P, P, P, P, P, I64, ctypes.c_double, ctypes.c_double]
    L.ncat_farpivots_lists.restype = P
    L.ncat_farpivots_lists.argtypes = [I64, P, P, P, P, P, P, P, P, P, P, I64, ctypes.c_double, ctypes.c_double]
    L.ncat_directional_fardata.restype = P
    L.ncat_directional_fardata.argtypes = [ctypes.c_int32, I64, P, P, P, P, P, P, P, P, P, D, D, D]
    L.ncat_dirdata_sizes.restype = None
    L.ncat_dirdata_sizes.argtypes = [P, P, P]
    L.ncat_dirdata_export.restype = None
    L.ncat_dirdata_export.argtypes = [P] * 6
    L.ncat_dirdata_free.restype = None
    L.ncat_dirdata_free.argtypes = [P]
    L.ncat_farfield_csr.restype = P
    L.ncat_farfield_csr.argtypes = [I64, P, P, P]
    L.ncat_dirfarfield.restype = P
    L.ncat_dirfarfield.argtypes = [I64, P, P, P, P, P, P, P]
    L.ncat_csr_nnz.restype = I64
    L.ncat_csr_nnz.argtypes = [P]
    L.ncat_csr_export.restype = None
    L.ncat_csr_export.argtypes = [P, P, P]
    L.ncat_csr_free.restype = None
    L.ncat_csr_free.argtypes = [P]
