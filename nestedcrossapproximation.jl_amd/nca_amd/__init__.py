"""MI355X-native batched bottom-up NCA compressor and H^2 matvec.

Host mirror of NestedCrossApproximation.jl's API for the hot path; all compute
runs in the HIP kernels of libnca_hip.so through the C-ABI of include/nca_hip.h.
"""
from ._abi import (NCA_OP_N, NCA_OP_T, NCA_OP_C, NCA_PREC_FP64, NCA_PREC_MIXED, NCAError,
                   SingularCrossMatrixError, DimensionMismatch, ArgumentError)
from ._native import hip_lib, tree_lib, NativeLibraryMissing
from .runtime import Context, default_context, HelmholtzKernel, HIPKernelMatrix
from .trees import (TwoNTree, KMeansTree, BlockTree, interactions, fardata, farfield, farpivots, FarData,
                    Partition, forest, forest_partition)
from .compressors import (ACA, IACA, TopDown, BatchedBottomUp, GeometricPivoting, assemble_level, testbases, trialbases,
                          assemble_couplingstore, concat_ranges)
from .nca import PetrovGalerkinNCA, PetrovGalerkinNCACore, H2Workspace, H2Graph, BatchedNCA, mul, transpose, \
    adjoint, ConcurrentWorkspaceError, h2desc
from . import geometry
