# MI355X backend of NestedCrossApproximation.jl: the batched bottom-up
# compressor (the seams assemble_level / compresslevel_test! /
# compresslevel_trial!, src/compressors/batchedbottomup.jl:80-244), the
# coupling store (assemble_couplingstore, src/couplingmatrices.jl:11-53) and
# the H^2 product behind LinearMaps._unsafe_mul! (src/nca/petrovgalerkinnca.jl:415-589)
# on the HIP kernels of libnca_hip.so, through plain ccall on pointers and
# sizes (include/nca_hip.h). Julia owns the host arrays and keeps them alive
# with GC.@preserve; the library owns device memory behind handles released
# by finalizers.
module NestedCrossApproximationHIPExt

using NestedCrossApproximation
const NCA = NestedCrossApproximation
using NestedCrossApproximation: SingularCrossMatrixError, BasisStore, TransferStore, CouplingStore,
    CouplingTraversalPlan, FarData
import NestedCrossApproximation: assemble_level, compresslevel_test!, compresslevel_trial!,
    assemble_couplingstore
using LinearMaps, H2Trees, AdaptiveCrossApproximation, StaticArrays
using NCAHIP: libnca_hip

include("runtime.jl")        # status codes, context, points, the far-field kernel matrix
include("compresslevel.jl")  # seams 1 and 2: level assembly, ACA / IACA + skeleton solve
include("couplingstore.jl")  # seam 3: coupling blocks on the device
include("operator.jl")       # seam 4: the device operator and mul!
include("directional.jl")    # directional (wideband) operators
include("device.jl")         # trees, partition and Phase 1 on the device
include("resident.jl")       # factors kept on the device: fused BottomUp side, TopDown, resident operator

end # module
