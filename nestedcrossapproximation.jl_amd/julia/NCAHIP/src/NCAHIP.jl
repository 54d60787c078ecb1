# The native library of the MI355X path: libnca_hip.so (HIP kernels for gfx950
# and the C-ABI of include/nca_hip.h), built by nestedcrossapproximation.jl_amd/build.py.
# Its location comes from NCA_HIP_LIB, else the build directory next to this package.
module NCAHIP

using Libdl

const libnca_hip = get(ENV, "NCA_HIP_LIB",
    normpath(joinpath(@__DIR__, "..", "..", "..", "lib", "libnca_hip.so")))

function __init__()
    isfile(libnca_hip) || error("NCAHIP: $(libnca_hip) is missing; build it with build.py or set NCA_HIP_LIB")
    Libdl.dlopen(libnca_hip)     # fail loudly here, not at the first ccall
    return nothing
end

end # module
