"""ctypes declarations of include/nca_hip.h."""
import ctypes

P = ctypes.c_void_p
PP = ctypes.POINTER(ctypes.c_void_p)
I32 = ctypes.c_int32
I64 = ctypes.c_int64
D = ctypes.c_double

NCA_OK = 0
NCA_ERR_ARGUMENT = 1
NCA_ERR_DIMENSION = 2
NCA_ERR_HIP = 3
NCA_ERR_SINGULAR = 4
NCA_ERR_LIMIT = 5

NCA_OP_N, NCA_OP_T, NCA_OP_C = 0, 1, 2
NCA_PREC_FP64, NCA_PREC_MIXED = 0, 1


class Helmholtz(ctypes.Structure):
    _fields_ = [("wavenumber", D), ("sign", I32), ("reserved", I32)]


class H2Desc(ctypes.Structure):
    _fields_ = [
        ("M", I64), ("N", I64), ("nn_test", I64), ("nn_trial", I64),
        ("test_perm", P), ("test_lo", P), ("test_hi", P),
        ("trial_perm", P), ("trial_lo", P), ("trial_hi", P),
        ("test_cptr", P), ("trial_cptr", P),
        ("nbasis_test", I64), ("nbasis_trial", I64),
        ("basis_test_nodes", P), ("basis_test_off", P), ("basis_trial_nodes", P), ("basis_trial_off", P),
        ("basis_test", P), ("basis_trial", P),
        ("nlev_test", I64), ("nlev_trial", I64),
        ("tt_levptr", P), ("tt_levidcs", P), ("tt_chdptr", P), ("tt_chds", P), ("tt_off", P),
        ("st_levptr", P), ("st_levidcs", P), ("st_chdptr", P), ("st_chds", P), ("st_off", P),
        ("tt_blocks", P), ("st_blocks", P),
        ("cp_ptr", P), ("cp_nodes", P),
        ("nearptr", P), ("nearnodes", P),
    ]


class DH2Desc(ctypes.Structure):
    _fields_ = [
        ("M", I64), ("N", I64), ("nn_test", I64), ("nn_trial", I64), ("ne_test", I64), ("ne_trial", I64),
        ("test_perm", P), ("test_lo", P), ("test_hi", P),
        ("trial_perm", P), ("trial_lo", P), ("trial_hi", P),
        ("test_dirptr", P), ("trial_dirptr", P),
        ("test_cptr", P), ("trial_cptr", P),
        ("nbnodes_test", I64), ("nbnodes_trial", I64),
        ("bt_nodes", P), ("bt_dirptr", P), ("bt_diridx", P), ("bt_off", P),
        ("bs_nodes", P), ("bs_dirptr", P), ("bs_diridx", P), ("bs_off", P),
        ("bt_blocks", P), ("bs_blocks", P),
        ("nlev_test", I64), ("nlev_trial", I64),
        ("tt_levptr", P), ("tt_levidcs", P), ("tt_chdptr", P), ("tt_chds", P), ("tt_off", P),
        ("st_levptr", P), ("st_levidcs", P), ("st_chdptr", P), ("st_chds", P), ("st_off", P),
        ("tt_blocks", P), ("st_blocks", P),
        ("cp_nodeptr", P), ("cp_tidcs", P), ("cp_sidcs", P),
        ("nearptr", P), ("nearnodes", P),
    ]


class H2Stats(ctypes.Structure):
    _fields_ = [(n, I64) for n in (
        "coupling_blocks", "coupling_entries", "basis_entries_test", "basis_entries_trial",
        "transfer_entries_test", "transfer_entries_trial", "near_pairs", "near_entries",
        "coeffs_test", "coeffs_trial", "plan_bytes")]


class H2Tuning(ctypes.Structure):
    _fields_ = [(n, I32) for n in ("near_k3", "leaf_order", "near_xcd_order", "couple_k3", "couple_tile_order",
                                   "couple_xcd_order", "candidates", "near_stagger")] + \
               [(n, ctypes.c_double) for n in ("near_ms", "couple_ms", "near_ms_start", "couple_ms_start")]


SIGNATURES = {
    "nca_last_error": (ctypes.c_char_p, []),
    "nca_abi_version": (I32, []),
    "nca_ctx_create": (I32, [I32, PP]),
    "nca_ctx_destroy": (I32, [P]),
    "nca_ctx_synchronize": (I32, [P]),
    "nca_ctx_stream": (I32, [P, PP]),
    "nca_ctx_device_name": (I32, [P, P, I64]),
    "nca_points_create": (I32, [P, P, I64, PP]),
    "nca_points_destroy": (I32, [P]),
    "nca_level_assemble": (I32, [P, P, P, P, I64, P, P, P, P, I64, I32, PP]),
    "nca_stack_upload": (I32, [P, P, I64, P, P, I64, PP]),
    "nca_stack_shape": (I32, [P, P, P, P]),
    "nca_stack_download": (I32, [P, P, P]),
    "nca_stack_destroy": (I32, [P]),
    "nca_aca_fullpiv": (I32, [P, P, D, I32, P, P, P]),
    "nca_skeleton_solve": (I32, [P, P, P, P, P, P, P, P]),
    "nca_compress_level": (I32, [P, P, D, I32, P, P, P, I64, P, P, P]),
    "nca_iaca": (I32, [P, P, D, P, P, P]),
    "nca_compress_level_iaca": (I32, [P, P, D, P, P, P, I64, P, P, P]),
    "nca_coupling_assemble": (I32, [P, P, P, P, I64, P, P, I64, P, P, P, P, PP]),
    "nca_farpivots": (I32, [P, P, I64, P, P, P, P, I64, P, P, P, I64, D, D, P, I64, P]),
    "nca_farpivots_tree": (I32, [P, P, I64, P, P, P, P, P, I64, P, P, P, I64, D, D, P, I64, P]),
    "nca_interactions": (I32, [P, I64, P, P, P, P, I64, P, P, P, P, I32, P, P, P, PP]),
    "nca_partition_export": (I32, [P, P, P, P, P]),
    "nca_partition_destroy": (I32, [P]),
    "nca_factors_create": (I32, [P, I64, PP]),
    "nca_factors_destroy": (I32, [P]),
    "nca_compress_level_resident": (I32, [P, P, D, I32, I32, P, P, P, P, P, P]),
    "nca_factors_size": (I32, [P, P]),
    "nca_factors_download": (I32, [P, P, P]),
    "nca_bottomup_side": (I32, [P, P, P, P, I32, I64, P, P, P, P, P, I64, P, P, P, P, I64, D, I32, I32, P, P, P, I64,
                                P, P]),
    "nca_topdown_side": (I32, [P, P, P, P, I32, I64, P, P, P, P, P, P, I64, P, P, P, P, P, I64, P, P, I64, D, P, P,
                               P, I64, P, P, P]),
    "nca_h2_create_resident": (I32, [P, P, P, P, P, P, I32, P, P, PP]),
    "nca_dh2_create_resident": (I32, [P, P, P, P, P, P, I32, P, P, PP]),
    "nca_tree_twon": (I32, [P, P, I64, D, P, PP]),
    "nca_tree_kmeans": (I32, [P, P, I64, I64, P, PP]),
    "nca_tree_export": (I32, [P, P, P, P, P, P, P, P, P, P]),
    "nca_tree_destroy": (I32, [P]),
    "nca_coupling_size": (I32, [P, P, P]),
    "nca_coupling_download": (I32, [P, P, P, P]),
    "nca_coupling_destroy": (I32, [P]),
    "nca_h2_create": (I32, [P, P, P, P, P, P, I32, PP]),
    "nca_dh2_create": (I32, [P, P, P, P, P, P, I32, PP]),
    "nca_h2_destroy": (I32, [P]),
    "nca_h2_set_precision": (I32, [P, I32]),
    "nca_h2_matvec": (I32, [P, I32, I64, P, I64, P, I64, P, P]),
    "nca_h2_matvec_device": (I32, [P, I32, I64, P, I64, P, I64, P, P]),
    "nca_h2_matvec_device_noalloc": (I32, [P, I32, I64, P, I64, P, I64, P, P]),
    "nca_h2_store_near": (I32, [P, I32]),
    "nca_h2_matvec_profile": (I32, [P, I32, I64, P, I64, P, I64, P, P, P]),
    "nca_h2_autotune": (I32, [P, I32, P]),
    "nca_h2_tuning_apply": (I32, [P, P]),
    "nca_h2_graph_create": (I32, [P, I32, I64, P, I64, P, I64, P, P, PP]),
    "nca_h2_graph_launch": (I32, [P]),
    "nca_h2_graph_destroy": (I32, [P]),
    "nca_h2_get_stats": (I32, [P, P]),
    "nca_h2_near_form": (I32, [P, P, P, P]),
    "nca_h2_couple_form": (I32, [P, P, P, P]),
    "nca_h2_rel32": (I32, [P, I32, P, P, P, P]),
}


def declare(lib):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


class NCAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class SingularCrossMatrixError(NCAError):
    pass


class DimensionMismatch(NCAError):
    pass


class ArgumentError(NCAError):
    pass


def check(lib, status):
    if status == NCA_OK:
        return
    msg = lib.nca_last_error().decode(errors="replace")
    cls = {NCA_ERR_SINGULAR: SingularCrossMatrixError, NCA_ERR_DIMENSION: DimensionMismatch,
           NCA_ERR_ARGUMENT: ArgumentError}.get(status, NCAError)
    raise cls(status, msg)
