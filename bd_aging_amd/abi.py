"""ctypes view of include/bd_abi.h and the loader of libbd_hip.so.

The library is built in-tree by ``bd_aging_amd.build`` (hipcc, gfx950).  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbd_hip.so")

BD_OK = 0
BD_E_ARG = 1
BD_E_HIP = 2
BD_E_CELL_OVERFLOW = 3
BD_E_NAN = 4
BD_E_STATE = 5
BD_E_NOMEM = 6

BD_POT_WCA = 0
BD_POT_WCA_MODIFIED = 1
BD_POT_HARMONIC = 2
BD_POT_WCA49_50 = 3
BD_POT_WCA49_50_AO = 4
BD_ORDER_REFERENCE = 1
BD_ORDER_SPLIT4 = 4
BD_ORDER_SPLIT8 = 8
BD_WALK_ROWS = 0
BD_WALK_SLABS = 1
BD_WALK_CHUNK2 = 2
BD_WALK_CHUNK4 = 3
BD_WALK_LIST = 4
BD_WALK_FLAT = 5
BD_TEMP_LINEAR = 0
BD_TEMP_SINE = 1


class BDParams(C.Structure):
    """bd_params (FP64 re-expression of parametros, encabezados.h:40-75)."""
    _fields_ = [
        ("n_big", C.c_int32), ("n_sml", C.c_int32), ("potential", C.c_int32),
        ("ntags_big", C.c_int32), ("ntags_sml", C.c_int32),
        ("temp_protocol", C.c_int32), ("rng_lanes", C.c_int32),
        ("reserved", C.c_int32),
        ("side", C.c_double), ("sigma_big", C.c_double), ("sigma_sml", C.c_double),
        ("dt", C.c_double), ("kappa", C.c_double),
        ("T0", C.c_double), ("Tf", C.c_double), ("t0", C.c_double), ("tf", C.c_double),
        ("period", C.c_double), ("seed", C.c_uint64), ("first_step", C.c_int64),
        ("pot_amp", C.c_double), ("pot_inv_t", C.c_double),
        ("sigma_pol", C.c_double), ("phi_pol", C.c_double),
    ]


class BDGeometry(C.Structure):
    """bd_geometry (main.cu:259-298)."""
    _fields_ = [
        ("ncell_big", C.c_int32), ("ncell_sml", C.c_int32),
        ("nrange_bb", C.c_int32), ("nrange_bs", C.c_int32),
        ("nrange_sb", C.c_int32), ("nrange_ss", C.c_int32),
        ("cell_side_big", C.c_double), ("cell_side_sml", C.c_double),
        ("rc_bb", C.c_double), ("rc_bs", C.c_double), ("rc_ss", C.c_double),
    ]


class BDObs(C.Structure):
    """bd_obs (fself.cu:3-21, main.cu:627-638,776-806)."""
    _fields_ = [
        ("fself_big", C.c_double), ("msd_big", C.c_double), ("msd_sml", C.c_double),
        ("epot_big", C.c_double), ("epot_sml", C.c_double), ("virial", C.c_double),
        ("z", C.c_double), ("temperature", C.c_double),
    ]


# symbol -> (restype, argtypes); this list is the full set declared in
# include/bd_abi.h and is checked against the header by the CPU tests.
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint64)
_F = C.POINTER(C.c_float)
SIGNATURES = {
    "bd_compute_geometry": (C.c_int, [C.POINTER(BDParams), C.POINTER(BDGeometry)]),
    "bd_create": (C.c_int, [C.POINTER(BDParams), C.c_int32, C.POINTER(C.c_void_p)]),
    "bd_destroy": (C.c_int, [_P]),
    "bd_num_replicas": (C.c_int, [_P, _I]),
    "bd_get_geometry": (C.c_int, [_P, C.c_int32, C.POINTER(BDGeometry)]),
    "bd_set_positions": (C.c_int, [_P, C.c_int32, _D]),
    "bd_get_positions": (C.c_int, [_P, C.c_int32, _D]),
    "bd_get_raw_positions": (C.c_int, [_P, C.c_int32, _D]),
    "bd_get_forces": (C.c_int, [_P, C.c_int32, _D, _D, _D]),
    "bd_set_forces": (C.c_int, [_P, C.c_int32, _D]),
    "bd_mark_origin": (C.c_int, [_P]),
    "bd_get_step": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "bd_cell_locate": (C.c_int, [_P]),
    "bd_get_cells": (C.c_int, [_P, C.c_int32, C.c_int32, _I, _I]),
    "bd_compute_forces": (C.c_int, [_P]),
    "bd_neighbours": (C.c_int, [_P, C.c_int32, C.c_int32, _I, _I]),
    "bd_update": (C.c_int, [_P]),
    "bd_step": (C.c_int, [_P, C.c_int32]),
    "bd_prepare_step": (C.c_int, [_P, C.c_int32]),
    "bd_autotune": (C.c_int, [_P, C.c_int32, C.c_char_p, _I]),
    "bd_set_force_order": (C.c_int, [_P, C.c_int32]),
    "bd_get_force_order": (C.c_int, [_P, _I]),
    "bd_set_persistent": (C.c_int, [_P, C.c_int32]),
    "bd_get_persistent": (C.c_int, [_P, _I]),
    "bd_set_graph_steps": (C.c_int, [_P, C.c_int32]),
    "bd_get_graph_steps": (C.c_int, [_P, _I]),
    "bd_set_walk": (C.c_int, [_P, C.c_int32]),
    "bd_get_walk": (C.c_int, [_P, _I]),
    "bd_set_flat_depth": (C.c_int, [_P, C.c_int32]),
    "bd_set_pass_split": (C.c_int, [_P, C.c_int32]),
    "bd_set_list_skin": (C.c_int, [_P, C.c_double]),
    "bd_get_list_skin": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "bd_get_pass_split": (C.c_int, [_P, _I]),
    "bd_set_fused_update": (C.c_int, [_P, C.c_int32]),
    "bd_get_fused_update": (C.c_int, [_P, _I]),
    "bd_set_defer_depth": (C.c_int, [_P, C.c_int32]),
    "bd_get_defer_depth": (C.c_int, [_P, _I]),
    "bd_set_brick": (C.c_int, [_P, C.c_int32]),
    "bd_get_brick": (C.c_int, [_P, _I]),
    "bd_kernel_times": (C.c_int, [_P, C.c_int32, _D]),
    "bd_kernel_times_fused": (C.c_int, [_P, C.c_int32, _D]),
    "bd_stream_bandwidth": (C.c_int, [C.c_int64, C.c_int32, _D]),
    "bd_grid_barrier_cost": (C.c_int, [C.c_int32, C.c_int32, _D, _I]),
    "bd_launch_floor": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _D]),
    "bd_checkpoint_save": (C.c_int, [_P, C.c_char_p]),
    "bd_checkpoint_load": (C.c_int, [_P, C.c_char_p]),
    "bd_get_rng_lane": (C.c_int, [_P, C.c_int32, C.c_int32, _U, _U, _U]),
    "bd_rng_stream": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, _U]),
    "bd_lane_seed": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "bd_random_positions": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, _D]),
    "bd_gder": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, _U]),
    "bd_max_occupancy": (C.c_int, [_P, C.c_int32, _I, _I]),
    "bd_observables": (C.c_int, [_P, C.c_int32, C.c_double, C.POINTER(BDObs)]),
    "bd_gder_all": (C.c_int, [_P, C.c_int32, C.c_double, _U]),
    "bd_gder_noalloc": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, _U]),
    "bd_set_origin": (C.c_int, [_P, C.c_int32]),
    "bd_fself_origins": (C.c_int, [_P, C.c_double, C.c_int32, _D]),
    "bd_cv3d_create": (C.c_int, [C.c_int32, C.c_double, C.POINTER(C.c_void_p)]),
    "bd_cv3d_destroy": (C.c_int, [_P]),
    "bd_cv3d_add": (C.c_int, [_P, _P, C.c_int32, C.c_int32]),
    "bd_cv3d_add_host": (C.c_int, [_P, _D, _D, C.c_int32, C.c_double]),
    "bd_cv3d_solve": (C.c_int, [_P]),
    "bd_cv3d_get": (C.c_int, [_P, _D]),
    "bd_cv3d_get_normal": (C.c_int, [_P, _D, _D]),
    "bd_cv3d_samples": (C.c_int, [_P, _I]),
    "bd_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                C.POINTER(C.c_void_p)]),
    "bd_mlp_destroy": (C.c_int, [_P]),
    "bd_mlp_num_params": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "bd_mlp_get_params": (C.c_int, [_P, _F]),
    "bd_mlp_set_params": (C.c_int, [_P, _F]),
    "bd_mlp_forward": (C.c_int, [_P, _F, C.c_int32, _F]),
    "bd_mlp_gradients": (C.c_int, [_P, _F, _F, C.c_int32, _F, _F]),
    "bd_mlp_trace": (C.c_int, [_P, C.c_int32, C.c_int32, _F]),
    "bd_mlp_train_step": (C.c_int, [_P, _F, _F, C.c_int32, C.c_float, _F]),
    "bd_mlp_set_data": (C.c_int, [_P, _F, _F, C.c_int32]),
    "bd_mlp_train": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _F]),
    "bd_mlp_predict": (C.c_int, [_P, _F, C.c_int32, _F, _F]),
    "bd_last_error": (C.c_char_p, []),
    "bd_strerror": (C.c_char_p, [C.c_int]),
}

_lib = None


def lib_exists():
    return os.path.exists(LIB_PATH)


def load():
    """Load libbd_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libbd_hip.so is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class BDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bd error {code}: {msg}")
        self.code = code


def check(code):
    if code != BD_OK:
        lib = load()
        raise BDError(code, lib.bd_last_error().decode(errors="replace"))
    return code
