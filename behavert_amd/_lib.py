"""ctypes binding of libbehavert_amd.so (the C ABI in include/behavert_amd.h).

The library is built in-tree by behavert_amd/build.py.  Loading fails loudly when
it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

from .types import BatchConfig, BcParams, FlockDesc, GridParams, ProxParams, SnapshotInfo, SteerParams, StepBuffers

# BRT_LIB names another build of the same library (e.g. one compiled with
# different launch constants); the default is the in-tree build.
LIB_PATH = os.environ.get("BRT_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                      "libbehavert_amd.so")

_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float
_I = C.c_int
_GP = C.POINTER(GridParams)
_PP = C.POINTER(ProxParams)
_SP = C.POINTER(SteerParams)

# name -> (restype, argtypes); every symbol include/behavert_amd.h declares
SIGNATURES = {
    "brt_context_create": (_I, [_I, C.POINTER(_P)]),
    "brt_context_destroy": (_I, [_P]),
    "brt_context_last_error": (C.c_char_p, [_P]),
    "brt_version": (C.c_char_p, []),
    "brt_context_set_option": (_I, [_P, _I, _I]),
    "brt_context_get_option": (_I, [_P, _I, C.POINTER(_I)]),
    "brt_context_step_shape": (_I, [_P, _PP, C.POINTER(_I), C.POINTER(_I)]),
    "brt_context_check": (_I, [_P]),
    "brt_grid_setup": (_I, [C.POINTER(_F), C.POINTER(_F), C.POINTER(_F), _U32, _GP]),
    "brt_list_stride": (_U32, [_U32]),
    "brt_radius_threshold": (_F, [_F]),
    "brt_envgrid_calc_hash": (_I, [_P, _GP, _P, _U32, _P, _P]),
    "brt_envgrid_sort": (_I, [_P, _P, _U32, _U32, _P]),
    "brt_envgrid_hash_sort": (_I, [_P, _GP, _P, _U32, _P, _P]),
    "brt_envgrid_reorder": (_I, [_P, _P, _U32, _U32, C.POINTER(_P), C.POINTER(_P), _P]),
    "brt_prox_cell_ranges": (_I, [_P, _P, _U32, _U32, _P, _P, _P]),
    "brt_envgrid_reorder_ranges": (_I, [_P, _P, _U32, _U32, C.POINTER(_P), C.POINTER(_P), _U32, _P, _P, _P]),
    "brt_prox_neighbors": (_I, [_P, _GP, _PP, _P, _P, _P, _P, _U32, _P, _P]),
    "brt_osw_steer_integrate": (_I, [_P, _GP, _PP, _SP, _F, _U32, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _P, _P]),
    "brt_bc_feature_update": (_I, [_P, _PP, _F, _U32, _P, _P, _P, _P, _P]),
    "brt_osw_steer_forces": (_I, [_P, _PP, _SP, _F, _U32, _U32, _P, _P, _P, _P, _P, _P, _P]),
    "brt_osw_apply_steering_force": (_I, [_P, _GP, _SP, _F, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "brt_bc_neigh_similarities": (_I, [_P, _PP, C.POINTER(BcParams), _U32, _P, _P, _P, _P, _P, _P, _P]),
    "brt_bc_steer_for_clustering": (_I, [_P, _PP, _SP, C.POINTER(BcParams), _U32, _P, _P, _P, _P, _P, _P, _P]),
    "brt_bc_similarities_steer": (_I, [_P, _PP, _SP, C.POINTER(BcParams), _U32, _P, _P, _P, _P, _P, _P, _P, _P,
                                       _P, _P, _P]),
    "brt_bc_cluster_membership": (_I, [_P, _PP, _U32, _U32, _P, _P, _P, _P, _P, _P, _P]),
    "brt_snapshot_save": (_I, [_P, C.c_char_p, _GP, _PP, _U32, _U32, C.POINTER(StepBuffers), _P]),
    "brt_snapshot_read_info": (_I, [C.c_char_p, C.POINTER(SnapshotInfo)]),
    "brt_snapshot_load": (_I, [_P, C.c_char_p, _U32, C.POINTER(ProxParams), C.POINTER(StepBuffers),
                              C.POINTER(SnapshotInfo), _P]),
    "brt_osw_steer_and_slow_to_avoid_neighbors": (_I, [_P, _PP, _SP, _U32, _P, _P, _P, _P, _P, _P, _P]),
    "brt_draw_neighborhood_color": (_I, [_P, _PP, C.POINTER(_F), _U32, _U32, _P, _P, _P, _P]),
    "brt_shapes_move_towards_target": (_I, [_P, _I, C.POINTER(_F), _U32, _P, _P, _P, _P, _P, _P]),
    "brt_bc_cluster_hit_rate": (_I, [_P, _PP, _F, _U32, _P, _P, _P, _P, _P, _P]),
    "brt_osw_adjoint_step": (_I, [_P, _GP, _PP, _SP, _F, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                  _P, _P]),
    "brt_step": (_I, [_P, _GP, _PP, _SP, _F, _F, _U32, C.POINTER(StepBuffers), _P]),
    "brt_step_noalloc": (_I, [_P, _GP, _PP, _SP, _F, _F, _U32, C.POINTER(StepBuffers), _P]),
    "brt_phase_timing_enable": (_I, [_P, _I]),
    "brt_phase_times": (_I, [_P, C.POINTER(C.c_double), C.POINTER(_U32)]),
    "brt_batch_create": (_I, [_P, C.POINTER(FlockDesc), _U32, _PP, C.POINTER(_P)]),
    "brt_batch_destroy": (_I, [_P]),
    "brt_batch_layout": (_I, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32)]),
    "brt_batch_set_config": (_I, [_P, C.POINTER(BatchConfig)]),
    "brt_batch_get_config": (_I, [_P, C.POINTER(BatchConfig)]),
    "brt_batch_step": (_I, [_P, _SP, _F, _F, _U32, C.POINTER(StepBuffers), _P]),
    "brt_batch_autotune": (_I, [_P, _SP, _F, C.POINTER(StepBuffers), _U32, C.c_char_p, C.POINTER(BatchConfig), _P]),
}

_lib = None


class BrtError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"behavert_amd error {code}: {message}")
        self.code = code


def is_built() -> bool:
    return os.path.exists(LIB_PATH)


def lib() -> C.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not is_built():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python behavert_amd/build.py` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
