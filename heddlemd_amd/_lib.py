"""ctypes binding of libheddle.so (the C ABI declared in include/heddle.h).

The library is built in-tree by ``heddlemd_amd/csrc/Makefile`` (hipcc,
gfx950).  There is no fallback: if the shared library is missing or fails to
load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libheddle.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "heddle.h")

_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_ctx = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "hd_version": (C.c_char_p, []),
    "hd_ctx_create": (C.c_int, [C.c_int, C.POINTER(_ctx)]),
    "hd_ctx_destroy": (None, [_ctx]),
    "hd_last_error": (C.c_char_p, [_ctx]),
    "hd_set_system": (C.c_int, [_ctx, C.c_uint64, _f32p, _f32p, _u32p, _vp]),
    "hd_set_lattice": (C.c_int, [_ctx, _f64p]),
    "hd_set_positions": (C.c_int, [_ctx, _f32p]),
    "hd_neighbor_pre_step": (C.c_int, [_ctx, C.POINTER(C.c_uint32)]),
    "hd_set_fused_settle": (C.c_int, [_ctx, C.c_int]),
    "hd_set_small_neighbor": (C.c_int, [_ctx, C.c_int]),
    "hd_set_velocities": (C.c_int, [_ctx, _f32p, _f32p, _f32p]),
    "hd_set_images": (C.c_int, [_ctx, _vp, _vp, _vp]),
    "hd_set_lj": (C.c_int, [_ctx, C.c_int, _f64p, _f64p, _f64p, _f64p]),
    "hd_set_spme": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_double]),
    "hd_set_spme_dft": (C.c_int, [_ctx, C.c_int]),
    "hd_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "hd_batch_destroy": (None, [C.c_void_p]),
    "hd_batch_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "hd_batch_set_lanes": (C.c_int, [C.c_void_p, C.c_int]),
    "hd_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "hd_set_exclusion_span": (C.c_int, [_ctx, C.c_uint32]),
    "hd_set_angles": (C.c_int, [_ctx, C.c_uint64, _u32p, _f64p, _f64p]),
    "hd_set_andersen": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_uint64]),
    "hd_set_csvr": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_uint64, C.c_uint64]),
    "hd_set_berendsen": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_uint64]),
    "hd_get_thermostat": (C.c_int, [_ctx, _f64p]),
    "hd_set_berendsen_barostat": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double]),
    "hd_minimize_begin": (C.c_int, [_ctx, C.c_double, _f64p]),
    "hd_minimize_step": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double, _f64p]),
    "hd_set_molecules": (C.c_int, [_ctx, C.c_uint64, _u32p, _u32p]),
    "hd_mc_barostat_move": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint64, _f64p]),
    "hd_set_bonds": (C.c_int, [_ctx, C.c_uint64, _u32p, C.POINTER(C.c_int), _f64p]),
    "hd_set_dihedrals": (C.c_int, [_ctx, C.c_uint64, _u32p, _f64p, _f64p, _f64p]),
    "hd_set_exclusions": (C.c_int, [_ctx, C.c_uint64, _u32p, _f64p, _f64p]),
    "hd_set_neighbor": (C.c_int, [_ctx, C.c_double]),
    "hd_set_pair_list": (C.c_int, [_ctx, C.c_int]),
    "hd_set_settle": (C.c_int, [_ctx, C.c_uint64, _u32p, C.c_double, C.c_double, C.c_double,
                                C.c_double]),
    "hd_set_integrator": (C.c_int, [_ctx, C.c_double, C.c_int]),
    "hd_set_nhc": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                             C.c_uint64]),
    "hd_set_crescale": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_uint64]),
    "hd_set_langevin": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_uint64]),
    "hd_set_shake": (C.c_int, [_ctx, C.c_uint64, _u32p, _u32p, _u32p, _u8p, _u8p, _f64p]),
    "hd_set_shake_tolerance": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_int]),
    "hd_get_constraint_stats": (C.c_int, [_ctx, _u64p]),
    "hd_get_langevin_counter": (C.c_int, [_ctx, C.POINTER(C.c_uint64)]),
    "hd_set_shake_predictor": (C.c_int, [_ctx, _vp]),
    "hd_get_shake_multipliers": (C.c_int, [_ctx, _vp, _vp]),
    "hd_shake_log_start": (C.c_int, [_ctx, C.c_uint64]),
    "hd_shake_log_stop": (C.c_int, [_ctx]),
    "hd_shake_log_size": (C.c_int, [_ctx, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "hd_shake_log_read": (C.c_int, [_ctx, C.c_uint64, C.c_uint64, _f32p, _f32p]),
    "hd_mlp_param_count": (C.c_int, []),
    "hd_mlp_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "hd_mlp_destroy": (None, [C.c_void_p]),
    "hd_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "hd_mlp_set_params": (C.c_int, [C.c_void_p, _f32p]),
    "hd_mlp_get_params": (C.c_int, [C.c_void_p, _f32p]),
    "hd_mlp_predict": (C.c_int, [C.c_void_p, _f32p, C.c_uint64, _f32p]),
    "hd_mlp_gradient": (C.c_int, [C.c_void_p, _f32p, _f32p, C.c_uint64, _f32p,
                                  C.POINTER(C.c_double)]),
    "hd_mlp_gradient_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                          C.c_void_p, C.c_void_p]),
    "hd_mlp_train": (C.c_int, [C.c_void_p, _f32p, _f32p, C.c_uint64, C.c_int, C.c_int,
                               C.c_float, C.c_uint64, _vp]),
    "hd_mlp_train_log": (C.c_int, [C.c_void_p, _ctx, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                   _vp]),
    "hd_setup": (C.c_int, [_ctx]),
    "hd_force_eval": (C.c_int, [_ctx, C.c_int]),
    "hd_constrain_velocities": (C.c_int, [_ctx]),
    "hd_step": (C.c_int, [_ctx, C.c_int, C.c_int]),
    "hd_negate_velocities": (C.c_int, [_ctx]),
    "hd_use_graph": (C.c_int, [_ctx, C.c_int]),
    "hd_sync": (C.c_int, [_ctx]),
    "hd_get_positions": (C.c_int, [_ctx, _f32p, _vp, _vp, _vp]),
    "hd_get_velocities": (C.c_int, [_ctx, _f32p, _f32p, _f32p]),
    "hd_get_forces": (C.c_int, [_ctx, _f32p, _f32p, _f32p]),
    "hd_get_per_atom": (C.c_int, [_ctx, _f32p, _f32p]),
    "hd_get_lattice": (C.c_int, [_ctx, _f64p]),
    "hd_get_scalars": (C.c_int, [_ctx, _f64p]),
    "hd_get_pairs": (C.c_int, [_ctx, _vp, C.c_uint64, C.POINTER(C.c_uint64)]),
    "hd_count_pairs": (C.c_int, [_ctx, C.POINTER(C.c_uint64)]),
    "hd_get_barostat": (C.c_int, [_ctx, _f64p]),
    "hd_get_neighbor_stats": (C.c_int, [_ctx, _u64p]),
    "hd_get_spme_grids": (C.c_int, [_ctx, _vp, _vp]),
    "hd_selftest_mfma": (C.c_int, [_ctx, _f32p]),
    "hd_selftest_mfma_k16": (C.c_int, [_ctx, C.c_int, _f32p]),
    "hd_set_mtk": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                             C.c_int, C.c_int, C.c_uint64]),
    "hd_get_mtk": (C.c_int, [_ctx, _f64p]),
    "hd_add_settle": (C.c_int, [_ctx, C.c_uint64, _u32p, C.c_double, C.c_double, C.c_double,
                                C.c_double]),
    "hd_selftest_dpp": (C.c_int, [_ctx, _i32p]),
    "hd_stage_times": (C.c_int, [_ctx, C.c_char_p, C.c_int, _f64p, C.c_int]),
    "hd_set_timing": (C.c_int, [_ctx, C.c_int]),
}


class HeddleError(RuntimeError):
    pass


def load():
    """Load libheddle.so (raises if it is missing: no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HeddleError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                          "(make -C heddlemd_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    lib = load()
    return [n for n in SIGNATURES if hasattr(lib, n)]
