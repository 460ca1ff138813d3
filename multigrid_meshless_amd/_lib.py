"""ctypes loader for libmgm.so (the HIP implementation).

Raises ImportError-like RuntimeError when the library is not built: there is no
CPU fallback for any operation exposed here.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmgm.so")
_LIB = None


class Config(C.Structure):
    """mgm_config (mirror of struct parameters, header_files/structures.h:15-41)."""
    _fields_ = [("dimension", C.c_int), ("poly_degree", C.c_int), ("coarse_poly_degree", C.c_int),
                ("phs_degree", C.c_int), ("cloud_size_multiplier", C.c_double),
                ("omega", C.c_float), ("num_relax", C.c_int), ("transfer_rows", C.c_int),
                ("sweep_order", C.c_int), ("boundary_cloud_size_multiplier", C.c_double)]


class LevelInfo(C.Structure):
    _fields_ = [("N", C.c_int), ("nb", C.c_int), ("m", C.c_int), ("n", C.c_int),
                ("poly_degree", C.c_int), ("num_colours", C.c_int), ("d_avg", C.c_double),
                ("d_avg_bbox", C.c_double), ("knn_ties", C.c_int), ("device_rows", C.c_int),
                ("slice_height", C.c_int), ("blocked", C.c_int), ("gather_sorted", C.c_int),
                ("small_relax", C.c_int), ("window_blocks", C.c_int), ("window_max", C.c_int),
                ("prolong_window_max", C.c_int)]


class NsParams(C.Structure):
    """mgm_ns_params: the scalars the two time steps read (struct parameters and FieldVariables.rho)."""
    _fields_ = [("dt", C.c_double), ("rho", C.c_double), ("mu", C.c_double), ("nu", C.c_double),
                ("facRe", C.c_double), ("facdt", C.c_double), ("field_rho", C.c_double),
                ("iter_momentum", C.c_int), ("iter_timple", C.c_int), ("num_vcycles", C.c_int)]


class TrainConfig(C.Structure):
    """mgm_train_config."""
    _fields_ = [("snapshots", C.c_int), ("samples_per_snapshot", C.c_int), ("epochs", C.c_int),
                ("batch", C.c_int), ("lr", C.c_float), ("seed", C.c_uint64)]


# (name, restype, argtypes) for every symbol declared in include/mgm.h
_vp, _i, _d, _f, _cp = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_char_p
SIGNATURES = [
    ("mgm_last_error", C.c_char_p, []),
    ("mgm_device_count", _i, []),
    ("mgm_init", _i, [_i]),
    ("mgm_config_default", None, [C.POINTER(Config)]),
    ("mgm_poly_terms", _i, [_i, _i, _vp, _vp, _vp]),
    ("mgm_points_read_gmsh", _i, [_cp, _i, C.POINTER(_vp)]),
    ("mgm_points_generate", _i, [_i, _i, C.c_uint64, _d, _i, _cp, C.POINTER(_vp)]),
    ("mgm_points_from_arrays", _i, [_i, _i, _i, _d] + [_vp] * 8 + [C.POINTER(_vp)]),
    ("mgm_points_info", _i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]),
    ("mgm_points_copy", _i, [_vp] + [_vp] * 8),
    ("mgm_points_free", None, [_vp]),
    ("mgm_splitmix64", C.c_uint64, [C.POINTER(C.c_uint64)]),
    ("mgm_problem_create", _i, [C.POINTER(Config), _i, C.POINTER(_vp)]),
    ("mgm_problem_destroy", None, [_vp]),
    ("mgm_problem_set_level", _i, [_vp, _i, _vp]),
    ("mgm_problem_setup", _i, [_vp]),
    ("mgm_level_info_get", _i, [_vp, _i, C.POINTER(LevelInfo)]),
    ("mgm_setup_timings", _i, [_vp, _vp]),
    ("mgm_rebuild_weights", _i, [_vp, _i, C.POINTER(_d), C.POINTER(C.c_long)]),
    ("mgm_weight_kernel_choice", _i, [_i, C.POINTER(_i), C.POINTER(_i)]),
    ("mgm_get_geometry", _i, [_vp, _i] + [_vp] * 7),
    ("mgm_get_cloud", _i, [_vp, _i, _vp]),
    ("mgm_get_points", _i, [_vp, _i, _i, _vp]),
    ("mgm_get_colours", _i, [_vp, _i, _vp]),
    ("mgm_get_matrix", _i, [_vp, _i, _i, _vp]),
    ("mgm_set_matrix", _i, [_vp, _i, _i, _vp]),
    ("mgm_dirichlet_rows", _i, [_vp]),
    ("mgm_get_field", _i, [_vp, _i, _i, _vp]),
    ("mgm_write_solution", _i, [_vp, _i, C.c_char_p, _i, _i, _vp]),
    ("mgm_write_history", _i, [C.c_char_p, _i, _vp, _i]),
    ("mgm_set_field", _i, [_vp, _i, _i, _vp]),
    ("mgm_zero_fields", _i, [_vp]),
    ("mgm_set_smoother", _i, [_vp, _f, _i]),
    ("mgm_relax", _i, [_vp, _i, _i]),
    ("mgm_residual", _i, [_vp, _i, _i]),
    ("mgm_restrict", _i, [_vp, _i, _i]),
    ("mgm_prolong", _i, [_vp, _i, _i]),
    ("mgm_vcycle", _i, [_vp, _i, _i]),
    ("mgm_spmv", _i, [_vp, _i, _i, _i, _i]),
    ("mgm_spmv_wide", _i, [_vp, _i, _i, _i, _i]),
    ("mgm_boundary_info", _i, [_vp, _i, _vp, _vp]),
    ("mgm_get_boundary_rows", _i, [_vp, _i, _i, _vp, _vp, _vp]),
    ("mgm_l2_norm_gen", _i, [_vp, _i, _i, _i, C.POINTER(_d)]),
    ("mgm_spmv_rows", _i, [_i, _i, _vp, _vp, _i, _vp, _vp]),
    ("mgm_spmv_rows_static", _i, [_i, _i, _vp, _vp, _i, _vp, _vp]),
    ("mgm_ns_set_params", _i, [_vp, _vp]),
    ("mgm_fs_step", _i, [_vp, C.POINTER(_d)]),
    ("mgm_timple_step", _i, [_vp, C.POINTER(_d)]),
    ("mgm_ns_march", _i, [_vp, _i, _i, _i, _d, _vp, C.POINTER(_i), C.POINTER(_d)]),
    ("mgm_mlp_inputs", _i, [_vp]),
    ("mgm_mlp_weight_count", C.c_size_t, [_i]),
    ("mgm_mlp_set_weights", _i, [_vp, _vp]),
    ("mgm_mlp_get_weights", _i, [_vp, _vp]),
    ("mgm_mlp_forward", _i, [_vp, _i, _vp, _vp]),
    ("mgm_mlp_gradient", _i, [_vp, _i, _vp, _vp, C.POINTER(_d), _vp]),
    ("mgm_mlp_features", _i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    ("mgm_mlp_fit", _i, [_vp, _i, _vp, _vp, _i, _i, _f, C.c_uint64, C.POINTER(_d), C.POINTER(_d)]),
    ("mgm_train_smoother", _i, [_vp, C.POINTER(TrainConfig), C.POINTER(_d), C.POINTER(_d)]),
    ("mgm_use_learned_smoother", _i, [_vp, _i]),
    ("mgm_learned_fell_back", _i, [_vp]),
    ("mgm_solve", _i, [_vp, _i, _i, _d, _i, C.POINTER(_i), _vp, C.POINTER(_d)]),
    ("mgm_tune_omega", _i, [_vp, _i, _i, _d, _i, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float),
                            C.POINTER(_i)]),
    ("mgm_tune_omega_levels", _i, [_vp, _i, _i, _d, _i, C.c_float, C.c_float, C.c_float, _i, _vp,
                                   C.POINTER(_i)]),
    ("mgm_set_smoother_levels", _i, [_vp, _vp, _i]),
    ("mgm_get_smoother_levels", _i, [_vp, _vp]),
    ("mgm_set_smoother_kinds", _i, [_vp, _vp, _i]),
    ("mgm_set_learned_levels", _i, [_vp, _vp, _i]),
    ("mgm_get_smoother_kinds", _i, [_vp, _vp]),
    ("mgm_tune_smoother", _i, [_vp, _i, _i, _d, _i, C.c_float, C.c_float, C.c_float, _i, _i, _vp, _vp,
                               C.POINTER(_i)]),
    ("mgm_stream_copy", _i, [C.c_size_t, _i, C.POINTER(_d)]),
    ("mgm_time_spmv", _i, [_vp, _i, _i, C.POINTER(_d)]),
]


def lib():
    """Load libmgm.so; raise if it is not built (no fallback exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run python multigrid_meshless_amd/_build.py")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class MgmError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().mgm_last_error().decode(errors="replace")
        raise MgmError(f"{what} failed ({rc}): {msg}")
