"""ctypes description of include/zombie_amd.h and the loader of libzombie_amd.so.

The library is loaded from the package directory (built in-tree by zombie_amd._build).
There is no fallback: a missing or unloadable library raises ZombieLibraryError.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libzombie_amd.so"

ZG_OK = 0
ZG_ERR_INVALID_ARGUMENT = 1
ZG_ERR_UNSUPPORTED = 2
ZG_ERR_DEVICE = 3
ZG_ERR_OUT_OF_MEMORY = 4
ZG_ERR_NO_DEVICE = 5

ZG_FN_ZERO = 0
ZG_FN_CONSTANT = 1
ZG_FN_EXP_COS = 2
ZG_FN_SIN_PRODUCT = 3
ZG_FN_CUBIC_HARMONIC = 4
ZG_FN_EXP_LINEAR = 5
ZG_FN_DENSE_GRID = 6

ZG_EVAL_VALUE = 0
ZG_EVAL_NORMAL_DERIVATIVE = 1
ZG_EVAL_ROBIN = 3
ZG_EVAL_SCREENED_SOURCE = 2

ZG_SAMPLE_IN_DOMAIN = 0
ZG_SAMPLE_ON_ABSORBING_BOUNDARY = 1
ZG_SAMPLE_ON_REFLECTING_BOUNDARY = 2
ZG_ESTIMATE_SOLUTION = 0
ZG_ESTIMATE_SOLUTION_AND_GRADIENT = 1
ZG_ESTIMATE_SKIP = 2

ZG_REACHED_ABSORBING_BOUNDARY = 0
ZG_TERMINATED_WITH_RUSSIAN_ROULETTE = 1
ZG_EXCEEDED_MAX_WALK_LENGTH = 2
ZG_ESCAPED_DOMAIN = 3

FLT_MAX = 3.4028234663852886e38


class zg_walk_settings(C.Structure):
    _fields_ = [
        ("epsilon_shell_for_absorbing_boundary", C.c_float),
        ("epsilon_shell_for_reflecting_boundary", C.c_float),
        ("silhouette_precision", C.c_float),
        ("russian_roulette_threshold", C.c_float),
        ("splitting_threshold", C.c_float),
        ("max_walk_length", C.c_int32),
        ("steps_before_applying_tikhonov", C.c_int32),
        ("steps_before_using_maximal_spheres", C.c_int32),
        ("solve_double_sided", C.c_uint8),
        ("use_gradient_control_variates", C.c_uint8),
        ("use_gradient_antithetic_variates", C.c_uint8),
        ("use_cosine_sampling_for_derivatives", C.c_uint8),
        ("ignore_absorbing_boundary_contribution", C.c_uint8),
        ("ignore_reflecting_boundary_contribution", C.c_uint8),
        ("ignore_source_contribution", C.c_uint8),
        ("print_logs", C.c_uint8),
    ]


class zg_function(C.Structure):
    _fields_ = [("kind", C.c_int32), ("eval", C.c_int32), ("params", C.c_float * 6), ("grid", C.c_void_p),
                ("shape", C.c_int32 * 3), ("interpolate", C.c_int32), ("grid_min", C.c_float * 3),
                ("grid_max", C.c_float * 3)]


class zg_pde(C.Structure):
    _fields_ = [
        ("absorption_coeff", C.c_float),
        ("is_source_constant", C.c_uint8),
        ("are_robin_conditions_pure_neumann", C.c_uint8),
        ("are_robin_coeffs_nonnegative", C.c_uint8),
        ("reserved", C.c_uint8),
        ("source", zg_function),
        ("dirichlet", zg_function),
        ("robin", zg_function),
        ("robin_coeff", zg_function),
    ]


class zg_build_options(C.Structure):
    _fields_ = [("bvh_width", C.c_int32), ("leaf_size", C.c_int32), ("sah_bins", C.c_int32), ("reserved", C.c_int32)]


class zg_scene_info(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("bvh_width", C.c_int32),
        ("n_absorbing_prims", C.c_int64),
        ("n_reflecting_prims", C.c_int64),
        ("n_absorbing_nodes", C.c_int64),
        ("n_reflecting_nodes", C.c_int64),
        ("absorbing_depth", C.c_int32),
        ("reflecting_depth", C.c_int32),
        ("node_bytes", C.c_int64),
        ("prim_bytes", C.c_int64),
        ("device_bytes", C.c_int64),
        ("n_silhouette_vertices", C.c_int64),
    ]


class zg_grid_field(C.Structure):
    _fields_ = [("data", C.c_void_p), ("res", C.c_int32 * 3), ("lo", C.c_float * 3), ("hi", C.c_float * 3)]


class zg_sample_points(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pt", C.c_void_p),
        ("normal", C.c_void_p),
        ("type", C.c_void_p),
        ("dist_to_absorbing_boundary", C.c_void_p),
        ("dist_to_reflecting_boundary", C.c_void_p),
        ("estimate_boundary_normal_aligned", C.c_void_p),
        ("first_sphere_radius", C.c_void_p),
        ("n_walks", C.c_void_p),
        ("cv_gradient", C.c_void_p),
        ("cv_gradient_grid", C.POINTER(zg_grid_field)),
        ("n_walks_max", C.c_int32),
        ("reserved", C.c_int32),
        ("estimation_quantity", C.c_void_p),
        ("direction_for_derivative", C.c_void_p),
    ]


class zg_statistics(C.Structure):
    _fields_ = [
        ("n_solution_estimates", C.c_void_p),
        ("solution_mean", C.c_void_p),
        ("solution_m2", C.c_void_p),
        ("total_walk_length", C.c_void_p),
        ("total_splits", C.c_void_p),
        ("n_gradient_estimates", C.c_void_p),
        ("gradient_mean", C.c_void_p),
        ("gradient_m2", C.c_void_p),
        ("total_first_source_contribution", C.c_void_p),
        ("total_derivative_contribution", C.c_void_p),
    ]


class zg_bvc_samples(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pt", C.c_void_p),
        ("normal", C.c_void_p),
        ("type", C.c_void_p),
        ("normal_aligned", C.c_void_p),
        ("pdf", C.c_void_p),
        ("solution", C.c_void_p),
        ("normal_derivative", C.c_void_p),
        ("contribution", C.c_void_p),
        ("robin_coeff", C.c_void_p),
    ]


class zg_bvc_eval_points(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pt", C.c_void_p),
        ("type", C.c_void_p),
        ("dist_to_absorbing_boundary", C.c_void_p),
        ("dist_to_reflecting_boundary", C.c_void_p),
    ]


class zg_bvc_statistics(C.Structure):
    _fields_ = [
        ("n", C.c_void_p),
        ("solution_mean", C.c_void_p),
        ("solution_m2", C.c_void_p),
        ("gradient_mean", C.c_void_p),
        ("gradient_m2", C.c_void_p),
    ]


ZG_BVC_CATEGORIES = 5


class zg_solve_options(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("stream_offset", C.c_uint32),
        ("walks_per_lane", C.c_int32),
        ("regroup_interval", C.c_int32),
        ("step_interval", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("count_visits", C.c_int32),
        ("distance_tolerance", C.c_float),
        ("wost_interleave_queries", C.c_int32),
    ]


class zg_solve_report(C.Structure):
    _fields_ = [
        ("completion_codes", C.c_int64 * 4),
        ("total_steps", C.c_int64),
        ("total_walks", C.c_int64),
        ("node_visits", C.c_int64),
        ("prim_visits", C.c_int64),
        ("queries", C.c_int64),
        ("kernel_ms", C.c_float),
        ("reserved", C.c_float),
        ("reflecting_node_visits", C.c_int64),
        ("reflecting_prim_visits", C.c_int64),
        ("silhouette_node_visits", C.c_int64),
        ("silhouette_prim_visits", C.c_int64),
    ]


class zg_mlp_config(C.Structure):
    _fields_ = [
        ("in_dim", C.c_int32),
        ("n_frequencies", C.c_int32),
        ("hidden", C.c_int32),
        ("n_hidden_layers", C.c_int32),
        ("frequency_sigma", C.c_float),
        ("reserved", C.c_float),
        ("seed", C.c_uint64),
    ]


# name -> (restype, argtypes); every symbol declared in include/zombie_amd.h
_P = C.c_void_p
SIGNATURES = {
    "zg_walk_settings_default": (None, [C.POINTER(zg_walk_settings), C.c_float, C.c_float, C.c_int32, C.c_int32]),
    "zg_solve_options_default": (None, [C.POINTER(zg_solve_options)]),
    "zg_scene_create": (C.c_int32, [C.c_int32, _P, C.c_int64, _P, C.c_int64, _P, _P, _P, C.c_int32, C.c_float,
                                    C.POINTER(zg_build_options), C.POINTER(C.c_void_p)]),
    "zg_scene_destroy": (C.c_int32, [_P]),
    "zg_scene_set_robin_coefficients": (C.c_int32, [_P, _P, _P]),
    "zg_sdf_scene_create": (C.c_int32, [C.c_int32, _P, _P, _P, _P, _P, _P, C.POINTER(C.c_void_p)]),
    "zg_scene_get_info": (C.c_int32, [_P, C.POINTER(zg_scene_info)]),
    "zg_closest_point": (C.c_int32, [_P, C.c_int32, C.c_int64, _P, _P, _P, _P, _P, _P]),
    "zg_inside_domain": (C.c_int32, [_P, C.c_int64, _P, _P, _P]),
    "zg_intersect_ray": (C.c_int32, [_P, C.c_int32, C.c_int64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "zg_star_radius": (C.c_int32, [_P, C.c_int64, _P, _P, _P, C.c_float, _P, _P, _P]),
    "zg_bvc_splat": (C.c_int32, [C.c_int32, C.c_float, C.POINTER(zg_bvc_samples), C.POINTER(zg_bvc_eval_points),
                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(zg_bvc_statistics),
                                 _P]),
    "zg_stratified_samples": (C.c_int32, [C.c_int32, C.c_int64, C.c_int32, C.c_uint64, C.c_uint32, _P, _P]),
    "zg_spatially_sorted_indices": (C.c_int32, [C.c_int32, _P, C.c_int64, _P]),
    "zg_evaluate_function": (C.c_int32, [C.c_int32, C.POINTER(zg_function), C.c_int64, _P, _P, _P, C.c_float, _P,
                                         _P]),
    "zg_probe_read_bandwidth": (C.c_int32, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_float), _P]),
    "zg_line_of_sight": (C.c_int32, [_P, C.c_int64, _P, _P, _P, _P]),
    "zg_sample_reflecting": (C.c_int32, [_P, C.c_int64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "zg_wos_solve": (C.c_int32, [_P, C.POINTER(zg_pde), C.POINTER(zg_walk_settings), C.POINTER(zg_sample_points),
                                 C.POINTER(zg_statistics), C.POINTER(zg_solve_options), C.POINTER(zg_solve_report), _P]),
    "zg_scene_copy_bvh": (C.c_int32, [_P, C.c_int32, _P, C.c_int64, _P, C.c_int64]),
    "zg_wos_solve_noalloc": (C.c_int32, [_P, C.POINTER(zg_pde), C.POINTER(zg_walk_settings),
                                         C.POINTER(zg_sample_points), C.POINTER(zg_statistics),
                                         C.POINTER(zg_solve_options), C.POINTER(zg_solve_report), _P]),
    "zg_wost_solve": (C.c_int32, [_P, C.POINTER(zg_pde), C.POINTER(zg_walk_settings), C.POINTER(zg_sample_points),
                                  C.POINTER(zg_statistics), C.POINTER(zg_solve_options), C.POINTER(zg_solve_report), _P]),
    "zg_batch_create": (C.c_int32, [C.c_int32, _P, C.POINTER(zg_pde), C.POINTER(C.c_void_p)]),
    "zg_batch_destroy": (C.c_int32, [_P]),
    "zg_batch_wos_solve": (C.c_int32, [_P, C.POINTER(zg_walk_settings), C.POINTER(zg_sample_points), _P,
                                       C.POINTER(zg_statistics), C.POINTER(zg_solve_options),
                                       C.POINTER(zg_solve_report), _P]),
    "zg_batch_wos_graph_create": (C.c_int32, [_P, C.POINTER(zg_walk_settings), C.POINTER(zg_sample_points), _P,
                                              C.POINTER(zg_statistics), C.POINTER(zg_solve_options),
                                              C.POINTER(C.c_void_p)]),
    "zg_graph_launch": (C.c_int32, [_P, _P]),
    "zg_graph_read_report": (C.c_int32, [_P, C.POINTER(zg_solve_report), _P]),
    "zg_graph_destroy": (C.c_int32, [_P]),
    "zg_mlp_create": (C.c_int32, [C.POINTER(zg_mlp_config), C.c_int64, C.POINTER(C.c_void_p)]),
    "zg_mlp_destroy": (C.c_int32, [_P]),
    "zg_mlp_parameter_count": (C.c_int32, [_P, C.POINTER(C.c_int64)]),
    "zg_mlp_get_parameters": (C.c_int32, [_P, _P, C.c_int64]),
    "zg_mlp_set_parameters": (C.c_int32, [_P, _P, C.c_int64]),
    "zg_mlp_get_frequencies": (C.c_int32, [_P, _P, C.c_int64]),
    "zg_mlp_gradients": (C.c_int32, [_P, _P, _P, C.c_int64, _P, _P, _P]),
    "zg_mlp_train_step": (C.c_int32, [_P, _P, _P, C.c_int64, C.c_float, _P, _P]),
    "zg_mlp_evaluate": (C.c_int32, [_P, _P, C.c_int64, _P, _P, _P]),
    "zg_device_count": (C.c_int32, [C.POINTER(C.c_int32)]),
    "zg_malloc": (C.c_int32, [C.POINTER(C.c_void_p), C.c_size_t]),
    "zg_free": (C.c_int32, [_P]),
    "zg_memcpy_htod": (C.c_int32, [_P, _P, C.c_size_t]),
    "zg_memcpy_dtoh": (C.c_int32, [_P, _P, C.c_size_t]),
    "zg_memset": (C.c_int32, [_P, C.c_int32, C.c_size_t]),
    "zg_synchronize": (C.c_int32, []),
    "zg_last_error": (C.c_int32, [C.c_char_p, C.c_int32]),
    "zg_version": (C.c_char_p, []),
}


class ZombieLibraryError(RuntimeError):
    """libzombie_amd.so is missing or cannot be loaded."""


class ZombieError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"zombie_amd error {status}: {message}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ZombieLibraryError(
                f"{LIB_PATH} is missing: build it with zombie_amd._build.build_library() "
                "(there is no CPU fallback)")
        try:
            handle = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        except OSError as e:
            raise ZombieLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().zg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != ZG_OK:
        raise ZombieError(status, last_error())
