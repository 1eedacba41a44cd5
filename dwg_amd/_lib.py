"""ctypes binding of libdwg_hip.so (declarations: include/dwg_hip.h).

The library must have been built (``__graft_entry__.build()``); importing it
on a machine without the built .so raises instead of falling back to any
CPU path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdwg_hip.so")

DWG_OK = 0
DWG_ERR_ARG = 1
DWG_ERR_HIP = 2
DWG_ERR_CAPACITY = 3
DWG_ERR_STATE = 4

DWG_SCREENED = 1
DWG_USE_RADII = 2
DWG_KAHAN = 4
DWG_FAR_BF16 = 8
DWG_NEAR_FP64 = 16
DWG_NOALLOC_MAX_SOURCES = 262144

EXPORTS = [
    "dwg_ctx_create", "dwg_ctx_destroy", "dwg_ctx_error", "dwg_ctx_sync", "dwg_cache_release", "dwg_ctx_profile",
    "dwg_ctx_profile_get", "dwg_ctx_profile_reset", "dwg_ctx_counter_get", "dwg_gwn_direct",
    "dwg_gwn_direct_noalloc", "dwg_octree_build", "dwg_octree_build_device",
    "dwg_octree_leaves", "dwg_octree_leaf_normals",
    "dwg_octree_destroy", "dwg_octree_info", "dwg_octree_point_order", "dwg_octree_update",
    "dwg_octree_set_point_areas", "dwg_gwn_tree", "dwg_gwn_tree_grad", "dwg_gwn_tree_self", "dwg_gwn_tree_grid",
    "dwg_gwn_tree_terms", "dwg_gwn_tree_grid_terms", "dwg_knn_build", "dwg_knn_destroy", "dwg_knn_query",
    "dwg_marching_cubes", "dwg_marching_cubes_any_order", "dwg_diffuse_normals", "dwg_radii", "dwg_point_areas", "dwg_pca_normals", "dwg_gauss_map", "dwg_transfer_normals",
    "dwg_outlier_filter", "dwg_normal_change", "dwg_reduce_sum", "dwg_bounds", "dwg_options_default", "dwg_launch",
    "dwg_free", "dwg_mlp_num_params", "dwg_mlp_create", "dwg_mlp_load", "dwg_mlp_save", "dwg_mlp_destroy",
    "dwg_mlp_get", "dwg_mlp_set", "dwg_mlp_patches", "dwg_mlp_patches_from", "dwg_mlp_init", "dwg_mlp_forward", "dwg_mlp_loss_grad", "dwg_mlp_adam",
    "dwg_mlp_augment", "dwg_gemm_bf16",
]


class DwgGrid(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("h", ctypes.c_float), ("res", ctypes.c_int32 * 3)]


class DwgOptions(ctypes.Structure):
    _fields_ = [("input_normal", ctypes.c_int32), ("grid_reso", ctypes.c_int32), ("final_reso", ctypes.c_int32),
                ("kw", ctypes.c_int32), ("final_kw", ctypes.c_int32), ("normal_kw", ctypes.c_int32),
                ("max_depth", ctypes.c_int32), ("sigma", ctypes.c_float), ("min_num", ctypes.c_int32),
                ("use_radii", ctypes.c_int32), ("com_As", ctypes.c_int32), ("As_threshold", ctypes.c_float),
                ("offset_1", ctypes.c_float), ("offset_2", ctypes.c_float), ("use_leaves_generate", ctypes.c_int32),
                ("total_iter", ctypes.c_int32), ("test_iter", ctypes.c_int32), ("output_mode", ctypes.c_int32),
                ("source_mode", ctypes.c_int32), ("leaf_size", ctypes.c_int32), ("beta", ctypes.c_float),
                ("precision", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("normalize_input", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("count_terms", ctypes.c_int32), ("mlp", ctypes.c_void_p),
                ("sigma_warm_iters", ctypes.c_int32), ("residual_tol", ctypes.c_float)]


class DwgResult(ctypes.Structure):
    _fields_ = [("normals", ctypes.c_void_p), ("want_mesh", ctypes.c_int32), ("mesh", ctypes.c_void_p),
                ("tri_count", ctypes.c_int64), ("n_sources", ctypes.c_int64), ("center", ctypes.c_double * 3),
                ("scale", ctypes.c_double), ("time_setup_s", ctypes.c_double), ("time_iter_s", ctypes.c_double),
                ("time_final_s", ctypes.c_double), ("time_total_s", ctypes.c_double), ("w_mean", ctypes.c_void_p),
                ("iter_tris", ctypes.c_void_p), ("iter_zeros", ctypes.c_void_p), ("terms_src", ctypes.c_int64),
                ("terms_grid", ctypes.c_int64), ("grid_points", ctypes.c_int64), ("source_normals", ctypes.c_void_p),
                ("residual", ctypes.c_void_p), ("iterations", ctypes.c_int32)]


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_u32 = ctypes.c_uint32
_f = ctypes.c_float
_lib = None


def load():
    """Load and prototype libdwg_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    # torch first: the library then binds to the HIP runtime torch loads.
    # Loaded the other way round, the process holds two HIP runtimes and the
    # library's dwg_ctx_create fails on a GPU
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    L.dwg_ctx_create.argtypes = [_int, _vp, P(_vp)]
    L.dwg_ctx_destroy.argtypes = [_vp]
    L.dwg_ctx_destroy.restype = None
    L.dwg_ctx_error.argtypes = [_vp]
    L.dwg_ctx_error.restype = ctypes.c_char_p
    L.dwg_ctx_sync.argtypes = [_vp]
    L.dwg_cache_release.argtypes = [P(_i64)]
    L.dwg_ctx_profile.argtypes = [_vp, _int]
    L.dwg_ctx_profile_get.argtypes = [_vp, ctypes.c_char_p, P(ctypes.c_double), P(_i64)]
    L.dwg_ctx_profile_reset.argtypes = [_vp]
    L.dwg_ctx_counter_get.argtypes = [_vp, ctypes.c_char_p, P(_i64)]
    L.dwg_gwn_direct.argtypes = [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _f, _u32, _vp, _vp]
    L.dwg_gwn_direct_noalloc.argtypes = [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _f, _u32, _vp, _vp]
    L.dwg_octree_build.argtypes = [_vp, _vp, _vp, _i64, _int, _int, P(_vp)]
    L.dwg_octree_build_device.argtypes = [_vp, _vp, _vp, _i64, _int, _int, P(_vp)]
    L.dwg_octree_destroy.argtypes = [_vp]
    L.dwg_octree_leaves.argtypes = [_vp, _vp, _int, _vp, _vp, P(_i64)]
    L.dwg_octree_leaf_normals.argtypes = [_vp, _vp, _int, _vp, _vp, _vp, P(_i64)]
    L.dwg_octree_destroy.restype = None
    L.dwg_octree_info.argtypes = [_vp, P(_i64), P(_i64), P(_int)]
    L.dwg_octree_point_order.argtypes = [_vp, _vp, _vp]
    L.dwg_octree_update.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.dwg_octree_set_point_areas.argtypes = [_vp, _vp, _vp, _vp]
    L.dwg_gwn_tree.argtypes = [_vp, _vp, _vp, _i64, _f, _f, _u32, _vp]
    L.dwg_gwn_tree_grad.argtypes = [_vp, _vp, _vp, _i64, _f, _f, _u32, _vp, _vp]
    L.dwg_gwn_tree_self.argtypes = [_vp, _vp, _f, _f, _u32, _vp]
    L.dwg_gwn_tree_grid.argtypes = [_vp, _vp, P(DwgGrid), _f, _f, _u32, _vp]
    L.dwg_gwn_tree_terms.argtypes = [_vp, _vp, _vp, _i64, _f, _vp]
    L.dwg_gwn_tree_grid_terms.argtypes = [_vp, _vp, P(DwgGrid), _f, P(_i64)]
    L.dwg_knn_build.argtypes = [_vp, _vp, _i64, _int, P(_vp)]
    L.dwg_knn_destroy.argtypes = [_vp]
    L.dwg_knn_destroy.restype = None
    L.dwg_knn_query.argtypes = [_vp, _vp, _vp, _i64, _int, _vp, _vp, _vp, _vp]
    L.dwg_marching_cubes.argtypes = [_vp, _vp, P(DwgGrid), _f, _i64, _vp, _vp, _vp, P(_i64)]
    L.dwg_marching_cubes_any_order.argtypes = [_vp, _vp, P(DwgGrid), _f, _i64, _vp, _vp, P(_i64)]
    L.dwg_diffuse_normals.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _int, _int, _int, _vp, P(_i64)]
    L.dwg_radii.argtypes = [_vp, _vp, _vp, _i64, _int, _f, _f, _vp]
    L.dwg_point_areas.argtypes = [_vp, _vp, _vp, _i64, _int, _vp, _vp]
    L.dwg_pca_normals.argtypes = [_vp, _vp, _vp, _i64, _int, _vp, _vp]
    L.dwg_gauss_map.argtypes = [_vp, _vp, _i64, P(_f), _vp]
    L.dwg_transfer_normals.argtypes = [_vp, _vp, _vp, _vp, _i64, _int, _vp]
    L.dwg_outlier_filter.argtypes = [_vp, _vp, _i64, _f]
    L.dwg_normal_change.argtypes = [_vp, _vp, _vp, _i64, P(ctypes.c_double)]
    L.dwg_reduce_sum.argtypes = [_vp, _vp, _i64, P(ctypes.c_double)]
    L.dwg_bounds.argtypes = [_vp, _vp, _i64, P(_f), P(_f)]
    L.dwg_options_default.argtypes = [P(DwgOptions)]
    L.dwg_options_default.restype = None
    L.dwg_launch.argtypes = [_vp, _vp, _vp, _i64, P(DwgOptions), ctypes.c_char_p, P(DwgResult)]
    L.dwg_free.argtypes = [_vp]
    L.dwg_free.restype = None
    L.dwg_mlp_num_params.argtypes = []
    L.dwg_mlp_num_params.restype = _i64
    L.dwg_mlp_create.argtypes = [_vp, ctypes.c_uint64, P(_vp)]
    L.dwg_mlp_load.argtypes = [_vp, ctypes.c_char_p, P(_vp)]
    L.dwg_mlp_save.argtypes = [_vp, _vp, ctypes.c_char_p]
    L.dwg_mlp_destroy.argtypes = [_vp]
    L.dwg_mlp_destroy.restype = None
    L.dwg_mlp_get.argtypes = [_vp, _vp, _vp, _vp]
    L.dwg_mlp_set.argtypes = [_vp, _vp, _vp]
    L.dwg_mlp_patches.argtypes = [_vp, _vp, _vp, _i64, _vp]
    L.dwg_mlp_forward.argtypes = [_vp, _vp, _vp, _i64, _int, _vp]
    L.dwg_mlp_patches_from.argtypes = [_vp, _vp, _vp, _i64, _vp, _i64, _vp]
    L.dwg_mlp_init.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp]
    L.dwg_mlp_loss_grad.argtypes = [_vp, _vp, _vp, _vp, _i64, _int, P(ctypes.c_double)]
    L.dwg_mlp_adam.argtypes = [_vp, _vp, _f, _f, _f, _f]
    L.dwg_mlp_augment.argtypes = [_vp, _vp, _vp, _vp, _i64, P(_f), _vp, _vp]
    L.dwg_gemm_bf16.argtypes = [_vp, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i64]
    _lib = L
    return L
