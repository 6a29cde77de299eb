"""ctypes binding of liboctohip.so (the C ABI declared in include/octohip.h).

The library is loaded from this directory only; if it is missing the import
of the front-end fails loudly — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboctohip.so")

OCT_OK = 0
OCT_ERR_ARGUMENT = -1
OCT_ERR_DIMENSION = -2
OCT_ERR_STATE = -3
OCT_ERR_DEVICE = -4
OCT_ERR_CAPACITY = -5

OCT_EDGES_EXACT = 0
OCT_EDGES_LEARNED = 1
OCT_ORDER_LEAF = 0
OCT_ORDER_RECEIVER = 1

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)


class OctEdges(ctypes.Structure):
    _fields_ = [
        ("senders", ctypes.c_void_p),
        ("receivers", ctypes.c_void_p),
        ("rel_displacement", ctypes.c_void_p),
        ("rel_dist_norm", ctypes.c_void_p),
        ("n_edges", ctypes.c_int64),
    ]


class OctTree(ctypes.Structure):
    _fields_ = [
        ("node_bounds_min", ctypes.c_void_p),
        ("node_bounds_max", ctypes.c_void_p),
        ("node_first", ctypes.c_void_p),
        ("node_last", ctypes.c_void_p),
        ("node_children", ctypes.c_void_p),
        ("permutation", ctypes.c_void_p),
        ("n_nodes", ctypes.c_int32),
        ("n_leaves", ctypes.c_int32),
        ("n_points", ctypes.c_int64),
        ("origin", ctypes.c_float * 3),
    ]


class OctView(ctypes.Structure):
    _fields_ = [
        ("coords_q", ctypes.c_void_p),
        ("coords_t", ctypes.c_void_p),
        ("bounds_min", ctypes.c_void_p),
        ("bounds_max", ctypes.c_void_p),
        ("node_first", ctypes.c_void_p),
        ("node_last", ctypes.c_void_p),
        ("node_children", ctypes.c_void_p),
        ("perm", ctypes.c_void_p),
        ("nodes", ctypes.c_void_p),
        ("sorted_t", ctypes.c_void_p),
        ("n_nodes", ctypes.c_int32),
        ("root", ctypes.c_int32),
        ("radius", ctypes.c_float),
        ("radius_sq", ctypes.c_float),
        ("n_q", ctypes.c_int64),
        ("index_base", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class OctBatchEdges(ctypes.Structure):
    _fields_ = [
        ("senders", ctypes.c_void_p),
        ("receivers", ctypes.c_void_p),
        ("rel_displacement", ctypes.c_void_p),
        ("rel_dist_norm", ctypes.c_void_p),
        ("n_edges", ctypes.c_int64),
        ("set_offsets", ctypes.c_void_p),
        ("n_sets", ctypes.c_int32),
        ("replayed", ctypes.c_int32),
        ("recounted_leaves", ctypes.c_int64),
        ("recorded_leaves", ctypes.c_int64),
        ("heavy_leaves", ctypes.c_int64),
    ]


class OctBuildInfo(ctypes.Structure):
    _fields_ = [
        ("changed_points", ctypes.c_int64),
        ("refit_used", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class OctNeighbors(ctypes.Structure):
    _fields_ = [
        ("offsets", ctypes.c_void_p),
        ("flat", ctypes.c_void_p),
        ("n_queries", ctypes.c_int64),
        ("total", ctypes.c_int64),
        ("materialized", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class OctEdgeStats(ctypes.Structure):
    _fields_ = [
        ("features", ctypes.c_void_p),
        ("predicted", ctypes.c_void_p),
        ("exact", ctypes.c_void_p),
        ("n_leaves", ctypes.c_int64),
        ("n_edges", ctypes.c_int64),
        ("recounted_leaves", ctypes.c_int64),
        ("recorded_leaves", ctypes.c_int64),
        ("predicted_total", ctypes.c_double),
        ("single_pass", ctypes.c_int32),
        ("used_predictor", ctypes.c_int32),
        ("attempts", ctypes.c_int32),
        ("heavy_leaves", ctypes.c_int32),
    ]


MLP_FEATURES = 29
MLP_HIDDEN = 64
MLP_PARAMS = 6203
MLP_TRAINED = MLP_PARAMS - 2 * MLP_FEATURES

_P = ctypes.c_void_p
# name -> (restype, argtypes); every function declared in include/octohip.h
SIGNATURES = {
    "oct_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "oct_destroy": (ctypes.c_int, [_P]),
    "oct_last_error": (ctypes.c_char_p, [_P]),
    "oct_set_stream": (ctypes.c_int, [_P, _P]),
    "oct_set_search_radius": (ctypes.c_int, [_P, ctypes.c_float]),
    "oct_set_refit_mode": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_set_leaf_size": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_add_point_set": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int)]),
    "oct_update_point_set": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int64]),
    "oct_resize_point_set": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int64]),
    "oct_set_active_search": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "oct_run": (ctypes.c_int, [_P]),
    "oct_get_tree": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(OctTree)]),
    "oct_build_edges": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(OctEdges)]),
    "oct_device_view": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OctView)]),
    "oct_materialize_neighbors": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OctNeighbors)]),
    "oct_get_neighbors": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OctNeighbors)]),
    "oct_apply_zsort": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_int64, ctypes.c_int]),
    "oct_apply_zsort_adjoint": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_int64, ctypes.c_int]),
    "oct_count_neighbors_device": (ctypes.c_int, [_P, ctypes.POINTER(OctView), _P]),
    "oct_build_edges_backward": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(OctEdges), _P, _P,
                                                ctypes.c_float, _P]),
    "oct_set_backward_form": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_create_typed": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int]),
    "oct_set_search_radius64": (ctypes.c_int, [_P, ctypes.c_double]),
    "oct_build_edges_backward64": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(OctEdges), _P, _P,
                                                  ctypes.c_double, _P]),
    "oct_profile": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_profile_phases": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "oct_profile_read": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int64)]),
    "oct_synchronize": (ctypes.c_int, [_P]),
    "oct_trim": (ctypes.c_int, [_P]),
    "oct_retired_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "oct_build_info_get": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(OctBuildInfo)]),
    "oct_set_predictor_policy": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_predictor_load": (ctypes.c_int, [_P, _P]),
    "oct_predictor_save": (ctypes.c_int, [_P, _P]),
    "oct_predictor_collect": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "oct_predictor_clear_data": (ctypes.c_int, [_P]),
    "oct_predictor_train": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "oct_edge_stats_get": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OctEdgeStats)]),
    "oct_edge_stats_peek": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OctEdgeStats)]),
    "oct_batch_create": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.POINTER(ctypes.c_int)]),
    "oct_batch_destroy": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_batch_config": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "oct_batch_run": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(OctBatchEdges)]),
    "oct_batch_autotune": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int, _P, ctypes.c_int, _P, ctypes.c_int,
                                          ctypes.c_int, _P, ctypes.POINTER(ctypes.c_int)]),
    "oct_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P, ctypes.c_int,
                                    _P, ctypes.c_int, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_int)]),
    "oct_set_run_leaves": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_set_run_chunk": (ctypes.c_int, [_P, ctypes.c_int]),
    "oct_stream_join": (ctypes.c_int, [_P, _P, ctypes.c_int]),
    "oct_mlp_forward": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P]),
    "oct_mlp_gradient": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int, _P, _P]),
    "oct_mlp_gradient_noalloc": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int, _P, _P]),
}

PROF_PHASES = ["bounds", "build", "leaf_lists", "count", "scan", "fill", "backward", "single", "predict"]

_lib = None


def load():
    """Load liboctohip.so and bind every exported entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"liboctohip.so not found at {LIB_PATH}; build it first (make -C octopus.jl_amd). "
            "There is no fallback path.")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
