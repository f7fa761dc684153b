"""ctypes binding of libjztree_amd.so (the C ABI declared in include/jztree_amd.h).

The product has no CPU fallback: importing this module on a machine without the built library
raises, and every entry point raises RuntimeError with the library's message on failure.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjztree_amd.so")
JZT_F32, JZT_F64, JZT_I32, JZT_I64 = 0, 1, 2, 3

_c = ctypes
_p = _c.c_void_p
_i64 = _c.c_int64
_i32 = _c.c_int
_f32 = _c.c_float
_f64 = _c.c_double
_u32 = _c.c_uint32
_sz = _c.c_size_t

_SIGNATURES = {
    "jzt_last_error": ([], _c.c_char_p),
    "jzt_abi_version": ([], _i32),
    "jzt_set_tuning": ([_i32, _i32], _i32),
    "jzt_get_tuning": ([_i32], _i32),
    "jzt_zsort_tmp_bytes": ([_i64, _i32, _i32], _sz),
    "jzt_pos_zorder_sort": ([_p, _p, _p, _p, _p, _sz, _i64, _i32, _i32], _i32),
    "jzt_pos_zorder_sort_noalloc": ([_p, _p, _p, _p, _i64, _i32, _i32], _i32),
    "jzt_zsort_segmented_tmp_bytes": ([_i64, _i64, _i32, _i32], _sz),
    "jzt_pos_zorder_sort_segmented": ([_p, _p, _p, _p, _p, _i64, _p, _sz, _i64, _i32, _i32], _i32),
    "jzt_segment_sort": ([_p, _p, _p, _p, _p, _p, _i64, _i64], _i32),
    "jzt_search_sorted_z": ([_p, _p, _p, _i64, _i64, _i32, _i32, _i32, _p], _i32),
    "jzt_flag_leaf_boundaries": ([_p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _i32, _i32, _i32], _i32),
    "jzt_find_node_boundaries": ([_p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _i32, _i32], _i32),
    "jzt_flag_leaf_boundaries_sets": ([_p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _i32, _i32], _i32),
    "jzt_find_node_boundaries_sets": ([_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _i32], _i32),
    "jzt_get_node_geometry": ([_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i32, _i32, _i32, _u32, _i32], _i32),
    "jzt_center_of_mass": ([_p, _p, _p, _p, _p, _i64, _i32, _i32, _i32], _i32),
    "jzt_geometry_to_f32": ([_p, _p, _p, _i64, _i32, _i32, _p, _p], _i32),
    "jzt_where_tmp_bytes": ([_i64], _sz),
    "jzt_where_flags": ([_p, _p, _i64, _p, _i64, _i32, _p, _p, _p, _sz], _i32),
    "jzt_count_not_nan": ([_p, _p, _i64, _i32, _i32, _p], _i32),
    "jzt_dense_ilist": ([_p, _p, _i32, _p, _p, _i64, _p, _p, _i64, _p], _i32),
    "jzt_dense_ilist_sets_tmp_bytes": ([_i64], _sz),
    "jzt_graph_count_nodes": ([_p, _p], _i32),
    "jzt_dense_ilist_sets": ([_p, _p, _p, _i64, _i32, _p, _p, _i64, _p, _p, _i64, _p, _p, _sz], _i32),
    "jzt_regularization_tmp_bytes": ([_i32, _i32], _sz),
    "jzt_regularization_level": ([_p, _p, _p, _i64, _p, _i32, _i32, _f32, _f32, _p, _sz], _i32),
    "jzt_knn_node2node": ([_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i32, _f32, _i32,
                           _i32], _i32),
    "jzt_knn_node2node_ex": ([_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i32,
                              _f32, _i32, _i32], _i32),
    "jzt_segment_max_f32": ([_p, _p, _p, _i64, _p, _i64, _f32], _i32),
    "jzt_flags_or_level": ([_p, _p, _p, _i64, _p, _p], _i32),
    "jzt_plane_cap": ([_p, _p, _p, _i64, _p, _p, _p, _i64, _i32, _i32, _p], _i32),
    "jzt_packed_get": ([_p, _p, _i64, _p, _i32, _i32, _p, _i32, _i32, _p, _i64], _i32),
    "jzt_knn_leaf2leaf_tmp_bytes": ([_i64, _i32, _i32], _sz),
    "jzt_knn_leaf2leaf": ([_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _i64, _i64, _i32, _f32,
                           _i32, _i32], _i32),
    "jzt_knn_leaf2leaf_ex": ([_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _i64, _i64,
                              _i64, _i32, _f32, _i32, _i32], _i32),
    "jzt_knn_fused": ([_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i32, _f32,
                       _i32, _i32], _i32),
    "jzt_knn_fused_retry": ([_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i32,
                             _f32, _i32, _i32], _i32),
    "jzt_proposer_num_params": ([], _i32),
    "jzt_leaf_features": ([_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _p], _i32),
    "jzt_mlp_infer": ([_p, _p, _p, _p, _p, _i64, _f32, _p, _p], _i32),
    "jzt_mlp_grad": ([_p, _p, _p, _p, _i64, _p, _p], _i32),
    "jzt_adam_step": ([_p, _p, _p, _p, _p, _p, _i64, _f32, _f32, _f32, _f32, _i32], _i32),
    "jzt_to_bf16": ([_p, _p, _p, _i64], _i32),
    "jzt_node_to_child_label": ([_p, _p, _p, _p, _p, _p, _i64, _i64, _f32, _i32, _i32], _i32),
    "jzt_fof_node2node": ([_p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _f32, _f32, _i32, _i32], _i32),
    "jzt_fof_leaf2leaf": ([_p, _p, _p, _i64, _p, _p, _p, _p, _i64, _i64, _f32, _f32, _i32, _i32], _i32),
    "jzt_fof_leaf2leaf_cells": ([_p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64, _i64, _f32, _f32, _i32, _i32], _i32),
    "jzt_node_to_child_label_sets": ([_p, _p, _p, _p, _p, _p, _i64, _i64, _i32, _i32], _i32),
    "jzt_fof_node2node_sets": ([_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _f32, _i32, _i32],
                               _i32),
    "jzt_fof_leaf2leaf_sets": ([_p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64, _i64, _f32, _i32, _i32], _i32),
    "jzt_fof_order_tmp_bytes": ([_i64], _sz),
    "jzt_fof_order": ([_p, _p, _p, _i64, _p, _p, _p, _sz], _i32),
    "jzt_fof_catalogue_tmp_bytes": ([_i64, _i64, _i32, _i32], _sz),
    "jzt_fof_catalogue": ([_p, _p, _p, _f64, _p, _p, _p, _p, _i64, _i64, _i32, _f64, _i32, _i32, _p, _p, _p, _p, _p,
                           _p, _p, _p, _p, _p, _sz], _i32),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python jztree_amd/build.py` (hipcc, gfx950). "
            "jztree_amd has no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


LIB = _load()


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = LIB.jzt_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def stream() -> int:
    """Handle of torch's current HIP stream (all ops are enqueued there)."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float32:
        return JZT_F32
    if dt == torch.float64:
        return JZT_F64
    raise TypeError(f"unsupported position dtype {dt}")


def sort_dtype_code(dt: torch.dtype) -> int:
    """Position dtype code of the sort and SearchSortedZ, which also take int32 / int64 positions."""
    if dt == torch.int32:
        return JZT_I32
    if dt == torch.int64:
        return JZT_I64
    return dtype_code(dt)


def call(name: str, *args):
    # only bound entry points: an unbound ctypes function would pass pointers as 32-bit ints
    if name not in _SIGNATURES:
        raise RuntimeError(f"{name} is not bound in jztree_amd._lib")
    check(getattr(LIB, name)(*args), name)
