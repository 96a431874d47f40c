"""ctypes binding of libyggdrax_amd.so (the HIP kernels behind the C ABI).

The library is loaded from the package directory. There is no fallback: if
the shared library is missing or a GPU is absent, every tree entry point
raises instead of computing anything on the CPU.
"""

from __future__ import annotations

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyggdrax_amd.so")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_INT = ctypes.c_int

ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "workspace too small",
          4: "unsupported size",
          5: "input larger than the fixed maximum of this entry point"}


class RadixTopologyPtrs(ctypes.Structure):
    _fields_ = [(n, _P) for n in (
        "parent", "left_child", "right_child", "left_is_leaf", "right_is_leaf",
        "node_ranges", "node_level", "level_offsets", "nodes_by_level", "leaf_codes",
        "scalars")]


class OctreeMetaPtrs(ctypes.Structure):
    _fields_ = [(n, _P) for n in (
        "oct_parent", "oct_children", "oct_child_counts", "oct_child_mask", "oct_valid_mask",
        "oct_node_codes", "oct_node_depths", "oct_node_ranges", "oct_nodes_by_level",
        "oct_level_offsets", "oct_leaf_mask", "oct_leaf_nodes", "radix_node_to_oct",
        "leaf_indices", "leaf_position", "scalars")]


class WalkParams(ctypes.Structure):
    _fields_ = [
        ("num_nodes", _I32), ("num_internal", _I32), ("num_leaves", _I32), ("root", _I32),
        ("mac_type", _I32), ("mode", _I32), ("collect_far", _I32), ("collect_near", _I32),
        ("theta_sq", _F), ("theta_root", _F), ("far_cap", _I32), ("near_cap", _I32),
        ("queue_cap", _I64), ("max_generations", _I32), ("pad0", _I32),
        ("nodes", _P), ("wf0", _P), ("wf1", _P), ("gen_sizes", _P), ("flags", _P),
        ("decisions", _P), ("far_counts", _P), ("near_counts", _P),
        ("far_row_start", _P), ("far_row_cap", _P), ("far_entries", _P),
        ("near_row_start", _P), ("near_row_cap", _P), ("near_entries", _P),
    ]


class FlatWalkParams(ctypes.Structure):
    _fields_ = [
        ("num_nodes", _I32), ("num_internal", _I32), ("num_leaves", _I32), ("root", _I32),
        ("mac_type", _I32), ("emit", _I32), ("max_generations", _I32), ("grid_blocks", _I32),
        ("theta_sq", _F), ("theta_root", _F), ("gen_hint", _I32), ("pairs_per_thread", _I32),
        ("queue_cap", _I64), ("far_cap", _I64),
        ("near_cap", _I64), ("nodes", _P), ("wf0", _P), ("wf1", _P), ("gen_sizes", _P),
        ("flags", _P), ("counters", _P), ("far_keys", _P), ("far_vals", _P),
        ("near_keys", _P), ("near_vals", _P), ("hist_gen", _I32), ("xcd_groups", _I32),
        ("node_hist", _P), ("roots", _P), ("num_roots", _I64), ("spread_counters", _P),
        ("spread_lanes", _I64), ("policy", _P), ("far_tags", _P), ("gen_fill", _P),
        ("start_generation", _I32), ("pair_records", _I32), ("gen_pairs_hint", _P),
    ]


class Batch(ctypes.Structure):
    _fields_ = [
        ("num_systems", _I32), ("pad0", _I32), ("leaf_size", _I64), ("point_offsets", _P),
        ("leaf_offsets", _P), ("num_points", _I64), ("num_leaves", _I64),
        ("max_leaves_per_system", _I64),
    ]


class PolicyDesc(ctypes.Structure):
    _fields_ = [
        ("accept", _I32), ("near", _I32), ("tag_metric", _I32), ("num_thresholds", _I32),
        ("accept_dist_sq", _F), ("tag_modulus", _I32), ("thresholds", _F * 8),
        ("tag_values", _I32 * 9),
    ]


class WidenItem(ctypes.Structure):
    _fields_ = [("src", _P), ("dst", _P), ("n", _I64)]


WIDEN_MAX_ITEMS = 32

_BP = ctypes.POINTER(Batch)

_SIGS = {
    "ygg_widen_i32_to_i64": (_INT, [_P, _INT, _P]),
    "ygg_walk_flat_fixed": (_INT, [ctypes.POINTER(FlatWalkParams), _P]),
    "ygg_batch_infer_bounds": (_INT, [_BP, _P, _P, _P]),
    "ygg_batch_morton_sort_workspace_bytes": (_SZ, [_BP]),
    "ygg_batch_morton_sort": (_INT, [_BP, _P, _P, _P, _P, _P, _SZ, _P]),
    "ygg_batch_reorder": (_INT, [_BP, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ygg_batch_radix_build_workspace_bytes": (_SZ, [_BP]),
    "ygg_batch_radix_build": (_INT, [_BP, _P, ctypes.POINTER(RadixTopologyPtrs), _P, _SZ, _P]),
    "ygg_batch_octree_meta_workspace_bytes": (_SZ, [_BP]),
    "ygg_batch_octree_meta": (_INT, [_BP, _P, _P, ctypes.POINTER(OctreeMetaPtrs), _P, _SZ, _P]),
    "ygg_batch_radix_leaf_order": (_INT, [_BP, _P, _P, _P]),
    "ygg_batch_geometry_workspace_bytes": (_SZ, [_BP]),
    "ygg_batch_geometry": (_INT, [_BP, _P, _P, _I64, _P, _P, _P, _P, _P, _SZ, _P]),
    "ygg_batch_walk_prepare_nodes": (_INT, [_BP, _P, _P, _P, _P, _P, _P, _P, _INT, _F, _P, _P,
                                            _P]),
    "ygg_batch_far_localize": (_INT, [_BP, _P, _P, _P, _I64, _INT, _P, _P, _P, _P, _P]),
    "ygg_batch_near_localize": (_INT, [_BP, _P, _I64, _INT, _P, _P, _P]),
    "ygg_batch_mlp_features": (_INT, [_BP] + [_P] * 11),
    "ygg_batch_mlp_totals": (_INT, [_BP, _P, _F, _F, _F, _P, _P]),
    "ygg_sort_workspace_bytes": (_SZ, [_I64, _INT, _INT, _INT, _INT]),
    "ygg_sort_default_digit_bits": (_INT, [_INT, _INT, _INT]),
    "ygg_sort_pairs_u64": (_INT, [_P, _P, _P, _P, _I64, _INT, _INT, _INT, _P, _SZ, _P]),
    "ygg_sort_pairs_u32": (_INT, [_P, _P, _P, _P, _I64, _INT, _INT, _INT, _P, _SZ, _P]),
    "ygg_bounds_workspace_bytes": (_SZ, []),
    "ygg_infer_bounds": (_INT, [_P, _I64, _P, _P, _SZ, _P]),
    "ygg_morton_encode": (_INT, [_P, _I64, _P, _P, _P]),
    "ygg_morton_sort_workspace_bytes": (_SZ, [_I64]),
    "ygg_morton_sort": (_INT, [_P, _I64, _P, _P, _P, _P, _SZ, _P]),
    "ygg_morton_sort_noalloc": (_INT, [_P, _I64, _P, _P, _P, _P]),
    "ygg_far_lists_sort_workspace_bytes": (_SZ, [_I64, _INT, _INT]),
    "ygg_far_lists_sort": (_INT, [_P, _P, _I64, _P, _I64, _INT, _INT, _P, _P, _P, _INT, _INT, _P,
                                  _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "ygg_near_lists_sort_workspace_bytes": (_SZ, [_I64, _INT, _INT]),
    "ygg_near_lists_sort": (_INT, [_P, _P, _I64, _P, _I64, _INT, _INT, _P, _P, _INT, _P, _P,
                                   _P, _P, _P, _P, _SZ, _P]),
    "ygg_near_dense_from_list": (_INT, [_I64, _P, _P, _P, _I64, _INT, _P, _P]),
    "ygg_target_block_offsets_workspace_bytes": (_SZ, [_I64]),
    "ygg_target_block_offsets": (_INT, [_I64, _P, _I64, _P, _P, _SZ, _P]),
    "ygg_target_blocks": (_INT, [_I64, _P, _P, _INT, _P, _I64, _P, _I64, _INT, _P, _P, _P,
                                 _P]),
    "ygg_morton_sort_reorder_workspace_bytes": (_SZ, [_I64]),
    "ygg_morton_sort_reorder": (_INT, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _INT, _P, _SZ, _P]),
    "ygg_morton_sort_reorder_async": (_INT, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _INT, _P, _SZ,
                                             _P, _P]),
    "ygg_reorder": (_INT, [_P, _P, _P, _I64, _P, _P, _P, _P]),
    "ygg_reorder_scatter": (_INT, [_P, _P, _P, _I64, _P, _P, _P, _P]),
    "ygg_radix_build_workspace_bytes": (_SZ, [_I64, _I64]),
    "ygg_radix_build": (_INT, [_P, _I64, _I64, ctypes.POINTER(RadixTopologyPtrs), _P, _SZ, _P]),
    "ygg_levels_pointer_doubling": (_INT, [ctypes.POINTER(RadixTopologyPtrs), _I64, _P, _SZ, _P]),
    "ygg_octree_meta_workspace_bytes": (_SZ, [_I64]),
    "ygg_octree_meta": (_INT, [_P, _P, _I64, ctypes.POINTER(OctreeMetaPtrs), _P, _SZ, _P]),
    "ygg_geometry_workspace_bytes": (_SZ, [_I64]),
    "ygg_geometry_level_major": (_INT, [_P] * 6 + [_I32, _I64] + [_P] * 5 + [_INT, _P]),
    "ygg_dense_interactions": (_INT, [_P] * 3 + [_INT, _I64, _I64] + [_P] * 2 +
                               [_I32, _I64, _I64] + [_P] * 7),
    "ygg_grouped_workspace_bytes": (_SZ, [_I64]),
    "ygg_tree_mass_moments_workspace_bytes": (_SZ, [_I64, _I64]),
    "ygg_tree_mass_moments": (_INT, [_P, _INT, _P, _INT, _I64, _P, _I64, _I64, _P, _P, _P, _SZ,
                                     _P]),
    "ygg_grouped_classify": (_INT, [_P, _P, _I64, _INT, _I64] + [_P] * 7 + [_P, _SZ, _P]),
    "ygg_grouped_emit": (_INT, [_I64, _INT] + [_P] * 10 + [_P, _SZ, _P]),
    "ygg_geometry": (_INT, [_P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _INT, _I32, _P, _P, _P,
                            _P, _P, _SZ, _P]),
    "ygg_radix_leaf_order": (_INT, [_I64, _P, _P, _P]),
    "ygg_walk_prepare_nodes": (_INT, [_P, _P, _P, _P, _P, _P, _P, _I64, _INT, _F, _P, _P]),
    "ygg_walk_flat": (_INT, [ctypes.POINTER(FlatWalkParams), _P, ctypes.POINTER(_I32), _P, _P]),
    "ygg_row_offsets": (_INT, [_P, _I64, _I64, _P, _P]),
    "ygg_far_lists_from_sorted": (_INT, [_I64, _P, _P, _I64, _P, _P, _P, _INT, _P, _P, _P, _P,
                                         _P, _P, _P]),
    "ygg_near_lists_from_sorted": (_INT, [_I64, _P, _P, _P, _I64, _INT, _P, _P, _P, _P, _P]),
    "ygg_walk_bfs": (_INT, [ctypes.POINTER(WalkParams), _P, ctypes.POINTER(_I32), _P, _P]),
    "ygg_scan_workspace_bytes": (_SZ, [_I64]),
    "ygg_exclusive_scan_i32": (_INT, [_P, _P, _P, _I64, _P, _SZ, _P]),
    "ygg_rows_from_offsets": (_INT, [_P, _P, _I64, _P, _P]),
    "ygg_cap_counts": (_INT, [_P, _I64, _I32, _P, _P]),
    "ygg_far_output": (_INT, [_I64, _P, _P, _P, _I32, _P, _P, _P, _INT, _P, _P, _P, _P, _P]),
    "ygg_near_output": (_INT, [_I64, _P, _P, _I32, _P, _P, _P, _I64, _INT, _P, _P, _P]),
    "ygg_policy_descriptor_actions": (_INT, [_P, _P, _I64, _I64, _INT, _F, _F,
                                             ctypes.POINTER(PolicyDesc)] + [_P] * 5),
    "ygg_policy_geometry": (_INT, [_P, _P, _I64, _I64, _INT, _F, _F] + [_P] * 10 + [_P]),
    "ygg_policy_apply": (_INT, [_P, _P, _I64, _I64] + [_P] * 5 + [_I64] + [_P] * 4 + [_I64] +
                         [_P] * 2 + [_I64] + [_P] * 3),
    "ygg_mlp_num_params": (_INT, []),
    "ygg_mlp_forward": (_INT, [_P, _P, _I64, _P, _P]),
    "ygg_mlp_train_workspace_floats": (_I64, [_I64]),
    "ygg_mlp_train_step": (_INT, [_P, _P, _P, _P, _I64, _F, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "ygg_mlp_params_to_bf16": (_INT, [_P, _P, _P]),
    "ygg_mlp_features": (_INT, [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ygg_mlp_totals": (_INT, [_P, _I64, _F, _F, _F, _P, _P]),
    "ygg_mlp_far_counts": (_INT, [_P, _I64, _F, _P, _P]),
}

_lib = None


def load():
    """Load the library (once). Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension first "
            "(python -m yggdrax_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols():
    return sorted(_SIGS)


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: {ERRORS.get(rc, rc)} (code {rc})")


def require_device(t: torch.Tensor) -> None:
    if not t.is_cuda:
        raise RuntimeError("yggdrax_amd runs on the GPU only; move inputs to a HIP device")


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
