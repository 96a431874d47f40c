"""Per-node geometry (yggdrax/geometry.py:18-75, _geometry_impl.py:176-338)."""

from __future__ import annotations

from typing import NamedTuple

import torch

from . import _lib
from .tree import device_topology


class TreeGeometry(NamedTuple):
    center: torch.Tensor
    half_extent: torch.Tensor
    radius: torch.Tensor
    max_extent: torch.Tensor


def compute_tree_geometry(tree, positions_sorted, *, max_leaf_size=None,
                          particle_radius=None) -> TreeGeometry:
    """Bounding boxes, centres, half extents (>= 1e-6), sphere radii and max
    half extents of every node, computed on the GPU."""
    if particle_radius is not None:
        raise NotImplementedError("particle_radius is not supported by the GPU geometry kernel")
    lib = _lib.load()
    dt = device_topology(tree)
    pos = torch.as_tensor(positions_sorted)
    _lib.require_device(pos)
    if pos.dtype != torch.float32:
        raise TypeError("positions_sorted must be float32")
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    if pos.shape[0] != dt.num_particles:
        raise ValueError("positions_sorted must match tree.num_particles")
    pos = pos.contiguous()
    cap = max_leaf_size if max_leaf_size is not None else (dt.leaf_size or dt.num_particles)
    T = dt.num_nodes
    dev = pos.device
    center = torch.empty((T, 3), dtype=torch.float32, device=dev)
    half = torch.empty((T, 3), dtype=torch.float32, device=dev)
    radius = torch.empty(T, dtype=torch.float32, device=dev)
    maxext = torch.empty(T, dtype=torch.float32, device=dev)
    ws = _lib.workspace(lib.ygg_geometry_workspace_bytes(T), dev)
    left = dt.left if dt.left.numel() else None
    right = dt.right if dt.right.numel() else None
    _lib.check(lib.ygg_geometry(
        pos.data_ptr(), dt.node_ranges.data_ptr(), _lib.ptr(left), _lib.ptr(right),
        dt.parent.data_ptr(), dt.node_level.data_ptr(), T, int(dt.leaf_size), int(cap),
        int(dt.malformed),
        int(dt.max_level), center.data_ptr(), half.data_ptr(), radius.data_ptr(),
        maxext.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev)), "geometry")
    return TreeGeometry(center, half, radius, maxext)


__all__ = ["TreeGeometry", "compute_tree_geometry"]
