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
    half extents of every node, computed on the GPU. ``particle_radius``
    (per particle, in the sorted order of ``positions_sorted``) bounds each
    particle as the cube [p - r, p + r] instead of the point p
    (_geometry_impl.py:107-112), so every node's box covers the balls its
    particles stand for."""
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
    prad = None
    if particle_radius is not None:
        prad = torch.as_tensor(particle_radius, device=pos.device).to(torch.float32).contiguous()
        if prad.shape != (pos.shape[0],):
            raise ValueError("particle_radius must have shape (n_particles,)")
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
        pos.data_ptr(), _lib.ptr(prad), dt.node_ranges.data_ptr(), _lib.ptr(left),
        _lib.ptr(right),
        dt.parent.data_ptr(), dt.node_level.data_ptr(), T, int(dt.leaf_size), int(cap),
        int(dt.malformed),
        int(dt.max_level), center.data_ptr(), half.data_ptr(), radius.data_ptr(),
        maxext.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev)), "geometry")
    return TreeGeometry(center, half, radius, maxext)


class LevelMajorTreeGeometry(NamedTuple):
    """Padded level-major geometry views (_geometry_impl.py:342-433)."""

    centers: torch.Tensor        # [num_levels, max_nodes, 3]
    half_extents: torch.Tensor   # [num_levels, max_nodes, 3]
    radii: torch.Tensor          # [num_levels, max_nodes]
    max_extents: torch.Tensor    # [num_levels, max_nodes]
    level_counts: torch.Tensor   # [num_levels]
    node_indices: torch.Tensor   # [num_levels, max_nodes], -1 padded


def geometry_to_level_major(tree, geometry: TreeGeometry) -> LevelMajorTreeGeometry:
    """Padded, level-major views of the node geometry
    (yggdrax/geometry.py:79-102, _geometry_impl.py:342-433), gathered by one
    kernel."""
    from .dtypes import INDEX_DTYPE
    lib = _lib.load()
    dt = device_topology(tree)
    num_levels = int(dt.num_levels)
    if num_levels <= 0:
        raise ValueError("tree must contain at least one populated level")
    dev = geometry.center.device
    if tuple(geometry.center.shape) != (dt.num_nodes, 3):
        raise ValueError(f"geometry.center must have shape ({dt.num_nodes}, 3)")
    off = dt.level_offsets[: num_levels + 1].to(torch.int32).contiguous()
    counts = (off[1:] - off[:-1]).to(INDEX_DTYPE)
    max_nodes = int(torch.clamp(counts, min=1).max())
    f32 = dict(dtype=torch.float32, device=dev)
    centers = torch.empty((num_levels, max_nodes, 3), **f32)
    halves = torch.empty((num_levels, max_nodes, 3), **f32)
    radii = torch.empty((num_levels, max_nodes), **f32)
    maxes = torch.empty((num_levels, max_nodes), **f32)
    idx = torch.empty((num_levels, max_nodes), dtype=INDEX_DTYPE, device=dev)
    g = geometry
    _lib.check(lib.ygg_geometry_level_major(
        g.center.contiguous().data_ptr(), g.half_extent.contiguous().data_ptr(),
        g.radius.contiguous().data_ptr(), g.max_extent.contiguous().data_ptr(),
        dt.nodes_by_level.data_ptr(), off.data_ptr(), num_levels, max_nodes, centers.data_ptr(),
        halves.data_ptr(), radii.data_ptr(), maxes.data_ptr(), idx.data_ptr(),
        int(INDEX_DTYPE == torch.int64), _lib.stream_ptr(dev)), "geometry level major")
    return LevelMajorTreeGeometry(centers, halves, radii, maxes, counts, idx)


__all__ = ["LevelMajorTreeGeometry", "TreeGeometry", "compute_tree_geometry",
           "geometry_to_level_major"]
