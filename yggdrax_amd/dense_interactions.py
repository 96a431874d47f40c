"""Dense level-major far-field buffers for a downward sweep
(yggdrax/dense_interactions.py).

The sparse level-major CSR far list is regrouped into padded
``(num_levels, max_nodes, max_interactions)`` buffers, one HIP kernel
writing every slot and entry (``ygg_dense_interactions``)."""

from __future__ import annotations

from typing import Callable, NamedTuple, Optional

import torch

from . import _lib
from .dtypes import INDEX_DTYPE
from .geometry import LevelMajorTreeGeometry, TreeGeometry, geometry_to_level_major
from .interactions import (DualTreeRetryEvent, DualTreeTraversalConfig, NodeInteractionList,
                           build_interactions_and_neighbors)
from .tree import device_topology, require_fmm_core_topology, resolve_tree_topology


class DenseInteractionBuffers(NamedTuple):
    """Dense level-major far-field pairs (dense_interactions.py:23-31)."""

    geometry: LevelMajorTreeGeometry
    m2l_sources: torch.Tensor
    m2l_displacements: torch.Tensor
    m2l_mask: torch.Tensor
    m2l_counts: torch.Tensor
    sparse_interactions: NodeInteractionList


def densify_interactions(tree, geometry: TreeGeometry,
                         interactions: NodeInteractionList) -> DenseInteractionBuffers:
    """Padded level-major M2L buffers of a far list (dense_interactions.py:
    143-187, _dense_m2l_buffers 35-140): for slot (level, j) holding target
    node t, entry k < counts[t] is the k-th source s of t's row, with mask
    True and displacement center[t] - center[s]; every other entry is -1,
    False and zeros. ``max_interactions`` is max(1, max count)."""
    topology = resolve_tree_topology(tree)
    require_fmm_core_topology(topology)
    level_geometry = geometry_to_level_major(topology, geometry)
    dt = device_topology(tree)
    num_levels, max_nodes = (int(s) for s in level_geometry.node_indices.shape)
    dev = geometry.center.device
    counts = interactions.counts.to(INDEX_DTYPE).contiguous()
    offsets = interactions.offsets.to(INDEX_DTYPE).contiguous()
    sources = interactions.sources.to(INDEX_DTYPE).contiguous()
    T = dt.num_nodes
    if counts.numel() != T or offsets.numel() < T:
        raise ValueError(f"interaction counts/offsets do not describe the tree's {T} nodes")
    num_far = int(sources.numel())
    if num_far == 0:
        sources = torch.zeros(1, dtype=INDEX_DTYPE, device=dev)
    k = max(1, int(counts.max()) if counts.numel() else 0)
    m2l_sources = torch.empty((num_levels, max_nodes, k), dtype=INDEX_DTYPE, device=dev)
    m2l_mask = torch.empty((num_levels, max_nodes, k), dtype=torch.bool, device=dev)
    m2l_counts = torch.empty((num_levels, max_nodes), dtype=INDEX_DTYPE, device=dev)
    m2l_disp = torch.empty((num_levels, max_nodes, k, 3), dtype=torch.float32, device=dev)
    err = torch.empty(1, dtype=torch.int32, device=dev)
    lib = _lib.load()
    _lib.check(lib.ygg_dense_interactions(
        offsets.data_ptr(), counts.data_ptr(), sources.data_ptr(),
        int(INDEX_DTYPE == torch.int64), T, num_far, dt.nodes_by_level.data_ptr(),
        dt.level_offsets.data_ptr(), num_levels, max_nodes, k,
        geometry.center.contiguous().data_ptr(), m2l_sources.data_ptr(), m2l_disp.data_ptr(),
        m2l_mask.data_ptr(), m2l_counts.data_ptr(), err.data_ptr(), _lib.stream_ptr(dev)),
        "dense interactions")
    if int(err.item()):
        raise ValueError("interaction offsets/counts/sources are not a CSR list over the tree's "
                         f"{T} nodes")
    return DenseInteractionBuffers(level_geometry, m2l_sources, m2l_disp, m2l_mask, m2l_counts,
                                   interactions)


def build_dense_interactions(
        tree, geometry: TreeGeometry, theta: float = 0.5, *,
        max_pair_queue: Optional[int] = None, process_block: Optional[int] = None,
        traversal_config: Optional[DualTreeTraversalConfig] = None,
        retry_logger: Optional[Callable[[DualTreeRetryEvent], object]] = None,
) -> DenseInteractionBuffers:
    """build_interactions_and_neighbors, then densify_interactions
    (dense_interactions.py:190-247)."""
    interactions, _ = build_interactions_and_neighbors(
        tree, geometry, theta=theta, max_pair_queue=max_pair_queue,
        process_block=process_block, traversal_config=traversal_config,
        retry_logger=retry_logger)
    return densify_interactions(tree, geometry, interactions)


__all__ = ["DenseInteractionBuffers", "build_dense_interactions", "densify_interactions"]
