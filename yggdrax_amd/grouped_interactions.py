"""Class-major grouping of far-field pairs (yggdrax/grouped_interactions.py).

Far pairs are keyed by (target level, source level, rint(delta / cell) per
axis), where delta is the float64 target-minus-source centre difference
and cell the target level's cell size, then ordered lexicographically by
key and, within a class, by input position. The reference does this on
the host with NumPy; here the keys, a stable radix sort of the packed keys,
the class boundaries and all outputs are HIP kernels
(``ygg_grouped_classify`` / ``ygg_grouped_emit``)."""

from __future__ import annotations

import ctypes
from typing import NamedTuple, Optional

import torch

from . import _lib
from .dtypes import INDEX_DTYPE
from .geometry import TreeGeometry
from .tree import device_topology


class GroupedInteractionBuffers(NamedTuple):
    """grouped_interactions.py:21-31."""

    class_keys: torch.Tensor           # (num_classes, 5) int32
    class_displacements: torch.Tensor  # (num_classes, 3) float32
    class_offsets: torch.Tensor        # (num_classes + 1,)
    class_sources: torch.Tensor        # (num_pairs,) in class order
    class_targets: torch.Tensor        # (num_pairs,) in class order
    class_ids: torch.Tensor            # (num_pairs,) in input order
    level_offsets: torch.Tensor
    level_nodes: torch.Tensor


def build_grouped_interactions(tree, geometry: TreeGeometry,
                               interactions) -> GroupedInteractionBuffers:
    """Group a NodeInteractionList's far pairs (grouped_interactions.py:41-71)."""
    return build_grouped_interactions_from_pairs(
        tree, geometry, interactions.sources, interactions.targets,
        level_offsets=getattr(interactions, "level_offsets", None))


def build_grouped_interactions_from_pairs(
        tree, geometry: TreeGeometry, sources: torch.Tensor, targets: torch.Tensor, *,
        level_offsets: Optional[torch.Tensor] = None) -> GroupedInteractionBuffers:
    """Displacement classes of far pairs (grouped_interactions.py:74-207).

    class_keys[c] = (target level, source level, ix, iy, iz) with
    i = rint((c_t - c_s) / cell) in float64, cell = max(extent / 2^level,
    1e-12) of the target level and extent = bounds_max - bounds_min; indices
    outside int32 become INT32_MIN, as NumPy's cast gives them on x86-64.
    class_displacements[c] = i * cell rounded to float32. Pairs are sorted
    stably by key: class_sources/targets list the pairs class by class,
    class_offsets is the class CSR and class_ids[p] the class of input pair p.
    """
    dt = device_topology(tree)
    dev = geometry.center.device
    src = torch.as_tensor(sources, device=dev).to(INDEX_DTYPE).contiguous().reshape(-1)
    tgt = torch.as_tensor(targets, device=dev).to(INDEX_DTYPE).contiguous().reshape(-1)
    n = int(src.numel())
    if int(tgt.numel()) != n:
        raise ValueError("sources and targets must have the same length")
    level_nodes = dt.nodes_by_level.to(INDEX_DTYPE)
    lo = (level_offsets.to(INDEX_DTYPE) if level_offsets is not None
          else torch.zeros(1, dtype=INDEX_DTYPE, device=dev))
    idx = dict(dtype=INDEX_DTYPE, device=dev)
    if n == 0:
        return GroupedInteractionBuffers(
            torch.zeros((0, 5), dtype=torch.int32, device=dev),
            torch.zeros((0, 3), dtype=geometry.center.dtype, device=dev),
            torch.zeros(1, **idx), torch.zeros(0, **idx), torch.zeros(0, **idx),
            torch.zeros(0, **idx), lo, level_nodes)
    bmin = torch.as_tensor(tree.bounds_min, device=dev)
    bmax = torch.as_tensor(tree.bounds_max, device=dev)
    if bmin.ndim > 1:
        bmin, bmax = bmin[0], bmax[0]
    bmin = bmin.to(torch.float32).contiguous()
    bmax = bmax.to(torch.float32).contiguous()
    center = geometry.center.to(torch.float32).contiguous()
    if center.shape != (dt.num_nodes, 3):
        raise ValueError(f"geometry.center must have shape ({dt.num_nodes}, 3)")
    lib = _lib.load()
    st = _lib.stream_ptr(dev)
    i64 = int(INDEX_DTYPE == torch.int64)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    scan = torch.empty(n + 1, dtype=torch.int64, device=dev)
    ws = _lib.workspace(lib.ygg_grouped_workspace_bytes(n), dev)
    num_classes = ctypes.c_int64(0)
    rc = lib.ygg_grouped_classify(
        src.data_ptr(), tgt.data_ptr(), n, i64, dt.num_nodes, dt.node_level.data_ptr(),
        center.data_ptr(), bmin.data_ptr(), bmax.data_ptr(), order.data_ptr(), scan.data_ptr(),
        ctypes.byref(num_classes), ws.data_ptr(), ws.numel(), st)
    if rc == 1:  # YGG_ERR_ARG: a pair names a node outside the tree
        raise IndexError(f"far-pair node index out of range [0, {dt.num_nodes})")
    _lib.check(rc, "grouped classify")
    c = int(num_classes.value)
    class_keys = torch.empty((c, 5), dtype=torch.int32, device=dev)
    class_disp = torch.empty((c, 3), dtype=torch.float32, device=dev)
    class_offsets = torch.empty(c + 1, **idx)
    class_sources = torch.empty(n, **idx)
    class_targets = torch.empty(n, **idx)
    class_ids = torch.empty(n, **idx)
    _lib.check(lib.ygg_grouped_emit(
        n, i64, bmin.data_ptr(), bmax.data_ptr(),
        order.data_ptr(), scan.data_ptr(), class_sources.data_ptr(), class_targets.data_ptr(),
        class_ids.data_ptr(), class_offsets.data_ptr(), class_keys.data_ptr(),
        class_disp.data_ptr(), ws.data_ptr(), ws.numel(), st), "grouped emit")
    return GroupedInteractionBuffers(class_keys, class_disp.to(geometry.center.dtype),
                                     class_offsets, class_sources, class_targets, class_ids,
                                     lo, level_nodes)


__all__ = ["GroupedInteractionBuffers", "build_grouped_interactions",
           "build_grouped_interactions_from_pairs"]
