"""Per-node mass moments (yggdrax/tree_moments.py:686-746) on the GPU.

``compute_tree_mass_moments`` gives every node's total mass and centre of
mass over its particle range in the Morton-sorted order. The reference
differences float prefix sums; here each node sum is accumulated in float64
by the ``ygg_tree_mass_moments`` kernels and rounded once, so the results
are the node sums of the reference's per-particle terms (m * x in the
product dtype) to within one rounding, without the prefix difference's
cancellation.
"""

from __future__ import annotations

from typing import NamedTuple

import torch

from . import _lib
from .tree import device_topology, require_fmm_core_topology, resolve_tree_topology


class TreeMassMoments(NamedTuple):
    """tree_moments.py:33-37."""

    mass: torch.Tensor            # [num_nodes], masses' dtype
    center_of_mass: torch.Tensor  # [num_nodes, 3]


def _validate_inputs(topology, positions_sorted, masses_sorted, num_nodes: int) -> None:
    """tree_moments.py:660-683."""
    if topology.node_ranges.shape[0] != num_nodes:
        raise ValueError("tree.node_ranges must align with tree.parent shape")
    num_particles = int(topology.num_particles)
    if positions_sorted.shape[0] != num_particles:
        raise ValueError("positions_sorted must match tree.num_particles")
    if masses_sorted.shape[0] != num_particles:
        raise ValueError("masses_sorted must match tree.num_particles")
    if positions_sorted.ndim != 2 or positions_sorted.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")


def compute_tree_mass_moments(tree, positions_sorted, masses_sorted) -> TreeMassMoments:
    """Total mass and centre of mass of every node (tree_moments.py:686-746):
    mass[t] = sum of m over node t's range, center_of_mass[t] = sum(m x) /
    mass[t] in the product dtype, 0 where mass[t] == 0."""
    topology = resolve_tree_topology(tree)
    require_fmm_core_topology(topology)
    dt = device_topology(tree)
    pos = torch.as_tensor(positions_sorted)
    mass = torch.as_tensor(masses_sorted)
    _lib.require_device(pos)
    _lib.require_device(mass)
    _validate_inputs(topology, pos, mass, dt.num_nodes)
    for name, t in (("positions_sorted", pos), ("masses_sorted", mass)):
        if t.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"{name} must be float32 or float64")
    pos = pos.contiguous()
    mass = mass.contiguous()
    pos64, mass64 = pos.dtype == torch.float64, mass.dtype == torch.float64
    dev = pos.device
    n, T = int(pos.shape[0]), dt.num_nodes
    mass_out = torch.empty(T, dtype=mass.dtype, device=dev)
    center = torch.empty((T, 3), dtype=torch.float64 if (pos64 or mass64) else torch.float32,
                         device=dev)
    if n == 0:
        return TreeMassMoments(mass_out.zero_(), center.zero_())
    chunk = int(dt.leaf_size) if dt.leaf_size else 64
    lib = _lib.load()
    ws = _lib.workspace(lib.ygg_tree_mass_moments_workspace_bytes(n, chunk), dev)
    _lib.check(lib.ygg_tree_mass_moments(
        pos.data_ptr(), int(pos64), mass.data_ptr(), int(mass64), n, dt.node_ranges.data_ptr(),
        T, chunk, mass_out.data_ptr(), center.data_ptr(), ws.data_ptr(), ws.numel(),
        _lib.stream_ptr(dev)), "tree mass moments")
    return TreeMassMoments(mass_out, center)


__all__ = ["TreeMassMoments", "compute_tree_mass_moments"]
