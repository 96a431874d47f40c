"""Batched mode: many independent particle systems built and traversed together.

The systems' points are concatenated (``offsets[s]:offsets[s+1]`` is system
s). Every stage launches once for the whole batch: per-system bounds and
Morton codes, one stable sort by (system, code), the Karras build, levels
and octree metadata over all systems at once (csrc/ygg_batch.hip), one
segment tree for every system's geometry, and ONE dual-tree walk over the
forest of all systems' trees (one seed pair per root). Each system's slice
of the outputs equals what build_octree / compute_tree_geometry /
build_interactions_and_neighbors return for that system alone
(``BatchedTrees.system(k)``, ``BatchedLists.system(k)``).

The reference has no batched mode: its entry points build one tree per call
(yggdrax/tree.py:1413-1549, interactions.py:338-448); per system the results
follow those functions exactly.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from types import SimpleNamespace
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .dtypes import INDEX_DTYPE
from .geometry import TreeGeometry
from .interactions import (FLAT_GRID_BLOCKS, MAC_TYPES, NodeInteractionList, NodeNeighborList,
                           _WalkContext, _ladder, _leaf_tables, _run_walk_flat)
from .tree import _assemble_tree, _ptrs, build_octree, build_tree


@dataclass
class BatchLayout:
    """Offsets of every system in the concatenated arrays (host copies and
    the device tables the kernels read)."""

    num_systems: int
    leaf_size: int
    pt_off: np.ndarray     # [S + 1] point offsets
    leaf_off: np.ndarray   # [S + 1] leaf offsets
    device: torch.device
    pt_off_d: torch.Tensor = None
    leaf_off_d: torch.Tensor = None
    struct: _lib.Batch = None

    @property
    def n(self) -> np.ndarray:
        return np.diff(self.pt_off)

    @property
    def L(self) -> np.ndarray:
        return np.diff(self.leaf_off)

    @property
    def node_off(self) -> np.ndarray:
        return 2 * self.leaf_off - np.arange(self.num_systems + 1)

    @property
    def ni_off(self) -> np.ndarray:
        return self.leaf_off - np.arange(self.num_systems + 1)

    @property
    def num_points(self) -> int:
        return int(self.pt_off[-1])

    @property
    def num_leaves(self) -> int:
        return int(self.leaf_off[-1])

    @property
    def num_nodes(self) -> int:
        return int(self.node_off[-1])

    @property
    def num_internal(self) -> int:
        return int(self.ni_off[-1])

    def ptr(self):
        return ctypes.byref(self.struct)


def make_layout(offsets, leaf_size: int, device) -> BatchLayout:
    pt = np.asarray(offsets.cpu() if torch.is_tensor(offsets) else offsets, dtype=np.int64)
    if pt.ndim != 1 or pt.size < 2 or pt[0] != 0 or np.any(np.diff(pt) < 1):
        raise ValueError("offsets must start at 0 and give every system at least one particle")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    if pt[-1] >= (1 << 31) or np.max(np.diff(pt)) >= (1 << 30):
        raise ValueError("a batch holds fewer than 2^31 particles, at most 2^30 - 1 per system")
    S = pt.size - 1
    L = (np.diff(pt) + leaf_size - 1) // leaf_size
    leaf_off = np.concatenate([[0], np.cumsum(L)]).astype(np.int64)
    lay = BatchLayout(S, int(leaf_size), pt, leaf_off, torch.device(device))
    lay.pt_off_d = torch.from_numpy(pt.astype(np.int32)).to(device)
    lay.leaf_off_d = torch.from_numpy(leaf_off.astype(np.int32)).to(device)
    b = _lib.Batch()
    b.num_systems, b.pad0, b.leaf_size = S, 0, int(leaf_size)
    b.point_offsets, b.leaf_offsets = lay.pt_off_d.data_ptr(), lay.leaf_off_d.data_ptr()
    b.num_points, b.num_leaves = int(pt[-1]), int(leaf_off[-1])
    b.max_leaves_per_system = int(L.max())
    lay.struct = b
    return lay


@dataclass
class BatchedTrees:
    layout: BatchLayout
    octree: bool
    rad: dict
    octa: Optional[dict]
    sorted_codes: torch.Tensor
    perm_local: torch.Tensor
    bounds6: torch.Tensor                 # [S, 6]
    scalars: np.ndarray                   # [S, 8] radix scalars
    oct_scalars: Optional[np.ndarray]     # [S, 4]
    positions_sorted: Optional[torch.Tensor]
    masses_sorted: Optional[torch.Tensor]
    inverse_permutation: Optional[torch.Tensor]
    fallback: List[int] = field(default_factory=list)  # systems rebuilt one at a time
    _views: dict = field(default_factory=dict)

    @property
    def num_systems(self) -> int:
        return self.layout.num_systems

    def _slices(self, k):
        lay = self.layout
        p0, p1 = int(lay.pt_off[k]), int(lay.pt_off[k + 1])
        l0, l1 = int(lay.leaf_off[k]), int(lay.leaf_off[k + 1])
        n0, n1 = int(lay.node_off[k]), int(lay.node_off[k + 1])
        i0, i1 = int(lay.ni_off[k]), int(lay.ni_off[k + 1])
        return p0, p1, l0, l1, n0, n1, i0, i1

    def radix_slices(self, k) -> dict:
        p0, p1, l0, l1, n0, n1, i0, i1 = self._slices(k)
        r = self.rad
        return dict(parent=r["parent"][n0:n1], left_child=r["left_child"][i0:i1],
                    right_child=r["right_child"][i0:i1], left_is_leaf=r["left_is_leaf"][i0:i1],
                    right_is_leaf=r["right_is_leaf"][i0:i1], node_ranges=r["node_ranges"][n0:n1],
                    node_level=r["node_level"][n0:n1],
                    level_offsets=r["level_offsets"][65 * k:65 * k + 65],
                    nodes_by_level=r["nodes_by_level"][n0:n1], leaf_codes=r["leaf_codes"][l0:l1],
                    scalars=r["scalars"][8 * k:8 * k + 8])

    def octree_slices(self, k) -> dict:
        p0, p1, l0, l1, n0, n1, i0, i1 = self._slices(k)
        o = self.octa
        out = {name: o[name][n0:n1] for name in (
            "oct_parent", "oct_children", "oct_child_counts", "oct_child_mask", "oct_valid_mask",
            "oct_node_codes", "oct_node_depths", "oct_node_ranges", "oct_nodes_by_level",
            "oct_leaf_mask", "oct_leaf_nodes", "radix_node_to_oct", "leaf_position")}
        out["oct_level_offsets"] = o["oct_level_offsets"][23 * k:23 * k + 23]
        out["leaf_indices"] = o["leaf_indices"][l0:l1]
        out["scalars"] = o["scalars"][4 * k:4 * k + 4]
        return out

    def system(self, k: int):
        """System k as the tree build_octree / build_tree returns for it."""
        k = int(k)
        if k in self._views:
            return self._views[k]
        p0, p1, l0, l1, n0, n1, i0, i1 = self._slices(k)
        rad = self.radix_slices(k)
        octa = self.octree_slices(k) if self.octree else None
        if octa is not None:
            leaf_indices, leaf_position = octa["leaf_indices"], octa["leaf_position"]
        else:
            leaf_indices = self.rad["leaf_indices"][l0:l1]
            leaf_position = self.rad["leaf_position"][n0:n1]
        sc = self.scalars[k]
        pos_s = None if self.positions_sorted is None else self.positions_sorted[p0:p1]
        mass_s = None if self.masses_sorted is None else self.masses_sorted[p0:p1]
        inv = None if self.inverse_permutation is None else self.inverse_permutation[p0:p1]
        t = _assemble_tree(rad, octa, self.perm_local[p0:p1], self.sorted_codes[p0:p1],
                           self.bounds6[k], p1 - p0, self.layout.leaf_size, int(sc[0]),
                           int(sc[1]), bool(sc[2]) or bool(sc[3]),
                           int(self.oct_scalars[k][0]) if self.octree else 0, leaf_indices,
                           leaf_position, pos_s, mass_s, inv, False)
        self._views[k] = t
        return t


def _copy_system(bt: BatchedTrees, k: int, tree) -> None:
    """Overwrite system k's slices with a single-system build's arrays."""
    dt = tree._device
    rs = bt.radix_slices(k)
    tp = tree.topology
    rs["parent"].copy_(dt.parent)
    rs["left_child"].copy_(dt.left)
    rs["right_child"].copy_(dt.right)
    rs["left_is_leaf"].copy_(tp.left_is_leaf.to(torch.uint8))
    rs["right_is_leaf"].copy_(tp.right_is_leaf.to(torch.uint8))
    rs["node_ranges"].copy_(dt.node_ranges)
    rs["node_level"].copy_(dt.node_level)
    rs["level_offsets"].copy_(dt.level_offsets)
    rs["nodes_by_level"].copy_(dt.nodes_by_level)
    rs["leaf_codes"].copy_(tp.leaf_codes.view(torch.int64))
    bt.scalars[k, 0] = int(tp.num_levels)
    bt.scalars[k, 1] = dt.max_level
    bt.scalars[k, 3] = int(dt.malformed)
    bt.scalars[k, 2] = 0
    if bt.octree:
        os_ = bt.octree_slices(k)
        for name in ("oct_parent", "oct_children", "oct_child_counts", "oct_node_depths",
                     "oct_node_ranges", "oct_nodes_by_level", "oct_leaf_nodes",
                     "radix_node_to_oct", "oct_level_offsets"):
            os_[name].copy_(getattr(tp, name).to(torch.int32))
        for name in ("oct_child_mask", "oct_valid_mask", "oct_leaf_mask"):
            os_[name].copy_(getattr(tp, name).to(torch.uint8))
        os_["oct_node_codes"].copy_(tp.oct_node_codes.view(torch.int64))
        os_["leaf_indices"].copy_(dt.leaf_indices)
        os_["leaf_position"].copy_(dt.leaf_position)
        bt.oct_scalars[k, 0] = int(tp.oct_num_levels)
    else:
        p0, p1, l0, l1, n0, n1, i0, i1 = bt._slices(k)
        bt.rad["leaf_indices"][l0:l1].copy_(dt.leaf_indices)
        bt.rad["leaf_position"][n0:n1].copy_(dt.leaf_position)
    bt._views.pop(k, None)


def _build_batched(positions, masses, offsets, *, leaf_size: int, octree: bool,
                   return_reordered: bool, layout: Optional[BatchLayout] = None,
                   sync: bool = True) -> BatchedTrees:
    """sync=False (stream capture): no host reads; the per-system scalars
    and the deep/malformed fallback are left to the caller."""
    lib = _lib.load()
    pos = positions
    _lib.require_device(pos)
    pos = pos.to(torch.float32).contiguous()
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    dev = pos.device
    lay = layout if layout is not None else make_layout(offsets, leaf_size, dev)
    if lay.num_points != pos.shape[0]:
        raise ValueError("offsets[-1] must equal the number of particles")
    st = _lib.stream_ptr(dev)
    S, N, Lt, T = lay.num_systems, lay.num_points, lay.num_leaves, lay.num_nodes
    NI = lay.num_internal
    b6 = torch.empty((S, 6), dtype=torch.float32, device=dev)
    _lib.check(lib.ygg_batch_infer_bounds(lay.ptr(), pos.data_ptr(), b6.data_ptr(), st),
               "batch bounds")
    sorted_codes = torch.empty(N, dtype=torch.int64, device=dev)
    perm = torch.empty(N, dtype=torch.int32, device=dev)
    ws = _lib.workspace(lib.ygg_batch_morton_sort_workspace_bytes(lay.ptr()), dev)
    _lib.check(lib.ygg_batch_morton_sort(lay.ptr(), pos.data_ptr(), b6.data_ptr(),
                                         sorted_codes.data_ptr(), perm.data_ptr(), ws.data_ptr(),
                                         ws.numel(), st), "batch morton sort")
    i32 = dict(dtype=torch.int32, device=dev)
    u8 = dict(dtype=torch.uint8, device=dev)
    rad = dict(
        parent=torch.empty(T, **i32), left_child=torch.empty(max(NI, 1), **i32),
        right_child=torch.empty(max(NI, 1), **i32), left_is_leaf=torch.empty(max(NI, 1), **u8),
        right_is_leaf=torch.empty(max(NI, 1), **u8), node_ranges=torch.empty((T, 2), **i32),
        node_level=torch.empty(T, **i32), level_offsets=torch.empty(65 * S, **i32),
        nodes_by_level=torch.empty(T, **i32), leaf_codes=torch.empty(Lt, dtype=torch.int64,
                                                                     device=dev),
        scalars=torch.zeros(8 * S, **i32))
    rws = _lib.workspace(lib.ygg_batch_radix_build_workspace_bytes(lay.ptr()), dev)
    _lib.check(lib.ygg_batch_radix_build(lay.ptr(), sorted_codes.data_ptr(),
                                         ctypes.byref(_ptrs(_lib.RadixTopologyPtrs, rad)),
                                         rws.data_ptr(), rws.numel(), st), "batch radix build")
    octa = None
    if octree:
        M = T
        octa = dict(
            oct_parent=torch.empty(M, **i32), oct_children=torch.empty((M, 8), **i32),
            oct_child_counts=torch.empty(M, **i32), oct_child_mask=torch.empty((M, 8), **u8),
            oct_valid_mask=torch.empty(M, **u8),
            oct_node_codes=torch.empty(M, dtype=torch.int64, device=dev),
            oct_node_depths=torch.empty(M, **i32), oct_node_ranges=torch.empty((M, 2), **i32),
            oct_nodes_by_level=torch.empty(M, **i32), oct_level_offsets=torch.empty(23 * S, **i32),
            oct_leaf_mask=torch.empty(M, **u8), oct_leaf_nodes=torch.empty(M, **i32),
            radix_node_to_oct=torch.empty(T, **i32), leaf_indices=torch.empty(Lt, **i32),
            leaf_position=torch.empty(T, **i32), scalars=torch.zeros(4 * S, **i32))
        ows = _lib.workspace(lib.ygg_batch_octree_meta_workspace_bytes(lay.ptr()), dev)
        _lib.check(lib.ygg_batch_octree_meta(lay.ptr(), sorted_codes.data_ptr(),
                                             rad["node_ranges"].data_ptr(),
                                             ctypes.byref(_ptrs(_lib.OctreeMetaPtrs, octa)),
                                             ows.data_ptr(), ows.numel(), st), "batch octree")
    else:
        rad["leaf_indices"] = torch.empty(Lt, **i32)
        rad["leaf_position"] = torch.empty(T, **i32)
        _lib.check(lib.ygg_batch_radix_leaf_order(lay.ptr(), rad["leaf_indices"].data_ptr(),
                                                  rad["leaf_position"].data_ptr(), st),
                   "batch leaf order")
    pos_sorted = mass_sorted = inv = None
    m = None
    if masses is not None:
        m = torch.as_tensor(masses, dtype=torch.float32, device=dev).contiguous()
    perm_local = torch.empty(N, **i32)
    if return_reordered:
        pos_sorted = torch.empty_like(pos)
        inv = torch.empty(N, **i32)
        if m is not None:
            mass_sorted = torch.empty_like(m)
    _lib.check(lib.ygg_batch_reorder(lay.ptr(), pos.data_ptr(), _lib.ptr(m), perm.data_ptr(),
                                     _lib.ptr(pos_sorted), _lib.ptr(mass_sorted),
                                     perm_local.data_ptr(), _lib.ptr(inv), st), "batch reorder")
    if not sync:
        return BatchedTrees(lay, octree, rad, octa, sorted_codes, perm_local, b6, None, None,
                            pos_sorted, mass_sorted, None if inv is None else inv)
    # one host read: every system's scalars
    scal = rad["scalars"].view(S, 8).cpu().numpy().copy()
    osc = octa["scalars"].view(S, 4).cpu().numpy().copy() if octree else None
    bt = BatchedTrees(lay, octree, rad, octa, sorted_codes, perm_local, b6, scal, osc,
                      pos_sorted, mass_sorted, None if inv is None else inv)
    # Deep (levels > 255) or malformed (duplicate leaf codes) trees: rebuild
    # those systems one at a time (exact pointer-doubling levels, level-
    # synchronous geometry).
    bad = np.nonzero((scal[:, 2] != 0) | (scal[:, 3] != 0))[0]
    for k in bad.tolist():
        p0, p1 = int(lay.pt_off[k]), int(lay.pt_off[k + 1])
        mk = None if m is None else m[p0:p1]
        build = build_octree if octree else build_tree
        t = build(pos[p0:p1], mk, leaf_size=leaf_size)
        _copy_system(bt, k, t)
        bt.fallback.append(k)
    return bt


def build_octree_batched(positions, masses, offsets, *, leaf_size: int = 8,
                         return_reordered: bool = True) -> BatchedTrees:
    """build_octree (yggdrax/tree.py:1482-1549) for every system of a batch."""
    return _build_batched(positions, masses, offsets, leaf_size=leaf_size, octree=True,
                          return_reordered=return_reordered)


def build_tree_batched(positions, masses, offsets, *, leaf_size: int = 8,
                       return_reordered: bool = True) -> BatchedTrees:
    """build_tree (yggdrax/tree.py:1413-1478) for every system of a batch."""
    return _build_batched(positions, masses, offsets, leaf_size=leaf_size, octree=False,
                          return_reordered=return_reordered)


@dataclass
class BatchedGeometry:
    layout: BatchLayout
    center: torch.Tensor
    half_extent: torch.Tensor
    radius: torch.Tensor
    max_extent: torch.Tensor

    def system(self, k: int) -> TreeGeometry:
        n0, n1 = int(self.layout.node_off[k]), int(self.layout.node_off[k + 1])
        return TreeGeometry(self.center[n0:n1], self.half_extent[n0:n1], self.radius[n0:n1],
                            self.max_extent[n0:n1])


def compute_tree_geometry_batched(trees: BatchedTrees, positions_sorted=None,
                                  max_leaf_size: Optional[int] = None) -> BatchedGeometry:
    """compute_tree_geometry (yggdrax/geometry.py:18-75) for every system."""
    from .geometry import compute_tree_geometry
    lib = _lib.load()
    lay = trees.layout
    pos = positions_sorted if positions_sorted is not None else trees.positions_sorted
    if pos is None:
        raise ValueError("positions_sorted is required")
    pos = pos.to(torch.float32).contiguous()
    dev = pos.device
    T = lay.num_nodes
    f32 = dict(dtype=torch.float32, device=dev)
    center, half = torch.empty((T, 3), **f32), torch.empty((T, 3), **f32)
    radius, maxext = torch.empty(T, **f32), torch.empty(T, **f32)
    cap = int(max_leaf_size) if max_leaf_size is not None else lay.leaf_size
    ws = _lib.workspace(lib.ygg_batch_geometry_workspace_bytes(lay.ptr()), dev)
    _lib.check(lib.ygg_batch_geometry(lay.ptr(), pos.data_ptr(),
                                      trees.rad["node_ranges"].data_ptr(), cap,
                                      center.data_ptr(), half.data_ptr(), radius.data_ptr(),
                                      maxext.data_ptr(), ws.data_ptr(), ws.numel(),
                                      _lib.stream_ptr(dev)), "batch geometry")
    geom = BatchedGeometry(lay, center, half, radius, maxext)
    for k in trees.fallback:  # malformed trees: the level-synchronous path
        p0, p1 = int(lay.pt_off[k]), int(lay.pt_off[k + 1])
        g = compute_tree_geometry(trees.system(k), pos[p0:p1], max_leaf_size=max_leaf_size)
        gk = geom.system(k)
        gk.center.copy_(g.center)
        gk.half_extent.copy_(g.half_extent)
        gk.radius.copy_(g.radius)
        gk.max_extent.copy_(g.max_extent)
    return geom


class _ForestContext(_WalkContext):
    """Walk state for the forest of a batch: node records with batch-global
    ids and one seed pair per system root."""

    def __init__(self, trees: BatchedTrees, geom: BatchedGeometry, theta: float, mac_type: str,
                 dehnen_radius_scale: float):
        from .interactions import NODE_RECORD_BYTES, MAX_GENERATIONS, WALK_TIMING
        lay = trees.layout
        self.lib = _lib.load()
        self.dev = geom.center.device
        self.stream = _lib.stream_ptr(self.dev)
        if mac_type not in MAC_TYPES:
            raise ValueError(f"Unknown mac_type: {mac_type}")
        self.mac = MAC_TYPES[mac_type]
        th = ctypes.c_float(float(theta)).value
        self.theta_sq = ctypes.c_float(ctypes.c_float(th * th).value).value
        self.theta_root = float(torch.sqrt(torch.tensor(self.theta_sq, dtype=torch.float32)))
        T = lay.num_nodes
        self.nodes = torch.empty(T * NODE_RECORD_BYTES, dtype=torch.uint8, device=self.dev)
        self.roots = torch.empty(lay.num_systems, dtype=torch.int32, device=self.dev)
        leaf_position = trees.octa["leaf_position"] if trees.octree else trees.rad["leaf_position"]
        _lib.check(self.lib.ygg_batch_walk_prepare_nodes(
            lay.ptr(), geom.center.data_ptr(), geom.radius.data_ptr(), geom.max_extent.data_ptr(),
            trees.rad["left_child"].data_ptr(), trees.rad["right_child"].data_ptr(),
            leaf_position.data_ptr(), trees.rad["nodes_by_level"].data_ptr(), self.mac,
            ctypes.c_float(float(dehnen_radius_scale)).value, self.nodes.data_ptr(),
            self.roots.data_ptr(), self.stream), "batch walk nodes")
        self.host_status = torch.zeros(MAX_GENERATIONS + 16, dtype=torch.int32, pin_memory=True)
        self.gen_sizes = torch.empty(MAX_GENERATIONS + 1, dtype=torch.int32, device=self.dev)
        self.flags = torch.empty(4, dtype=torch.int32, device=self.dev)
        self.decisions = torch.empty(3, dtype=torch.int64, device=self.dev)
        self.timing = WALK_TIMING["enabled"]
        max_levels = int(trees.scalars[:, 0].max())
        self.gen_hint = int(1.5 * max(1, max_levels)) + 4
        self.grid_blocks = FLAT_GRID_BLOCKS
        from .interactions import WALK_PAIRS_PER_THREAD
        self.pairs_per_thread = WALK_PAIRS_PER_THREAD
        self.kernel_log = WALK_TIMING["log"]
        self.dt = SimpleNamespace(num_nodes=T, num_internal=max(1, lay.num_internal),
                                  num_leaves=lay.num_leaves, num_levels=max_levels)
        self.profile_key = (T, lay.num_internal, lay.num_leaves, lay.num_systems, self.mac,
                            self.theta_sq)


@dataclass
class BatchedLists:
    """Far and near lists of every system; `system(k)` gives the reference's
    NodeInteractionList / NodeNeighborList of system k."""

    layout: BatchLayout
    trees: BatchedTrees
    sources: torch.Tensor          # concatenated, local node ids
    targets: torch.Tensor
    target_levels: torch.Tensor
    far_counts: torch.Tensor       # [sum T]
    far_offsets: torch.Tensor      # [sum (T_s + 1)]
    far_level_offsets: torch.Tensor  # [S, 65]
    neighbors: torch.Tensor        # concatenated, local leaf node ids
    near_counts: torch.Tensor      # [sum L]
    near_offsets: torch.Tensor     # [sum (L_s + 1)]
    stats: dict
    far_base: np.ndarray = None    # [S] first far entry of each system
    near_base: np.ndarray = None   # [S] first near entry of each system
    _base_idx: tuple = None

    def set_bases(self) -> None:
        """Host copies of every system's first far / near entry (one read)."""
        base = torch.stack(list(self._base_idx)).cpu().numpy()
        self.far_base, self.near_base = base[0], base[1]

    def system(self, k: int):
        lay = self.layout
        k = int(k)
        n0, n1 = int(lay.node_off[k]), int(lay.node_off[k + 1])
        l0, l1 = int(lay.leaf_off[k]), int(lay.leaf_off[k + 1])
        fo = self.far_offsets[n0 + k:n1 + k + 1]
        no = self.near_offsets[l0 + k:l1 + k + 1]
        fbase = int(self.far_base[k])
        nbase = int(self.near_base[k])
        fend, nend = fbase + int(fo[-1]), nbase + int(no[-1])
        nl = int(self.trees.scalars[k, 0])
        inter = NodeInteractionList(fo, self.sources[fbase:fend], self.targets[fbase:fend],
                                    self.far_counts[n0:n1],
                                    self.far_level_offsets[k, :nl + 1],
                                    self.target_levels[fbase:fend])
        tree = self.trees.system(k)
        dt = tree._device
        idx = INDEX_DTYPE
        dev = self.neighbors.device
        L = l1 - l0
        leaf_indices, po, p2n = _leaf_tables(dt)
        nbrs = NodeNeighborList(
            offsets=no, neighbors=self.neighbors[nbase:nend], leaf_indices=leaf_indices,
            counts=self.near_counts[l0:l1], particle_order_leaf_indices=po,
            particle_order_to_native_leaf=p2n,
            neighbor_leaf_positions=torch.empty((L, 0), dtype=idx, device=dev),
            target_block_leaf_ids=torch.zeros(0, dtype=idx, device=dev),
            target_block_source_leaf_ids=torch.zeros((0, 0), dtype=idx, device=dev),
            target_block_valid_mask=torch.zeros((0, 0), dtype=torch.bool, device=dev),
            target_block_offsets=torch.zeros(L + 1, dtype=idx, device=dev), target_block_size=0)
        return inter, nbrs


def build_interactions_and_neighbors_batched(trees: BatchedTrees, geom: BatchedGeometry,
                                             theta: float = 0.5, *, mac_type: str = "bh",
                                             dehnen_radius_scale: float = 1.0,
                                             capacity_hint: Optional[dict] = None
                                             ) -> BatchedLists:
    """build_interactions_and_neighbors (interactions.py:338-448) for every
    system of a batch, by one dual-tree walk over the forest. With
    `capacity_hint` ({"far", "near", "queue"} for the whole batch) the walk
    starts at those sizes: far and near streams grow if they turn out too
    small, and a pair queue that overflows reruns the walk without it."""
    lib = _lib.load()
    lay = trees.layout
    if dehnen_radius_scale <= 0:
        raise ValueError("dehnen_radius_scale must be > 0")
    ctx = _ForestContext(trees, geom, theta, mac_type, dehnen_radius_scale)
    dt = ctx.dt
    lad = _ladder(dt, mac_type, max_interactions_per_node=None, max_neighbors_per_leaf=2048,
                  max_pair_queue=None, traversal_config=None)
    li = _LayoutIndex(lay, geom.center.device)
    stats = {}

    def lists(em):
        out, fk, far_off, near_off = _lists_from_streams(
            trees, li, em["far_keys"], em["far_vals"], em["near_keys"], em["near_vals"],
            em["far_total"], em["near_total"], stats,
            pair_records=bool(em.get("pair_records")))
        return out, None, fk, far_off, near_off
    fr = _run_walk_flat(ctx, dt, lad, sizes=capacity_hint, strict_int32=False, need_max=False,
                        fused_lists=lists)
    out = fr.inter
    stats.update(fr.stats)
    out.set_bases()
    return out


class _LayoutIndex:
    """Device index tensors that depend on the layout only (built before any
    stream capture: they need host-to-device copies)."""

    def __init__(self, lay: BatchLayout, dev):
        S = lay.num_systems
        self.node_off = torch.from_numpy(lay.node_off[:-1].astype(np.int64)).to(dev)
        self.leaf_off = torch.from_numpy(lay.leaf_off[:-1].astype(np.int64)).to(dev)
        self.sys_of_node = torch.repeat_interleave(
            torch.arange(S, device=dev), torch.from_numpy(np.diff(lay.node_off)).to(dev))
        # each level-major rank's and each leaf row's system node offset: what
        # turns batch-global node ids into the system's own
        self.node_sub = self.node_off[self.sys_of_node].contiguous()
        self.leaf_sub = torch.repeat_interleave(
            self.node_off, torch.from_numpy(np.diff(lay.leaf_off)).to(dev)).contiguous()
        # node_sub over the whole key range of node ids (the near lists of
        # pair records are keyed by node id; ids past the last node are no
        # key and take 0)
        keys = 1 << max(1, (lay.num_nodes - 1).bit_length())
        self.node_sub_keys = torch.cat([self.node_sub, torch.zeros(
            keys - self.node_sub.numel(), dtype=self.node_sub.dtype, device=dev)])


def _lists_from_streams(trees: BatchedTrees, li: _LayoutIndex, far_keys, far_vals, near_keys,
                        near_vals, nf: int, nn: int, stats: dict, far_count=None,
                        near_count=None, concurrent: bool = True, pair_records: bool = False):
    """Per-system far/near lists from the forest walk's emission streams.
    ygg_far_lists_sort (on a side stream) and ygg_near_lists_sort (on the
    current one) group the entries by row into the lists; the bucketing
    writes every system's local ids itself (each row's system node offset
    taken off), the stable sort writes batch-global ids that the localize
    kernels rewrite, and the localize kernels write every system's CSR
    offsets. `far_count` / `near_count` (int64
    device scalars) hold the valid entry counts of streams allocated at
    capacities nf / nn; without them all nf / nn entries are valid. With
    `concurrent` False everything runs on the current stream (stream
    capture). With `pair_records` the streams hold the walk's pair records
    (ceil(nf / 2) and ceil(nn / 2) slots, the counts counting records),
    expanded by the bucketing through the batch-global row key -> node
    tables.
    Returns (BatchedLists, the far lists' key scratch (contents
    unspecified), far row offsets, near row offsets)."""
    from .interactions import _side_stream, row_lists_digit_bits
    lib = _lib.load()
    lay = trees.layout
    dev = far_keys.device
    idx = INDEX_DTYPE
    i64 = int(idx == torch.int64)
    T, Lt, S = lay.num_nodes, lay.num_leaves, lay.num_systems
    nbl_global = (trees.rad["nodes_by_level"].long() + li.node_sub).to(torch.int32)
    rf, rn = nf, nn  # stream slots
    leaf_nodes = None
    if pair_records:
        # record slots, room for both entries of every record, and each
        # batch leaf row's global node (the near rows' order); records are
        # keyed by global node id, so the near values' system offsets are
        # taken by node
        rf, rn = -(-nf // 2), -(-nn // 2)
        nf, nn = 2 * rf, 2 * rn
        leaf_indices = (trees.octa if trees.octree else trees.rad)["leaf_indices"]
        leaf_nodes = (leaf_indices.long() + li.leaf_sub).to(torch.int32)
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev) if concurrent else main
    if concurrent:
        side.wait_stream(main)
    with torch.cuda.stream(side):
        fend = max(1, (T - 1).bit_length())
        fbits = row_lists_digit_bits(fend, nf)
        sources = torch.empty(nf, dtype=idx, device=dev)
        targets = torch.empty(nf, dtype=idx, device=dev)
        levels = torch.empty(nf, dtype=idx, device=dev)
        counts = torch.empty(T, dtype=idx, device=dev)
        counts32 = torch.empty(T, dtype=torch.int32, device=dev)
        node_off_glob = torch.empty(T + 1, dtype=idx, device=dev)
        far_off = torch.empty(T + 1, dtype=torch.int64, device=dev)
        fsorted = torch.empty(max(1, nf), dtype=torch.int32, device=dev)
        fws = _lib.workspace(lib.ygg_far_lists_sort_workspace_bytes(nf, fend, fbits), dev)
        # the bucketing writes system-local sources and targets itself
        far_local = fbits == 0
        _lib.check(lib.ygg_far_lists_sort(
            far_keys.data_ptr(), far_vals.data_ptr(), rf, _lib.ptr(far_count), T, fend, fbits,
            nbl_global.data_ptr(), trees.rad["node_level"].data_ptr(),
            li.node_sub.data_ptr() if far_local else None, int(pair_records), i64,
            fsorted.data_ptr(),
            far_off.data_ptr(), sources.data_ptr(), targets.data_ptr(), levels.data_ptr(),
            counts.data_ptr(), node_off_glob.data_ptr(), counts32.data_ptr(), fws.data_ptr(),
            fws.numel(), _lib.stream_ptr(dev)), "batch far lists sort")
        far_offsets = torch.empty(T + S, dtype=idx, device=dev)
        far_lvl = torch.empty((S, 65), dtype=idx, device=dev)
        _lib.check(lib.ygg_batch_far_localize(
            lay.ptr(), far_off.data_ptr(), nbl_global.data_ptr(),
            trees.rad["level_offsets"].data_ptr(), 0 if far_local else nf, i64,
            sources.data_ptr(), targets.data_ptr(), far_offsets.data_ptr(), far_lvl.data_ptr(),
            _lib.stream_ptr(dev)), "batch far localize")
    nend = max(1, ((T if pair_records else Lt) - 1).bit_length())
    nbits = row_lists_digit_bits(nend, nn)
    neighbors = torch.empty(nn, dtype=idx, device=dev)
    ncounts = torch.empty(Lt, dtype=idx, device=dev)
    noff_glob = torch.empty(Lt + 1, dtype=idx, device=dev)
    near_off = torch.empty(Lt + 1, dtype=torch.int64, device=dev)
    nsorted = torch.empty(max(1, nn), dtype=torch.int32, device=dev)
    nws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(nn, nend, nbits), dev)
    st = _lib.stream_ptr(dev)
    near_local = nbits == 0  # system-local neighbours from the bucketing itself
    _lib.check(lib.ygg_near_lists_sort(
        near_keys.data_ptr(), near_vals.data_ptr(), rn, _lib.ptr(near_count), Lt, nend, nbits,
        (li.node_sub_keys if pair_records else li.leaf_sub).data_ptr() if near_local else None,
        _lib.ptr(leaf_nodes), i64,
        nsorted.data_ptr(),
        near_off.data_ptr(), neighbors.data_ptr(), ncounts.data_ptr(), noff_glob.data_ptr(),
        nws.data_ptr(), nws.numel(), st), "batch near lists sort")
    near_offsets = torch.empty(Lt + S, dtype=idx, device=dev)
    _lib.check(lib.ygg_batch_near_localize(lay.ptr(), near_off.data_ptr(),
                                           0 if near_local else nn, i64, neighbors.data_ptr(),
                                           near_offsets.data_ptr(), st), "batch near localize")
    if concurrent:
        main.wait_stream(side)
        for t in (sources, targets, levels, counts, counts32, node_off_glob, far_off, fsorted,
                  fws, far_offsets, far_lvl, nbl_global):
            t.record_stream(main)
    out = BatchedLists(lay, trees, sources, targets, levels, counts, far_offsets, far_lvl,
                       neighbors, ncounts, near_offsets, stats)
    out._base_idx = (far_off[li.node_off], near_off[li.leaf_off])
    return out, fsorted, far_off, near_off


def _graph_walk(trees: BatchedTrees, geom: BatchedGeometry, theta: float, mac_type: str,
                dehnen_radius_scale: float, caps: dict, generations: int, grid_blocks: int,
                pairs_per_thread: int, pair_records: bool = False):
    """Forest walk with fixed stream capacities and generation count and no
    host synchronisation (stream-capturable). Returns the raw emission
    streams and the walk's device-side totals, which bound the row sorts.
    With `pair_records` the streams hold one record per pair (capacities
    caps / 2, rounded up; the totals count records)."""
    from .interactions import NODE_RECORD_BYTES
    lib = _lib.load()
    lay = trees.layout
    dev = trees.rad["parent"].device
    st = _lib.stream_ptr(dev)
    T, Lt, S = lay.num_nodes, lay.num_leaves, lay.num_systems
    mac = MAC_TYPES[mac_type]
    th = ctypes.c_float(float(theta)).value
    theta_sq = ctypes.c_float(ctypes.c_float(th * th).value).value
    theta_root = float(np.sqrt(np.float32(theta_sq)))
    nodes = torch.empty(T * NODE_RECORD_BYTES, dtype=torch.uint8, device=dev)
    roots = torch.empty(S, dtype=torch.int32, device=dev)
    leaf_position = trees.octa["leaf_position"] if trees.octree else trees.rad["leaf_position"]
    _lib.check(lib.ygg_batch_walk_prepare_nodes(
        lay.ptr(), geom.center.data_ptr(), geom.radius.data_ptr(), geom.max_extent.data_ptr(),
        trees.rad["left_child"].data_ptr(), trees.rad["right_child"].data_ptr(),
        leaf_position.data_ptr(), trees.rad["nodes_by_level"].data_ptr(), mac,
        ctypes.c_float(float(dehnen_radius_scale)).value, nodes.data_ptr(), roots.data_ptr(), st),
        "graph walk nodes")
    Q, FC, NC, G = int(caps["queue"]), int(caps["far"]), int(caps["near"]), int(generations)
    if pair_records:
        FC, NC = -(-FC // 2), -(-NC // 2)
    wf = torch.empty((2, Q), dtype=torch.int64, device=dev)
    counters = torch.empty(2, dtype=torch.int64, device=dev)
    gen_sizes = torch.empty(G + 1, dtype=torch.int32, device=dev)
    flags = torch.empty(4, dtype=torch.int32, device=dev)
    fk = torch.empty(FC, dtype=torch.int32, device=dev)
    fv = torch.empty(FC, dtype=torch.int32, device=dev)
    nk = torch.empty(NC, dtype=torch.int32, device=dev)
    nv = torch.empty(NC, dtype=torch.int32, device=dev)
    p = _lib.FlatWalkParams()
    p.num_nodes, p.num_internal, p.num_leaves, p.root = T, max(1, lay.num_internal), Lt, 0
    p.mac_type, p.emit, p.max_generations, p.grid_blocks = mac, 1, G, int(grid_blocks)
    p.theta_sq, p.theta_root = theta_sq, theta_root
    p.queue_cap, p.far_cap, p.near_cap = Q, FC, NC
    p.nodes, p.wf0, p.wf1 = nodes.data_ptr(), wf[0].data_ptr(), wf[1].data_ptr()
    p.gen_sizes, p.flags, p.counters = gen_sizes.data_ptr(), flags.data_ptr(), counters.data_ptr()
    p.far_keys, p.far_vals = fk.data_ptr(), fv.data_ptr()
    p.near_keys, p.near_vals = nk.data_ptr(), nv.data_ptr()
    p.hist_gen, p.node_hist = -1, None
    p.roots, p.num_roots = roots.data_ptr(), S
    from .interactions import WALK_XCD_GROUPS
    p.xcd_groups = WALK_XCD_GROUPS
    p.pairs_per_thread = int(pairs_per_thread)
    p.pair_records = int(pair_records)
    _lib.check(lib.ygg_walk_flat_fixed(ctypes.byref(p), st), "graph walk")
    return dict(far_keys=fk, far_vals=fv, near_keys=nk, near_vals=nv, counters=counters,
                gen_sizes=gen_sizes, flags=flags, nodes=nodes, wf=wf)


class BatchedGraph:
    """The whole batched pipeline -- build_octree, compute_tree_geometry and
    the forest walk with its lists -- captured once as a HIP graph for a
    fixed batch layout. `replay(positions, masses)` reruns it on new inputs
    of the same layout (e.g. every step of a simulation) with one launch and
    one small status read.

    Stream capacities and the generation count are fixed at capture. With a
    capacity model the far, near and queue capacities are the model's
    predicted totals for the capture inputs; without one they are the exact
    totals of a first eager run times `slack`. `capacities` ({"far", "near",
    "queue"}) and `generations` override both. The generation count is the
    eager run's times `slack`, plus 4. A replay whose walk overflows, does
    not finish within the captured generations, or meets a system that needs
    the single-system path returns the eager pipeline's result instead
    (`last_replay_fell_back`), so results never depend on the capacities.

    With `tune` the autotuner times its candidates during the eager run (the
    forest walk's grid and queue tile, the row sorts' digit widths, also at
    the graph's capacity-sized stream lengths) and the captured graph uses
    the chosen plans (`plans`).

    A replay that does not fall back returns the graph's own output buffers:
    the next replay overwrites them in place, so a caller that keeps a
    result across replays clones its tensors first."""

    def __init__(self, positions, masses, offsets, *, leaf_size: int = 32, theta: float = 0.6,
                 mac_type: str = "dehnen", dehnen_radius_scale: float = 1.0, model=None,
                 slack: float = 1.1, grid_blocks: int = FLAT_GRID_BLOCKS,
                 capacities: Optional[dict] = None, generations: Optional[int] = None,
                 tune: bool = False):
        from . import autotune
        from .interactions import WALK_PAIRS_PER_THREAD, pair_records_ok, row_lists_digit_bits
        _lib.require_device(positions)
        dev = positions.device
        self.theta, self.mac_type, self.scale = float(theta), mac_type, float(dehnen_radius_scale)
        self.leaf_size, self.grid_blocks = int(leaf_size), int(grid_blocks)
        self.offsets = np.asarray(offsets, dtype=np.int64)
        self.pos = positions.to(torch.float32).contiguous().clone()
        self.mass = None if masses is None else torch.as_tensor(
            masses, dtype=torch.float32, device=dev).contiguous().clone()
        self.lay = make_layout(self.offsets, leaf_size, dev)
        self.li = _LayoutIndex(self.lay, dev)
        # eager run (learned capacities when a model is given): exact stream
        # totals, generation count and a warm-up of every kernel. The graph's
        # streams are sized from those totals times `slack`; the row sorts
        # stop at the walk's device-side totals, so capacity costs memory,
        # not sort work.
        was = autotune.enabled()
        if tune:
            autotune.enable(True)
        try:
            lists = self._eager(model)
        finally:
            autotune.enable(was)
        gens = sum(1 for v in lists.stats["gen_sizes"] if v > 0)
        if capacities is not None:
            self.caps = {k: max(1, int(capacities[k])) for k in ("far", "near", "queue")}
        elif self.hint is not None:
            self.caps = {k: max(1, int(self.hint[k])) for k in ("far", "near", "queue")}
        else:
            caps = {"far": lists.stats["far_total"], "near": lists.stats["near_total"],
                    "queue": lists.stats["peak"]}
            self.caps = {k: int(v * slack) + 1024 for k, v in caps.items()}
        self.generations = (max(1, int(generations)) if generations is not None
                            else int(gens * slack) + 4)
        T, Lt = self.lay.num_nodes, self.lay.num_leaves
        self.pair_records = pair_records_ok(T)
        if tune:
            # the graph looks its row-sort plans up by the stream capacities
            autotune.enable(True)
            try:
                row_lists_digit_bits(max(1, (T - 1).bit_length()), self.caps["far"])
                row_lists_digit_bits(max(1, (Lt - 1).bit_length()), self.caps["near"])
            finally:
                autotune.enable(was)
        self.walk_grid, self.walk_tile = autotune.walk_plan(T, self.grid_blocks,
                                                            WALK_PAIRS_PER_THREAD)
        self.plans = autotune.plans()
        was = autotune.enabled()
        autotune.enable(False)
        try:
            torch.cuda.synchronize(dev)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.out = self._pipeline()
        finally:
            autotune.enable(was)
        self.last_replay_fell_back = False

    def _eager(self, model):
        bt = _build_batched(self.pos, self.mass, self.offsets, leaf_size=self.leaf_size,
                            octree=True, return_reordered=True, layout=self.lay)
        geom = compute_tree_geometry_batched(bt)
        self.hint = model.capacity_hint_batched(bt, geom) if model is not None else None
        from .interactions import FLAT_STATS
        miss0 = FLAT_STATS["hint_overflows"] + FLAT_STATS["hint_growths"]
        lists = build_interactions_and_neighbors_batched(
            bt, geom, self.theta, mac_type=self.mac_type, dehnen_radius_scale=self.scale,
            capacity_hint=self.hint)
        lists.stats["hinted"] = (self.hint is not None and FLAT_STATS["hint_overflows"]
                                 + FLAT_STATS["hint_growths"] == miss0)
        return lists

    def _pipeline(self) -> dict:
        bt = _build_batched(self.pos, self.mass, None, leaf_size=self.leaf_size, octree=True,
                            return_reordered=True, layout=self.lay, sync=False)
        geom = compute_tree_geometry_batched(bt)
        w = _graph_walk(bt, geom, self.theta, self.mac_type, self.scale, self.caps,
                        self.generations, self.walk_grid, self.walk_tile, self.pair_records)
        # the sorts and list kernels stop at the walk's device-side totals
        lists, _, far_off, near_off = _lists_from_streams(
            bt, self.li, w["far_keys"], w["far_vals"], w["near_keys"], w["near_vals"],
            self.caps["far"], self.caps["near"], {}, far_count=w["counters"][0:1],
            near_count=w["counters"][1:2], concurrent=False, pair_records=self.pair_records)
        lay = self.lay
        T, Lt = lay.num_nodes, lay.num_leaves
        status = torch.cat([
            w["flags"][:3].long(), w["gen_sizes"][self.generations:].long(), w["counters"],
            far_off[T:T + 1], near_off[Lt:Lt + 1],
            bt.rad["scalars"].long(), bt.octa["scalars"].long(),
            lists._base_idx[0], lists._base_idx[1]])
        return dict(trees=bt, geom=geom, lists=lists, walk=w, status=status)

    def replay(self, positions=None, masses=None):
        """(trees, geometry, lists) for the current (or given) inputs; the
        results alias the graph's buffers (see the class notes)."""
        if masses is not None and self.mass is None:
            raise ValueError("this graph was captured without masses; masses cannot be "
                             "replayed")
        if positions is not None:
            self.pos.copy_(positions)
        if masses is not None:
            self.mass.copy_(torch.as_tensor(masses, dtype=torch.float32))
        self.graph.replay()
        S = self.lay.num_systems
        st = self.out["status"].cpu().numpy()
        flags, unfinished = st[:3], st[3]
        far_total, near_total = int(st[4]), int(st[5])     # stream totals of the walk
        if self.pair_records:  # records: two entries each
            far_total, near_total = 2 * far_total, 2 * near_total
        rows_far, rows_near = int(st[6]), int(st[7])       # entries the row sorts listed
        scal = st[8:8 + 8 * S].reshape(S, 8)
        osc = st[8 + 8 * S:8 + 12 * S].reshape(S, 4)
        fbase, nbase = st[8 + 12 * S:8 + 13 * S], st[8 + 13 * S:8 + 14 * S]
        bad = bool(np.any((scal[:, 2] != 0) | (scal[:, 3] != 0)))
        if (flags.any() or unfinished != 0 or bad or far_total != rows_far
                or near_total != rows_near):
            self.last_replay_fell_back = True
            bt = _build_batched(self.pos, self.mass, self.offsets, leaf_size=self.leaf_size,
                                octree=True, return_reordered=True, layout=self.lay)
            geom = compute_tree_geometry_batched(bt)
            return bt, geom, build_interactions_and_neighbors_batched(
                bt, geom, self.theta, mac_type=self.mac_type, dehnen_radius_scale=self.scale)
        self.last_replay_fell_back = False
        bt, geom, lists = self.out["trees"], self.out["geom"], self.out["lists"]
        bt.scalars, bt.oct_scalars = scal.astype(np.int32), osc.astype(np.int32)
        bt._views = {}
        lists.far_base, lists.near_base = fbase, nbase
        lists.stats = dict(far_total=far_total, near_total=near_total)
        return bt, geom, lists


__all__ = ["BatchLayout", "BatchedGeometry", "BatchedGraph", "BatchedLists", "BatchedTrees",
           "build_interactions_and_neighbors_batched", "build_octree_batched",
           "build_tree_batched", "compute_tree_geometry_batched", "make_layout"]
