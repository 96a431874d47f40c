"""Tree construction: the drop-in surface of yggdrax/tree.py and morton.py.

``build_tree`` (radix/LBVH, yggdrax/tree.py:1413-1478) and ``build_octree``
(radix topology plus explicit octree metadata, yggdrax/tree.py:1482-1549)
run on the GPU: bounds, Morton encode fused with the sort's digit
histograms, a stable radix sort, the Karras build with node levels and the
octree metadata kernels. Topology fields carry the reference's names and
dtypes (indices widened to INDEX_DTYPE, Morton codes as uint64).
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Callable, NamedTuple, Optional

import torch

from . import _lib
from .dtypes import INDEX_DTYPE

MAX_TREE_LEVELS = 64
# Particle reorder by scattering (coalesced reads) instead of gathering.
REORDER_SCATTER = os.environ.get("YGGDRAX_REORDER_SCATTER", "0") not in ("", "0")


class RadixTreeTopology(NamedTuple):
    """Fields of the reference's RadixTree (yggdrax/_tree_impl.py:24-71)."""

    parent: torch.Tensor
    left_child: torch.Tensor
    right_child: torch.Tensor
    left_is_leaf: torch.Tensor
    right_is_leaf: torch.Tensor
    particle_indices: torch.Tensor
    morton_codes: torch.Tensor
    node_ranges: torch.Tensor
    num_particles: int
    num_internal_nodes: int
    node_level: torch.Tensor
    level_offsets: torch.Tensor
    nodes_by_level: torch.Tensor
    num_levels: int
    bounds_min: torch.Tensor
    bounds_max: torch.Tensor
    leaf_codes: torch.Tensor
    leaf_depths: torch.Tensor
    use_morton_geometry: torch.Tensor
    leaf_size: Optional[int]


class OctreeTopology(NamedTuple):
    """Radix fields plus explicit octree metadata (yggdrax/octree.py:73-110)."""

    parent: torch.Tensor
    left_child: torch.Tensor
    right_child: torch.Tensor
    left_is_leaf: torch.Tensor
    right_is_leaf: torch.Tensor
    particle_indices: torch.Tensor
    morton_codes: torch.Tensor
    node_ranges: torch.Tensor
    num_particles: int
    num_internal_nodes: int
    node_level: torch.Tensor
    level_offsets: torch.Tensor
    nodes_by_level: torch.Tensor
    num_levels: int
    bounds_min: torch.Tensor
    bounds_max: torch.Tensor
    leaf_codes: torch.Tensor
    leaf_depths: torch.Tensor
    use_morton_geometry: torch.Tensor
    leaf_size: Optional[int]
    oct_parent: torch.Tensor
    oct_children: torch.Tensor
    oct_child_counts: torch.Tensor
    oct_child_mask: torch.Tensor
    oct_valid_mask: torch.Tensor
    oct_node_codes: torch.Tensor
    oct_node_depths: torch.Tensor
    oct_node_ranges: torch.Tensor
    oct_nodes_by_level: torch.Tensor
    oct_level_offsets: torch.Tensor
    oct_num_levels: int
    oct_leaf_mask: torch.Tensor
    oct_leaf_nodes: torch.Tensor
    radix_node_to_oct: torch.Tensor
    radix_leaf_to_oct: torch.Tensor


@dataclass
class DeviceTopology:
    """int32 device arrays the kernels consume (kept beside the public fields)."""

    num_particles: int
    num_nodes: int
    num_internal: int
    num_leaves: int
    leaf_size: int
    parent: torch.Tensor
    left: torch.Tensor
    right: torch.Tensor
    node_ranges: torch.Tensor
    node_level: torch.Tensor
    nodes_by_level: torch.Tensor
    level_offsets: torch.Tensor
    leaf_indices: torch.Tensor
    leaf_position: torch.Tensor
    num_levels: int
    max_level: int
    malformed: bool
    is_octree: bool


@dataclass(frozen=True)
class TreeBuildConfig:
    """yggdrax/tree.py:28-43."""

    leaf_size: int = 8
    return_reordered: bool = False
    workspace: object = None
    return_workspace: bool = False


@dataclass(frozen=True)
class RadixTreeWorkspace:
    """Accepted for signature compatibility; device buffers are pooled by torch."""

    parent: object = None


@dataclass(frozen=True)
class Tree:
    """Public base container (yggdrax/tree.py:184-330)."""

    @property
    def num_nodes(self) -> int:
        return int(self.topology.parent.shape[0])

    @property
    def num_particles(self) -> int:
        return int(self.topology.num_particles)

    @property
    def num_leaves(self) -> int:
        return int(self.topology.parent.shape[0]) - int(self.topology.num_internal_nodes)

    def __getattr__(self, name):
        if name in ("topology", "_device"):
            raise AttributeError(name)
        return getattr(self.topology, name)


@dataclass(frozen=True)
class RadixTree(Tree):
    topology: RadixTreeTopology
    build_mode: str = "adaptive"
    positions_sorted: Optional[torch.Tensor] = None
    masses_sorted: Optional[torch.Tensor] = None
    inverse_permutation: Optional[torch.Tensor] = None
    workspace: object = None
    _device: Optional[DeviceTopology] = None

    @property
    def tree_type(self) -> str:
        return "radix"

    @classmethod
    def from_particles(cls, positions, masses, *, build_mode="adaptive", bounds=None,
                       return_reordered=True, leaf_size=8, **_ignored):
        if build_mode != "adaptive":
            raise ValueError(f"Unsupported build_mode '{build_mode}'. Supported: ('adaptive',)")
        out = build_tree(positions, masses, bounds, return_reordered=return_reordered,
                         leaf_size=leaf_size)
        return out[0] if isinstance(out, tuple) else out


@dataclass(frozen=True)
class OctreeTree(RadixTree):
    @property
    def tree_type(self) -> str:
        return "octree"

    @property
    def oct_num_nodes(self) -> int:
        return int(self.topology.oct_valid_mask.sum())

    @property
    def oct_num_leaf_nodes(self) -> int:
        return int(self.topology.oct_leaf_mask.sum())

    @classmethod
    def from_particles(cls, positions, masses, *, build_mode="adaptive", bounds=None,
                       return_reordered=True, leaf_size=8, **_ignored):
        if build_mode != "adaptive":
            raise ValueError(f"Unsupported build_mode '{build_mode}'. Supported: ('adaptive',)")
        out = build_octree(positions, masses, bounds, return_reordered=return_reordered,
                           leaf_size=leaf_size)
        return out[0] if isinstance(out, tuple) else out


@dataclass(frozen=True)
class TreeBuildRequest:
    """yggdrax/tree.py:110-130 (adaptive subset)."""

    positions: object
    masses: object
    build_mode: str
    bounds: object
    return_reordered: bool
    leaf_size: int


TreeBuilder = Callable[[TreeBuildRequest], Tree]


def _as_positions(positions) -> torch.Tensor:
    t = torch.as_tensor(positions)
    if not t.is_cuda:
        if not torch.cuda.is_available():
            raise RuntimeError("yggdrax_amd runs on the GPU only: no HIP device is available")
        t = t.to("cuda")
    if t.dtype != torch.float32:
        raise TypeError("positions must be float32")
    if t.ndim != 2 or t.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    return t.contiguous()


def infer_bounds(positions) -> tuple[torch.Tensor, torch.Tensor]:
    """yggdrax/bounds.py:11-33 on the GPU."""
    lib = _lib.load()
    pos = _as_positions(positions)
    b6 = torch.empty(6, dtype=torch.float32, device=pos.device)
    ws = _lib.workspace(lib.ygg_bounds_workspace_bytes(), pos.device)
    _lib.check(lib.ygg_infer_bounds(pos.data_ptr(), pos.shape[0], b6.data_ptr(), ws.data_ptr(),
                                    ws.numel(), _lib.stream_ptr(pos.device)), "infer_bounds")
    return b6[:3], b6[3:]


def _bounds6(pos: torch.Tensor, bounds) -> torch.Tensor:
    if bounds is None:
        lo, hi = infer_bounds(pos)
        return torch.cat([lo, hi])
    lo = torch.as_tensor(bounds[0], dtype=torch.float32, device=pos.device).reshape(3)
    hi = torch.as_tensor(bounds[1], dtype=torch.float32, device=pos.device).reshape(3)
    return torch.cat([lo, hi]).contiguous()


def morton_encode(positions, bounds) -> torch.Tensor:
    """yggdrax/morton.py:43-80: uint64 codes (21 bits per axis)."""
    lib = _lib.load()
    pos = _as_positions(positions)
    b6 = _bounds6(pos, bounds)
    codes = torch.empty(pos.shape[0], dtype=torch.int64, device=pos.device)
    _lib.check(lib.ygg_morton_encode(pos.data_ptr(), pos.shape[0], b6.data_ptr(),
                                     codes.data_ptr(), _lib.stream_ptr(pos.device)),
               "morton_encode")
    return codes.view(torch.uint64)


class _Scratch:
    """Pinned host words for the few scalar reads a build needs."""

    def __init__(self):
        self.buf = torch.zeros(16, dtype=torch.int32, pin_memory=True)


def _alloc_radix(L: int, device, scalars: Optional[torch.Tensor] = None):
    T = 2 * L - 1
    NI = L - 1
    i32 = dict(dtype=torch.int32, device=device)
    return dict(
        parent=torch.empty(T, **i32),
        left_child=torch.empty(max(NI, 0), **i32),
        right_child=torch.empty(max(NI, 0), **i32),
        left_is_leaf=torch.empty(max(NI, 0), dtype=torch.uint8, device=device),
        right_is_leaf=torch.empty(max(NI, 0), dtype=torch.uint8, device=device),
        node_ranges=torch.empty((T, 2), **i32),
        node_level=torch.empty(T, **i32),
        level_offsets=torch.empty(65, **i32),
        nodes_by_level=torch.empty(T, **i32),
        leaf_codes=torch.empty(L, dtype=torch.int64, device=device),
        scalars=torch.zeros(8, **i32) if scalars is None else scalars,
    )


def _ptrs(struct_cls, arrays: dict):
    s = struct_cls()
    for name, _ in struct_cls._fields_:
        t = arrays.get(name)
        setattr(s, name, None if t is None or t.numel() == 0 else t.data_ptr())
    return s


def _alloc_octree(L: int, device, scalars: Optional[torch.Tensor] = None):
    M = max(1, 2 * L - 1)
    T = 2 * L - 1
    i32 = dict(dtype=torch.int32, device=device)
    u8 = dict(dtype=torch.uint8, device=device)
    return dict(
        oct_parent=torch.empty(M, **i32), oct_children=torch.empty((M, 8), **i32),
        oct_child_counts=torch.empty(M, **i32), oct_child_mask=torch.empty((M, 8), **u8),
        oct_valid_mask=torch.empty(M, **u8),
        oct_node_codes=torch.empty(M, dtype=torch.int64, device=device),
        oct_node_depths=torch.empty(M, **i32), oct_node_ranges=torch.empty((M, 2), **i32),
        oct_nodes_by_level=torch.empty(M, **i32), oct_level_offsets=torch.empty(23, **i32),
        oct_leaf_mask=torch.empty(M, **u8), oct_leaf_nodes=torch.empty(M, **i32),
        radix_node_to_oct=torch.empty(T, **i32), leaf_indices=torch.empty(L, **i32),
        leaf_position=torch.empty(T, **i32),
        scalars=torch.zeros(4, **i32) if scalars is None else scalars,
    )


def _build(positions, masses, bounds, *, leaf_size: int, return_reordered: bool,
           octree: bool):
    lib = _lib.load()
    pos = _as_positions(positions)
    dev = pos.device
    n = int(pos.shape[0])
    if n < 1:
        raise ValueError("Need at least one particle")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    if n >= (1 << 30):
        raise ValueError("at most 2^30 - 1 particles per tree")
    st = _lib.stream_ptr(dev)
    b6 = _bounds6(pos, bounds)
    sorted_codes = torch.empty(n, dtype=torch.int64, device=dev)
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    pos_sorted = mass_sorted = inv = m = None
    if masses is not None:
        m = torch.as_tensor(masses, dtype=torch.float32, device=dev).contiguous()
    reorder_side = None
    if return_reordered and not REORDER_SCATTER:
        # sort on this stream; the gather of packed (x, y, z, m) records and
        # the inverse run on a side stream, under the radix build below,
        # which needs only the sorted codes
        from .interactions import _side_stream
        pos_sorted = torch.empty_like(pos)
        mass_sorted = torch.empty_like(m) if m is not None else None
        inv = torch.empty(n, dtype=INDEX_DTYPE, device=dev)
        ws = _lib.workspace(lib.ygg_morton_sort_reorder_workspace_bytes(n), dev)
        reorder_side = _side_stream(dev)
        _lib.check(lib.ygg_morton_sort_reorder_async(
            pos.data_ptr(), _lib.ptr(m), n, b6.data_ptr(), sorted_codes.data_ptr(),
            perm.data_ptr(), pos_sorted.data_ptr(), _lib.ptr(mass_sorted), inv.data_ptr(),
            int(INDEX_DTYPE == torch.int64), ws.data_ptr(), ws.numel(), st,
            reorder_side.cuda_stream), "morton_sort")
        ws.record_stream(reorder_side)  # the side stream still reads the packed records
    else:
        ws = _lib.workspace(lib.ygg_morton_sort_workspace_bytes(n), dev)
        _lib.check(lib.ygg_morton_sort(pos.data_ptr(), n, b6.data_ptr(), sorted_codes.data_ptr(),
                                       perm.data_ptr(), ws.data_ptr(), ws.numel(), st),
                   "morton_sort")
    L = (n + leaf_size - 1) // leaf_size
    T = 2 * L - 1
    NI = L - 1
    # both builds' scalar words in one buffer: one host copy reads them
    scalars = torch.zeros(12, dtype=torch.int32, device=dev)
    rad = _alloc_radix(L, dev, scalars[:8])
    rptr = _ptrs(_lib.RadixTopologyPtrs, rad)
    rws = _lib.workspace(lib.ygg_radix_build_workspace_bytes(n, leaf_size), dev)
    _lib.check(lib.ygg_radix_build(sorted_codes.data_ptr(), n, leaf_size, ctypes.byref(rptr),
                                   rws.data_ptr(), rws.numel(), st), "radix_build")
    if octree:
        octa = _alloc_octree(L, dev, scalars[8:])
        optr = _ptrs(_lib.OctreeMetaPtrs, octa)
        ows = _lib.workspace(lib.ygg_octree_meta_workspace_bytes(L), dev)
        _lib.check(lib.ygg_octree_meta(sorted_codes.data_ptr(), rad["node_ranges"].data_ptr(), L,
                                       ctypes.byref(optr), ows.data_ptr(), ows.numel(), st),
                   "octree_meta")
        leaf_indices, leaf_position = octa["leaf_indices"], octa["leaf_position"]
    else:
        leaf_indices = torch.empty(L, dtype=torch.int32, device=dev)
        leaf_position = torch.empty(T, dtype=torch.int32, device=dev)
        _lib.check(lib.ygg_radix_leaf_order(L, leaf_indices.data_ptr(), leaf_position.data_ptr(),
                                            st), "leaf_order")
    if return_reordered and REORDER_SCATTER:
        pos_sorted = torch.empty_like(pos)
        inv = torch.empty(n, dtype=torch.int32, device=dev)
        if m is not None:
            mass_sorted = torch.empty_like(m)
        reorder = lib.ygg_reorder_scatter
        _lib.check(reorder(pos.data_ptr(), _lib.ptr(m), perm.data_ptr(), n,
                           pos_sorted.data_ptr(), _lib.ptr(mass_sorted), inv.data_ptr(), st),
                   "reorder")
    # One host read of the build's scalar words (level count, depth flags and
    # the octree level count), queued before the index widening so that the
    # host waits only for the build; the widening runs meanwhile.
    main = torch.cuda.current_stream(dev)
    host_scalars = _read_scalars_async(rad["scalars"], octa["scalars"] if octree else None)
    if reorder_side is not None:  # reordered particles and inverse are ready after this
        main.wait_stream(reorder_side)
    wide = _widen_fields(rad, octa if octree else None, perm, inv)
    scal, oscal = host_scalars()
    if scal[2] != 0:  # levels beyond the fast walk: exact pointer doubling
        _lib.check(lib.ygg_levels_pointer_doubling(ctypes.byref(rptr), T, rws.data_ptr(),
                                                   rws.numel(), st), "levels")
        wide.update(_widen_index_arrays(
            {k: rad[k] for k in ("node_level", "level_offsets", "nodes_by_level")}))
        scal, _ = _read_scalars(rad["scalars"], None)
    num_levels = scal[0]
    max_level = scal[1]
    malformed = bool(scal[3]) or bool(scal[2])
    if malformed and max_level >= 65536:
        raise RuntimeError("tree depth exceeds 65535 levels (cyclic parent links from "
                           "duplicate Morton codes)")
    oct_num_levels = oscal[0] if octree else 0
    return _assemble_tree(rad, octa if octree else None, perm, sorted_codes, b6, n, leaf_size,
                          num_levels, max_level, malformed, oct_num_levels, leaf_indices,
                          leaf_position, pos_sorted, mass_sorted, inv, return_reordered,
                          wide=wide)


def _read_scalars(radix_scalars: torch.Tensor, octree_scalars: Optional[torch.Tensor]):
    """Both builds' int32 scalar words in one device-to-host copy."""
    return _read_scalars_async(radix_scalars, octree_scalars)()


def _read_scalars_async(radix_scalars: torch.Tensor, octree_scalars: Optional[torch.Tensor]):
    """Queues the copy of both builds' int32 scalar words to pinned host
    memory (one copy when they are adjacent in one buffer) and returns a
    function that waits for it and gives (radix words, octree words)."""
    nr = radix_scalars.numel()
    no = octree_scalars.numel() if octree_scalars is not None else 0
    host = torch.empty(nr + no, dtype=torch.int32, pin_memory=True)
    adjacent = (no and octree_scalars.untyped_storage().data_ptr()
                == radix_scalars.untyped_storage().data_ptr()
                and octree_scalars.data_ptr() == radix_scalars.data_ptr() + 4 * nr)
    if adjacent:
        host.copy_(torch.as_strided(radix_scalars, (nr + no,), (1,)), non_blocking=True)
    else:
        host[:nr].copy_(radix_scalars, non_blocking=True)
        if no:
            host[nr:].copy_(octree_scalars, non_blocking=True)
    done = torch.cuda.Event()
    done.record(torch.cuda.current_stream(radix_scalars.device))

    def wait():
        done.synchronize()
        words = host.tolist()
        return words[:nr], words[nr:]
    return wait


def _widen_index_arrays(arrays: dict) -> dict:
    """INDEX_DTYPE versions of int32 device arrays: contiguous int32 GPU
    tensors are widened to int64 together in one kernel launch
    (ygg_widen_i32_to_i64) into one allocation, each output a view of its
    own 256-byte aligned range; anything else goes through .to(). The
    launch goes out before the views are made, so the kernel runs while
    the host builds them."""
    idx = INDEX_DTYPE
    out, fast = {}, []
    for name, t in arrays.items():
        if (idx == torch.int64 and t.dtype == torch.int32 and t.is_cuda and t.is_contiguous()
                and len(fast) < _lib.WIDEN_MAX_ITEMS):
            fast.append((name, t, t.numel()))
        else:
            out[name] = t.to(idx)
    if not fast:
        return out
    spans = [(k + 31) // 32 * 32 for _, _, k in fast]
    dev = fast[0][1].device
    buf = torch.empty(sum(spans), dtype=torch.int64, device=dev)
    base = buf.data_ptr()
    arr = (_lib.WidenItem * len(fast))()
    at = 0
    for j, ((_, t, k), sp) in enumerate(zip(fast, spans)):
        arr[j].src = t.data_ptr() if k else None
        arr[j].dst = base + 8 * at if k else None
        arr[j].n = k
        at += sp
    _lib.check(_lib.load().ygg_widen_i32_to_i64(arr, len(fast), _lib.stream_ptr(dev)),
               "widen index arrays")
    at = 0
    for (name, t, k), sp in zip(fast, spans):
        out[name] = buf[at:at + k].view(t.shape)
        at += sp
    return out


def _widen_fields(rad, octa, perm, inv) -> dict:
    """INDEX_DTYPE copies of every index array of the public topology."""
    NI = int(rad["leaf_codes"].shape[0]) - 1
    wide = dict(parent=rad["parent"], left_child=rad["left_child"],
                right_child=rad["right_child"], particle_indices=perm,
                node_ranges=rad["node_ranges"], node_level=rad["node_level"],
                level_offsets=rad["level_offsets"], nodes_by_level=rad["nodes_by_level"])
    if octa is not None:
        wide.update({k: octa[k] for k in (
            "oct_parent", "oct_children", "oct_child_counts", "oct_node_depths",
            "oct_node_ranges", "oct_nodes_by_level", "oct_level_offsets", "oct_leaf_nodes",
            "radix_node_to_oct")})
        wide["radix_leaf_to_oct"] = octa["radix_node_to_oct"][NI:]
    if inv is not None:
        wide["inverse_permutation"] = inv
    return _widen_index_arrays(wide)


def _assemble_tree(rad, octa, perm, sorted_codes, b6, n, leaf_size, num_levels, max_level,
                   malformed, oct_num_levels, leaf_indices, leaf_position, pos_sorted,
                   mass_sorted, inv, return_reordered, wide: Optional[dict] = None):
    """Public topology (reference field names and dtypes) and the kernels'
    int32 view from the build's device arrays; also used for one system's
    slices of a batched build. `wide` holds the already widened index
    arrays (_widen_fields), computed here when absent."""
    octree = octa is not None
    dev = rad["parent"].device
    L = int(rad["leaf_codes"].shape[0])
    T = 2 * L - 1
    NI = L - 1
    idx = INDEX_DTYPE
    if wide is None:
        wide = _widen_fields(rad, octa, perm, inv)
    fields = dict(
        parent=wide["parent"],
        left_child=wide["left_child"],
        right_child=wide["right_child"],
        left_is_leaf=rad["left_is_leaf"].view(torch.bool),
        right_is_leaf=rad["right_is_leaf"].view(torch.bool),
        particle_indices=wide["particle_indices"],
        morton_codes=sorted_codes.view(torch.uint64),
        node_ranges=wide["node_ranges"],
        num_particles=n,
        num_internal_nodes=NI,
        node_level=wide["node_level"],
        level_offsets=wide["level_offsets"],
        nodes_by_level=wide["nodes_by_level"],
        num_levels=num_levels,
        bounds_min=b6[:3],
        bounds_max=b6[3:],
        leaf_codes=rad["leaf_codes"].view(torch.uint64),
        leaf_depths=torch.full((L,), -1, dtype=idx, device=dev),
        use_morton_geometry=torch.zeros((), dtype=torch.bool, device=dev),
        leaf_size=int(leaf_size),
    )
    if octree:
        fields.update(
            oct_parent=wide["oct_parent"],
            oct_children=wide["oct_children"],
            oct_child_counts=wide["oct_child_counts"],
            oct_child_mask=octa["oct_child_mask"].view(torch.bool),
            oct_valid_mask=octa["oct_valid_mask"].view(torch.bool),
            oct_node_codes=octa["oct_node_codes"].view(torch.uint64),
            oct_node_depths=wide["oct_node_depths"],
            oct_node_ranges=wide["oct_node_ranges"],
            oct_nodes_by_level=wide["oct_nodes_by_level"],
            oct_level_offsets=wide["oct_level_offsets"],
            oct_num_levels=int(oct_num_levels),
            oct_leaf_mask=octa["oct_leaf_mask"].view(torch.bool),
            oct_leaf_nodes=wide["oct_leaf_nodes"],
            radix_node_to_oct=wide["radix_node_to_oct"],
            radix_leaf_to_oct=wide["radix_leaf_to_oct"],
        )
        topo = OctreeTopology(**fields)
    else:
        topo = RadixTreeTopology(**fields)
    dt = DeviceTopology(
        num_particles=n, num_nodes=T, num_internal=NI, num_leaves=L, leaf_size=int(leaf_size),
        parent=rad["parent"], left=rad["left_child"], right=rad["right_child"],
        node_ranges=rad["node_ranges"], node_level=rad["node_level"],
        nodes_by_level=rad["nodes_by_level"], level_offsets=rad["level_offsets"],
        leaf_indices=leaf_indices, leaf_position=leaf_position, num_levels=num_levels,
        max_level=max_level, malformed=malformed, is_octree=octree)
    cls = OctreeTree if octree else RadixTree
    tree = cls(topology=topo, build_mode="adaptive", positions_sorted=pos_sorted,
               masses_sorted=mass_sorted,
               inverse_permutation=wide.get("inverse_permutation"), _device=dt)
    if return_reordered:
        return tree, pos_sorted, mass_sorted, tree.inverse_permutation
    return tree


def build_tree(positions, masses=None, bounds=None, *, return_reordered=False, leaf_size=8,
               workspace=None, return_workspace=False, config: Optional[TreeBuildConfig] = None):
    """Adaptive radix (LBVH) tree (yggdrax/tree.py:1413-1478)."""
    if config is not None:
        return_reordered = config.return_reordered
        leaf_size = config.leaf_size
        return_workspace = config.return_workspace
    out = _build(positions, masses, bounds, leaf_size=leaf_size,
                 return_reordered=return_reordered, octree=False)
    if return_workspace:
        return (*out, RadixTreeWorkspace()) if isinstance(out, tuple) else (out, RadixTreeWorkspace())
    return out


def build_octree(positions, masses=None, bounds=None, *, return_reordered=False, leaf_size=8,
                 workspace=None, return_workspace=False, config: Optional[TreeBuildConfig] = None):
    """Radix topology augmented with explicit octree metadata
    (yggdrax/tree.py:1482-1549)."""
    if config is not None:
        return_reordered = config.return_reordered
        leaf_size = config.leaf_size
        return_workspace = config.return_workspace
    out = _build(positions, masses, bounds, leaf_size=leaf_size,
                 return_reordered=return_reordered, octree=True)
    if return_workspace:
        return (*out, RadixTreeWorkspace()) if isinstance(out, tuple) else (out, RadixTreeWorkspace())
    return out


def _radix_builder(req: TreeBuildRequest) -> Tree:
    return RadixTree.from_particles(req.positions, req.masses, bounds=req.bounds,
                                    return_reordered=req.return_reordered,
                                    leaf_size=req.leaf_size)


def _octree_builder(req: TreeBuildRequest) -> Tree:
    return OctreeTree.from_particles(req.positions, req.masses, bounds=req.bounds,
                                     return_reordered=req.return_reordered,
                                     leaf_size=req.leaf_size)


_TREE_BUILDERS: dict[str, TreeBuilder] = {"radix": _radix_builder, "octree": _octree_builder}


def register_tree_builder(tree_type: str, builder: TreeBuilder, *, overwrite: bool = False) -> None:
    """yggdrax/tree.py:1308-1321."""
    name = tree_type.strip()
    if not name:
        raise ValueError("tree_type must be a non-empty string")
    if name in _TREE_BUILDERS and not overwrite:
        raise ValueError(f"tree_type '{name}' is already registered; pass overwrite=True to replace it")
    _TREE_BUILDERS[name] = builder


def available_tree_types() -> tuple[str, ...]:
    return tuple(sorted(_TREE_BUILDERS))


def tree_from_particles(positions, masses, *, tree_type="radix", build_mode="adaptive",
                        bounds=None, return_reordered=True, leaf_size=8) -> Tree:
    """Tree.from_particles dispatch through the builder registry (tree.py:245-330)."""
    builder = _TREE_BUILDERS.get(tree_type)
    if builder is None:
        supported = ", ".join(f"'{k}'" for k in sorted(_TREE_BUILDERS))
        raise ValueError(f"Unsupported tree_type '{tree_type}'. Supported: ({supported})")
    return builder(TreeBuildRequest(positions, masses, build_mode, bounds, return_reordered,
                                    leaf_size))


def resolve_tree_topology(tree):
    topo = getattr(tree, "topology", None)
    return tree if topo is None else topo


FMM_CORE_REQUIRED_FIELDS = ("parent", "left_child", "right_child", "node_ranges",
                            "num_particles", "use_morton_geometry")


def require_fmm_core_topology(tree_or_topology) -> None:
    """ValueError naming the FMM-core fields a topology lacks
    (yggdrax/tree.py:135-142, 1085-1095, 1138-1145)."""
    topo = resolve_tree_topology(tree_or_topology)
    missing = [f for f in FMM_CORE_REQUIRED_FIELDS if not hasattr(topo, f)]
    if missing:
        tree_type = getattr(tree_or_topology, "tree_type", None)
        prefix = f"tree_type='{tree_type}' " if tree_type is not None else ""
        raise ValueError(f"{prefix}topology is missing FMM-core-required fields: "
                         + ", ".join(missing))


def device_topology(tree) -> DeviceTopology:
    """int32 device view of a tree or a bare topology."""
    dt = getattr(tree, "_device", None)
    if dt is not None:
        return dt
    topo = resolve_tree_topology(tree)
    lib = _lib.load()
    dev = topo.parent.device
    T = int(topo.parent.shape[0])
    NI = int(topo.num_internal_nodes)
    L = T - NI
    i32 = lambda t: torch.as_tensor(t, device=dev).to(torch.int32).contiguous()  # noqa: E731
    leaf_indices = torch.empty(L, dtype=torch.int32, device=dev)
    leaf_position = torch.empty(T, dtype=torch.int32, device=dev)
    octree = hasattr(topo, "radix_leaf_to_oct")
    if octree:
        order = torch.argsort(i32(topo.radix_leaf_to_oct), stable=True).to(torch.int32)
        leaf_indices = order + NI
        leaf_position = torch.full((T,), -1, dtype=torch.int32, device=dev)
        leaf_position[leaf_indices.long()] = torch.arange(L, dtype=torch.int32, device=dev)
    else:
        _lib.check(lib.ygg_radix_leaf_order(L, leaf_indices.data_ptr(), leaf_position.data_ptr(),
                                            _lib.stream_ptr(dev)), "leaf_order")
    lvl = i32(topo.node_level)
    return DeviceTopology(
        num_particles=int(topo.num_particles), num_nodes=T, num_internal=NI, num_leaves=L,
        leaf_size=int(topo.leaf_size or 0), parent=i32(topo.parent), left=i32(topo.left_child),
        right=i32(topo.right_child), node_ranges=i32(topo.node_ranges), node_level=lvl,
        nodes_by_level=i32(topo.nodes_by_level), level_offsets=i32(topo.level_offsets),
        leaf_indices=leaf_indices, leaf_position=leaf_position,
        num_levels=int(topo.num_levels), max_level=int(lvl.max()), malformed=True,
        is_octree=octree)


__all__ = [
    "MAX_TREE_LEVELS", "OctreeTopology", "OctreeTree", "RadixTree", "RadixTreeTopology",
    "RadixTreeWorkspace", "Tree", "TreeBuildConfig", "TreeBuildRequest", "available_tree_types",
    "build_octree", "build_tree", "device_topology", "infer_bounds", "morton_encode",
    "register_tree_builder", "require_fmm_core_topology", "resolve_tree_topology",
    "tree_from_particles",
]
