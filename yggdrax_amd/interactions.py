"""Dual-tree interaction and neighbour lists (yggdrax/interactions.py:338-448,
yggdrax/_interactions_impl.py).

The walk runs on the GPU (ygg_walk_bfs); this module restates the
reference's host-side control: capacity ladders and retries
(_run_dual_tree_walk_raw, 3840-4440), the queue-capacity LRU cache
(245-293), the count-pass sizing path (3989-4299) and the output shaping
(_result_to_interactions / _result_to_neighbors, 3561-3837).
"""

from __future__ import annotations

import contextvars
import ctypes
import logging
import os
import time
from collections import OrderedDict
from dataclasses import dataclass
from typing import Callable, List, NamedTuple, Optional

import numpy as np
import torch

from . import _lib
from .dtypes import INDEX_DTYPE
from .geometry import TreeGeometry
from .tree import DeviceTopology, device_topology, resolve_tree_topology

logger = logging.getLogger(__name__)

_DEFAULT_MAX_INTERACTIONS = 512
_DEFAULT_MAX_NEIGHBORS = 2048
DEFAULT_PAIR_QUEUE_MULTIPLIER = 128
_COUNT_PASS_FILL_INTERACTION_CAP = 4096
_AUTO_INTERACTION_MAX = 16384
MAX_GENERATIONS = 1024
# 64-bit words per generation record of a resumable walk
# (YGG_WALK_GEN_RECORD_WORDS of the C ABI)
GEN_RECORD_WORDS = 16
# Opt-in per-generation kernel timing (used by the benchmark's roofline leg).
WALK_TIMING = {"enabled": False, "log": []}
# Walks run by the flat path: counting walks, emission walks without a
# capacity hint, hinted walks, hinted walks whose pair queue overflowed
# (rerun without the hint) or whose far/near streams had to grow, and
# stream growths of any resumable walk.
FLAT_STATS = {"count_walks": 0, "plain_walks": 0, "hinted_walks": 0, "hint_overflows": 0,
              "hint_growths": 0, "stream_growths": 0}
# Stream sizing of a walk without a capacity hint. "grow": one resumable
# emission walk whose far and near streams start at the reference's default
# per-row capacities in aggregate (max_interactions_per_node = 512 entries
# per node, max_neighbors_per_leaf = 2048 per leaf;
# _interactions_impl.py:47-48), at most STREAM_START_MAX_BYTES, and grow
# when a generation might not fit. "count": a counting walk, then an
# emission walk at the exact totals (the reference's count pass,
# _interactions_impl.py:3989-4299). Both give the same lists.
STREAM_SIZING = os.environ.get("YGGDRAX_STREAM_SIZING", "grow")
STREAM_START_MAX_BYTES = 16 << 30
# How the far and near row lists are grouped by row. "bucket": a split by
# the row key's high digit, then per-bucket counts and a scatter by the low
# digit, ranked with LDS atomics (row keys of at most 22 bits). "sort": the
# stable radix sort with the autotuned digit width, its last pass writing
# the lists. The order inside a row is the walk's emission order in both,
# which is not specified; the rows are the same.
ROW_LISTS = os.environ.get("YGGDRAX_ROW_LISTS", "bucket")
ROW_BUCKET_MAX_BITS = 22
ROW_BUCKET_DIRECT_BITS = 11  # row keys of at most this many bits are one bucket
# Stream format of the emission walk when the bucketing groups the lists:
# one record per accepted or near pair (both row keys), which the row-list
# split expands into the pair's two entries through a row key -> node table,
# instead of the two entries themselves. The walk writes, and the split
# reads, half the bytes; the lists are the same.
PAIR_RECORDS = os.environ.get("YGGDRAX_PAIR_RECORDS", "1") != "0"


def pair_records_ok(num_nodes: int) -> bool:
    """Whether a walk whose lists are grouped by ygg_far_lists_sort /
    ygg_near_lists_sort may emit pair records: only the bucketing lists
    them, keyed by node id (at most ROW_BUCKET_MAX_BITS bits). Trees of
    at most 2^ROW_BUCKET_DIRECT_BITS nodes keep entries: their row keys
    are one bucket, which the bucketing reads in place, while records
    would add the histogram and split passes."""
    bits = max(1, (num_nodes - 1).bit_length())
    return (PAIR_RECORDS and ROW_LISTS == "bucket"
            and ROW_BUCKET_DIRECT_BITS < bits <= ROW_BUCKET_MAX_BITS)
# Sizing of the last flat walk: how its streams were sized ("hint", "grow"
# or "count"), their final capacities and the number of growths.
LAST_WALK: dict = {}
# Launch shapes of the flat walk from the last walk on a tree of the same
# shape (node and leaf counts, roots, MAC, opening angle): its pairs per
# generation size each generation's grid (ygg_walk_flat's gen_pairs_hint)
# and the first batch of generation launches, so the tail generations do
# not dispatch thousands of workgroups that exit at once and the walk needs
# no second host round trip. Only launch costs depend on it: any grid covers
# a generation.
WALK_GEN_PROFILE = os.environ.get("YGGDRAX_WALK_GEN_PROFILE", "1") != "0"
_GEN_PROFILES: "OrderedDict[tuple, list]" = OrderedDict()
_GEN_PROFILES_MAX = 32
# Walk chunks split into this many contiguous ranges, one per XCD (MI355X
# has 8; 0 or 1 = plain grid-stride order).
WALK_XCD_GROUPS = int(os.environ.get("YGGDRAX_WALK_XCD_GROUPS", "0"))
# Queue tile of the flat walk: pairs per thread of one 256-thread chunk
# (4, 8 or 16); the autotuner may override per size class.
WALK_PAIRS_PER_THREAD = int(os.environ.get("YGGDRAX_WALK_PAIRS_PER_THREAD", "8"))
MAC_TYPES = {"bh": 0, "dehnen": 1, "engblom": 2}
FLAT_GRID_BLOCKS = 2048
# Counter lanes of the counting walk's far/near totals (workgroup b adds to
# lane b % SPREAD_LANES).
SPREAD_LANES = 64
# Forces the attempt-by-attempt walk with per-row slots (the reference's
# fixed-capacity method) instead of the flat walk plus ladder emulation.
EXACT_ATTEMPTS = os.environ.get("YGGDRAX_EXACT_ATTEMPTS", "0") == "1"
NODE_RECORD_BYTES = 32


def _env_flag(name: str, default: bool = False) -> bool:
    raw = os.environ.get(name)
    if raw is None:
        return bool(default)
    return str(raw).strip().lower() in ("1", "true", "yes", "on")


# Diagnostic lines of the traversal driver on stdout, as the reference's
# YGGDRAX_TRAVERSAL_DIAGNOSTICS (_interactions_impl.py:401-413).
_TRAVERSAL_DIAGNOSTICS = _env_flag("YGGDRAX_TRAVERSAL_DIAGNOSTICS", False)


def _traversal_diag(message: str) -> None:
    if _TRAVERSAL_DIAGNOSTICS:
        print(f"[yggdrax.traversal] {message}", flush=True)


# Stage-timing hook of build_compact_far_pairs_and_leaf_neighbor_lists
# (timing_callback, _interactions_impl.py:5100-5102): set for the duration of
# one call, called with (stage name, seconds).
_STAGE_CALLBACK: "contextvars.ContextVar[Optional[Callable[[str, float], None]]]" = \
    contextvars.ContextVar("ygg_stage_callback", default=None)


def _record_stage(name: str, start: float, device) -> None:
    cb = _STAGE_CALLBACK.get()
    if cb is not None:
        torch.cuda.synchronize(device)
        cb(name, float(time.perf_counter() - start))


@dataclass(frozen=True)
class DualTreeTraversalConfig:
    """Fixed traversal parameters (yggdrax/_interactions_impl.py:235-242)."""

    max_pair_queue: int
    process_block: int
    max_interactions_per_node: int
    max_neighbors_per_leaf: int = _DEFAULT_MAX_NEIGHBORS


_GLOBAL_DUAL_TREE_CONFIG: Optional[DualTreeTraversalConfig] = None
_MAX_DUAL_TREE_QUEUE_CACHE_ENTRIES = 128
_DUAL_TREE_QUEUE_CACHE: "OrderedDict[tuple, int]" = OrderedDict()


def set_default_dual_tree_config(config: Optional[DualTreeTraversalConfig]) -> None:
    if config is not None:
        for name in ("max_pair_queue", "process_block", "max_interactions_per_node",
                     "max_neighbors_per_leaf"):
            if getattr(config, name) < 1:
                raise ValueError(f"{name} must be >= 1")
    global _GLOBAL_DUAL_TREE_CONFIG
    _GLOBAL_DUAL_TREE_CONFIG = config


def _get_cached_queue_capacity(key):
    v = _DUAL_TREE_QUEUE_CACHE.get(key)
    if v is None:
        return None
    _DUAL_TREE_QUEUE_CACHE.move_to_end(key)
    return int(v)


def _store_cached_queue_capacity(key, cap: int) -> None:
    _DUAL_TREE_QUEUE_CACHE[key] = int(cap)
    _DUAL_TREE_QUEUE_CACHE.move_to_end(key)
    while len(_DUAL_TREE_QUEUE_CACHE) > _MAX_DUAL_TREE_QUEUE_CACHE_ENTRIES:
        _DUAL_TREE_QUEUE_CACHE.popitem(last=False)


class DualTreeRetryEvent(NamedTuple):
    attempt: int
    queue_capacity: int
    interaction_capacity: int
    status: str
    far_pair_count: int
    near_pair_count: int


class NodeInteractionList(NamedTuple):
    offsets: torch.Tensor
    sources: torch.Tensor
    targets: torch.Tensor
    counts: torch.Tensor
    level_offsets: torch.Tensor
    target_levels: torch.Tensor


class NodeNeighborList(NamedTuple):
    offsets: torch.Tensor
    neighbors: torch.Tensor
    leaf_indices: torch.Tensor
    counts: torch.Tensor
    particle_order_leaf_indices: torch.Tensor
    particle_order_to_native_leaf: torch.Tensor
    neighbor_leaf_positions: torch.Tensor
    target_block_leaf_ids: torch.Tensor
    target_block_source_leaf_ids: torch.Tensor
    target_block_valid_mask: torch.Tensor
    target_block_offsets: torch.Tensor
    target_block_size: int


class CompactTaggedFarPairs(NamedTuple):
    """Exact-length far pairs with per-pair tags (_interactions_impl.py:634-644):
    the walk's far entries in level-major order (the order of
    NodeInteractionList.sources/targets), tags -1 without a pair policy."""

    sources: torch.Tensor
    targets: torch.Tensor
    tags: torch.Tensor
    far_pair_count: torch.Tensor


class DualTreeWalkResult(NamedTuple):
    interaction_offsets: torch.Tensor
    interaction_sources: torch.Tensor
    interaction_targets: torch.Tensor
    interaction_tags: torch.Tensor
    interaction_counts: torch.Tensor
    neighbor_offsets: torch.Tensor
    neighbor_indices: torch.Tensor
    neighbor_counts: torch.Tensor
    leaf_indices: torch.Tensor
    far_pair_count: int
    near_pair_count: int
    queue_overflow: bool
    far_overflow: bool
    near_overflow: bool
    accept_decisions: int
    near_decisions: int
    refine_decisions: int


# ---- capacity ladders (_interactions_impl.py:296-316, 451-563) -------------

def _queue_candidates_bounded(*, max_capacity: int, process_block: int) -> list[int]:
    qmax = max(4, int(max_capacity))
    base = min(max(1024, int(process_block) * 16), qmax)
    c = [base]
    while c[-1] < qmax:
        nxt = min(qmax, c[-1] * 2)
        if nxt == c[-1]:
            break
        c.append(nxt)
    return c


def _auto_pair_queue_candidates(total_nodes: int, num_internal: int) -> List[int]:
    conservative = DEFAULT_PAIR_QUEUE_MULTIPLIER * max(4, total_nodes)
    start = min(max(4096, num_internal * 2), conservative)
    if start >= conservative:
        return [int(conservative)]
    c: List[int] = []
    cap = start
    while cap < conservative:
        c.append(int(cap))
        cap = max(cap * 2, cap + 1024)
    c.append(int(conservative))
    return list(dict.fromkeys(c))


def _interaction_capacity_candidates(requested, total_nodes):
    if requested is not None:
        if requested < 1:
            raise ValueError("max_interactions_per_node must be >= 1")
        return [int(requested)], False
    scaled = min(max(total_nodes // 4, 2048), _AUTO_INTERACTION_MAX)
    p2 = 1 << (max(1, scaled) - 1).bit_length()
    return [int(min(_AUTO_INTERACTION_MAX, max(2048, p2)))], False


def _neighbor_capacity_candidates(requested: int, *, num_leaves: int) -> list[int]:
    base = max(1, int(requested))
    hard = max(1, ((1 << 31) - 1) // max(1, int(num_leaves)))
    base = min(base, hard)
    c = [int(base)]
    while c[-1] < hard:
        nxt = min(hard, max(c[-1] * 2, c[-1] + 1024))
        if nxt == c[-1]:
            break
        c.append(int(nxt))
    return list(dict.fromkeys(c))


QUEUE_ERR_USER = "Pair queue capacity exceeded; increase max_pair_queue and rebuild."
QUEUE_ERR_AUTO = ("Pair queue capacity exceeded; pass a larger max_pair_queue value to "
                  "override auto-scaling.")
FAR_ERR = ("Interaction list capacity exceeded; increase max_interactions_per_node and "
           "rebuild.")
NEAR_ERR = "Neighbor list capacity exceeded; increase max_neighbors_per_leaf and rebuild."


# ---- one GPU walk attempt ---------------------------------------------------

@dataclass
class WalkOutcome:
    """Device results of one walk attempt plus the host-visible status."""

    far_counts: torch.Tensor        # int32 [T] (raw, may exceed capacity on overflow)
    near_counts: torch.Tensor       # int32 [L]
    far_entries: Optional[torch.Tensor]
    near_entries: Optional[torch.Tensor]
    far_layout: tuple               # (fixed_cap, row_start or None)
    near_layout: tuple
    gen_sizes: list                 # wavefront size per generation (clipped to queue cap)
    queue_overflow: bool
    far_overflow: bool
    near_overflow: bool
    decisions: tuple                # accept, near, refine
    queue_capacity: int

    @property
    def max_wavefront(self) -> int:
        return max([1] + list(self.gen_sizes))


def _start_capacities(num_nodes: int, num_leaves: int) -> tuple:
    """Starting far and near stream capacities of a walk without a hint:
    the reference's default per-row capacities summed over the rows, scaled
    down together to at most STREAM_START_MAX_BYTES (8 bytes per entry)."""
    far = _DEFAULT_MAX_INTERACTIONS * max(1, num_nodes)
    near = _DEFAULT_MAX_NEIGHBORS * max(1, num_leaves)
    limit = STREAM_START_MAX_BYTES // 8
    if far + near > limit:
        far, near = far * limit // (far + near), near * limit // (far + near)
    return max(1, far), max(1, near)


def _host_totals(hs: list, G: int) -> tuple:
    """Far and near totals of a flat walk from its host status words
    (two 64-bit words after the generation sizes and the four flags)."""
    w = [v & 0xFFFFFFFF for v in hs[G + 5:G + 9]]
    return w[0] | (w[1] << 32), w[2] | (w[3] << 32)


def _grown(t: Optional[torch.Tensor], cap: int, keep: int) -> Optional[torch.Tensor]:
    """A stream of `cap` entries holding the first `keep` entries of `t`."""
    if t is None:
        return None
    out = torch.empty(max(1, cap), dtype=t.dtype, device=t.device)
    if keep:
        out[:keep].copy_(t[:keep])
    return out


class _WalkContext:
    """Per (tree, geometry, MAC) state shared by all attempts of one call."""

    def __init__(self, dt: DeviceTopology, geometry: TreeGeometry, theta: float, mac_type: str,
                 dehnen_radius_scale: float):
        self.lib = _lib.load()
        self.dt = dt
        self.dev = geometry.center.device
        self.stream = _lib.stream_ptr(self.dev)
        if mac_type not in MAC_TYPES:
            raise ValueError(f"Unknown mac_type: {mac_type}")
        self.mac = MAC_TYPES[mac_type]
        th = ctypes.c_float(float(theta)).value
        th_sq = ctypes.c_float(ctypes.c_float(th * th).value).value
        self.theta_sq = th_sq
        self.theta_root = float(np.sqrt(np.float32(th_sq)))  # float32 square root
        self.nodes = torch.empty(dt.num_nodes * NODE_RECORD_BYTES, dtype=torch.uint8,
                                 device=self.dev)
        g = geometry
        left = dt.left if dt.left.numel() else None
        right = dt.right if dt.right.numel() else None
        _lib.check(self.lib.ygg_walk_prepare_nodes(
            g.center.contiguous().data_ptr(), g.radius.contiguous().data_ptr(),
            g.max_extent.contiguous().data_ptr(), _lib.ptr(left), _lib.ptr(right),
            dt.leaf_position.data_ptr(), dt.nodes_by_level.data_ptr(), dt.num_nodes, self.mac,
            ctypes.c_float(float(dehnen_radius_scale)).value, self.nodes.data_ptr(),
            self.stream), "walk_prepare_nodes")
        self.host_status = torch.zeros(MAX_GENERATIONS + 16, dtype=torch.int32, pin_memory=True)
        self.gen_sizes = torch.empty(MAX_GENERATIONS + 1, dtype=torch.int32, device=self.dev)
        self.flags = torch.empty(4, dtype=torch.int32, device=self.dev)
        self.decisions = torch.empty(3, dtype=torch.int64, device=self.dev)
        self.timing = WALK_TIMING["enabled"]
        self.gen_hint = int(1.5 * max(1, dt.num_levels)) + 4
        self.profile_key = (dt.num_nodes, dt.num_internal, dt.num_leaves, self.mac,
                            self.theta_sq)
        self.grid_blocks = FLAT_GRID_BLOCKS
        self.pairs_per_thread = WALK_PAIRS_PER_THREAD
        self.kernel_log = WALK_TIMING["log"]

    def run_flat(self, *, emit: bool, queue_cap: int, far_cap: int = 0, near_cap: int = 0,
                 gen_hint: int = 0, hist_gen: int = -1, node_hist=None, grid_blocks: int = 0,
                 policy=None, grow: bool = False):
        """One flat-emission walk over the whole tree (ygg_walk_flat).

        ``policy`` (a _lib.PolicyDesc) replaces the built-in accept / near
        rules; the emission walk then also writes a tag per far entry.
        ``grow`` makes the emission walk resumable: a generation that might
        not fit the far or near stream (two entries per pair) pauses the
        walk before it runs, the streams are reallocated larger with their
        entries copied over, and the walk continues from that generation,
        so the capacities are starting sizes and never overflow.
        With the context's ``pair_records`` set (and no policy) the
        emission walk writes one record per pair instead of two entries:
        capacities, totals and decisions stay in entries, the streams hold
        records (``pair_records`` True in the result)."""
        dt = self.dt
        dev = self.dev
        roots = getattr(self, "roots", None)
        # generation 0 holds every root pair, so the queue holds at least those
        Q = max(4, int(queue_cap), 0 if roots is None else int(roots.numel()))
        wf = torch.empty((2, Q), dtype=torch.int64, device=dev)
        counters = torch.empty(2, dtype=torch.int64, device=dev)
        # the counting walk adds its totals into spread lanes instead of
        # reserving per chunk on two contended addresses
        spread = None if emit else torch.empty(2 * SPREAD_LANES, dtype=torch.int64, device=dev)
        fk = fv = nk = nv = ftag = None
        pairs = bool(emit and policy is None and getattr(self, "pair_records", False))
        per = 1 if pairs else 2  # stream slots per accepted or near pair
        if pairs:  # entry capacities to record capacities
            far_cap, near_cap = -(-int(far_cap) // 2), -(-int(near_cap) // 2)
        if emit:
            fk = torch.empty(max(1, far_cap), dtype=torch.int32, device=dev)
            fv = torch.empty(max(1, far_cap), dtype=torch.int32, device=dev)
            if policy is not None:
                ftag = torch.empty(max(1, far_cap), dtype=torch.int32, device=dev)
            nk = torch.empty(max(1, near_cap), dtype=torch.int32, device=dev)
            nv = torch.empty(max(1, near_cap), dtype=torch.int32, device=dev)
        p = _lib.FlatWalkParams()
        p.num_nodes, p.num_internal, p.num_leaves, p.root = (dt.num_nodes, dt.num_internal,
                                                             dt.num_leaves, 0)
        p.mac_type, p.emit, p.max_generations = self.mac, int(emit), MAX_GENERATIONS
        p.grid_blocks = int(grid_blocks) if grid_blocks else self.grid_blocks
        p.gen_hint = int(gen_hint) if gen_hint else int(self.gen_hint)
        p.theta_sq, p.theta_root = self.theta_sq, self.theta_root
        p.queue_cap, p.far_cap, p.near_cap = Q, int(far_cap), int(near_cap)
        p.nodes = self.nodes.data_ptr()
        p.wf0, p.wf1 = wf[0].data_ptr(), wf[1].data_ptr()
        p.gen_sizes, p.flags = self.gen_sizes.data_ptr(), self.flags.data_ptr()
        p.counters = counters.data_ptr()
        p.far_keys, p.far_vals = _lib.ptr(fk), _lib.ptr(fv)
        p.near_keys, p.near_vals = _lib.ptr(nk), _lib.ptr(nv)
        p.hist_gen, p.node_hist = int(hist_gen), _lib.ptr(node_hist)
        p.xcd_groups = WALK_XCD_GROUPS
        p.pairs_per_thread = int(self.pairs_per_thread)
        p.pair_records = int(pairs)
        if spread is not None:
            p.spread_counters, p.spread_lanes = spread.data_ptr(), SPREAD_LANES
        if policy is not None:
            p.policy = ctypes.cast(ctypes.pointer(policy), ctypes.c_void_p)
            p.far_tags = _lib.ptr(ftag)
        if roots is not None:
            p.roots, p.num_roots = roots.data_ptr(), int(roots.numel())
        G = MAX_GENERATIONS
        if emit and grow:
            gen_fill = torch.empty(G * GEN_RECORD_WORDS, dtype=torch.int64, device=dev)
            p.gen_fill = gen_fill.data_ptr()
        pkey = (self.profile_key, 0 if roots is None else int(roots.numel()), policy is None)
        prof = _GEN_PROFILES.get(pkey) if WALK_GEN_PROFILE else None
        if prof:
            # (its last generation is the empty one that ended the walk)
            p.gen_pairs_hint = ctypes.cast(prof[1], ctypes.c_void_p)
            if not gen_hint:  # one batch: every generation of the profile
                p.gen_hint = max(int(p.gen_hint), min(len(prof[0]), G))
        ngen = ctypes.c_int32(0)
        gen_ms = (ctypes.c_float * MAX_GENERATIONS)() if self.timing else None
        growths = 0
        while True:
            _lib.check(self.lib.ygg_walk_flat(ctypes.byref(p), self.host_status.data_ptr(),
                                              ctypes.byref(ngen), gen_ms, self.stream),
                       "walk_flat")
            hs = self.host_status[:G + 9].tolist()  # pinned words the walk wrote
            paused = hs[G + 4] if emit and grow else 0
            if not paused:
                break
            # generation g did not run: grow the streams to hold its worst
            # case after the entries written so far, and continue from it
            g = paused - 1
            filled = _host_totals(hs, G)
            size = min(hs[g] & 0xFFFFFFFF, Q)
            need_f, need_n = filled[0] + per * size, filled[1] + per * size
            if need_f <= p.far_cap and need_n <= p.near_cap:
                raise RuntimeError(f"walk paused at generation {g} with room in both streams")
            if need_f > p.far_cap:
                cap = max(need_f, 3 * p.far_cap // 2)
                fk, fv, ftag = (_grown(t, cap, filled[0]) for t in (fk, fv, ftag))
                p.far_cap = cap
                p.far_keys, p.far_vals, p.far_tags = _lib.ptr(fk), _lib.ptr(fv), _lib.ptr(ftag)
            if need_n > p.near_cap:
                cap = max(need_n, 3 * p.near_cap // 2)
                nk, nv = (_grown(t, cap, filled[1]) for t in (nk, nv))
                p.near_cap = cap
                p.near_keys, p.near_vals = _lib.ptr(nk), _lib.ptr(nv)
            p.start_generation = g
            growths += 1
            FLAT_STATS["stream_growths"] += 1
        raw_gens = [v & 0xFFFFFFFF for v in hs[:min(ngen.value + 2, G + 1)]]
        if WALK_GEN_PROFILE and not (hs[G + 1] or hs[G + 2] or hs[G + 3]):
            sizes = [min(v, Q) for v in raw_gens]
            if not prof or prof[0] != sizes:
                # generations past the profile are unknown: their grids from the bound
                arr = (ctypes.c_uint32 * G)(*sizes[:G])
                for k in range(len(sizes), G):
                    arr[k] = 0xFFFFFFFF
                _GEN_PROFILES[pkey] = (sizes, arr)
            _GEN_PROFILES.move_to_end(pkey)
            while len(_GEN_PROFILES) > _GEN_PROFILES_MAX:
                _GEN_PROFILES.popitem(last=False)
        slots = _host_totals(hs, G)
        # every evaluated pair is exactly one of accept (2 far entries, or
        # one record), near (2 near entries, or one record) or refine
        c = (2 * slots[0], 2 * slots[1]) if pairs else slots
        evaluated = sum(min(v, Q) for v in raw_gens)
        dec = (c[0] // 2, c[1] // 2, evaluated - c[0] // 2 - c[1] // 2)
        self.gen_hint = ngen.value + 1
        if gen_ms is not None:
            self.kernel_log.append(dict(kernel="k_walk_flat", emit=bool(emit),
                                        gen_ms=list(gen_ms)[: ngen.value], decisions=dec,
                                        gen_sizes=raw_gens, far_total=c[0], near_total=c[1]))
        return dict(far_total=c[0], near_total=c[1], decisions=dec,
                    gen_sizes=raw_gens, peak=max(raw_gens + [1]),
                    queue_overflow=bool(hs[G + 1]), far_overflow=bool(hs[G + 2]),
                    near_overflow=bool(hs[G + 3]), far_keys=fk, far_vals=fv,
                    near_keys=nk, near_vals=nv, far_tags=ftag, queue_capacity=Q,
                    far_capacity=int(p.far_cap) * (2 // per),
                    near_capacity=int(p.near_cap) * (2 // per),
                    stream_bytes=(8 * (int(p.far_cap) + int(p.near_cap))
                                  + (4 * int(p.far_cap) if ftag is not None else 0)) if emit else 0,
                    stream_growths=growths, pair_records=pairs)

    def run(self, *, mode: int, queue_cap: int, far_cap: int = 0, near_cap: int = 0,
            far_row_start=None, far_row_cap=None, near_row_start=None, near_row_cap=None,
            far_capacity_total: int = 0, near_capacity_total: int = 0,
            collect_far: bool = True, collect_near: bool = True) -> WalkOutcome:
        dt = self.dt
        dev = self.dev
        Q = max(4, int(queue_cap))
        wf = torch.empty((2, Q), dtype=torch.int64, device=dev)
        far_counts = torch.empty(dt.num_nodes, dtype=torch.int32, device=dev)
        near_counts = torch.empty(dt.num_leaves, dtype=torch.int32, device=dev)
        far_entries = near_entries = None
        if mode == 1:
            far_entries = torch.empty(max(1, dt.num_nodes * far_cap), dtype=torch.int32, device=dev)
            near_entries = torch.empty(max(1, dt.num_leaves * near_cap), dtype=torch.int32,
                                       device=dev)
        elif mode == 2:
            far_entries = torch.empty(max(1, far_capacity_total), dtype=torch.int32, device=dev)
            near_entries = torch.empty(max(1, near_capacity_total), dtype=torch.int32, device=dev)
        p = _lib.WalkParams()
        p.num_nodes, p.num_internal, p.num_leaves = dt.num_nodes, dt.num_internal, dt.num_leaves
        p.root = 0
        p.mac_type, p.mode = self.mac, mode
        p.collect_far, p.collect_near = int(collect_far), int(collect_near)
        p.theta_sq, p.theta_root = self.theta_sq, self.theta_root
        p.far_cap, p.near_cap = int(far_cap), int(near_cap)
        p.queue_cap = Q
        p.max_generations = MAX_GENERATIONS
        p.nodes = self.nodes.data_ptr()
        p.wf0, p.wf1 = wf[0].data_ptr(), wf[1].data_ptr()
        p.gen_sizes, p.flags = self.gen_sizes.data_ptr(), self.flags.data_ptr()
        p.decisions = self.decisions.data_ptr()
        p.far_counts, p.near_counts = far_counts.data_ptr(), near_counts.data_ptr()
        p.far_row_start, p.far_row_cap = _lib.ptr(far_row_start), _lib.ptr(far_row_cap)
        p.near_row_start, p.near_row_cap = _lib.ptr(near_row_start), _lib.ptr(near_row_cap)
        p.far_entries, p.near_entries = _lib.ptr(far_entries), _lib.ptr(near_entries)
        ngen = ctypes.c_int32(0)
        gen_ms = (ctypes.c_float * MAX_GENERATIONS)() if self.timing else None
        _lib.check(self.lib.ygg_walk_bfs(ctypes.byref(p), self.host_status.data_ptr(),
                                         ctypes.byref(ngen), gen_ms, self.stream), "walk")
        if gen_ms is not None:
            self.kernel_log.append(dict(mode=mode, gen_ms=list(gen_ms)[: ngen.value],
                                        decisions=tuple(int(v) for v in self.decisions.cpu()),
                                        gen_sizes=[min(int(self.host_status[k]), Q)
                                                   for k in range(ngen.value + 1)]))
        hs = self.host_status
        gens = [min(int(hs[k]) & 0xFFFFFFFF, Q) for k in range(ngen.value + 1)]
        G = MAX_GENERATIONS
        fl = [int(hs[G + 1 + k]) for k in range(3)]
        dec = tuple(int(v) for v in self.decisions.cpu())
        layout_far = (far_cap if mode == 1 else 0, far_row_start)
        layout_near = (near_cap if mode == 1 else 0, near_row_start)
        return WalkOutcome(far_counts, near_counts, far_entries, near_entries, layout_far,
                           layout_near, gens, bool(fl[0]), bool(fl[1]), bool(fl[2]), dec, Q)


def _scan(counts: torch.Tensor, gather: Optional[torch.Tensor] = None) -> torch.Tensor:
    lib = _lib.load()
    n = counts.numel() if gather is None else gather.numel()
    out = torch.empty(n + 1, dtype=torch.int64, device=counts.device)
    ws = _lib.workspace(lib.ygg_scan_workspace_bytes(n), counts.device)
    _lib.check(lib.ygg_exclusive_scan_i32(counts.data_ptr(), _lib.ptr(gather), out.data_ptr(), n,
                                          ws.data_ptr(), ws.numel(),
                                          _lib.stream_ptr(counts.device)), "scan")
    return out


def _capped(counts: torch.Tensor, fixed_cap: int, caps: Optional[torch.Tensor]) -> torch.Tensor:
    c = counts.clone()
    if fixed_cap > 0 or caps is not None:
        _lib.check(_lib.load().ygg_cap_counts(c.data_ptr(), c.numel(), int(fixed_cap),
                                              _lib.ptr(caps), _lib.stream_ptr(c.device)), "cap")
    return c


class _RawWalk(NamedTuple):
    outcome: WalkOutcome
    far_counts: torch.Tensor   # int32, capped
    near_counts: torch.Tensor  # int32, capped
    far_caps: Optional[torch.Tensor]
    near_caps: Optional[torch.Tensor]


def _emit(retry_logger, status, attempt, q, k, res: _RawWalk):
    if retry_logger is None:
        return
    ev = DualTreeRetryEvent(attempt=int(attempt), queue_capacity=int(q),
                            interaction_capacity=int(k), status=status,
                            far_pair_count=int(res.far_counts.sum()),
                            near_pair_count=int(res.near_counts.sum()))
    try:
        retry_logger(ev)
    except Exception:  # pragma: no cover - defensive, as the reference
        logger.exception("retry_logger raised", exc_info=True)


class _Ladder(NamedTuple):
    queue: list
    inter: list
    nbr: list
    user_q: bool
    cache_key: Optional[tuple]
    use_count: bool


def _ladder(dt: DeviceTopology, mac_type, *, max_interactions_per_node, max_neighbors_per_leaf,
            max_pair_queue, traversal_config, collect_far=True, collect_near=True) -> _Ladder:
    """Capacity candidates of _run_dual_tree_walk_raw (3917-4003)."""
    T, NI, L = dt.num_nodes, dt.num_internal, dt.num_leaves
    config = traversal_config if traversal_config is not None else _GLOBAL_DUAL_TREE_CONFIG
    if config is not None:
        qmax = max(4, int(config.max_pair_queue))
        inter = [int(config.max_interactions_per_node)]
        nlimit = int(config.max_neighbors_per_leaf)
        nbr = _neighbor_capacity_candidates(nlimit, num_leaves=L)
        pov = int(config.process_block)
        cache_key = (T, NI, inter[0], nlimit, qmax, pov, bool(collect_far), bool(collect_near),
                     str(mac_type))
        cached = _get_cached_queue_capacity(cache_key)
        qc = [cached] if cached is not None else _queue_candidates_bounded(
            max_capacity=qmax, process_block=pov)
        for cap in _auto_pair_queue_candidates(T, NI):
            if cap > qc[-1]:
                qc.append(cap)
        return _Ladder(list(dict.fromkeys(qc)), inter, nbr, True, cache_key, False)
    user_q = max_pair_queue is not None
    qc = [max(4, int(max_pair_queue))] if user_q else _auto_pair_queue_candidates(T, NI)
    inter, _ = _interaction_capacity_candidates(max_interactions_per_node, T)
    nbr = _neighbor_capacity_candidates(max_neighbors_per_leaf, num_leaves=L)
    use_count = (max_interactions_per_node is None and max_pair_queue is None
                 and traversal_config is None)
    return _Ladder(qc, inter, nbr, user_q, None, use_count)


def _run_walk_raw(tree, geometry: TreeGeometry, theta: float, mac_type: str, *,
                  max_interactions_per_node, max_neighbors_per_leaf, max_pair_queue,
                  traversal_config, dehnen_radius_scale, process_block, retry_logger,
                  collect_far=True, collect_near=True, ctx=None) -> Optional[_RawWalk]:
    """_run_dual_tree_walk_raw (3840-4440), attempt by attempt, over the GPU walk
    with per-row slots (the reference's fixed-capacity method)."""
    dt = device_topology(tree)
    if float(dehnen_radius_scale) <= 0.0:
        raise ValueError("dehnen_radius_scale must be > 0")
    T, NI, L = dt.num_nodes, dt.num_internal, dt.num_leaves
    if NI == 0:
        return None
    if ctx is None:
        ctx = _WalkContext(dt, geometry, theta, mac_type, dehnen_radius_scale)
    lad = _ladder(dt, mac_type, max_interactions_per_node=max_interactions_per_node,
                  max_neighbors_per_leaf=max_neighbors_per_leaf, max_pair_queue=max_pair_queue,
                  traversal_config=traversal_config, collect_far=collect_far,
                  collect_near=collect_near)
    qc, inter, nbr, cache_key = lad.queue, lad.inter, lad.nbr, lad.cache_key
    qerr = QUEUE_ERR_USER if lad.user_q else QUEUE_ERR_AUTO
    if lad.use_count:
        return _count_then_fill(ctx, dt, retry_logger, collect_far, collect_near)

    result: Optional[_RawWalk] = None
    success = False
    attempt = 0
    for Q in qc:
        queue_retry = False
        for K in inter:
            for Kn in nbr:
                attempt += 1
                o = ctx.run(mode=1, queue_cap=Q, far_cap=K, near_cap=Kn,
                            collect_far=collect_far, collect_near=collect_near)
                result = _RawWalk(o, _capped(o.far_counts, K, None),
                                  _capped(o.near_counts, Kn, None), None, None)
                if o.queue_overflow:
                    _emit(retry_logger, "queue_overflow", attempt, Q, K, result)
                    queue_retry = True
                    break
                if o.far_overflow:
                    raise RuntimeError(FAR_ERR)
                if collect_near and o.near_overflow:
                    _emit(retry_logger, "neighbor_overflow", attempt, Q, K, result)
                    continue
                success = True
                if attempt > 1:
                    _emit(retry_logger, "success", attempt, Q, K, result)
                break
            if queue_retry or success:
                break
        else:
            continue
        if queue_retry and not success:
            continue
        if success:
            if cache_key is not None:
                _store_cached_queue_capacity(cache_key, Q)
            break
    else:
        if result is not None and result.outcome.queue_overflow:
            raise RuntimeError(qerr)
    assert result is not None
    if collect_far and result.outcome.far_overflow:
        raise RuntimeError(FAR_ERR)
    if collect_near and result.outcome.near_overflow:
        raise RuntimeError(NEAR_ERR)
    return result


# ---- flat-emission path -----------------------------------------------------

class _FlatResult(NamedTuple):
    far_keys: torch.Tensor     # sorted level-major row keys (uint32 as int32); the sort's
                               # key scratch when the sorts wrote the lists themselves
    far_vals: Optional[torch.Tensor]
    near_vals: Optional[torch.Tensor]    # sorted by native leaf row
    far_off: torch.Tensor      # int64 [T + 1] level-major row offsets
    near_off: torch.Tensor     # int64 [L + 1]
    stats: dict
    # the list outputs, when the sorts wrote them directly (fused last pass)
    inter: Optional["NodeInteractionList"] = None
    near: Optional[tuple] = None       # (neighbors, counts, offsets)


def row_lists_digit_bits(end_bit: int, n: int) -> int:
    """digit_bits argument of ygg_far_lists_sort / ygg_near_lists_sort for
    n entries with end_bit-bit row keys: 0 selects the bucketing kernels,
    8..11 the stable sort with that digit width."""
    if ROW_LISTS == "bucket" and end_bit <= ROW_BUCKET_MAX_BITS:
        return 0
    from . import autotune
    return autotune.sort_digit_bits(4, end_bit, n) if n else 8


def _fused_lists(em: dict, dt: DeviceTopology):
    """Far and near row lists in the reference's layouts, grouped by row
    by ygg_far_lists_sort / ygg_near_lists_sort (bucketing kernels, or the
    row-key sort whose last pass writes the lists). The near lists, the
    longer of the two, are launched first on the current stream; the far
    lists follow on a side stream, ordered after the walk only. Returns the
    NodeInteractionList, the near (neighbors, counts, offsets) and both
    int64 row offset arrays."""
    lib = _lib.load()
    dev = dt.parent.device
    idx = INDEX_DTYPE
    i64 = int(idx == torch.int64)
    T, L = dt.num_nodes, dt.num_leaves
    nf, nn = em["far_total"], em["near_total"]
    # pair records: half as many stream slots, keyed by node id; the near
    # rows follow leaf_indices (leaf order -> node)
    pairs = bool(em.get("pair_records"))
    n_near, n_far = (nn // 2, nf // 2) if pairs else (nn, nf)
    leaf_nodes = dt.leaf_indices.data_ptr() if pairs else None
    main = torch.cuda.current_stream(dev)
    walked = torch.cuda.Event()
    walked.record(main)
    nend = max(1, ((T if pairs else L) - 1).bit_length())
    nbits = row_lists_digit_bits(nend, nn)
    neighbors = torch.empty(nn, dtype=idx, device=dev)
    ncounts = torch.empty(L, dtype=idx, device=dev)
    noff = torch.empty(L + 1, dtype=idx, device=dev)
    near_off = torch.empty(L + 1, dtype=torch.int64, device=dev)
    nsorted = torch.empty(max(1, nn), dtype=torch.int32, device=dev)
    nws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(nn, nend, nbits), dev)
    _lib.check(lib.ygg_near_lists_sort(
        em["near_keys"].data_ptr(), em["near_vals"].data_ptr(), n_near, None, L, nend, nbits,
        None, leaf_nodes, i64, nsorted.data_ptr(), near_off.data_ptr(), neighbors.data_ptr(),
        ncounts.data_ptr(),
        noff.data_ptr(), nws.data_ptr(), nws.numel(), _lib.stream_ptr(dev)), "near lists sort")
    side = _side_stream(dev)
    side.wait_event(walked)
    with torch.cuda.stream(side):
        fend = max(1, (T - 1).bit_length())
        fbits = row_lists_digit_bits(fend, nf)
        sources = torch.empty(nf, dtype=idx, device=dev)
        targets = torch.empty(nf, dtype=idx, device=dev)
        levels = torch.empty(nf, dtype=idx, device=dev)
        counts = torch.empty(T, dtype=idx, device=dev)
        counts32 = torch.empty(T, dtype=torch.int32, device=dev)
        node_off = torch.empty(T + 1, dtype=idx, device=dev)
        far_off = torch.empty(T + 1, dtype=torch.int64, device=dev)
        fsorted = torch.empty(max(1, nf), dtype=torch.int32, device=dev)
        fws = _lib.workspace(lib.ygg_far_lists_sort_workspace_bytes(nf, fend, fbits), dev)
        _lib.check(lib.ygg_far_lists_sort(
            em["far_keys"].data_ptr(), em["far_vals"].data_ptr(), n_far, None, T, fend, fbits,
            dt.nodes_by_level.data_ptr(), dt.node_level.data_ptr(), None, int(pairs), i64,
            fsorted.data_ptr(),
            far_off.data_ptr(), sources.data_ptr(), targets.data_ptr(), levels.data_ptr(),
            counts.data_ptr(), node_off.data_ptr(), counts32.data_ptr(), fws.data_ptr(),
            fws.numel(), _lib.stream_ptr(dev)), "far lists sort")
    # the walk's streams are read on the side stream too
    for t in (em["far_keys"], em["far_vals"]):
        t.record_stream(side)
    main.wait_stream(side)
    for t in (sources, targets, levels, counts, counts32, node_off, far_off, fsorted, fws):
        t.record_stream(main)
    nl = int(dt.num_levels)
    level_offsets = far_off[dt.level_offsets[: nl + 1].long()].to(idx)
    if nf == 0:
        node_off = torch.zeros(T + 1, dtype=idx, device=dev)
        level_offsets = torch.zeros(nl + 1, dtype=idx, device=dev)
    inter = NodeInteractionList(node_off, sources, targets, counts, level_offsets, levels)
    return inter, (neighbors, ncounts, noff), fsorted, far_off, near_off


def _sort_entries(keys, vals, n, rows, exact_keys=False):
    """Stable radix sort of (row key, value) pairs by row key; the digit
    width comes from the autotuner's plan for this size. Keys below `rows`
    need (rows - 1).bit_length() bits (`exact_keys`); streams padded with
    0xFFFFFFFF fillers sort on rows.bit_length() bits so every filler lands
    after the last row."""
    from . import autotune
    lib = _lib.load()
    dev = keys.device
    end_bit = max(1, (int(rows) - 1 if exact_keys else int(rows)).bit_length())
    ko = torch.empty(max(1, n), dtype=torch.int32, device=dev)
    vo = torch.empty(max(1, n), dtype=torch.int32, device=dev)
    if n > 0:
        bits = autotune.sort_digit_bits(4, end_bit, n)
        ws = _lib.workspace(lib.ygg_sort_workspace_bytes(n, 4, 0, end_bit, bits), dev)
        _lib.check(lib.ygg_sort_pairs_u32(keys.data_ptr(), vals.data_ptr(), ko.data_ptr(),
                                          vo.data_ptr(), n, 0, end_bit, bits, ws.data_ptr(),
                                          ws.numel(), _lib.stream_ptr(dev)), "sort entries")
    off = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    _lib.check(lib.ygg_row_offsets(ko.data_ptr(), n, rows, off.data_ptr(), _lib.stream_ptr(dev)),
               "row offsets")
    return ko, vo, off


_SIDE_STREAMS: dict = {}


def _side_stream(dev) -> torch.cuda.Stream:
    """A second stream of the device (default priority) for independent work."""
    key = torch.device(dev).index
    st = _SIDE_STREAMS.get(key)
    if st is None:
        st = _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return st


def _sort_far_and_near(fk, fv, nf, T, nk, nv, nn, L):
    """The far and near row-key sorts are independent: the far one runs on a
    side stream while the near one runs on the current stream, so their
    passes (each limited by its tile look-back chain) overlap."""
    dev = fk.device
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        far = _sort_entries(fk, fv, nf, T, exact_keys=True)
    near = _sort_entries(nk, nv, nn, L, exact_keys=True)
    main.wait_stream(side)
    for t in far:
        t.record_stream(main)
    return far, near


def _emulate_ladder(lad: _Ladder, peak: int, max_far: int, max_near: int):
    """Outcome of the reference's retry loop from full-walk statistics.

    Returns the (queue, interaction, neighbour) capacities of the successful
    attempt, raises the reference's error, or returns None when the outcome
    depends on the generation at which two overflows happen (then the
    attempt-by-attempt walk decides)."""
    success = None
    for Q in lad.queue:
        q_over = peak > Q
        for K in lad.inter:
            f_over = max_far > K
            if q_over and f_over:
                return None
            if f_over:
                raise RuntimeError(FAR_ERR)
            for Kn in lad.nbr:
                n_over = max_near > Kn
                if q_over or n_over:
                    continue
                success = (Q, K, Kn)
                break
            if success is not None:
                break
        if success is not None:
            break
    if success is None:
        if peak > lad.queue[-1]:
            raise RuntimeError(QUEUE_ERR_USER if lad.user_q else QUEUE_ERR_AUTO)
        raise RuntimeError(NEAR_ERR)
    if lad.cache_key is not None:
        _store_cached_queue_capacity(lad.cache_key, success[0])
    return success


def _run_walk_flat(ctx: _WalkContext, dt: DeviceTopology, lad: _Ladder,
                   sizes: Optional[dict] = None, strict_int32: bool = True,
                   need_max: bool = True, fused_lists: bool = False) -> Optional[_FlatResult]:
    """Full walk with flat emission. `sizes` (queue, far, near capacities)
    are the starting sizes of one resumable emission walk: far and near
    streams that turn out too small grow during the walk, and a pair queue
    that overflows reruns the walk without the hint. Without `sizes` the
    walk follows STREAM_SIZING: one resumable walk from the default
    capacities, or a counting walk and an emission walk at the exact
    totals (also the path after a queue overflow). The largest far and
    near row counts (one host read after the sorts) are taken when
    `need_max`, or when the int32 guard or the retry-ladder emulation needs
    them; otherwise stats hold None for them. With `fused_lists` the row
    sorts write the list outputs themselves (result fields inter, near);
    a callable `fused_lists(em)` builds them from the emission streams
    instead and returns (inter, near, sorted far keys, far row offsets,
    near row offsets)."""
    T, NI, L = dt.num_nodes, dt.num_internal, dt.num_leaves
    qbig = max(T * DEFAULT_PAIR_QUEUE_MULTIPLIER, _auto_pair_queue_candidates(T, NI)[-1])
    from . import autotune
    # the fused list builders expand pair records; the host sort path and
    # the raw result read entries
    ctx.pair_records = bool(fused_lists) and pair_records_ok(T)

    def emission_walk(grid_blocks=0):
        # the walk the plain path runs: resumable, from the default capacities
        f0, n0 = _start_capacities(T, L)
        ctx.run_flat(emit=True, queue_cap=qbig, far_cap=f0, near_cap=n0, grow=True,
                     grid_blocks=grid_blocks)
    ctx.grid_blocks = autotune.walk_grid_blocks(T, lambda b: emission_walk(b), FLAT_GRID_BLOCKS)

    def with_tile(per):
        old, ctx.pairs_per_thread = ctx.pairs_per_thread, per
        try:
            emission_walk()
        finally:
            ctx.pairs_per_thread = old
    ctx.pairs_per_thread = autotune.walk_tile(T, with_tile, ctx.pairs_per_thread)
    em = None
    t_stage = time.perf_counter()
    hinted = False
    c0 = FLAT_STATS["count_walks"]
    if sizes is not None:
        FLAT_STATS["hinted_walks"] += 1
        em = ctx.run_flat(emit=True, queue_cap=sizes["queue"], far_cap=sizes["far"],
                          near_cap=sizes["near"], grow=True)
        if em["queue_overflow"]:
            FLAT_STATS["hint_overflows"] += 1
            em = None
        else:
            hinted = True
            if em["stream_growths"]:
                FLAT_STATS["hint_growths"] += 1
    if em is None:
        FLAT_STATS["plain_walks"] += 1
        if STREAM_SIZING == "grow":
            f0, n0 = _start_capacities(T, L)
            em = ctx.run_flat(emit=True, queue_cap=qbig, far_cap=f0, near_cap=n0, grow=True)
            if em["queue_overflow"]:
                em = None  # the counting walk below reports it
    if em is None:
        FLAT_STATS["count_walks"] += 1
        tot = ctx.run_flat(emit=False, queue_cap=qbig)
        if tot["queue_overflow"]:
            if lad.use_count:
                raise RuntimeError("Count-pass traversal wavefront overflowed its conservative "
                                   "buffer.")
            return None
        if (strict_int32 and lad.use_count
                and max(tot["far_total"], tot["near_total"]) > (1 << 31) - 1):
            raise RuntimeError("Dual-tree traversal pair totals exceed signed int32 capacity; "
                               "refine the tree or use a lower-memory traversal strategy.")
        _traversal_diag(f"count pass far_total={tot['far_total']} "
                        f"near_total={tot['near_total']} max_wf={tot['peak']} "
                        f"generations={len(tot['gen_sizes'])}")
        _record_stage("dual_split_shared_count", t_stage, ctx.dev)
        t_stage = time.perf_counter()
        em = ctx.run_flat(emit=True, queue_cap=tot["peak"], far_cap=tot["far_total"],
                          near_cap=tot["near_total"])
        assert not (em["queue_overflow"] or em["far_overflow"] or em["near_overflow"])
    else:
        # one walk gave the totals and the entries: it is the counting stage
        # of the reference's timing and diagnostics, and the row sorts below
        # are the fill
        _traversal_diag(f"count pass far_total={em['far_total']} "
                        f"near_total={em['near_total']} max_wf={em['peak']} "
                        f"generations={len(em['gen_sizes'])}")
        _record_stage("dual_split_shared_count", t_stage, ctx.dev)
        t_stage = time.perf_counter()
    nf, nn = em["far_total"], em["near_total"]
    LAST_WALK.clear()
    LAST_WALK.update(sizing="hint" if hinted else ("count" if FLAT_STATS["count_walks"] > c0
                                                   else "grow"),
                     far_capacity=em["far_capacity"], near_capacity=em["near_capacity"],
                     queue_capacity=em["queue_capacity"], stream_growths=em["stream_growths"],
                     far_total=nf, near_total=nn, stream_bytes=em["stream_bytes"],
                     stream_bytes_per_entry=4 if em["pair_records"] else 8)
    if strict_int32 and lad.use_count and max(nf, nn) > (1 << 31) - 1:
        raise RuntimeError("Dual-tree traversal pair totals exceed signed int32 capacity; "
                           "refine the tree or use a lower-memory traversal strategy.")
    _traversal_diag(f"compact fill sizing queue={em['queue_capacity']} "
                    f"total_far_pairs={nf} total_near_pairs={nn} hinted={hinted} "
                    f"stream_growths={em.get('stream_growths', 0)}")
    inter = near = None
    if callable(fused_lists):  # the caller's own list layout (batched mode)
        inter, near, fk, far_off, near_off = fused_lists(em)
        fv = nv = None
    elif fused_lists:
        inter, near, fk, far_off, near_off = _fused_lists(em, dt)
        fv = nv = None
    else:
        (fk, fv, far_off), (_, nv, near_off) = _sort_far_and_near(
            em["far_keys"], em["far_vals"], nf, T, em["near_keys"], em["near_vals"], nn, L)
    _record_stage("dual_split_shared_combined_fill", t_stage, ctx.dev)
    max_far = max_near = None
    if need_max or strict_int32 or not lad.use_count:
        mx = torch.stack([torch.diff(far_off).max(), torch.diff(near_off).max()]).tolist()
        max_far, max_near = int(mx[0]), int(mx[1])
    stats = dict(peak=em["peak"], gen_sizes=em["gen_sizes"], decisions=em["decisions"],
                 max_far=max_far, max_near=max_near, far_total=nf, near_total=nn)
    if lad.use_count:
        if strict_int32 and max_near > max(1, ((1 << 31) - 1) // max(1, L)):
            raise RuntimeError(
                "Suggested per-leaf neighbor capacity is too large to safely allocate on the "
                "host (would overflow int32). Enable local refinement or pass explicit "
                "capacities (max_neighbors_per_leaf / max_interactions_per_node).")
    else:
        caps = _emulate_ladder(lad, em["peak"], max_far, max_near)
        if caps is None:
            return None
        stats["queue_capacity"] = caps[0]
        stats["attempt_caps"] = caps
    return _FlatResult(fk, fv, nv, far_off, near_off, stats, inter, near)


def _flat_outputs(fr: _FlatResult, dt: DeviceTopology, topo, neighbor_positions: bool = True):
    lib = _lib.load()
    dev = dt.parent.device
    idx = INDEX_DTYPE
    T, L = dt.num_nodes, dt.num_leaves
    i64 = int(idx == torch.int64)
    st = _lib.stream_ptr(dev)
    if fr.inter is not None:
        neighbors, ncounts, noff = fr.near
        maxc = fr.stats["max_near"] if neighbor_positions else 0
        nlp = torch.empty((L, maxc), dtype=idx, device=dev)
        if maxc > 0:
            _lib.check(lib.ygg_near_dense_from_list(L, fr.near_off.data_ptr(),
                                                    neighbors.data_ptr(),
                                                    dt.leaf_position.data_ptr(), maxc, i64,
                                                    nlp.data_ptr(), st), "near dense")
        leaf_indices, po, p2n = _leaf_tables(dt)
        return fr.inter, _with_target_blocks(NodeNeighborList(
            offsets=noff, neighbors=neighbors, leaf_indices=leaf_indices, counts=ncounts,
            particle_order_leaf_indices=po, particle_order_to_native_leaf=p2n,
            neighbor_leaf_positions=nlp,
            target_block_leaf_ids=torch.zeros(0, dtype=idx, device=dev),
            target_block_source_leaf_ids=torch.zeros((0, 0), dtype=idx, device=dev),
            target_block_valid_mask=torch.zeros((0, 0), dtype=torch.bool, device=dev),
            target_block_offsets=torch.zeros(L + 1, dtype=idx, device=dev),
            target_block_size=0), dt)
    nf = fr.stats["far_total"]
    nn = fr.stats["near_total"]
    sources = torch.empty(nf, dtype=idx, device=dev)
    targets = torch.empty(nf, dtype=idx, device=dev)
    levels = torch.empty(nf, dtype=idx, device=dev)
    counts = torch.empty(T, dtype=idx, device=dev)
    counts32 = torch.empty(T, dtype=torch.int32, device=dev)
    node_off = torch.empty(T + 1, dtype=idx, device=dev)
    _lib.check(lib.ygg_far_lists_from_sorted(
        T, fr.far_keys.data_ptr(), fr.far_vals.data_ptr(), nf, fr.far_off.data_ptr(),
        dt.nodes_by_level.data_ptr(), dt.node_level.data_ptr(), i64, sources.data_ptr(),
        targets.data_ptr(), levels.data_ptr(), counts.data_ptr(), node_off.data_ptr(),
        counts32.data_ptr(), st), "far lists")
    nl = int(topo.num_levels)
    level_offsets = fr.far_off[dt.level_offsets[: nl + 1].long()].to(idx)
    if nf == 0:
        node_off = torch.zeros(T + 1, dtype=idx, device=dev)
        level_offsets = torch.zeros(nl + 1, dtype=idx, device=dev)
    inter = NodeInteractionList(node_off, sources, targets, counts, level_offsets, levels)
    maxc = fr.stats["max_near"] if neighbor_positions else 0
    neighbors = torch.empty(nn, dtype=idx, device=dev)
    nlp = torch.empty((L, maxc), dtype=idx, device=dev)
    ncounts = torch.empty(L, dtype=idx, device=dev)
    noff = torch.empty(L + 1, dtype=idx, device=dev)
    _lib.check(lib.ygg_near_lists_from_sorted(
        L, fr.near_off.data_ptr(), fr.near_vals.data_ptr(), dt.leaf_position.data_ptr(), maxc,
        i64, neighbors.data_ptr(), nlp.data_ptr() if maxc > 0 else None, ncounts.data_ptr(),
        noff.data_ptr(), st), "near lists")
    leaf_indices, po, p2n = _leaf_tables(dt)
    nbrs = NodeNeighborList(
        offsets=noff, neighbors=neighbors, leaf_indices=leaf_indices, counts=ncounts,
        particle_order_leaf_indices=po, particle_order_to_native_leaf=p2n,
        neighbor_leaf_positions=nlp, target_block_leaf_ids=torch.zeros(0, dtype=idx, device=dev),
        target_block_source_leaf_ids=torch.zeros((0, 0), dtype=idx, device=dev),
        target_block_valid_mask=torch.zeros((0, 0), dtype=torch.bool, device=dev),
        target_block_offsets=torch.zeros(L + 1, dtype=idx, device=dev), target_block_size=0)
    return inter, _with_target_blocks(nbrs, dt)


def _count_then_fill(ctx: _WalkContext, dt: DeviceTopology, retry_logger, collect_far,
                     collect_near) -> _RawWalk:
    """Count pass, then an exact-length fill (3989-4299)."""
    T, NI, L = dt.num_nodes, dt.num_internal, dt.num_leaves
    qcount = max(T * DEFAULT_PAIR_QUEUE_MULTIPLIER, _auto_pair_queue_candidates(T, NI)[-1])
    cnt = ctx.run(mode=0, queue_cap=qcount, collect_far=collect_far, collect_near=collect_near)
    if cnt.queue_overflow:
        raise RuntimeError("Count-pass traversal wavefront overflowed its conservative buffer.")
    return fill_from_counts(ctx, dt, cnt.far_counts, cnt.near_counts, cnt.max_wavefront,
                            retry_logger=retry_logger, collect_far=collect_far,
                            collect_near=collect_near)


def fill_from_counts(ctx: _WalkContext, dt: DeviceTopology, far_caps: torch.Tensor,
                     near_caps: torch.Tensor, max_wavefront: int, *, retry_logger=None,
                     collect_far=True, collect_near=True, queue_candidates=None) -> _RawWalk:
    """Exact-length fill at level-major far offsets and native near offsets,
    sized by per-row capacities (exact counts or predicted capacities)."""
    T, NI, L = dt.num_nodes, dt.num_internal, dt.num_leaves
    far_lm = _scan(far_caps, dt.nodes_by_level)
    near_off = _scan(near_caps)
    totals = torch.stack([far_lm[-1], near_off[-1]]).cpu()
    far_total, near_total = int(totals[0]), int(totals[1])
    max_near = int(near_caps.max()) if L else 0
    if max_near > max(1, ((1 << 31) - 1) // max(1, L)):
        raise RuntimeError(
            "Suggested per-leaf neighbor capacity is too large to safely allocate on the host "
            "(would overflow int32). Enable local refinement or pass explicit capacities "
            "(max_neighbors_per_leaf / max_interactions_per_node).")
    if far_total > (1 << 31) - 1 or near_total > (1 << 31) - 1:
        raise RuntimeError("Dual-tree traversal pair totals exceed signed int32 capacity; "
                           "refine the tree or use a lower-memory traversal strategy.")
    far_start = torch.empty(T, dtype=torch.int64, device=ctx.dev)
    _lib.check(ctx.lib.ygg_rows_from_offsets(dt.nodes_by_level.data_ptr(), far_lm.data_ptr(), T,
                                             far_start.data_ptr(), ctx.stream), "rows")
    if queue_candidates is None:
        q_suggest = max(4, max_wavefront + max_wavefront // 2)
        auto = _auto_pair_queue_candidates(T, NI)
        queue_candidates = [c for c in auto if c >= q_suggest]
        if not queue_candidates or queue_candidates[0] != q_suggest:
            queue_candidates = [q_suggest, *queue_candidates]
        queue_candidates = list(dict.fromkeys(queue_candidates))
    k_suggest = max(1, int(far_caps.max())) if T else 1
    result = None
    for attempt, Q in enumerate(queue_candidates, start=1):
        o = ctx.run(mode=2, queue_cap=Q, far_row_start=far_start, far_row_cap=far_caps,
                    near_row_start=near_off, near_row_cap=near_caps,
                    far_capacity_total=far_total, near_capacity_total=near_total,
                    collect_far=collect_far, collect_near=collect_near)
        result = _RawWalk(o, _capped(o.far_counts, 0, far_caps), _capped(o.near_counts, 0, near_caps),
                          far_caps, near_caps)
        if o.queue_overflow:
            _emit(retry_logger, "queue_overflow", attempt, Q, k_suggest, result)
            continue
        if attempt > 1:
            _emit(retry_logger, "success", attempt, Q, k_suggest, result)
        break
    assert result is not None
    if result.outcome.queue_overflow:
        raise RuntimeError(QUEUE_ERR_USER)
    if collect_far and result.outcome.far_overflow:
        raise RuntimeError(FAR_ERR)
    if collect_near and result.outcome.near_overflow:
        raise RuntimeError(NEAR_ERR)
    return result


def _empty_outputs(dt: DeviceTopology, topo):
    dev = dt.parent.device
    idx = INDEX_DTYPE
    T, L = dt.num_nodes, dt.num_leaves
    nl = int(topo.num_levels)
    inter = NodeInteractionList(
        offsets=torch.zeros(T + 1, dtype=idx, device=dev), sources=torch.zeros(0, dtype=idx, device=dev),
        targets=torch.zeros(0, dtype=idx, device=dev), counts=torch.zeros(T, dtype=idx, device=dev),
        level_offsets=torch.zeros(nl + 1, dtype=idx, device=dev),
        target_levels=torch.zeros(0, dtype=idx, device=dev))
    return inter


def _to_interactions(raw: _RawWalk, dt: DeviceTopology, topo) -> NodeInteractionList:
    """_result_to_interactions (3561-3637)."""
    lib = _lib.load()
    dev = dt.parent.device
    idx = INDEX_DTYPE
    T = dt.num_nodes
    counts = raw.far_counts
    lm = _scan(counts, dt.nodes_by_level)
    total = int(lm[-1])
    nl = int(topo.num_levels)
    counts_out = counts.to(idx)
    if total == 0:
        e = _empty_outputs(dt, topo)
        return e._replace(counts=counts_out)
    sources = torch.empty(total, dtype=idx, device=dev)
    targets = torch.empty(total, dtype=idx, device=dev)
    levels = torch.empty(total, dtype=idx, device=dev)
    offsets = torch.empty(T + 1, dtype=idx, device=dev)
    fixed_cap, row_start = raw.outcome.far_layout
    _lib.check(lib.ygg_far_output(T, dt.nodes_by_level.data_ptr(), lm.data_ptr(), counts.data_ptr(),
                                  int(fixed_cap), _lib.ptr(row_start),
                                  raw.outcome.far_entries.data_ptr(), dt.node_level.data_ptr(),
                                  int(idx == torch.int64), sources.data_ptr(), targets.data_ptr(),
                                  levels.data_ptr(), offsets.data_ptr(), _lib.stream_ptr(dev)),
               "far_output")
    level_idx = dt.level_offsets[: nl + 1].long()
    level_offsets = lm[level_idx].to(idx)
    return NodeInteractionList(offsets, sources, targets, counts_out, level_offsets, levels)


def _leaf_tables(dt: DeviceTopology):
    dev = dt.parent.device
    idx = INDEX_DTYPE
    L, NI = dt.num_leaves, dt.num_internal
    po = torch.arange(NI, NI + L, dtype=idx, device=dev)
    p2n = torch.empty(L, dtype=idx, device=dev)
    native = dt.leaf_indices.long() - NI
    p2n[native] = torch.arange(L, dtype=idx, device=dev)
    return dt.leaf_indices.to(idx), po, p2n


def _target_block_size() -> int:
    try:
        return max(0, int(os.environ.get("YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE", "0")))
    except Exception:
        return 0


def _with_target_blocks(nbrs: NodeNeighborList, dt: DeviceTopology) -> NodeNeighborList:
    """Fill the target-block fields of a near list (_build_target_blocks,
    _interactions_impl.py:3689-3770) when YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE
    is positive: each native leaf row is cut into ceil(count / k) blocks of
    k consecutive neighbours, each slot holding the neighbour's native leaf
    row (0 and invalid past the row's end)."""
    k = _target_block_size()
    if k <= 0:
        return nbrs
    lib = _lib.load()
    dev = dt.parent.device
    idx = INDEX_DTYPE
    L = dt.num_leaves
    if L == 0 or int(nbrs.neighbors.numel()) == 0:
        # no neighbour at all: the reference's empty blocks
        return nbrs._replace(
            target_block_leaf_ids=torch.zeros(0, dtype=idx, device=dev),
            target_block_source_leaf_ids=torch.zeros((0, k if L else 0), dtype=idx, device=dev),
            target_block_valid_mask=torch.zeros((0, k if L else 0), dtype=torch.bool, device=dev),
            target_block_offsets=torch.zeros(L + 1, dtype=idx, device=dev), target_block_size=k)
    st = _lib.stream_ptr(dev)
    off = nbrs.offsets.to(torch.int64).contiguous()
    boff = torch.empty(L + 1, dtype=torch.int64, device=dev)
    ws = _lib.workspace(lib.ygg_target_block_offsets_workspace_bytes(L), dev)
    _lib.check(lib.ygg_target_block_offsets(L, off.data_ptr(), k, boff.data_ptr(), ws.data_ptr(),
                                            ws.numel(), st), "target block offsets")
    B = int(boff[-1])
    rows = torch.empty(B, dtype=idx, device=dev)
    src = torch.empty((B, k), dtype=idx, device=dev)
    valid = torch.empty((B, k), dtype=torch.uint8, device=dev)
    nb = nbrs.neighbors.contiguous()
    _lib.check(lib.ygg_target_blocks(L, off.data_ptr(), nb.data_ptr(),
                                     int(nb.dtype == torch.int64), dt.leaf_position.data_ptr(), k,
                                     boff.data_ptr(), B, int(idx == torch.int64),
                                     rows.data_ptr(), src.data_ptr(), valid.data_ptr(), st),
               "target blocks")
    return nbrs._replace(target_block_leaf_ids=rows, target_block_source_leaf_ids=src,
                         target_block_valid_mask=valid.bool(),
                         target_block_offsets=boff.to(idx), target_block_size=k)


def _to_neighbors(raw: Optional[_RawWalk], dt: DeviceTopology) -> NodeNeighborList:
    """_result_to_neighbors (3640-3837)."""
    lib = _lib.load()
    dev = dt.parent.device
    idx = INDEX_DTYPE
    L = dt.num_leaves
    leaf_indices, po, p2n = _leaf_tables(dt)
    if raw is None:
        counts = torch.zeros(L, dtype=torch.int32, device=dev)
    else:
        counts = raw.near_counts
    off = _scan(counts)
    stats = torch.stack([off[-1], counts.max().to(torch.int64)]).cpu() if L else torch.zeros(2)
    total, maxc = int(stats[0]), int(stats[1])
    neighbors = torch.empty(total, dtype=idx, device=dev)
    nlp = torch.empty((L, maxc), dtype=idx, device=dev)
    if total > 0:
        fixed_cap, row_start = raw.outcome.near_layout
        _lib.check(lib.ygg_near_output(L, off.data_ptr(), counts.data_ptr(), int(fixed_cap),
                                       _lib.ptr(row_start), raw.outcome.near_entries.data_ptr(),
                                       dt.leaf_position.data_ptr(), maxc, int(idx == torch.int64),
                                       neighbors.data_ptr(), nlp.data_ptr(), _lib.stream_ptr(dev)),
                   "near_output")
    return _with_target_blocks(NodeNeighborList(
        offsets=off.to(idx), neighbors=neighbors, leaf_indices=leaf_indices,
        counts=counts.to(idx), particle_order_leaf_indices=po,
        particle_order_to_native_leaf=p2n, neighbor_leaf_positions=nlp,
        target_block_leaf_ids=torch.zeros(0, dtype=idx, device=dev),
        target_block_source_leaf_ids=torch.zeros((0, 0), dtype=idx, device=dev),
        target_block_valid_mask=torch.zeros((0, 0), dtype=torch.bool, device=dev),
        target_block_offsets=torch.zeros(L + 1, dtype=idx, device=dev), target_block_size=0), dt)


def _node_order_rows(values_lm: torch.Tensor, inter: NodeInteractionList,
                     node_off: torch.Tensor) -> torch.Tensor:
    """Rows of a level-major far array rearranged into node order."""
    F = values_lm.numel()
    if F == 0:
        return values_lm.clone()
    tgt = inter.targets.long()
    rank = torch.arange(F, dtype=torch.int64, device=values_lm.device)
    pos = node_off[tgt].long() + (rank - inter.offsets[tgt].long())
    out = torch.empty_like(values_lm)
    out[pos] = values_lm
    return out


def _padded(values: torch.Tensor, length: int) -> torch.Tensor:
    out = torch.full((max(int(length), 1),), -1, dtype=values.dtype, device=values.device)
    out[:values.numel()] = values
    return out


def _to_result(raw, dt: DeviceTopology, inter: NodeInteractionList, nbrs: NodeNeighborList,
               tags: Optional[torch.Tensor] = None) -> DualTreeWalkResult:
    """The reference's raw DualTreeWalkResult (_raw_to_result, 4606-4627).

    Fixed-capacity path (the caller gave capacities): interaction_offsets is
    the node-order exclusive cumsum of the counts while sources, targets and
    tags are level-major (quirk B9, 1787-1847), padded with -1 to
    T * max_interactions_per_node; neighbour indices in native leaf-row
    order padded with -1 to L * max_neighbors_per_leaf (1206-1236). Count
    path: node-order offsets with node-order rows of exact length
    (_dual_tree_walk_compact_fill_impl, 3089-3111, 3540-3558). ``tags`` are
    level-major far tags (-1 without them)."""
    dev = dt.parent.device
    idx = INDEX_DTYPE
    T, L = dt.num_nodes, dt.num_leaves
    counts = inter.counts
    node_off = torch.zeros(T + 1, dtype=idx, device=dev)
    node_off[1:] = torch.cumsum(counts, 0)
    caps = None
    if isinstance(raw, _FlatResult):
        dec = raw.stats["decisions"]
        caps = raw.stats.get("attempt_caps")
    elif raw is not None:
        dec = raw.outcome.decisions
        if raw.outcome.far_layout[0] > 0:
            caps = (raw.outcome.queue_capacity, raw.outcome.far_layout[0],
                    raw.outcome.near_layout[0])
    else:
        dec = (0, 0, 0)
    F = int(inter.sources.numel())
    if tags is None:
        tags = torch.full((F,), -1, dtype=idx, device=dev)
    tags = tags.to(idx)
    if caps is not None:
        _, K, Kn = caps
        src, tgt = _padded(inter.sources, T * K), _padded(inter.targets, T * K)
        tags = _padded(tags, T * K)
        nb = _padded(nbrs.neighbors, L * Kn)
    elif F:
        src = _node_order_rows(inter.sources, inter, node_off)
        tgt = _node_order_rows(inter.targets, inter, node_off)
        tags = _node_order_rows(tags, inter, node_off)
        nb = nbrs.neighbors
    else:
        src, tgt, nb = inter.sources, inter.targets, nbrs.neighbors
    return DualTreeWalkResult(
        interaction_offsets=node_off, interaction_sources=src, interaction_targets=tgt,
        interaction_tags=tags, interaction_counts=counts, neighbor_offsets=nbrs.offsets,
        neighbor_indices=nb, neighbor_counts=nbrs.counts, leaf_indices=nbrs.leaf_indices,
        far_pair_count=F, near_pair_count=int(nbrs.neighbors.numel()),
        queue_overflow=False, far_overflow=False, near_overflow=False,
        accept_decisions=int(dec[0]), near_decisions=int(dec[1]), refine_decisions=int(dec[2]))


def build_interactions_and_neighbors(
    tree, geometry: TreeGeometry, theta: float = 0.5,
    max_interactions_per_node: Optional[int] = None,
    max_neighbors_per_leaf: int = _DEFAULT_MAX_NEIGHBORS,
    max_pair_queue: Optional[int] = None, process_block: Optional[int] = None,
    traversal_config: Optional[DualTreeTraversalConfig] = None,
    retry_logger: Optional[Callable[[DualTreeRetryEvent], object]] = None,
    mac_type: str = "bh", dehnen_radius_scale: float = 1.0, pair_policy=None,
    policy_state: object = None, *, return_result: bool = False,
    return_compact_far_pairs: bool = False, return_interactions: bool = True,
    return_grouped: bool = False, capacity_hint: Optional[dict] = None,
    neighbor_positions: bool = True, strict_int32: bool = True):
    """Far-field interaction lists and near-field neighbour lists from one
    dual-tree walk (yggdrax/interactions.py:338-448).

    ``capacity_hint`` ({"queue", "far", "near"} stream capacities, e.g. from
    the learned capacity model) lets the walk size its buffers in one pass;
    an overflow falls back to counting first, so outputs never depend on it.

    ``neighbor_positions=False`` leaves NodeNeighborList.neighbor_leaf_positions
    as an [L, 0] array instead of the dense -1 padded [L, max_count] matrix,
    and ``strict_int32=False`` drops the reference's host-allocation guards
    (per-leaf neighbour count <= (2^31 - 1) / L, totals <= 2^31 - 1) of the
    count-pass path, which exist for its dense (L x capacity) buffers; the
    CSR lists need neither. Both default to the reference's behaviour.

    ``return_grouped`` appends the displacement classes of the far pairs
    (grouped_interactions.build_grouped_interactions_from_pairs over the
    level-major far list, _interactions_impl.py:6109-6124) as the last
    element of the returned tuple."""
    if return_grouped:
        if not return_interactions:
            raise ValueError("grouped interaction return requires return_interactions=True")
        from .grouped_interactions import build_grouped_interactions_from_pairs
        out = build_interactions_and_neighbors(
            tree, geometry, theta, max_interactions_per_node, max_neighbors_per_leaf,
            max_pair_queue, process_block, traversal_config, retry_logger, mac_type,
            dehnen_radius_scale, pair_policy, policy_state, return_result=return_result,
            return_compact_far_pairs=return_compact_far_pairs, return_interactions=True,
            capacity_hint=capacity_hint, neighbor_positions=neighbor_positions,
            strict_int32=strict_int32)
        inter = out[0]
        grouped = build_grouped_interactions_from_pairs(
            tree, geometry, inter.sources, inter.targets, level_offsets=inter.level_offsets)
        return (*out, grouped)
    if return_compact_far_pairs:
        # the compact pairs come from the result's level-major far arrays
        inter, nbrs, res = build_interactions_and_neighbors(
            tree, geometry, theta, max_interactions_per_node, max_neighbors_per_leaf,
            max_pair_queue, process_block, traversal_config, retry_logger, mac_type,
            dehnen_radius_scale, pair_policy, policy_state, return_result=True,
            return_interactions=True, capacity_hint=capacity_hint,
            neighbor_positions=neighbor_positions, strict_int32=strict_int32)
        compact = _compact_far_pairs(inter, res)
        out_inter = inter if return_interactions else None
        if return_result:
            return out_inter, nbrs, res, compact
        return out_inter, nbrs, compact
    if pair_policy is not None:
        from .policy import build_with_pair_policy  # noqa: E402
        return build_with_pair_policy(
            tree, geometry, theta, mac_type, dehnen_radius_scale, pair_policy, policy_state,
            max_interactions_per_node=max_interactions_per_node,
            max_neighbors_per_leaf=max_neighbors_per_leaf, max_pair_queue=max_pair_queue,
            traversal_config=traversal_config, return_result=return_result)
    dt = device_topology(tree)
    topo = resolve_tree_topology(tree)
    if float(dehnen_radius_scale) <= 0.0:
        raise ValueError("dehnen_radius_scale must be > 0")
    kw = dict(max_interactions_per_node=max_interactions_per_node,
              max_neighbors_per_leaf=max_neighbors_per_leaf, max_pair_queue=max_pair_queue,
              traversal_config=traversal_config)
    use_count = (max_interactions_per_node is None and max_pair_queue is None
                 and traversal_config is None)
    _traversal_diag(
        f"walk start total_nodes={dt.num_nodes} num_internal={dt.num_internal} "
        f"use_count_pass={use_count} "
        f"max_interactions_per_node={max_interactions_per_node} "
        f"max_neighbors_per_leaf={int(max_neighbors_per_leaf)} "
        f"max_pair_queue={max_pair_queue} traversal_config={traversal_config} "
        f"process_block={process_block} mac_type={mac_type}")
    if dt.num_internal == 0:
        inter, nbrs, raw = _empty_outputs(dt, topo), _to_neighbors(None, dt), None
    else:
        ctx = _WalkContext(dt, geometry, theta, mac_type, dehnen_radius_scale)
        fr = None
        if retry_logger is None and not EXACT_ATTEMPTS:
            fr = _run_walk_flat(ctx, dt, _ladder(dt, mac_type, **kw), sizes=capacity_hint,
                                strict_int32=strict_int32, need_max=neighbor_positions,
                                fused_lists=True)
        if fr is not None:
            inter, nbrs = _flat_outputs(fr, dt, topo, neighbor_positions)
            raw = fr
        else:
            raw = _run_walk_raw(tree, geometry, theta, mac_type,
                                dehnen_radius_scale=dehnen_radius_scale,
                                process_block=process_block, retry_logger=retry_logger, ctx=ctx,
                                **kw)
            inter = _to_interactions(raw, dt, topo)
            nbrs = _to_neighbors(raw, dt)
    out_inter = inter if return_interactions else None
    if return_result:
        return out_inter, nbrs, _to_result(raw, dt, inter, nbrs)
    return out_inter, nbrs


def _compact_far_pairs(inter: NodeInteractionList, res) -> CompactTaggedFarPairs:
    """Compact pairs in the level-major order of ``inter``. The result's tags
    are level-major on the fixed-capacity path and node-order on the count
    path; where the two row orders differ, each level-major entry reads its
    tag at its node-order position."""
    n = int(inter.sources.numel())
    dev = inter.sources.device
    if res is None:
        tags = torch.full((n,), -1, dtype=INDEX_DTYPE, device=dev)
    elif n == 0 or torch.equal(res.interaction_targets[:n], inter.targets):
        tags = res.interaction_tags[:n]
    else:
        tgt = inter.targets.long()
        rank = torch.arange(n, dtype=torch.int64, device=dev)
        pos = res.interaction_offsets[tgt].long() + (rank - inter.offsets[tgt].long())
        tags = res.interaction_tags[pos]
    return CompactTaggedFarPairs(inter.sources, inter.targets, tags.to(INDEX_DTYPE),
                                 torch.tensor(n, dtype=INDEX_DTYPE,
                                              device=inter.sources.device))


def build_compact_far_pairs_and_leaf_neighbor_lists(
        tree, geometry: TreeGeometry, theta: float = 0.5,
        max_neighbors_per_leaf: int = _DEFAULT_MAX_NEIGHBORS,
        max_interactions_per_node: Optional[int] = None, mac_type: str = "bh", *,
        pair_policy=None, policy_state: object = None, max_pair_queue: Optional[int] = None,
        process_block: Optional[int] = None,
        traversal_config: Optional[DualTreeTraversalConfig] = None,
        retry_logger: Optional[Callable[[DualTreeRetryEvent], object]] = None,
        dehnen_radius_scale: float = 1.0, timing_callback=None,
        compact_far_pair_capacity: Optional[int] = None, capacity_hint: Optional[dict] = None):
    """Compact far pairs and leaf neighbour lists from one walk
    (_interactions_impl.py:5819-5915). With ``compact_far_pair_capacity`` the
    arrays are padded with -1 to that length and a larger count raises the
    reference's error (_interactions_impl.py:5346-5372). ``timing_callback``
    receives (stage, seconds) for the counting walk
    ("dual_split_shared_count") and the emission walk with its row sorts
    ("dual_split_shared_combined_fill"), the stages of the reference's shared
    count pass (5100-5102, 5131-5296)."""
    token = _STAGE_CALLBACK.set(timing_callback)
    try:
        _, nbrs, compact = build_interactions_and_neighbors(
            tree, geometry, theta, max_interactions_per_node, max_neighbors_per_leaf,
            max_pair_queue, process_block, traversal_config, retry_logger, mac_type,
            dehnen_radius_scale, pair_policy, policy_state, return_compact_far_pairs=True,
            return_interactions=False, capacity_hint=capacity_hint)
    finally:
        _STAGE_CALLBACK.reset(token)
    if compact_far_pair_capacity is not None:
        cap = int(compact_far_pair_capacity)
        n = int(compact.far_pair_count)
        if n > cap:
            raise RuntimeError("strict fused compact far-pair cap exceeded: "
                               f"cap={cap}. Increase "
                               "JACCPOT_STATIC_STRICT_FUSED_COMPACT_FAR_PAIR_CAP.")

        def pad(x):
            out = torch.full((cap,), -1, dtype=x.dtype, device=x.device)
            out[:n] = x[:n]
            return out
        compact = CompactTaggedFarPairs(pad(compact.sources), pad(compact.targets),
                                        pad(compact.tags), compact.far_pair_count)
    return compact, nbrs


def build_compact_far_pairs(tree, geometry: TreeGeometry, theta: float = 0.5,
                            max_interactions_per_node: Optional[int] = None,
                            mac_type: str = "bh", *, pair_policy=None,
                            policy_state: object = None, max_pair_queue: Optional[int] = None,
                            process_block: Optional[int] = None,
                            traversal_config: Optional[DualTreeTraversalConfig] = None,
                            retry_logger=None, dehnen_radius_scale: float = 1.0,
                            capacity_hint: Optional[dict] = None) -> CompactTaggedFarPairs:
    """Exact-length tagged far pairs (_interactions_impl.py:5765-5816)."""
    far, _ = build_compact_far_pairs_and_leaf_neighbor_lists(
        tree, geometry, theta, _DEFAULT_MAX_NEIGHBORS, max_interactions_per_node, mac_type,
        pair_policy=pair_policy, policy_state=policy_state, max_pair_queue=max_pair_queue,
        process_block=process_block, traversal_config=traversal_config,
        retry_logger=retry_logger, dehnen_radius_scale=dehnen_radius_scale,
        capacity_hint=capacity_hint)
    return far


def build_leaf_neighbor_lists(tree, geometry: TreeGeometry, theta: float = 0.5,
                              max_neighbors_per_leaf: int = _DEFAULT_MAX_NEIGHBORS,
                              max_interactions_per_node: Optional[int] = None,
                              mac_type: str = "bh", *, pair_policy=None,
                              policy_state: object = None, max_pair_queue: Optional[int] = None,
                              process_block: Optional[int] = None,
                              traversal_config: Optional[DualTreeTraversalConfig] = None,
                              retry_logger=None, dehnen_radius_scale: float = 1.0,
                              capacity_hint: Optional[dict] = None) -> NodeNeighborList:
    """Near-field adjacency of the leaves (_interactions_impl.py:5917-5968)."""
    _, nbrs = build_interactions_and_neighbors(
        tree, geometry, theta, max_interactions_per_node, max_neighbors_per_leaf,
        max_pair_queue, process_block, traversal_config, retry_logger, mac_type,
        dehnen_radius_scale, pair_policy, policy_state, capacity_hint=capacity_hint)
    return nbrs


def interactions_for_node(data: NodeInteractionList, node: int) -> torch.Tensor:
    """Source nodes of ``node``'s far list (_interactions_impl.py:6147-6154)."""
    start = int(data.offsets[node])
    return data.sources[start:start + int(data.counts[node])]


def neighbors_for_leaf(data: NodeNeighborList, leaf_node: int) -> torch.Tensor:
    """Neighbour leaf nodes of ``leaf_node``; ValueError when it is not a
    leaf of the list (_interactions_impl.py:6157-6170)."""
    m = (data.leaf_indices == int(leaf_node)).nonzero()
    if m.numel() == 0:
        raise ValueError("leaf_node not present in neighbor list")
    pos = int(m[0, 0])
    start = int(data.offsets[pos])
    return data.neighbors[start:start + int(data.counts[pos])]


__all__ = [
    "DEFAULT_PAIR_QUEUE_MULTIPLIER", "DualTreeRetryEvent", "DualTreeTraversalConfig",
    "DualTreeWalkResult", "NodeInteractionList", "NodeNeighborList",
    "build_interactions_and_neighbors", "interactions_for_node", "neighbors_for_leaf",
    "set_default_dual_tree_config",
]
