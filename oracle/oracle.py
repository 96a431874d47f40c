"""Python driver of the CPU oracle (test infrastructure only).

Wraps ``libygg_oracle.so`` (a C++ restatement of the reference algorithm)
and restates, in plain Python/numpy, the reference's host-side pieces:
topology assembly (yggdrax/tree.py:375-423, 1482-1549), MAC extents
(yggdrax/_interactions_impl.py:993-1061), leaf ordering (1071-1103), the
capacity/retry driver (3840-4440) and output shaping (3561-3837), and
the downstream far-field layouts: level-major geometry
(yggdrax/_geometry_impl.py:342-433), dense M2L buffers
(yggdrax/dense_interactions.py:35-187) and displacement classes
(yggdrax/grouped_interactions.py:74-207), and per-node mass moments
(yggdrax/tree_moments.py:686-746). Outputs are dicts of numpy arrays whose
keys are the reference's NamedTuple field names; indices are int64 as under
the reference's default INDEX_DTYPE.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libygg_oracle.so")


def _load():
    if not os.path.exists(_LIB_PATH):
        import subprocess

        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return ctypes.CDLL(_LIB_PATH)


_P = ctypes.c_void_p
_I = ctypes.c_int64


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def configure(handle):
    """Declare the ctypes signatures this driver relies on for a loaded
    oracle library (the default one, or another build of the same source)."""
    handle.ygo_walk_run.restype = ctypes.c_void_p
    handle.ygo_walk_run.argtypes = [_I, _I, _P, _P, _P, _P, _P, _I, ctypes.c_float,
                                    ctypes.c_int, _I, _I, _I, ctypes.c_int, ctypes.c_int]
    handle.ygo_walk_sizes.argtypes = [_P, _P]
    handle.ygo_walk_copy.argtypes = [_P, _P, _P, _P, _P, _P]
    handle.ygo_walk_free.argtypes = [_P]
    handle.ygo_geometry.restype = ctypes.c_int
    return handle


lib = configure(_load())

MAC_CODES = {"bh": 0, "dehnen": 1, "engblom": 2}


def infer_bounds(positions):
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib.ygo_infer_bounds(_ptr(pos), _I(pos.shape[0]), _ptr(lo), _ptr(hi))
    return lo, hi


def morton_encode(positions, bounds):
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    lo = np.ascontiguousarray(bounds[0], dtype=np.float32)
    hi = np.ascontiguousarray(bounds[1], dtype=np.float32)
    codes = np.zeros(pos.shape[0], np.uint64)
    lib.ygo_morton_encode(_ptr(pos), _I(pos.shape[0]), _ptr(lo), _ptr(hi), _ptr(codes))
    return codes


def stable_argsort_u64(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.zeros(keys.shape[0], np.int64)
    lib.ygo_stable_argsort_u64(_ptr(keys), _I(keys.shape[0]), _ptr(perm))
    return perm


def _radix_from_sorted(sorted_codes, leaf_size):
    n = sorted_codes.shape[0]
    L = (n + leaf_size - 1) // leaf_size
    T = 2 * L - 1
    NI = L - 1
    out = dict(
        parent=np.zeros(T, np.int64),
        left_child=np.zeros(NI, np.int64),
        right_child=np.zeros(NI, np.int64),
        left_is_leaf=np.zeros(NI, np.uint8),
        right_is_leaf=np.zeros(NI, np.uint8),
        node_ranges=np.zeros((T, 2), np.int64),
        node_level=np.zeros(T, np.int64),
        level_offsets=np.zeros(65, np.int64),
        nodes_by_level=np.zeros(T, np.int64),
        leaf_codes=np.zeros(L, np.uint64),
    )
    scal = np.zeros(4, np.int64)
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
    lib.ygo_radix_build(
        _ptr(sc), _I(n), _I(leaf_size), _ptr(out["parent"]), _ptr(out["left_child"]),
        _ptr(out["right_child"]), _ptr(out["left_is_leaf"]), _ptr(out["right_is_leaf"]),
        _ptr(out["node_ranges"]), _ptr(out["node_level"]), _ptr(out["level_offsets"]),
        _ptr(out["nodes_by_level"]), _ptr(scal), _ptr(out["leaf_codes"]))
    out["left_is_leaf"] = out["left_is_leaf"].astype(bool)
    out["right_is_leaf"] = out["right_is_leaf"].astype(bool)
    out["num_levels"] = int(scal[0])
    out["num_internal_nodes"] = NI
    out["num_particles"] = n
    out["leaf_depths"] = np.full(L, -1, np.int64)
    out["use_morton_geometry"] = False
    out["leaf_size"] = int(leaf_size)
    return out


def build_tree(positions, masses=None, bounds=None, *, leaf_size=8,
               lexsort=False):
    """tree.py:1413-1478 -> _tree_impl.py:355-420 (radix adaptive build)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    n = pos.shape[0]
    if n < 1:
        raise ValueError("Need at least one particle")
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    b = infer_bounds(pos) if bounds is None else (
        np.asarray(bounds[0], np.float32), np.asarray(bounds[1], np.float32))
    codes = morton_encode(pos, b)
    perm = stable_argsort_u64(codes)
    sorted_codes = codes[perm]
    topo = _radix_from_sorted(sorted_codes, leaf_size)
    topo["particle_indices"] = perm
    topo["morton_codes"] = sorted_codes
    topo["bounds_min"] = b[0].copy()
    topo["bounds_max"] = b[1].copy()
    topo["positions_sorted"] = pos[perm]
    if masses is not None:
        topo["masses_sorted"] = np.asarray(masses, np.float32)[perm]
    inv = np.zeros(n, np.int64)
    inv[perm] = np.arange(n)
    topo["inverse_permutation"] = inv
    return topo


def octree_metadata(topo):
    """octree.py:479-574 on a radix topology with adaptive leaves."""
    T = topo["parent"].shape[0]
    L = (T + 1) // 2
    M = max(1, 2 * L - 1)
    out = dict(
        oct_parent=np.zeros(M, np.int64),
        oct_children=np.zeros((M, 8), np.int64),
        oct_child_counts=np.zeros(M, np.int64),
        oct_child_mask=np.zeros((M, 8), np.uint8),
        oct_valid_mask=np.zeros(M, np.uint8),
        oct_node_codes=np.zeros(M, np.uint64),
        oct_node_depths=np.zeros(M, np.int64),
        oct_node_ranges=np.zeros((M, 2), np.int64),
        oct_nodes_by_level=np.zeros(M, np.int64),
        oct_level_offsets=np.zeros(23, np.int64),
        oct_leaf_mask=np.zeros(M, np.uint8),
        oct_leaf_nodes=np.zeros(M, np.int64),
        radix_node_to_oct=np.zeros(T, np.int64),
        radix_leaf_to_oct=np.zeros(L, np.int64),
    )
    resolved_codes = np.zeros(L, np.uint64)
    resolved_depths = np.zeros(L, np.int64)
    scal = np.zeros(4, np.int64)
    sc = np.ascontiguousarray(topo["morton_codes"], np.uint64)
    nr = np.ascontiguousarray(topo["node_ranges"], np.int64)
    lib.ygo_octree_meta(
        _ptr(sc), _ptr(nr), _I(L), _ptr(out["oct_parent"]), _ptr(out["oct_children"]),
        _ptr(out["oct_child_counts"]), _ptr(out["oct_child_mask"]), _ptr(out["oct_valid_mask"]),
        _ptr(out["oct_node_codes"]), _ptr(out["oct_node_depths"]), _ptr(out["oct_node_ranges"]),
        _ptr(out["oct_nodes_by_level"]), _ptr(out["oct_level_offsets"]), _ptr(scal),
        _ptr(out["oct_leaf_mask"]), _ptr(out["oct_leaf_nodes"]), _ptr(out["radix_node_to_oct"]),
        _ptr(out["radix_leaf_to_oct"]), _ptr(resolved_codes), _ptr(resolved_depths))
    for k in ("oct_child_mask", "oct_valid_mask", "oct_leaf_mask"):
        out[k] = out[k].astype(bool)
    out["oct_num_levels"] = int(scal[0])
    out["_resolved_codes"] = resolved_codes
    out["_resolved_depths"] = resolved_depths
    return out


def build_octree(positions, masses=None, bounds=None, *, leaf_size=8):
    """tree.py:1482-1549 -> _build_octree_result (375-423) + octree.py:553-574.

    The octree path sorts with lexsort((orig_idx, codes)), which is the same
    stable order as the radix path's stable argsort.
    """
    topo = build_tree(positions, masses, bounds, leaf_size=leaf_size)
    topo.update(octree_metadata(topo))
    return topo


def compute_tree_geometry(topo, positions_sorted, max_leaf_size=None, particle_radius=None):
    """geometry.py:18-75 -> _geometry_impl.py:176-338; particle_radius (sorted
    order) bounds each particle as [p - r, p + r] (107-112)."""
    pos = np.ascontiguousarray(positions_sorted, dtype=np.float32)
    prad = (None if particle_radius is None
            else np.ascontiguousarray(particle_radius, dtype=np.float32))
    if prad is not None and prad.shape != (pos.shape[0],):
        raise ValueError("particle_radius must have shape (n_particles,)")
    T = topo["parent"].shape[0]
    cap = max_leaf_size if max_leaf_size is not None else topo.get("leaf_size")
    if cap is None:
        cap = int((topo["node_ranges"][:, 1] - topo["node_ranges"][:, 0]).max()) + 1
    center = np.zeros((T, 3), np.float32)
    half = np.zeros((T, 3), np.float32)
    radius = np.zeros(T, np.float32)
    maxext = np.zeros(T, np.float32)
    NI = (T - 1) // 2
    left = np.ascontiguousarray(topo["left_child"], np.int64) if NI else np.zeros(1, np.int64)
    right = np.ascontiguousarray(topo["right_child"], np.int64) if NI else np.zeros(1, np.int64)
    st = lib.ygo_geometry(_ptr(pos), None if prad is None else _ptr(prad),
                          _ptr(np.ascontiguousarray(topo["node_ranges"], np.int64)),
                          _ptr(left), _ptr(right),
                          _ptr(np.ascontiguousarray(topo["parent"], np.int64)), _I(T), _I(int(cap)),
                          _ptr(center), _ptr(half), _ptr(radius), _ptr(maxext))
    if st != 0:
        raise RuntimeError("geometry depth pointer doubling did not converge")
    return dict(center=center, half_extent=half, radius=radius, max_extent=maxext)


def geometry_to_level_major(topo, geom):
    """_geometry_impl.py:342-433: padded [num_levels, max_nodes] level-major
    copies; slot j of level l is nodes_by_level[level_offsets[l] + j] while
    j < count(l), else zeros and node index -1."""
    nl = int(topo["num_levels"])
    off = np.asarray(topo["level_offsets"], np.int64)[: nl + 1]
    counts = off[1:] - off[:-1]
    M = int(np.maximum(counts, 1).max())
    nbl = np.asarray(topo["nodes_by_level"], np.int64)
    out = dict(centers=np.zeros((nl, M, 3), np.float32),
               half_extents=np.zeros((nl, M, 3), np.float32),
               radii=np.zeros((nl, M), np.float32), max_extents=np.zeros((nl, M), np.float32),
               level_counts=counts, node_indices=np.full((nl, M), -1, np.int64))
    for level in range(nl):
        for j in range(int(counts[level])):
            node = int(nbl[off[level] + j])
            out["centers"][level, j] = geom["center"][node]
            out["half_extents"][level, j] = geom["half_extent"][node]
            out["radii"][level, j] = geom["radius"][node]
            out["max_extents"][level, j] = geom["max_extent"][node]
            out["node_indices"][level, j] = node
    return out


def densify_interactions(topo, geom, inter):
    """dense_interactions.py:143-187 with _dense_m2l_buffers (35-140): for
    each level-major slot holding node t, the first counts[t] entries are
    t's far sources in CSR row order, mask True, displacement
    center[t] - center[s] in float32; every other entry is -1 / False / 0.
    The entry axis has max(1, max count) entries."""
    lm = geometry_to_level_major(topo, geom)
    nodes = lm["node_indices"]
    nl, M = nodes.shape
    counts = np.asarray(inter["counts"], np.int64)
    offsets = np.asarray(inter["offsets"], np.int64)
    sources = np.asarray(inter["sources"], np.int64)
    K = max(1, int(counts.max()) if counts.size else 0)
    center = np.asarray(geom["center"], np.float32)
    out = dict(geometry=lm, m2l_sources=np.full((nl, M, K), -1, np.int64),
               m2l_mask=np.zeros((nl, M, K), bool), m2l_counts=np.zeros((nl, M), np.int64),
               m2l_displacements=np.zeros((nl, M, K, 3), np.float32))
    for level in range(nl):
        for j in range(M):
            t = int(nodes[level, j])
            if t < 0:
                continue
            c = int(counts[t])
            row = sources[offsets[t]:offsets[t] + c]
            out["m2l_sources"][level, j, :c] = row
            out["m2l_mask"][level, j, :c] = True
            out["m2l_counts"][level, j] = c
            out["m2l_displacements"][level, j, :c] = lm["centers"][level, j] - center[row]
    return out


def compute_tree_mass_moments(topo, positions_sorted, masses_sorted):
    """compute_tree_mass_moments (tree_moments.py:686-746): prefix sums of
    the masses and of the per-particle terms m * x (rounded to the product
    dtype, as masses[:, None] * positions is), differenced at every node's
    range, then center = moment / mass in the product dtype, 0 where the
    mass is 0. The prefix sums run in 80-bit extended precision instead of
    the reference's float dtype, so they approximate the exact node sums
    (error <= n * 2^-64 * prefix magnitude) rather than reproducing the
    float cumsum's rounding."""
    pos = np.asarray(positions_sorted)
    m = np.asarray(masses_sorted)
    w = np.result_type(pos.dtype, m.dtype)
    terms = m.astype(w)[:, None] * pos.astype(w)
    ld = np.longdouble
    mp = np.concatenate([np.zeros(1, ld), np.cumsum(m.astype(ld))])
    wp = np.concatenate([np.zeros((1, 3), ld), np.cumsum(terms.astype(ld), axis=0)])
    r = np.asarray(topo["node_ranges"], np.int64)
    s, e = r[:, 0], r[:, 1] + 1
    mass = (mp[e] - mp[s]).astype(m.dtype)
    msum = (wp[e] - wp[s]).astype(w)
    safe = np.where(mass == 0, 1, mass).astype(w)
    with np.errstate(invalid="ignore", divide="ignore"):
        center = msum / safe[:, None]
    center = np.where(mass[:, None] == 0, 0, center).astype(w)
    return dict(mass=mass, center_of_mass=center)


def _cell_size(extent, level):
    """grouped_interactions.py:34-38: max(extent / max(2^level, 1), 1e-12)
    per axis, float64."""
    denom = np.power(2.0, np.asarray(level, np.float64))[:, None]
    return np.maximum(np.asarray(extent, np.float64)[None, :] / np.maximum(denom, 1.0), 1e-12)


def to_int32_like_numpy_cast(r):
    """NumPy's float64 -> int32 cast as it behaves on x86-64: integral
    values inside the int32 range convert, anything else (out of range,
    NaN) becomes INT32_MIN. Stated explicitly because the C cast underneath
    is undefined for those values."""
    r = np.asarray(r, np.float64)
    ok = (r >= -2147483648.0) & (r <= 2147483647.0)
    return np.where(ok, r, -2147483648.0).astype(np.int32)


def grouped_interactions_from_pairs(topo, geom, sources, targets, level_offsets=None):
    """build_grouped_interactions_from_pairs (grouped_interactions.py:74-207):
    class key (target level, source level, rint(delta / cell)) per pair with
    float64 delta = c_t - c_s and the target level's cell, a stable
    lexicographic sort of the keys, class starts where the key changes,
    class displacements index * cell rounded to float32."""
    src = np.asarray(sources, np.int64).reshape(-1)
    tgt = np.asarray(targets, np.int64).reshape(-1)
    nbl = np.asarray(topo["nodes_by_level"], np.int64)
    lo = (np.zeros(1, np.int64) if level_offsets is None
          else np.asarray(level_offsets, np.int64))
    n = src.size
    if n == 0:
        return dict(class_keys=np.zeros((0, 5), np.int32),
                    class_displacements=np.zeros((0, 3), np.float32),
                    class_offsets=np.zeros(1, np.int64), class_sources=np.zeros(0, np.int64),
                    class_targets=np.zeros(0, np.int64), class_ids=np.zeros(0, np.int64),
                    level_offsets=lo, level_nodes=nbl)
    center = np.asarray(geom["center"], np.float64)
    level = np.asarray(topo["node_level"], np.int64)
    extent = (np.asarray(topo["bounds_max"], np.float64).reshape(-1, 3)[0]
              - np.asarray(topo["bounds_min"], np.float64).reshape(-1, 3)[0])
    tl, sl = level[tgt], level[src]
    idx = to_int32_like_numpy_cast(np.rint((center[tgt] - center[src]) / _cell_size(extent, tl)))
    keys = np.column_stack([tl, sl, idx.astype(np.int64)])
    order = np.lexsort(tuple(keys[:, k] for k in range(4, -1, -1)))  # stable, column 0 major
    ks = keys[order]
    start = np.ones(n, bool)
    start[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    starts = np.flatnonzero(start)
    class_ids = np.empty(n, np.int64)
    class_ids[order] = np.cumsum(start) - 1
    ck = ks[starts]
    disp = (ck[:, 2:5].astype(np.float64) * _cell_size(extent, ck[:, 0])).astype(np.float32)
    return dict(class_keys=ck.astype(np.int32), class_displacements=disp,
                class_offsets=np.concatenate([starts, [n]]).astype(np.int64),
                class_sources=src[order], class_targets=tgt[order], class_ids=class_ids,
                level_offsets=lo, level_nodes=nbl)


def mac_extents(geom, mac_type, dehnen_radius_scale=1.0):
    """_build_mac_extents (993-1061). Zero-extent propagation is the identity
    here because every half extent is clamped to >= 1e-6."""
    if mac_type in ("dehnen", "engblom"):
        e = np.asarray(geom["radius"], np.float32)
    else:
        e = np.asarray(geom["max_extent"], np.float32)
    if mac_type == "dehnen":
        e = (np.float32(dehnen_radius_scale) * e).astype(np.float32)
    return np.ascontiguousarray(e, np.float32)


def leaf_ordering(topo):
    """_resolve_leaf_ordering (1071-1103)."""
    T = topo["parent"].shape[0]
    NI = (T - 1) // 2
    L = T - NI
    po = np.arange(NI, T, dtype=np.int64)
    native = np.arange(L, dtype=np.int64)
    if "radix_leaf_to_oct" in topo:
        native = np.argsort(np.asarray(topo["radix_leaf_to_oct"], np.int64), kind="stable")
    leaf_indices = po[native]
    p2n = np.zeros(L, np.int64)
    p2n[native] = np.arange(L)
    leaf_pos = np.full(T, -1, np.int64)
    leaf_pos[leaf_indices] = np.arange(L)
    return leaf_indices, leaf_pos, po, p2n


def walk(topo, geom, theta, mac_type="bh", dehnen_radius_scale=1.0, *,
         queue_cap, far_cap, near_cap, collect_far=True, collect_near=True):
    """One fixed-capacity attempt of _dual_tree_walk_impl (1131-1905)."""
    T = topo["parent"].shape[0]
    NI = (T - 1) // 2
    L = T - NI
    leaf_indices, leaf_pos, _, _ = leaf_ordering(topo)
    ext = mac_extents(geom, mac_type, dehnen_radius_scale)
    center = np.ascontiguousarray(geom["center"], np.float32)
    root = int(np.argmin(topo["parent"]))
    left = np.ascontiguousarray(topo["left_child"], np.int64) if NI else np.zeros(1, np.int64)
    right = np.ascontiguousarray(topo["right_child"], np.int64) if NI else np.zeros(1, np.int64)
    h = lib.ygo_walk_run(_I(T), _I(NI), _ptr(left), _ptr(right), _ptr(center), _ptr(ext),
                         _ptr(leaf_pos), _I(root), ctypes.c_float(theta),
                         ctypes.c_int(MAC_CODES[mac_type]), _I(int(queue_cap)), _I(int(far_cap)),
                         _I(int(near_cap)), ctypes.c_int(int(collect_far)),
                         ctypes.c_int(int(collect_near)))
    try:
        sizes = np.zeros(9, np.int64)
        lib.ygo_walk_sizes(ctypes.c_void_p(h), _ptr(sizes))
        fc = np.zeros(T, np.int64)
        nc = np.zeros(L, np.int64)
        fe = np.zeros(max(1, sizes[0]), np.int64)
        ne = np.zeros(max(1, sizes[1]), np.int64)
        gs = np.zeros(max(1, sizes[2]), np.int64)
        lib.ygo_walk_copy(ctypes.c_void_p(h), _ptr(fc), _ptr(fe), _ptr(nc), _ptr(ne), _ptr(gs))
    finally:
        lib.ygo_walk_free(ctypes.c_void_p(h))
    return dict(far_counts=fc, far_entries=fe[: sizes[0]], near_counts=nc,
                near_entries=ne[: sizes[1]], gen_sizes=gs[: sizes[2]],
                queue_overflow=bool(sizes[3]), far_overflow=bool(sizes[4]),
                near_overflow=bool(sizes[5]), accept_decisions=int(sizes[6]),
                near_decisions=int(sizes[7]), refine_decisions=int(sizes[8]),
                leaf_indices=leaf_indices, leaf_position=leaf_pos)


# --- capacity ladders (_interactions_impl.py:296-316, 451-563) -------------

def _queue_candidates_bounded(max_capacity, process_block):
    qmax = max(4, int(max_capacity))
    base = min(max(1024, int(process_block) * 16), qmax)
    c = [base]
    while c[-1] < qmax:
        nxt = min(qmax, c[-1] * 2)
        if nxt == c[-1]:
            break
        c.append(nxt)
    return c


def _auto_pair_queue_candidates(total_nodes, num_internal):
    conservative = 128 * max(4, total_nodes)
    start = min(max(4096, num_internal * 2), conservative)
    if start >= conservative:
        return [int(conservative)]
    c = []
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
        return [int(requested)]
    scaled = min(max(total_nodes // 4, 2048), 16384)
    p2 = 1 << (max(1, scaled) - 1).bit_length()
    return [int(min(16384, max(2048, p2)))]


def _neighbor_capacity_candidates(requested, num_leaves):
    base = max(1, int(requested))
    hard = max(1, ((1 << 31) - 1) // max(1, int(num_leaves)))
    base = min(base, hard)
    c = [base]
    while c[-1] < hard:
        nxt = min(hard, max(c[-1] * 2, c[-1] + 1024))
        if nxt == c[-1]:
            break
        c.append(nxt)
    return list(dict.fromkeys(c))


QUEUE_ERR_USER = "Pair queue capacity exceeded; increase max_pair_queue and rebuild."
QUEUE_ERR_AUTO = ("Pair queue capacity exceeded; pass a larger max_pair_queue "
                  "value to override auto-scaling.")
FAR_ERR = "Interaction list capacity exceeded; increase max_interactions_per_node and rebuild."
NEAR_ERR = "Neighbor list capacity exceeded; increase max_neighbors_per_leaf and rebuild."


def run_walk_with_retries(topo, geom, theta, mac_type="bh", dehnen_radius_scale=1.0, *,
                          max_interactions_per_node=None, max_neighbors_per_leaf=2048,
                          max_pair_queue=None, traversal_config=None, events=None,
                          queue_cache=None):
    """_run_dual_tree_walk_raw (3840-4440), eager mode, default pair policy.

    ``events`` collects the retry events the reference hands to its
    retry_logger (4031-4052, 4339-4381) as (status, attempt, queue_capacity,
    interaction_capacity, far_pair_count, near_pair_count) tuples; the pair
    counts are the attempt's stored entries. ``queue_cache`` (a dict) plays
    the module-level queue-capacity cache (245-293, 3927-3952) for calls with
    a traversal_config. The returned raw dict carries ``path`` ("count" or
    "fixed") and, on the fixed path, the successful ``attempt_caps``."""
    T = topo["parent"].shape[0]
    NI = (T - 1) // 2
    L = T - NI
    if dehnen_radius_scale <= 0.0:
        raise ValueError("dehnen_radius_scale must be > 0")
    kw = dict(theta=theta, mac_type=mac_type, dehnen_radius_scale=dehnen_radius_scale)
    cfg = traversal_config
    cache_key = None
    if cfg is not None:
        qmax = max(4, int(cfg.max_pair_queue))
        inter = [int(cfg.max_interactions_per_node)]
        nbr = _neighbor_capacity_candidates(int(cfg.max_neighbors_per_leaf), L)
        user_q = True
        cache_key = (T, NI, inter[0], int(cfg.max_neighbors_per_leaf), qmax,
                     int(cfg.process_block), True, True, str(mac_type))
        if queue_cache is not None and cache_key in queue_cache:
            qc = [int(queue_cache[cache_key])]
        else:
            qc = _queue_candidates_bounded(qmax, int(cfg.process_block))
        for cap in _auto_pair_queue_candidates(T, NI):
            if cap > qc[-1]:
                qc.append(cap)
        qc = list(dict.fromkeys(qc))
    else:
        user_q = max_pair_queue is not None
        qc = [max(4, int(max_pair_queue))] if user_q else _auto_pair_queue_candidates(T, NI)
        inter = _interaction_capacity_candidates(max_interactions_per_node, T)
        nbr = _neighbor_capacity_candidates(max_neighbors_per_leaf, L)
    use_count = max_interactions_per_node is None and max_pair_queue is None and cfg is None
    qerr = QUEUE_ERR_USER if user_q else QUEUE_ERR_AUTO

    if use_count:
        cap = max(T * 128, _auto_pair_queue_candidates(T, NI)[-1])
        cnt = walk(topo, geom, queue_cap=cap, far_cap=1 << 62, near_cap=1 << 62, **kw)
        if cnt["queue_overflow"]:
            raise RuntimeError("Count-pass traversal wavefront overflowed its conservative buffer.")
        if int(cnt["near_counts"].max(initial=0)) > max(1, ((1 << 31) - 1) // max(1, L)):
            raise RuntimeError("Suggested per-leaf neighbor capacity is too large to safely "
                               "allocate on the host (would overflow int32). Enable local "
                               "refinement or pass explicit capacities "
                               "(max_neighbors_per_leaf / max_interactions_per_node).")
        if cnt["far_counts"].sum() > (1 << 31) - 1 or cnt["near_counts"].sum() > (1 << 31) - 1:
            raise RuntimeError("Dual-tree traversal pair totals exceed signed int32 capacity; "
                               "refine the tree or use a lower-memory traversal strategy.")
        cnt["max_wf"] = int(max(1, cnt["gen_sizes"].max(initial=1)))
        # The exact-length fill that follows (4226-4290) runs with a queue of
        # 1.5x the observed peak, so its first attempt succeeds and emits no
        # event; its lists are the counting walk's.
        cnt["path"] = "count"
        return cnt

    def emit(status, attempt, Q, K, res):
        if events is not None:
            events.append((status, attempt, Q, K, int(res["far_counts"].sum()),
                           int(res["near_counts"].sum())))

    result = None
    success = False
    attempt = 0
    for Q in qc:
        queue_retry = False
        for K in inter:
            for Kn in nbr:
                attempt += 1
                result = walk(topo, geom, queue_cap=Q, far_cap=K, near_cap=Kn, **kw)
                result["attempt_caps"] = (Q, K, Kn)
                result["path"] = "fixed"
                if result["queue_overflow"]:
                    emit("queue_overflow", attempt, Q, K, result)
                    queue_retry = True
                    break
                if result["far_overflow"]:
                    raise RuntimeError(FAR_ERR)
                if result["near_overflow"]:
                    emit("neighbor_overflow", attempt, Q, K, result)
                    continue
                success = True
                if attempt > 1:
                    emit("success", attempt, Q, K, result)
                break
            if queue_retry or success:
                break
        else:
            continue
        if queue_retry and not success:
            continue
        if success:
            if queue_cache is not None and cache_key is not None:
                queue_cache[cache_key] = Q
            break
    else:
        if result["queue_overflow"]:
            raise RuntimeError(qerr)
    if not success and result["far_overflow"]:
        raise RuntimeError(FAR_ERR)
    if result["far_overflow"]:
        raise RuntimeError(FAR_ERR)
    if result["near_overflow"]:
        raise RuntimeError(NEAR_ERR)
    return result


def to_result(raw, topo):
    """The raw DualTreeWalkResult the reference returns with
    return_result=True (_raw_to_result, 4606-4627).

    Fixed-capacity path (_dual_tree_walk_impl, 1787-1905): interaction_offsets
    is the node-order exclusive cumsum of the counts while the sources and
    targets are laid out level-major (quirk B9), padded with -1 to
    T * max_interactions_per_node; tags are all -1; neighbour indices are in
    native leaf-row order, padded with -1 to L * max_neighbors_per_leaf.
    Count path (_dual_tree_walk_compact_fill_impl, 2942-3558): node-order
    offsets and node-order rows of exact length. Decision counters are those
    of the returned attempt; the flags are False after a successful call."""
    T = topo["parent"].shape[0]
    fc = np.asarray(raw["far_counts"], np.int64)
    nc = np.asarray(raw["near_counts"], np.int64)
    L = nc.shape[0]
    F, M = int(fc.sum()), int(nc.sum())
    node_off = np.concatenate([[0], np.cumsum(fc)]).astype(np.int64)
    near_off = np.concatenate([[0], np.cumsum(nc)]).astype(np.int64)
    fe = np.asarray(raw["far_entries"], np.int64)
    ne = np.asarray(raw["near_entries"], np.int64)
    node_tgt = np.repeat(np.arange(T, dtype=np.int64), fc)
    if raw["path"] == "fixed":
        _, K, Kn = raw["attempt_caps"]
        nbl = np.asarray(topo["nodes_by_level"], np.int64)
        src = np.full(max(T * K, 1), -1, np.int64)
        tgt = np.full(max(T * K, 1), -1, np.int64)
        o = 0
        for node in nbl:
            c = int(fc[node])
            src[o:o + c] = fe[node_off[node]:node_off[node] + c]
            tgt[o:o + c] = node
            o += c
        tags = np.full(max(T * K, 1), -1, np.int64)
        nbrs = np.full(max(L * Kn, 1), -1, np.int64)
        nbrs[:M] = ne
    else:
        src, tgt = fe.copy(), node_tgt
        tags = np.full(F, -1, np.int64)
        nbrs = ne.copy()
    leaf_indices, _, _, _ = leaf_ordering(topo)
    return dict(interaction_offsets=node_off, interaction_sources=src, interaction_targets=tgt,
                interaction_tags=tags, interaction_counts=fc, neighbor_offsets=near_off,
                neighbor_indices=nbrs, neighbor_counts=nc, leaf_indices=leaf_indices,
                far_pair_count=F, near_pair_count=M, queue_overflow=False, far_overflow=False,
                near_overflow=False, accept_decisions=raw["accept_decisions"],
                near_decisions=raw["near_decisions"], refine_decisions=raw["refine_decisions"])


def to_interactions(raw, topo):
    """_result_to_interactions (3561-3637)."""
    T = topo["parent"].shape[0]
    counts = raw["far_counts"]
    total = int(counts.sum())
    num_levels = int(topo["num_levels"])
    level_idx = np.asarray(topo["level_offsets"], np.int64)[: num_levels + 1]
    nbl = np.asarray(topo["nodes_by_level"], np.int64)
    cbl = counts[nbl]
    obl = np.concatenate([[0], np.cumsum(cbl)]).astype(np.int64)
    node_off = np.zeros(T, np.int64)
    node_off[nbl] = obl[:-1]
    node_off = np.concatenate([node_off, [total]]).astype(np.int64)
    if total == 0:
        return dict(offsets=np.zeros(T + 1, np.int64), sources=np.zeros(0, np.int64),
                    targets=np.zeros(0, np.int64), counts=counts,
                    level_offsets=np.zeros(num_levels + 1, np.int64),
                    target_levels=np.zeros(0, np.int64))
    # raw entries are node-id ordered rows; regroup rows in level-major order.
    row_start = np.concatenate([[0], np.cumsum(counts)])
    src = np.empty(total, np.int64)
    tgt = np.empty(total, np.int64)
    for node in range(T):
        c = counts[node]
        if c:
            o = node_off[node]
            src[o:o + c] = raw["far_entries"][row_start[node]:row_start[node] + c]
            tgt[o:o + c] = node
    levels = np.asarray(topo["node_level"], np.int64)[tgt]
    return dict(offsets=node_off, sources=src, targets=tgt, counts=counts,
                level_offsets=obl[level_idx], target_levels=levels)


def to_neighbors(raw, topo, target_block_size=None):
    """_result_to_neighbors (3640-3837); the target block size defaults to
    YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE as in the reference (0 = none)."""
    counts = raw["near_counts"]
    L = counts.shape[0]
    leaf_indices, _, po, p2n = leaf_ordering(topo)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    nbrs = np.asarray(raw["near_entries"], np.int64)
    T = topo["parent"].shape[0]
    mx = int(counts.max(initial=0))
    if mx == 0:
        nlp = np.zeros((L, 0), np.int64)
    else:
        lookup = np.full(T, -1, np.int64)
        lookup[leaf_indices] = np.arange(L)
        cols = np.arange(mx)
        valid = cols[None, :] < counts[:, None]
        idx = np.where(valid, offsets[:-1, None] + cols[None, :], 0)
        nlp = np.where(valid, lookup[nbrs[idx]] if nbrs.size else -1, -1)
    k = target_block_size
    if k is None:
        try:
            k = max(0, int(os.environ.get("YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE", "0")))
        except ValueError:
            k = 0
    blocks = target_blocks(counts, offsets, nlp.astype(np.int64), k)
    return dict(offsets=offsets, neighbors=nbrs, leaf_indices=leaf_indices, counts=counts,
                particle_order_leaf_indices=po, particle_order_to_native_leaf=p2n,
                neighbor_leaf_positions=nlp.astype(np.int64),
                target_block_leaf_ids=blocks[0], target_block_source_leaf_ids=blocks[1],
                target_block_valid_mask=blocks[2], target_block_offsets=blocks[3],
                target_block_size=int(k))


def target_blocks(counts, offsets, nlp, k):
    """_build_target_blocks (_interactions_impl.py:3689-3770): row r's
    neighbour columns in blocks of k, from the dense neighbour-leaf-position
    matrix; invalid slots hold 0."""
    L = counts.shape[0]
    if k <= 0 or L == 0:
        return (np.zeros(0, np.int64), np.zeros((0, 0), np.int64), np.zeros((0, 0), bool),
                np.zeros(L + 1, np.int64))
    max_nbr = nlp.shape[1]
    if max_nbr == 0:
        return (np.zeros(0, np.int64), np.zeros((0, k), np.int64), np.zeros((0, k), bool),
                np.zeros(L + 1, np.int64))
    per = (counts + (k - 1)) // k
    boff = np.concatenate([[0], np.cumsum(per)]).astype(np.int64)
    B = int(boff[-1])
    ids = np.arange(B, dtype=np.int64)
    row = np.searchsorted(boff[1:], ids, side="right")
    local = ids - boff[row]
    start = offsets[row] + local * k
    stop = offsets[row + 1]
    j = np.arange(k, dtype=np.int64)
    edge_valid = (start[:, None] + j[None, :]) < stop[:, None]
    lidx = local[:, None] * k + j[None, :]
    lvalid = lidx < max_nbr
    src = nlp[row[:, None], np.where(edge_valid & lvalid, lidx, 0)]
    ok = edge_valid & lvalid & (src >= 0)
    return row.astype(np.int64), np.where(ok, src, 0).astype(np.int64), ok, boff


def build_interactions_and_neighbors(topo, geom, theta=0.5, max_interactions_per_node=None,
                                     max_neighbors_per_leaf=2048, max_pair_queue=None,
                                     process_block=None, traversal_config=None,
                                     mac_type="bh", dehnen_radius_scale=1.0, events=None,
                                     queue_cache=None):
    """_interactions_impl.py:6032-6144; the third element is the walk's raw
    outputs (to_result gives the reference's DualTreeWalkResult)."""
    raw = run_walk_with_retries(
        topo, geom, theta, mac_type, dehnen_radius_scale,
        max_interactions_per_node=max_interactions_per_node,
        max_neighbors_per_leaf=max_neighbors_per_leaf, max_pair_queue=max_pair_queue,
        traversal_config=traversal_config, events=events, queue_cache=queue_cache)
    return to_interactions(raw, topo), to_neighbors(raw, topo), raw


def walk_with_policy(topo, geom, theta, mac_type, dehnen_radius_scale, policy, policy_state):
    """Generation-synchronous walk with a custom pair policy evaluated in both
    directions (_resolve_pair_actions, 731-810, inside 1131-1785). Returns
    node-order rows of (source, tag) far entries and near rows."""
    T = topo["parent"].shape[0]
    NI = (T - 1) // 2
    L = T - NI
    leaf_indices, leaf_pos, _, _ = leaf_ordering(topo)
    ext = mac_extents(geom, mac_type, dehnen_radius_scale)
    c = np.asarray(geom["center"], np.float32)
    th = np.float32(theta)
    th2 = np.float32(th * th)
    left = np.asarray(topo["left_child"], np.int64)
    right = np.asarray(topo["right_child"], np.int64)
    far = [[] for _ in range(T)]
    near = [[] for _ in range(L)]
    wt = np.array([0], np.int64)
    ws = np.array([0], np.int64)
    gens = [1]
    while wt.size:
        d = (c[wt] - c[ws]).astype(np.float32)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)
        d2 = (d2 + d[:, 2] * d[:, 2]).astype(np.float32)
        et, es = ext[wt], ext[ws]
        diff = wt != ws
        if mac_type == "engblom":
            rb, rs = np.maximum(et, es), np.minimum(et, es)
            lhs = (rb + np.float32(np.sqrt(th2)) * rs).astype(np.float32)
            ok = diff & (d2 > 0) & (lhs * lhs <= th2 * d2)
        else:
            sm = (et + es).astype(np.float32)
            ok = diff & (d2 > 0) & (sm * sm <= th2 * d2)
        tl, sl = wt >= NI, ws >= NI
        valid = np.ones(wt.size, bool)
        common = dict(valid_pairs=valid, mac_ok=ok, different_nodes=diff, same_node=~diff,
                      dist_sq=d2)
        fa, ft = policy(policy_state, target_leaf=tl, source_leaf=sl, target_nodes=wt,
                        source_nodes=ws, center_target=c[wt], center_source=c[ws],
                        extent_target=et, extent_source=es, **common)
        ra, rt = policy(policy_state, target_leaf=sl, source_leaf=tl, target_nodes=ws,
                        source_nodes=wt, center_target=c[ws], center_source=c[wt],
                        extent_target=es, extent_source=et, **common)
        fa, ra, ft, rt = (np.asarray(x, np.int64) for x in (fa, ra, ft, rt))
        acc = (fa == 0) & (ra == 0)
        nr = (fa == 1) & (ra == 1) & ~acc
        nt_, ns_ = [], []
        for p in range(wt.size):
            t, s = int(wt[p]), int(ws[p])
            if acc[p]:
                far[t].append((s, int(ft[p])))
                far[s].append((t, int(rt[p])))
            elif nr[p]:
                near[leaf_pos[t]].append(s)
                near[leaf_pos[s]].append(t)
            else:
                same = t == s
                ti, si = t < NI, s < NI
                st = ti and (same or not si or et[p] >= es[p])
                ss = si and (same or not ti or es[p] > et[p])
                if st and ss and same:
                    kids = [(left[t], left[t]), (left[t], right[t]), (right[t], right[t])]
                elif st and ss:
                    kids = [(left[t], left[s]), (left[t], right[s]), (right[t], left[s]),
                            (right[t], right[s])]
                elif st:
                    kids = [(left[t], s), (right[t], s)]
                elif ss:
                    kids = [(t, left[s]), (t, right[s])]
                else:
                    kids = []
                for a, b in kids:
                    nt_.append(min(a, b))
                    ns_.append(max(a, b))
        wt = np.array(nt_, np.int64)
        ws = np.array(ns_, np.int64)
        gens.append(int(wt.size))
    return far, near, gens
