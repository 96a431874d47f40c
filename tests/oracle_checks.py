"""Brute-force and hand-derived checks of the oracle's octree fields and of
the walk's split tie-break (CPU only).

Each check restates one rule of the reference in a different form from the
oracle's (enumeration and counting instead of binary search, a hand-derived
known answer instead of the walk), so a one-character slip in the oracle
(a wrong search side, a shift clamp of 63 instead of 62, a strict instead of
a non-strict comparison) fails it. The checks raise AssertionError; the
tests in test_oracle_pins_octree.py run them on the default oracle and
test_oracle_mutants.py runs them on deliberately mutated oracle builds.
"""

from __future__ import annotations

import numpy as np

import oracle

MORTON_BITS = 63
MAX_LEVEL = 21


def common_depth(a: int, b: int) -> int:
    """Octree depth of the common prefix cell of two 63-bit codes
    (octree.py:150-165: common leading bits minus the unused top bit, // 3)."""
    x = int(a) ^ int(b)
    bits = MORTON_BITS if x == 0 else max(64 - x.bit_length() - 1, 0)
    return bits // 3


def prefix(code: int, depth: int) -> int:
    if depth <= 0:
        return 0
    s = max(0, MORTON_BITS - 3 * depth)
    return (int(code) >> s) << s


def address(code: int, depth: int) -> int:
    """Level-order cell address: cells of shallower depth first, then by code."""
    return (8 ** depth - 1) // 7 + (int(code) >> (MORTON_BITS - 3 * depth))


def resolved_cells(t):
    """Per radix leaf: (depth, code) of the common-prefix cell of its first
    and last particle code (octree.py:168-186, leaf_depths all -1)."""
    NI = int(t["num_internal_nodes"])
    sc = t["morton_codes"]
    out = []
    for a, b in t["node_ranges"][NI:]:
        d = common_depth(sc[a], sc[b])
        out.append((d, prefix(sc[a], d)))
    return out


def valid_slot_cells(t):
    nu = int(np.count_nonzero(t["oct_valid_mask"]))
    return [(int(t["oct_node_depths"][k]), int(t["oct_node_codes"][k])) for k in range(nu)]


def check_oct_node_ranges(t):
    """octree.py:284-313. For a valid slot (code C, depth d) the reference
    covers the leaves whose resolved code lies in [C, C | mask] with
    mask = 2^min(63 - 3d, 62) - 1 (the root's shift is clamped to 62), found
    by lower/upper bound on the resolved leaf codes, clipped to [0, L-1], and
    maps them to (first leaf's start, last leaf's end). Restated here by
    enumerating the leaves in the interval, which equals the bounds on sorted
    codes; unsorted codes go to check_oct_node_ranges_unsorted. Invalid slots
    hold (0, -1)."""
    NI = int(t["num_internal_nodes"])
    cells = resolved_cells(t)
    codes = [c for _, c in cells]
    if codes != sorted(codes):
        check_oct_node_ranges_unsorted(t)
        return
    L = len(codes)
    leaf_ranges = t["node_ranges"][NI:]
    slots = valid_slot_cells(t)
    ranges = t["oct_node_ranges"]
    for k, (d, c) in enumerate(slots):
        shift = min(MORTON_BITS - 3 * d, MORTON_BITS - 1)
        end = c | ((1 << shift) - 1)
        members = [j for j in range(L) if c <= codes[j] <= end]
        if members:
            assert members == list(range(members[0], members[-1] + 1))
            first, last = members[0], members[-1]
        else:
            first = sum(1 for x in codes if x < c)
            last = first - 1
        first = min(max(first, 0), L - 1)
        last = min(max(last, 0), L - 1)
        exp = (int(leaf_ranges[first][0]), int(leaf_ranges[last][1]))
        assert (int(ranges[k][0]), int(ranges[k][1])) == exp, (k, d, hex(c))
    for k in range(len(slots), ranges.shape[0]):
        assert tuple(int(v) for v in ranges[k]) == (0, -1)


def _is_transition(codes, h, pred):
    """h is a transition point of pred over codes: pred holds at h (or
    h == len) and fails at h - 1 (or h == 0)."""
    n = len(codes)
    return (h == n or pred(codes[h])) and (h == 0 or not pred(codes[h - 1]))


def check_oct_node_ranges_unsorted(t):
    """Resolved leaf codes can be unsorted (a leaf whose particles span a
    coarse cell is followed by one inside a finer cell). The reference then
    takes jnp.searchsorted's fixed-round bisection on unsorted input. Any
    bisection with a consistent predicate stops at a transition point of it,
    which is what is checked here: the first leaf index is a transition of
    ``C <= code`` (side='left') and last + 1 one of ``end < code``
    (side='right'), up to the clip to [0, L-1]. Which transition the
    bisection picks is not pinned by this (it follows the bisection order)."""
    NI = int(t["num_internal_nodes"])
    codes = [c for _, c in resolved_cells(t)]
    L = len(codes)
    leaf_ranges = t["node_ranges"][NI:]
    start_of = {int(a): j for j, (a, _) in enumerate(leaf_ranges)}
    end_of = {int(b): j for j, (_, b) in enumerate(leaf_ranges)}
    for k, (d, c) in enumerate(valid_slot_cells(t)):
        shift = min(MORTON_BITS - 3 * d, MORTON_BITS - 1)
        end = c | ((1 << shift) - 1)
        f = start_of[int(t["oct_node_ranges"][k][0])]
        last = end_of[int(t["oct_node_ranges"][k][1])]
        left = lambda x, c=c: c <= x  # noqa: E731
        right = lambda x, e=end: e < x  # noqa: E731
        assert _is_transition(codes, f, left) or (f == L - 1 and _is_transition(codes, L, left)), k
        ok_last = _is_transition(codes, last + 1, right)
        if last == 0:
            ok_last = ok_last or _is_transition(codes, 0, right)
        assert ok_last, k


def check_radix_node_to_oct(t):
    """octree.py:518-531. Every radix node (internal: the common-prefix cell
    of its first and last particle code; leaf: its resolved cell) maps to the
    lower bound of its cell address among the slot addresses, invalid slots
    reading as the largest u64. Restated by counting the valid slots with a
    smaller address; where the node's cell is itself a slot, that slot."""
    NI = int(t["num_internal_nodes"])
    sc = t["morton_codes"]
    slots = valid_slot_cells(t)
    addrs = [address(c, d) for d, c in slots]
    assert addrs == sorted(addrs) and len(set(addrs)) == len(addrs)
    lookup = {dc: k for k, dc in enumerate(slots)}
    cells = resolved_cells(t)
    T = t["parent"].shape[0]
    r2o = t["radix_node_to_oct"]
    for i in range(T):
        if i >= NI:
            d, c = cells[i - NI]
        else:
            a, b = t["node_ranges"][i]
            d = common_depth(sc[a], sc[b])
            c = prefix(sc[a], d)
        A = address(c, d)
        exp = sum(1 for x in addrs if x < A)
        assert int(r2o[i]) == exp, (i, d, hex(c))
        if (d, c) in lookup:
            assert int(r2o[i]) == lookup[(d, c)]
    assert np.array_equal(t["radix_leaf_to_oct"], r2o[NI:])


def check_oct_levels(t):
    """octree.py:315-337: level_offsets[23] = [0, cumsum of the valid slots'
    depth histogram over 22 bins]; num_levels = deepest valid depth + 1."""
    slots = valid_slot_cells(t)
    hist = [0] * (MAX_LEVEL + 1)
    for d, _ in slots:
        hist[d] += 1
    exp = [0]
    for h in hist:
        exp.append(exp[-1] + h)
    assert [int(v) for v in t["oct_level_offsets"]] == exp
    assert int(t["oct_num_levels"]) == max(d for d, _ in slots) + 1


def tie_break_case():
    """Four points on the x axis, mirror-symmetric about 0, one per leaf.

    The radix tree is root 0 over internal nodes 1 = {-3, -1} and
    2 = {1, 3}; leaves 3..6 in x order. Nodes 1 and 2 have identical MAC
    extents (radius of half extents (1, 1e-6, 1e-6) = 1 in float32), centres
    -2 and 2, so at theta = 0.45 the pair (1, 2) fails the dehnen MAC
    ((1 + 1)^2 = 4 > 0.2025 * 16) and is refined with e_t == e_s. The
    reference splits the target, the lower node id, on a tie
    (_interactions_impl.py:1444-1454: e_t >= e_s for the target, e_s > e_t
    for the source). Hand-derived lists:
      (1,1) -> (3,3), (3,4), (4,4): (3,4) accepted (two points 2 apart);
      (1,2) -> (3,2), (4,2): accepted (d = 5 and 3 against extent 1);
      (2,2) -> (5,5), (5,6), (6,6): (5,6) accepted.
    Far: 2:{3,4}, 3:{2,4}, 4:{2,3}, 5:{6}, 6:{5}; no near entries. Splitting
    the source instead would give 1:{5,6}, 5:{1,6}, 6:{1,5}, 3:{4}, 4:{3};
    splitting neither would lose the (1,2) subtree's pairs.
    """
    pos = np.array([[-3, 0, 0], [-1, 0, 0], [1, 0, 0], [3, 0, 0]], np.float32)
    mass = np.ones(4, np.float32)
    far = {2: [3, 4], 3: [2, 4], 4: [2, 3], 5: [6], 6: [5]}
    return pos, mass, 0.45, far


def check_tie_break(far_lists, near_total, counts):
    """far_lists: {node: sorted sources}; compares with the hand-derived
    answer of tie_break_case."""
    _, _, _, far = tie_break_case()
    got = {k: v for k, v in far_lists.items() if v}
    assert got == far, got
    assert near_total == 0
    assert [int(c) for c in counts] == [len(far.get(k, [])) for k in range(7)]


def run_tie_break_on_oracle():
    pos, mass, theta, _ = tie_break_case()
    t = oracle.build_tree(pos, mass, leaf_size=1)
    assert [int(v) for v in t["left_child"]] == [1, 3, 5]
    assert [int(v) for v in t["right_child"]] == [2, 4, 6]
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])
    inter, nbrs, _ = oracle.build_interactions_and_neighbors(t, g, theta=theta,
                                                             mac_type="dehnen")
    lists = {}
    for k in range(7):
        o, n = int(inter["offsets"][k]), int(inter["counts"][k])
        lists[k] = sorted(int(s) for s in inter["sources"][o:o + n])
    check_tie_break(lists, int(nbrs["counts"].sum()), inter["counts"])


def b3_root_clamp_case():
    """Eight points, one per leaf: four with z = -0.5 and four with z = +0.5.
    With bounds padded symmetrically about 0 the upper-z points have the top
    code bit (bit 62, the z bit of level 1) set and sort last. The root cell
    (depth 0) has shift 63 - 0 = 63, clamped to 62 (octree.py:292-294), so
    its interval [0, 2^62 - 1] holds only the four lower-z leaves: the
    reference's root range is (0, 3), not the geometric (0, 7)."""
    pos = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [-0.5, 0.5, -0.5], [0.5, 0.5, -0.5],
                    [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [-0.5, 0.5, 0.5], [0.5, 0.5, 0.5]],
                   np.float32)
    return pos, (0, 3)


def mapping_case():
    """tests/unit/test_octree_topology.py:65-93: eight points on the diagonal,
    build_octree leaf_size=2."""
    return np.array([[0.05, 0.10, 0.15], [0.18, 0.21, 0.22], [0.31, 0.30, 0.35],
                     [0.44, 0.41, 0.48], [0.55, 0.62, 0.58], [0.68, 0.70, 0.66],
                     [0.82, 0.84, 0.79], [0.93, 0.92, 0.95]], np.float32)


def bfs_generations(t, g, theta, mac, scale=1.0):
    """A plain-Python generation-by-generation dual-tree walk (the visiting
    rule of _interactions_impl.py:1131-1785: MAC per pair, accept, near leaf
    pairs, split rule 1444-1454, children sorted as (min, max)), written
    independently of the oracle's C++ walk. Returns, per generation, the
    number of pairs, the per-node far and per-leaf near entries it adds
    (both directions) and the number of child pairs it pushes."""
    from collections import Counter

    from oracle import oracle as O
    NI = int(t["num_internal_nodes"])
    ext = O.mac_extents(g, mac, scale)
    c = np.asarray(g["center"], np.float32)
    th2 = np.float32(np.float32(theta) * np.float32(theta))
    Lc, Rc = t["left_child"], t["right_child"]
    wave = [(0, 0)]
    gens = []
    while wave:
        far, near = Counter(), Counter()
        nxt = []
        for a, b in wave:
            d = c[a] - c[b]
            d2 = np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
            ea, eb = ext[a], ext[b]
            if mac == "engblom":
                rb, rs = max(ea, eb), min(ea, eb)
                lhs = np.float32(rb + np.float32(np.sqrt(th2)) * rs)
                ok = a != b and d2 > 0 and lhs * lhs <= th2 * d2
            else:
                s = np.float32(ea + eb)
                ok = a != b and d2 > 0 and s * s <= th2 * d2
            if ok:
                far[a] += 1
                far[b] += 1
                continue
            if a >= NI and b >= NI:
                if a != b:
                    near[a] += 1
                    near[b] += 1
                continue
            same = a == b
            st = a < NI and (same or b >= NI or ea >= eb)
            ss = b < NI and (same or a >= NI or eb > ea)
            if st and ss and same:
                kids = [(Lc[a], Lc[a]), (Lc[a], Rc[a]), (Rc[a], Rc[a])]
            elif st and ss:
                kids = [(x, y) for x in (Lc[a], Rc[a]) for y in (Lc[b], Rc[b])]
            elif st:
                kids = [(Lc[a], b), (Rc[a], b)]
            else:
                kids = [(a, Lc[b]), (a, Rc[b])]
            nxt.extend((min(int(x), int(y)), max(int(x), int(y))) for x, y in kids)
        gens.append(dict(pairs=len(wave), far=far, near=near, pushes=len(nxt)))
        wave = nxt
    return gens


def simulate_attempt(gens, Q, K, Kn):
    """One fixed-capacity attempt over the generations: entries beyond a
    row's capacity are dropped and raise its flag, pushes beyond the queue
    raise the queue flag, and the walk stops after the first generation
    that raised any flag. Returns (queue, far, near) flags and the stored
    far and near totals."""
    from collections import Counter
    fc, nc = Counter(), Counter()
    qo = fo = no = False
    for gen in gens:
        for k, v in gen["far"].items():
            fo |= fc[k] + v > K
            fc[k] = min(fc[k] + v, K)
        for k, v in gen["near"].items():
            no |= nc[k] + v > Kn
            nc[k] = min(nc[k] + v, Kn)
        qo |= gen["pushes"] > Q
        if qo or fo or no:
            break
    return (qo, fo, no), sum(fc.values()), sum(nc.values())
