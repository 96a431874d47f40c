"""CPU pins of the oracle (no GPU): known answers, library routines and
brute-force restatements chosen so a dropped term, wrong index or swapped
operand fails."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import oracle as O
from yggdrax_amd import inputs


# ---- input generators: Random123 threefry2x32-20 known-answer vectors -------

@pytest.mark.parametrize("key,ctr,expected", [
    ((0, 0), (0, 0), (0x6B200159, 0x99BA4EFE)),
    ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
    ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
])
def test_threefry_random123_kat(key, ctr, expected):
    x0, x1 = inputs.threefry2x32(key[0], key[1], np.array([ctr[0]], np.uint32),
                                 np.array([ctr[1]], np.uint32))
    assert (int(x0[0]), int(x1[0])) == expected


def _getting_started():
    pos, mass = inputs.uniform_cube(256, 7)
    topo = oracle.build_tree(pos, mass, leaf_size=16)
    geom = oracle.compute_tree_geometry(topo, topo["positions_sorted"])

    class Cfg:
        max_pair_queue, process_block = 8192, 256
        max_interactions_per_node, max_neighbors_per_leaf = 2048, 2048
    inter, nbrs, raw = oracle.build_interactions_and_neighbors(
        topo, geom, theta=0.8, mac_type="dehnen", traversal_config=Cfg)
    return pos, topo, geom, inter, nbrs


def test_getting_started_notebook_known_answer():
    """examples/getting_started.ipynb:33-52 (cells 2-3): PRNGKey(7), 256
    points, build_tree leaf 16, dehnen theta 0.8 -> 31 nodes, 16 leaves,
    4 far entries, 236 near entries."""
    pos, topo, _, inter, nbrs = _getting_started()
    assert topo["parent"].shape[0] == 31
    assert topo["parent"].shape[0] - topo["num_internal_nodes"] == 16
    assert int(inter["counts"].sum()) == 4
    assert int(nbrs["counts"].sum()) == 236


def test_getting_started_morton_neighbour_distance():
    """Same notebook, cell 8: mean Morton-order NN distance 0.2072429060935974
    and loss 0.015111171640455723 pin the particle order and the jax PRNG."""
    pos, topo, *_ = _getting_started()
    ps = pos[topo["particle_indices"]]
    n = ps.shape[0]
    idx = np.arange(n)
    cands = []
    for off in range(1, 9):
        for j, ok in ((np.clip(idx + off, 0, n - 1), idx + off < n),
                      (np.clip(idx - off, 0, n - 1), idx - off >= 0)):
            d = np.sqrt(((ps - ps[j]) ** 2).sum(1) + np.float32(1e-12))
            cands.append(np.where(ok, d, 1e6))
    nn = np.min(np.stack(cands, 1), 1)
    assert abs(float(nn.mean()) - 0.2072429060935974) < 2e-7
    assert abs(float(((nn - 0.12) ** 2).mean()) - 0.015111171640455723) < 2e-8


# ---- bounds and Morton (tests/unit/test_bounds.py, test_morton.py) ---------

def test_bounds_padding():
    p = np.array([[-1.0, 2.0, 0.5], [3.0, -2.0, 1.5], [0.5, 1.0, -4.0]], np.float32)
    lo, hi = oracle.infer_bounds(p)
    assert np.all(lo < p.min(0)) and np.all(hi > p.max(0))
    np.testing.assert_array_equal(lo, p.min(0) - np.float32(0.05) * (p.max(0) - p.min(0)))
    d = np.full((2, 3), 2.0, np.float32)
    lo, hi = oracle.infer_bounds(d)
    assert np.all(hi - lo >= 1.9e-6)


def _naive_interleave(cx, cy, cz):
    out = 0
    for b in range(21):
        out |= ((cx >> b) & 1) << (3 * b)
        out |= ((cy >> b) & 1) << (3 * b + 1)
        out |= ((cz >> b) & 1) << (3 * b + 2)
    return out


def test_morton_interleave_matches_naive_bit_loop():
    rng = np.random.default_rng(0)
    pos = rng.uniform(-1, 1, (200, 3)).astype(np.float32)
    lo = np.full(3, -1.0, np.float32)
    hi = np.full(3, 1.0, np.float32)
    codes = oracle.morton_encode(pos, (lo, hi))
    for i in range(pos.shape[0]):
        q = [int(np.rint(np.float32(np.clip((pos[i, d] - lo[d]) / (hi[d] - lo[d]), 0, 1))
                         * np.float32(2097151))) for d in range(3)]
        assert int(codes[i]) == _naive_interleave(*q)
    assert np.all(codes < (1 << 63))


def _decode(codes, lo, hi):
    def compact(x):
        out = np.zeros_like(x)
        for b in range(21):
            out |= ((x >> np.uint64(3 * b)) & np.uint64(1)) << np.uint64(b)
        return out
    c = [compact(codes >> np.uint64(k)) for k in range(3)]
    q = np.stack(c, 1).astype(np.float32) / np.float32(2097151)
    return q * (hi - lo) + lo


def test_morton_clamp_and_roundtrip():
    lo = np.full(3, -1.0, np.float32)
    hi = np.full(3, 1.0, np.float32)
    p = np.array([[-2, -2, -2], [2, 2, 2]], np.float32)
    dec = _decode(oracle.morton_encode(p, (lo, hi)), lo, hi)
    np.testing.assert_allclose(dec[0], lo, atol=1e-4)
    np.testing.assert_allclose(dec[1], hi, atol=1e-4)
    q = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [1, 1, 1], [-0.5, 0.2, 0.8]], np.float32)
    np.testing.assert_allclose(_decode(oracle.morton_encode(q, (lo, hi)), lo, hi), q, atol=1e-4)


def test_morton_round_half_even():
    # norm * 2097151 = k + 0.5 exactly for norm = (2k+1)/(2*2097151) representable?
    lo = np.zeros(3, np.float32)
    hi = np.full(3, 2097151.0, np.float32)
    p = np.array([[2.5, 3.5, 4.5]], np.float32)
    c = oracle.morton_encode(p, (lo, hi))
    assert int(c[0]) == _naive_interleave(2, 4, 4)


# ---- sort ---------------------------------------------------------------

def test_stable_argsort_matches_numpy():
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 5000).astype(np.uint64) << np.uint64(40)
    assert np.array_equal(O.stable_argsort_u64(keys), np.argsort(keys, kind="stable"))


def test_identical_points_keep_input_order():
    """tests/unit/test_tree.py:193-214."""
    topo = oracle.build_tree(np.full((40, 3), 0.3, np.float32), None, leaf_size=4)
    assert np.array_equal(topo["particle_indices"], np.arange(40))


# ---- Karras build: textbook recursive split for distinct leaf codes ---------

def _textbook_tree(leaf_codes):
    """Top-down split at the highest differing bit; returns the set of
    (first, last, split) internal ranges."""
    out = set()

    def rec(a, b):
        if a == b:
            return
        common = 64 if leaf_codes[a] == leaf_codes[b] else 64 - int(leaf_codes[a] ^ leaf_codes[b]).bit_length()
        split = a
        for j in range(a, b):
            x = int(leaf_codes[a] ^ leaf_codes[j + 1])
            if (64 - x.bit_length() if x else 64) > common:
                split = j + 1
        # split = last index sharing more than `common` bits with a
        out.add((a, b, split))
        rec(a, split)
        rec(split + 1, b)
    rec(0, len(leaf_codes) - 1)
    return out


@pytest.mark.parametrize("n,leaf,seed", [(300, 4, 0), (1000, 7, 1), (64, 1, 2)])
def test_karras_matches_textbook_split_for_distinct_codes(n, leaf, seed):
    pos = np.random.default_rng(seed).normal(size=(n, 3)).astype(np.float32)
    topo = oracle.build_tree(pos, None, leaf_size=leaf)
    lc = topo["leaf_codes"]
    if len(np.unique(lc)) != len(lc):
        pytest.skip("duplicate leaf codes")
    L = len(lc)
    NI = L - 1
    got = set()
    for i in range(NI):
        # recover leaf range of internal node i from its particle range
        a, b = topo["node_ranges"][i]
        first, last = a // leaf, b // leaf
        lcld, rcld = topo["left_child"][i], topo["right_child"][i]
        split = (lcld - NI) if topo["left_is_leaf"][i] else None
        if split is None:
            split = int(topo["node_ranges"][lcld][1]) // leaf
        got.add((int(first), int(last), int(split)))
    assert got == _textbook_tree([int(c) for c in lc])


@pytest.mark.parametrize("n,leaf", [(500, 3), (2048, 16)])
def test_topology_invariants(n, leaf):
    pos = np.random.default_rng(n).uniform(-1, 1, (n, 3)).astype(np.float32)
    t = oracle.build_tree(pos, None, leaf_size=leaf)
    T = t["parent"].shape[0]
    NI = t["num_internal_nodes"]
    L = T - NI
    assert NI == -(-n // leaf) - 1
    assert np.all(np.diff(t["morton_codes"].astype(np.float64)) >= 0)
    for i in range(NI):  # parent / child back pointers and flags
        l, r = t["left_child"][i], t["right_child"][i]
        assert t["parent"][l] == i and t["parent"][r] == i
        assert t["left_is_leaf"][i] == (l >= NI) and t["right_is_leaf"][i] == (r >= NI)
        assert t["node_ranges"][i][0] == t["node_ranges"][l][0]
        assert t["node_ranges"][i][1] == t["node_ranges"][r][1]
        assert t["node_ranges"][l][1] + 1 == t["node_ranges"][r][0]
    order = []  # DFS leaf order == sorted order (test_tree.py:217-274)
    stack = [0]
    while stack:
        k = stack.pop()
        if k >= NI:
            order.append(k - NI)
        else:
            stack += [t["right_child"][k], t["left_child"][k]]
    assert order == list(range(L))
    # levels == BFS depth; offsets and stable level order
    depth = np.zeros(T, np.int64)
    q = [0]
    while q:
        k = q.pop()
        if k < NI:
            for c in (t["left_child"][k], t["right_child"][k]):
                depth[c] = depth[k] + 1
                q.append(c)
    assert np.array_equal(t["node_level"], depth)
    assert np.array_equal(t["nodes_by_level"], np.argsort(depth, kind="stable"))
    cnt = np.bincount(depth, minlength=64)[:64]
    assert np.array_equal(t["level_offsets"], np.concatenate([[0], np.cumsum(cnt)]))
    assert t["num_levels"] == depth.max() + 1


# ---- octree metadata: brute force ------------------------------------------

def _cd(a, b):
    x = int(a) ^ int(b)
    bits = 63 if x == 0 else max(64 - x.bit_length() - 1, 0)
    return bits // 3


def _pre(c, d):
    if d <= 0:
        return 0
    s = max(0, 63 - 3 * d)
    return (int(c) >> s) << s


def _addr(c, d):
    return (8 ** d - 1) // 7 + (int(c) >> (63 - 3 * d))


@pytest.mark.parametrize("n,leaf,gen", [(700, 8, "uniform"), (3000, 16, "plummer"), (2, 1, "uniform"), (9, 9, "uniform")])
def test_octree_metadata_brute_force(n, leaf, gen):
    if gen == "uniform":
        pos = np.random.default_rng(n).uniform(-1, 1, (n, 3)).astype(np.float32)
    else:
        pos, _ = inputs.plummer_sphere(n, 7, a=0.2)
    t = oracle.build_octree(pos, None, leaf_size=leaf)
    NI = t["num_internal_nodes"]
    T = t["parent"].shape[0]
    L = T - NI
    sc = t["morton_codes"]
    cells = []
    for j in range(L):
        a, b = t["node_ranges"][NI + j]
        d = _cd(sc[a], sc[b])
        cells.append((d, _pre(sc[a], d)))
    cand = set(cells)
    for j in range(L - 1):
        d = _cd(cells[j][1], cells[j + 1][1])
        cand.add((d, _pre(cells[j][1], d)))
    uniq = sorted(cand, key=lambda dc: _addr(dc[1], dc[0]))
    nu = len(uniq)
    valid = t["oct_valid_mask"]
    assert valid.sum() == nu and np.all(valid[:nu]) and not np.any(valid[nu:])
    assert [(int(t["oct_node_depths"][k]), int(t["oct_node_codes"][k])) for k in range(nu)] == uniq
    # parent = deepest strict ancestor cell present (brute force)
    lookup = {dc: k for k, dc in enumerate(uniq)}
    for k, (d, c) in enumerate(uniq):
        exp = -1
        for pd in range(d - 1, -1, -1):
            key = (pd, _pre(c, pd))
            if key in lookup:
                exp = lookup[key]
                break
        assert t["oct_parent"][k] == exp
    # children / counts / leaf mask consistent with parents
    kids = {}
    for k in range(nu):
        p = int(t["oct_parent"][k])
        if p >= 0:
            d, c = uniq[k]
            o = (c >> (63 - 3 * d)) & 7 if d > 0 else 0
            kids[(p, o)] = max(kids.get((p, o), -1), k)
    for k in range(nu):
        row = [kids.get((k, o), -1) for o in range(8)]
        assert list(t["oct_children"][k]) == row
        assert t["oct_leaf_mask"][k] == (all(v < 0 for v in row))
    # leaf mapping lands on each leaf's own cell
    for j in range(L):
        assert uniq[t["radix_leaf_to_oct"][j]] == cells[j]
    assert np.array_equal(t["oct_nodes_by_level"],
                          np.argsort(np.where(valid, t["oct_node_depths"], 22), kind="stable"))


def test_fixed_round_search_equals_numpy_on_sorted():
    """The restated jnp.searchsorted(method='scan') is lower/upper bound on
    sorted input (checked against numpy), for sizes around powers of two."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4, 5, 7, 8, 9, 31, 32, 33, 100):
        a = np.sort(rng.integers(0, 40, n)).astype(np.uint64)
        qs = np.arange(-1, 42).astype(np.uint64)
        lo = np.array([O.lib and _search(a, q, True) for q in qs])
        hi = np.array([_search(a, q, False) for q in qs])
        assert np.array_equal(lo, np.searchsorted(a, qs, "left"))
        assert np.array_equal(hi, np.searchsorted(a, qs, "right"))


def _search(a, q, left):
    n = len(a)
    low, high = 0, n
    for _ in range(n.bit_length()):
        mid = (low + high) // 2
        v = a[min(mid, n - 1)]
        go = (q <= v) if left else (q < v)
        low, high = (low, mid) if go else (mid, high)
    return high


# ---- geometry ------------------------------------------------------------

@pytest.mark.parametrize("n,leaf", [(1000, 8), (333, 1)])
def test_geometry_is_direct_min_max(n, leaf):
    pos = np.random.default_rng(n).normal(size=(n, 3)).astype(np.float32)
    t = oracle.build_tree(pos, None, leaf_size=leaf)
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])
    ps = t["positions_sorted"]
    for k in range(t["parent"].shape[0]):
        a, b = t["node_ranges"][k]
        mn, mx = ps[a:b + 1].min(0), ps[a:b + 1].max(0)
        np.testing.assert_array_equal(g["center"][k], np.float32(0.5) * (mn + mx))
        h = np.maximum(np.float32(0.5) * (mx - mn), np.float32(1e-6))
        np.testing.assert_array_equal(g["half_extent"][k], h)
        assert g["radius"][k] == np.sqrt(np.float32(h[0] * h[0] + h[1] * h[1]) + h[2] * h[2])
        assert g["max_extent"][k] == h.max()
    if leaf == 1:  # leaf centres are the points, half extent the clamp
        NI = t["num_internal_nodes"]
        np.testing.assert_array_equal(g["center"][NI:], ps)
        assert np.all(g["half_extent"][NI:] == np.float32(1e-6))


# ---- walk: independent recursive restatement + brute-force invariants -----

def _recursive_walk(t, g, theta, mac, scale=1.0):
    NI = t["num_internal_nodes"]
    ext = O.mac_extents(g, mac, scale)
    c = g["center"]
    th2 = np.float32(theta) * np.float32(theta)
    far, near = {}, {}

    def rec(a, b):
        a, b = min(a, b), max(a, b)
        d = c[a] - c[b]
        d2 = np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        if mac == "engblom":
            rb, rs = max(ext[a], ext[b]), min(ext[a], ext[b])
            lhs = np.float32(rb + np.float32(np.sqrt(th2)) * rs)
            ok = a != b and d2 > 0 and lhs * lhs <= th2 * d2
        else:
            s = np.float32(ext[a] + ext[b])
            ok = a != b and d2 > 0 and s * s <= th2 * d2
        if ok:
            far.setdefault(a, []).append(b)
            far.setdefault(b, []).append(a)
            return
        if a >= NI and b >= NI:
            if a != b:
                near.setdefault(a, []).append(b)
                near.setdefault(b, []).append(a)
            return
        same = a == b
        st = a < NI and (same or b >= NI or ext[a] >= ext[b])
        ss = b < NI and (same or a >= NI or ext[b] > ext[a])
        L_, R_ = t["left_child"], t["right_child"]
        if st and ss and same:
            rec(L_[a], L_[a]); rec(L_[a], R_[a]); rec(R_[a], R_[a])
        elif st and ss:
            for x in (L_[a], R_[a]):
                for y in (L_[b], R_[b]):
                    rec(x, y)
        elif st:
            rec(L_[a], b); rec(R_[a], b)
        elif ss:
            rec(a, L_[b]); rec(a, R_[b])
    rec(0, 0)
    return far, near


@pytest.mark.parametrize("n,leaf,theta,mac", [(800, 8, 0.6, "dehnen"), (600, 4, 0.5, "bh"),
                                             (500, 8, 0.7, "engblom")])
def test_walk_equals_recursive_restatement_and_partitions_pairs(n, leaf, theta, mac):
    pos, _ = inputs.plummer_sphere(n, 3, a=0.5)
    t = oracle.build_tree(pos, None, leaf_size=leaf)
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])
    inter, nbrs, raw = oracle.build_interactions_and_neighbors(t, g, theta=theta, mac_type=mac)
    far, near = _recursive_walk(t, g, theta, mac)
    T = t["parent"].shape[0]
    NI = t["num_internal_nodes"]
    for k in range(T):
        got = np.sort(inter["sources"][inter["offsets"][k]:inter["offsets"][k] + inter["counts"][k]])
        assert np.array_equal(got, np.sort(far.get(k, [])))
    li = nbrs["leaf_indices"]
    for r in range(T - NI):
        got = np.sort(nbrs["neighbors"][nbrs["offsets"][r]:nbrs["offsets"][r + 1]])
        assert np.array_equal(got, np.sort(near.get(int(li[r]), [])))
    # every ordered leaf pair is covered exactly once (far via an ancestor pair, or near)
    par = t["parent"]

    def chain(x):
        out = [x]
        while par[out[-1]] >= 0:
            out.append(par[out[-1]])
        return out
    fset = {(a, b) for a, lst in far.items() for b in lst}
    leaves = range(NI, T)
    chains = {x: chain(x) for x in leaves}
    for A in leaves:
        for B in leaves:
            if A == B:
                continue
            hits = sum((x, y) in fset for x in chains[A] for y in chains[B])
            isnear = B in near.get(A, [])
            assert hits + isnear == 1
    # CSR contract and level order
    assert inter["offsets"][-1] == len(inter["sources"])
    assert np.all(np.diff(inter["target_levels"]) >= 0)
    assert raw["accept_decisions"] * 2 == inter["counts"].sum()
    assert raw["near_decisions"] * 2 == nbrs["counts"].sum()


def test_capacity_ladder_errors():
    pos, _ = inputs.uniform_cube(512, 0)
    t = oracle.build_tree(pos, None, leaf_size=8)
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])

    class Cfg:
        max_pair_queue, process_block = 8192, 256
        max_interactions_per_node, max_neighbors_per_leaf = 1, 2048
    with pytest.raises(RuntimeError, match="Interaction list capacity exceeded"):
        oracle.build_interactions_and_neighbors(t, g, theta=0.6, mac_type="dehnen",
                                                traversal_config=Cfg)
    Cfg.max_interactions_per_node, Cfg.max_neighbors_per_leaf = 2048, 2
    ev = []
    a, b, raw = oracle.build_interactions_and_neighbors(t, g, theta=0.6, mac_type="dehnen",
                                                        traversal_config=Cfg, events=ev)
    full, fb, _ = oracle.build_interactions_and_neighbors(t, g, theta=0.6, mac_type="dehnen")
    assert np.array_equal(b["counts"], fb["counts"]) and ev[0][0] == "neighbor_overflow"
    with pytest.raises(RuntimeError, match="Pair queue capacity exceeded"):
        oracle.build_interactions_and_neighbors(t, g, theta=0.6, mac_type="dehnen",
                                                max_pair_queue=4, max_interactions_per_node=64)


# ---- capacity-model reference: backprop pinned by finite differences --------

def test_mlp_reference_gradient_matches_finite_differences():
    from oracle import mlp_ref
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, 9091)
    x = rng.normal(0, 1, (16, 8))
    t = rng.poisson(3.0, (16, 3)).astype(np.float64)
    m = (rng.uniform(size=(16, 3)) > 0.3).astype(np.float64)
    loss, g = mlp_ref.loss_and_grad(p, x, t, m, rounding=False)
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp, _ = mlp_ref.loss_and_grad(p + e, x, t, m, rounding=False)
        lm, _ = mlp_ref.loss_and_grad(p - e, x, t, m, rounding=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])


def test_mlp_reference_poisson_optimum_is_the_masked_mean():
    """With every weight zero the output is b4; the Poisson NLL is stationary
    exactly where 2^b4 equals the masked mean count (the MLE of a rate)."""
    from oracle import mlp_ref
    rng = np.random.default_rng(1)
    t = rng.poisson(7.0, (64, 3)).astype(np.float64)
    m = (rng.uniform(size=(64, 3)) > 0.4).astype(np.float64)
    p = np.zeros(9091)
    mean = (t * m).sum(0) / m.sum(0)
    p[-3:] = np.log2(mean)
    _, g = mlp_ref.loss_and_grad(p, rng.normal(size=(64, 8)), t, m, rounding=False)
    assert np.allclose(g[-3:], 0.0, atol=1e-12)
    p[-3:] += 0.1
    _, g = mlp_ref.loss_and_grad(p, rng.normal(size=(64, 8)), t, m, rounding=False)
    assert np.all(g[-3:] > 0)


def test_oracle_policy_walk_with_mac_policy_equals_walk():
    from oracle import oracle as O
    pos, _ = inputs.plummer_sphere(1500, 9, a=0.4)
    t = oracle.build_octree(pos, None, leaf_size=8)
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])

    def mac_policy(state, *, valid_pairs, mac_ok, different_nodes, target_leaf, source_leaf,
                   same_node, target_nodes, source_nodes, center_target, center_source, dist_sq,
                   extent_target, extent_source):
        near = valid_pairs & ~mac_ok & target_leaf & source_leaf & different_nodes
        return np.where(mac_ok, 0, np.where(near, 1, 2)), np.where(mac_ok, 1, -1)
    far, near, _ = O.walk_with_policy(t, g, 0.55, "dehnen", 1.0, mac_policy, None)
    inter, nbrs, _ = oracle.build_interactions_and_neighbors(t, g, theta=0.55, mac_type="dehnen")
    for k in range(len(far)):
        row = inter["sources"][inter["offsets"][k]:inter["offsets"][k] + inter["counts"][k]]
        assert sorted(s for s, _ in far[k]) == sorted(row.tolist())
    for r in range(len(near)):
        assert sorted(near[r]) == sorted(nbrs["neighbors"][nbrs["offsets"][r]:nbrs["offsets"][r + 1]].tolist())


def test_level_major_geometry_is_a_per_level_listing():
    """Every node appears exactly once, in the row of its own level, in
    nodes_by_level order; padding is zeros with node index -1."""
    pos, mass = inputs.plummer_sphere(3000, 4, a=0.3)
    t = oracle.build_octree(pos, mass, leaf_size=8)
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])
    lm = oracle.geometry_to_level_major(t, g)
    lvl = np.asarray(t["node_level"])
    idx = lm["node_indices"]
    valid = idx >= 0
    assert np.array_equal(np.sort(idx[valid]), np.arange(lvl.size))
    for level in range(idx.shape[0]):
        row = idx[level][idx[level] >= 0]
        assert np.all(lvl[row] == level)
        assert np.array_equal(row, np.asarray(t["nodes_by_level"])[lvl[np.asarray(
            t["nodes_by_level"])] == level])
        assert np.array_equal(lm["radii"][level][: row.size], g["radius"][row])
    assert np.all(lm["centers"][~valid] == 0) and np.all(lm["radii"][~valid] == 0)
    assert np.array_equal(lm["level_counts"], np.bincount(lvl, minlength=idx.shape[0]))


# ---- downstream layouts: dense level-major buffers and displacement classes ---

def _plummer_lists(n=3000, seed=4, leaf=8, theta=0.6):
    pos, mass = inputs.plummer_sphere(n, seed, a=0.3)
    t = oracle.build_octree(pos, mass, leaf_size=leaf)
    g = oracle.compute_tree_geometry(t, t["positions_sorted"])
    inter, _, _ = oracle.build_interactions_and_neighbors(t, g, theta=theta, mac_type="dehnen")
    return t, g, inter


def test_dense_buffers_round_trip_to_the_sparse_list():
    """Reading the dense buffers slot by slot in level-major order gives back
    the sparse CSR list exactly; every displacement is c_t - c_s within one
    float32 rounding, with the sign of target minus source."""
    t, g, inter = _plummer_lists()
    d = oracle.densify_interactions(t, g, inter)
    nodes = d["geometry"]["node_indices"]
    src, mask, cnt, disp = (d["m2l_sources"], d["m2l_mask"], d["m2l_counts"],
                            d["m2l_displacements"])
    assert src.shape[2] == max(1, int(inter["counts"].max()))
    assert np.array_equal(mask, src >= 0)
    assert np.array_equal(cnt, mask.sum(-1))
    assert np.all(mask[..., 0] | (cnt == 0))  # valid entries form a prefix
    rows_t, rows_s = [], []
    for level in range(nodes.shape[0]):
        for j in range(nodes.shape[1]):
            node = nodes[level, j]
            if node < 0:
                assert cnt[level, j] == 0
                continue
            k = int(cnt[level, j])
            rows_t.append(np.full(k, node))
            rows_s.append(src[level, j, :k])
    assert np.array_equal(np.concatenate(rows_t), inter["targets"])
    assert np.array_equal(np.concatenate(rows_s), inter["sources"])
    c = np.asarray(g["center"], np.float64)
    lv, jj, kk = np.nonzero(mask)
    exact = c[nodes[lv, jj]] - c[src[lv, jj, kk]]
    err = np.abs(disp[lv, jj, kk].astype(np.float64) - exact)
    assert np.all(err <= np.abs(exact) * 2.0 ** -24)
    assert np.all(disp[~mask] == 0)


def test_dense_buffers_getting_started_known_answer():
    """examples/getting_started.ipynb:33-52: 4 far pairs in total."""
    _, topo, geom, inter, _ = _getting_started()
    d = oracle.densify_interactions(topo, geom, inter)
    assert int(d["m2l_mask"].sum()) == 4 == int(d["m2l_counts"].sum())


def test_int32_cast_matches_numpy_astype():
    """The reference casts rint(delta / cell) with NumPy's astype(int32);
    the oracle's explicit rule must give what NumPy gives on this host, for
    small (scalar loop) and large (vector loop) arrays."""
    vals = np.array([0.0, 2.0, -3.0, 2147483647.0, -2147483648.0, 2147483648.0,
                     -2147483649.0, 1e20, -1e20, np.inf, -np.inf, np.nan, 4.0e9, -7.0])
    for arr in (vals, np.tile(vals, 1000)):
        with np.errstate(invalid="ignore"):
            expect = arr.astype(np.int32)
        assert np.array_equal(oracle.to_int32_like_numpy_cast(arr), expect)


def test_grouped_known_answer_grid():
    """Hand-computed classes: bounds [0, 8]^3, so the cell is 8 / 2^level
    (1 at level 3, 4 at level 1, 2^-37 at level 40). Covers the field order
    (target level first), the sign (target minus source), round-half-even,
    the out-of-range index, stability inside a class and the class ids."""
    INT_MIN = -2 ** 31
    topo = dict(node_level=np.array([3, 3, 3, 3, 1, 40]),
                nodes_by_level=np.arange(6), bounds_min=np.zeros(3, np.float32),
                bounds_max=np.full(3, 8.0, np.float32))
    geom = dict(center=np.array([[0.5, 0.5, 0.5], [2.5, 0.5, 0.5], [3.0, 3.0, 4.0],
                                 [-1.0, 3.5, 0.0], [2.0, 2.0, 2.0], [1.5, 0.5, 0.5]],
                                np.float32))
    tgt = np.array([1, 2, 0, 4, 5, 0, 1, 2])
    src = np.array([0, 0, 3, 0, 0, 2, 0, 1])
    r = oracle.grouped_interactions_from_pairs(topo, geom, src, tgt)
    assert r["class_keys"].tolist() == [
        [1, 3, 0, 0, 0], [3, 3, -2, -2, -4], [3, 3, 0, 2, 4], [3, 3, 2, -3, 0],
        [3, 3, 2, 0, 0], [3, 3, 2, 2, 4], [40, 3, INT_MIN, 0, 0]]
    assert r["class_offsets"].tolist() == [0, 1, 2, 3, 4, 6, 7, 8]
    assert r["class_ids"].tolist() == [4, 5, 3, 0, 6, 1, 4, 2]
    assert r["class_targets"].tolist() == [4, 0, 2, 0, 1, 1, 2, 5]
    assert r["class_sources"].tolist() == [0, 2, 1, 3, 0, 0, 0, 0]
    assert r["class_displacements"].tolist() == [
        [0, 0, 0], [-2, -2, -4], [0, 2, 4], [2, -3, 0], [2, 0, 0], [2, 2, 4],
        [-2.0 ** -6, 0, 0]]
    assert r["level_offsets"].tolist() == [0]
    empty = oracle.grouped_interactions_from_pairs(topo, geom, src[:0], tgt[:0])
    assert empty["class_keys"].shape == (0, 5) and empty["class_offsets"].tolist() == [0]


def test_grouped_classes_reconstruct_the_pairs():
    """On real far lists: classes are strictly increasing keys; each class's
    members, taken in input order through class_ids, are exactly its CSR
    slice (stable order); every member's displacement in cells is within
    1/2 of its key; class displacements are key * cell in float32."""
    t, g, inter = _plummer_lists()
    src, tgt = inter["sources"], inter["targets"]
    r = oracle.grouped_interactions_from_pairs(t, g, src, tgt, inter["level_offsets"])
    keys = [tuple(k) for k in r["class_keys"].tolist()]
    assert all(a < b for a, b in zip(keys, keys[1:]))
    off = r["class_offsets"]
    assert off[0] == 0 and off[-1] == src.size and np.all(np.diff(off) > 0)
    ids = r["class_ids"]
    lvl = np.asarray(t["node_level"])
    ext = [float(t["bounds_max"][a]) - float(t["bounds_min"][a]) for a in range(3)]
    for c in range(len(keys)):
        members = np.flatnonzero(ids == c)
        assert np.array_equal(r["class_sources"][off[c]:off[c + 1]], src[members])
        assert np.array_equal(r["class_targets"][off[c]:off[c + 1]], tgt[members])
        tl, sl = keys[c][0], keys[c][1]
        assert np.all(lvl[tgt[members]] == tl) and np.all(lvl[src[members]] == sl)
        for a in range(3):
            cell = max(ext[a] / 2.0 ** tl, 1e-12)
            delta = (g["center"][tgt[members], a].astype(np.float64)
                     - g["center"][src[members], a].astype(np.float64)) / cell
            assert np.all(np.abs(delta - keys[c][2 + a]) <= 0.5)
            want = keys[c][2 + a] * cell
            got = float(r["class_displacements"][c, a])
            assert abs(got - want) <= abs(want) * 2.0 ** -24


# ---- per-node mass moments ---------------------------------------------------

def _fsum_node(vals, s, e):
    import math
    return math.fsum(vals[s:e + 1].tolist())


def test_mass_moments_reference_sample():
    """tests/unit/test_tree_moments.py:18-40, 69-83, 135-160: three particles,
    leaf size 1: leaves hold their particle's mass and position, the root the
    total mass and the mass-weighted mean position."""
    pos = np.array([[-0.5, -0.5, -0.5], [0.1, 0.0, 0.3], [0.6, 0.4, -0.2]], np.float32)
    mass = np.array([1.0, 2.0, 3.0], np.float32)
    t = oracle.build_tree(pos, mass, (np.full(3, -1.0, np.float32), np.ones(3, np.float32)),
                          leaf_size=1)
    ps, ms = t["positions_sorted"], t["masses_sorted"]
    mm = oracle.compute_tree_mass_moments(t, ps, ms)
    ni = int(t["num_internal_nodes"])
    assert mm["mass"].dtype == np.float32 and mm["center_of_mass"].dtype == np.float32
    assert np.array_equal(mm["mass"][ni:], ms)
    assert np.allclose(mm["center_of_mass"][ni:], ps, rtol=1e-6, atol=0)
    root = int(np.flatnonzero(np.asarray(t["parent"]) == -1)[0])
    assert mm["mass"][root] == 6.0
    want = (mass[:, None].astype(np.float64) * pos).sum(0) / 6.0
    assert np.allclose(mm["center_of_mass"][root], want, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mass_moments_are_the_exact_node_sums(dtype):
    """Brute force: every node's mass and moment is math.fsum of its terms
    (the exactly rounded sum) to within one unit in the last place, and the
    centre is the moment divided by the mass in the product dtype."""
    pos, mass = inputs.plummer_sphere(3000, 8, a=0.3)
    t = oracle.build_octree(pos, mass, leaf_size=8)
    ps = t["positions_sorted"].astype(dtype)
    ms = t["masses_sorted"].astype(dtype)
    mm = oracle.compute_tree_mass_moments(t, ps, ms)
    terms = ms[:, None] * ps
    r = np.asarray(t["node_ranges"])
    for node in range(0, r.shape[0], 7):
        s, e = int(r[node, 0]), int(r[node, 1])
        exact_m = dtype(_fsum_node(ms, s, e))
        assert abs(mm["mass"][node] - exact_m) <= np.spacing(exact_m)
        for a in range(3):
            exact_x = dtype(_fsum_node(terms[:, a], s, e))
            com = mm["center_of_mass"][node, a]
            assert abs(com - exact_x / exact_m) <= 2 * np.spacing(abs(exact_x / exact_m)) + (
                np.spacing(abs(exact_x)) / exact_m)


def test_mass_moment_dtypes_follow_the_inputs():
    """mass has the masses' dtype; the centre the product dtype."""
    pos, mass = inputs.uniform_cube(200, 3)
    t = oracle.build_tree(pos, mass, leaf_size=4)
    ps, ms = t["positions_sorted"], t["masses_sorted"]
    mm = oracle.compute_tree_mass_moments(t, ps, ms.astype(np.float64))
    assert mm["mass"].dtype == np.float64 and mm["center_of_mass"].dtype == np.float64
    mm = oracle.compute_tree_mass_moments(t, ps.astype(np.float64), ms)
    assert mm["mass"].dtype == np.float32 and mm["center_of_mass"].dtype == np.float64


def test_pair_policy_descriptor_validation():
    """Descriptor fields are checked when it is made (no device needed)."""
    from yggdrax_amd import PairPolicyDescriptor as D
    d = D(accept="distance", accept_dist_sq=0.5, tag="dist_sq", tag_thresholds=(0.1, 0.2),
          tag_values=(1, 2, 3))
    c = d.to_c()
    assert (c.accept, c.near, c.tag_metric, c.num_thresholds) == (1, 0, 1, 2)
    assert list(c.tag_values)[:3] == [1, 2, 3] and abs(c.thresholds[1] - 0.2) < 1e-7
    for bad in (dict(accept="sometimes"), dict(near="far"), dict(tag="colour"),
                dict(tag="dist_sq", tag_thresholds=(0.2, 0.1), tag_values=(1, 2, 3)),
                dict(tag="dist_sq", tag_thresholds=(0.1,), tag_values=(1,)),
                dict(tag="dist_sq", tag_thresholds=tuple(range(9)), tag_values=tuple(range(10))),
                dict(tag="target_mod", tag_modulus=0)):
        with pytest.raises(ValueError):
            D(**bad)
