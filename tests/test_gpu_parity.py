"""GPU parity: the HIP pipeline against the CPU oracle on identical inputs.

Integer outputs (codes, permutation, topology, octree tables, counts,
offsets, per-row interaction/neighbour sets) must be identical; float32
geometry must match within 1e-6 relative (and is checked bit-exact).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


def _cfg(c):
    if c is None:
        return None
    from yggdrax_amd import DualTreeTraversalConfig
    return DualTreeTraversalConfig(*c)


class _OCfg:
    def __init__(self, c):
        self.max_pair_queue, self.process_block, self.max_interactions_per_node, \
            self.max_neighbors_per_leaf = c


CASES = {
    "c1_uniform512_octree": (lambda: inputs.uniform_cube(512, 0), "octree", 16, 0.6, "dehnen",
                             (8192, 256, 2048, 2048)),
    "getting_started_radix": (lambda: inputs.uniform_cube(256, 7), "radix", 16, 0.8, "dehnen",
                              (8192, 256, 2048, 2048)),
    "plummer_16k_radix": (lambda: inputs.plummer_sphere(1 << 14, 2), "radix", 32, 0.5, "dehnen",
                          None),
    "clustered_32k_octree": (lambda: inputs.clustered(1 << 15, 3), "octree", 64, 0.5, "dehnen",
                             None),
    "uniform4k_bh_radix": (lambda: inputs.uniform_cube(4096, 11), "radix", 8, 0.7, "bh", None),
    "plummer4k_engblom_octree": (lambda: inputs.plummer_sphere(4096, 5, a=0.3), "octree", 8, 0.5,
                                 "engblom", None),
}


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _run_both(ygg, name):
    gen, kind, leaf, theta, mac, cfg = CASES[name]
    print(f"[parity] case={name}", flush=True)
    pos, mass = gen()
    build = oracle.build_octree if kind == "octree" else oracle.build_tree
    ref = build(pos, mass, leaf_size=leaf)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    rint, rnb, _ = oracle.build_interactions_and_neighbors(
        ref, rgeom, theta=theta, mac_type=mac, traversal_config=None if cfg is None else _OCfg(cfg))
    gbuild = ygg.build_octree if kind == "octree" else ygg.build_tree
    tree, pos_sorted, _, _ = gbuild(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, pos_sorted)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=theta, mac_type=mac,
                                                       traversal_config=_cfg(cfg))
    torch.cuda.synchronize()
    return (ref, rgeom, rint, rnb), (tree, pos_sorted, geom, inter, nbrs), kind


@pytest.mark.parametrize("name", list(CASES))
def test_pipeline_parity(ygg, name):
    (ref, rgeom, rint, rnb), (tree, pos_sorted, geom, inter, nbrs), kind = _run_both(ygg, name)
    parity.assert_topology_equal(tree, ref, octree=(kind == "octree"))
    assert np.array_equal(parity.np_(pos_sorted), ref["positions_sorted"])
    parity.assert_geometry_close(geom, rgeom)
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)


def test_getting_started_known_answer(ygg):
    """examples/getting_started.ipynb:33-52: 31 nodes, 16 leaves, 4 far, 236 near."""
    pos, mass = inputs.uniform_cube(256, 7)
    tree = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(), leaf_size=16)
    ps = torch.from_numpy(pos).cuda()[tree.particle_indices]
    geom = ygg.compute_tree_geometry(tree, ps)
    cfg = ygg.DualTreeTraversalConfig(8192, 256, 2048, 2048)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=0.8, mac_type="dehnen",
                                                       traversal_config=cfg)
    assert tree.num_nodes == 31 and tree.num_leaves == 16
    assert int(inter.counts.sum()) == 4
    assert int(nbrs.counts.sum()) == 236


@pytest.mark.parametrize("n", [1 << 20])
def test_sort_large_uniform(ygg, n):
    rng = np.random.default_rng(5)
    pos = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    lo, hi = oracle.infer_bounds(pos)
    codes = oracle.morton_encode(pos, (lo, hi))
    perm = np.argsort(codes, kind="stable")
    tree = ygg.build_tree(torch.from_numpy(pos).cuda(), None, leaf_size=32)
    assert np.array_equal(parity.np_(tree.particle_indices), perm)
    assert np.array_equal(parity.np_(tree.morton_codes), codes[perm])


@pytest.mark.parametrize("key_bytes,range_bits,n", [(4, 18, 300007), (4, 20, 1 << 20),
                                                    (4, 32, 70001), (8, 63, 500009),
                                                    (8, 21, 4097)])
@pytest.mark.parametrize("digit_bits", [8, 9, 10, 11])
def test_radix_sort_every_digit_width(ygg, key_bytes, range_bits, n, digit_bits):
    """Every digit width gives numpy's stable argsort (heavy duplicates in
    the narrow-range cases exercise stability)."""
    from yggdrax_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(range_bits * 7 + digit_bits)
    hi = 1 << range_bits
    if key_bytes == 8:
        keys = rng.integers(0, hi, n, dtype=np.uint64)
        kt = torch.from_numpy(keys.view(np.int64)).cuda()
        fn = lib.ygg_sort_pairs_u64
    else:
        keys = rng.integers(0, min(hi, 1 << 32), n, dtype=np.uint64).astype(np.uint32)
        kt = torch.from_numpy(keys.view(np.int32)).cuda()
        fn = lib.ygg_sort_pairs_u32
    ko = torch.empty_like(kt)
    vo = torch.empty(n, dtype=torch.int32, device="cuda")
    ws = _lib.workspace(lib.ygg_sort_workspace_bytes(n, key_bytes, 0, range_bits, digit_bits),
                        kt.device)
    _lib.check(fn(kt.data_ptr(), None, ko.data_ptr(), vo.data_ptr(), n, 0, range_bits,
                  digit_bits, ws.data_ptr(), ws.numel(), _lib.stream_ptr(kt.device)), "sort")
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(vo.cpu().numpy().astype(np.int64), perm)
    out = ko.cpu().numpy().view(np.uint64 if key_bytes == 8 else np.uint32)
    assert np.array_equal(out, keys[perm])


def test_identical_points_stable(ygg):
    """tests/unit/test_tree.py:193-214: identical points keep input order."""
    pos = np.full((64, 3), 0.25, np.float32)
    tree = ygg.build_tree(torch.from_numpy(pos).cuda(), None, leaf_size=4)
    assert np.array_equal(parity.np_(tree.particle_indices), np.arange(64))
    ref = oracle.build_tree(pos, None, leaf_size=4)
    parity.assert_topology_equal(tree, ref)


@pytest.mark.parametrize("n,leaf", [(1, 8), (2, 1), (17, 4), (33, 32), (1000, 1)])
def test_small_sizes(ygg, n, leaf):
    rng = np.random.default_rng(n)
    pos = rng.normal(size=(n, 3)).astype(np.float32)
    ref = oracle.build_octree(pos, None, leaf_size=leaf)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), None, leaf_size=leaf,
                                      return_reordered=True)
    parity.assert_topology_equal(tree, ref, octree=True)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    geom = ygg.compute_tree_geometry(tree, ps)
    parity.assert_geometry_close(geom, rgeom)
    rint, rnb, _ = oracle.build_interactions_and_neighbors(ref, rgeom, theta=0.5,
                                                           mac_type="dehnen")
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=0.5, mac_type="dehnen")
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)


def test_overflow_errors_match(ygg):
    """Tiny capacities raise the reference's errors (queue ladder and far cap)."""
    pos, mass = inputs.uniform_cube(512, 0)
    ref = oracle.build_tree(pos, mass, leaf_size=8)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=8, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    with pytest.raises(RuntimeError, match="Interaction list capacity exceeded"):
        oracle.build_interactions_and_neighbors(ref, rgeom, theta=0.6, mac_type="dehnen",
                                                traversal_config=_OCfg((8192, 256, 1, 2048)))
    with pytest.raises(RuntimeError, match="Interaction list capacity exceeded"):
        ygg.build_interactions_and_neighbors(
            tree, geom, theta=0.6, mac_type="dehnen",
            traversal_config=ygg.DualTreeTraversalConfig(8192, 256, 1, 2048))
    # tiny neighbour capacity: both grow it through the ladder and succeed
    rint, rnb, _ = oracle.build_interactions_and_neighbors(
        ref, rgeom, theta=0.6, mac_type="dehnen", traversal_config=_OCfg((8192, 256, 2048, 2)))
    inter, nbrs = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen",
        traversal_config=ygg.DualTreeTraversalConfig(8192, 256, 2048, 2))
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)
    # explicit tiny queue without config: queue error
    with pytest.raises(RuntimeError, match="Pair queue capacity exceeded"):
        ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen",
                                             max_pair_queue=4, max_interactions_per_node=64)


def test_gpu_matches_c1_fixture(ygg):
    """The HIP path reproduces the committed C1 fixture (the reference's
    CPU-runnable configuration)."""
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "c1_uniform512.npz"), allow_pickle=False)
    pos, mass = z["positions"], z["masses"]
    tree, ps, ms, inv = ygg.build_octree(torch.from_numpy(pos).cuda(),
                                         torch.from_numpy(mass).cuda(), leaf_size=16,
                                         return_reordered=True)
    ref = {k[5:]: z[k] for k in z.files if k.startswith("topo_")}
    parity.assert_topology_equal(tree, ref, octree=True)
    assert np.array_equal(parity.np_(ps), ref["positions_sorted"])
    assert np.array_equal(parity.np_(inv), ref["inverse_permutation"])
    geom = ygg.compute_tree_geometry(tree, ps)
    parity.assert_geometry_close(geom, {k[5:]: z[k] for k in z.files if k.startswith("geom_")})
    cfg = ygg.DualTreeTraversalConfig(8192, 256, 2048, 2048)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen",
                                                       traversal_config=cfg)
    parity.assert_interactions_equal(inter, {k[4:]: z[k] for k in z.files if k.startswith("far_")})
    parity.assert_neighbors_equal(nbrs, {k[5:]: z[k] for k in z.files if k.startswith("near_")})


def test_compact_far_pairs_match_oracle_and_split_builders(ygg):
    """build_compact_far_pairs_and_leaf_neighbor_lists (tests/unit/
    test_interactions_adapter.py:147-188): the compact pairs are the walk's
    far entries in level-major order with tags -1; shared and split builders
    give the same pair sets and neighbour lists; per-target sets equal the
    oracle's."""
    pos, mass = inputs.uniform_cube(2048, 21)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=8, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    cfg = ygg.DualTreeTraversalConfig(8192, 256, 2048, 2048)
    far, nbrs = ygg.build_compact_far_pairs_and_leaf_neighbor_lists(
        tree, geom, theta=0.6, mac_type="dehnen", traversal_config=cfg)
    split = ygg.build_compact_far_pairs(tree, geom, theta=0.6, mac_type="dehnen",
                                        traversal_config=cfg)
    snb = ygg.build_leaf_neighbor_lists(tree, geom, theta=0.6, mac_type="dehnen",
                                        traversal_config=cfg)
    n = int(far.far_pair_count)
    assert far.sources.numel() == n and bool((far.tags == -1).all())

    def pairset(f):
        return sorted(zip(parity.np_(f.targets).tolist(), parity.np_(f.sources).tolist()))
    assert pairset(far) == pairset(split)
    assert torch.equal(nbrs.offsets, snb.offsets) and torch.equal(nbrs.counts, snb.counts)
    ref = oracle.build_tree(pos, mass, leaf_size=8)
    rg = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    ri, rn, _ = oracle.build_interactions_and_neighbors(ref, rg, theta=0.6, mac_type="dehnen",
                                                        traversal_config=_OCfg((8192, 256, 2048,
                                                                                2048)))
    exp = sorted(zip(np.asarray(ri["targets"]).tolist(), np.asarray(ri["sources"]).tolist()))
    assert pairset(far) == exp
    # level-major order: targets' levels never decrease
    lv = parity.np_(tree.topology.node_level)[parity.np_(far.targets)]
    assert np.all(np.diff(lv) >= 0)
    padded, _ = ygg.build_compact_far_pairs_and_leaf_neighbor_lists(
        tree, geom, theta=0.6, mac_type="dehnen", traversal_config=cfg,
        compact_far_pair_capacity=n + 5)
    assert padded.sources.numel() == n + 5 and bool((padded.sources[n:] == -1).all())
    with pytest.raises(RuntimeError, match="compact far-pair cap exceeded"):
        ygg.build_compact_far_pairs_and_leaf_neighbor_lists(
            tree, geom, theta=0.6, mac_type="dehnen", traversal_config=cfg,
            compact_far_pair_capacity=max(0, n - 1))


@pytest.mark.parametrize("name", ["clustered_32k_octree", "getting_started_radix"])
def test_geometry_to_level_major_matches_oracle(ygg, name):
    (ref, rgeom, _, _), (tree, _, geom, _, _), _ = _run_both(ygg, name)
    lm = ygg.geometry_to_level_major(tree, geom)
    rl = oracle.geometry_to_level_major(ref, rgeom)
    for f in ("centers", "half_extents", "radii", "max_extents", "level_counts",
              "node_indices"):
        got, exp = parity.np_(getattr(lm, f)), np.asarray(rl[f])
        assert got.shape == exp.shape, f
        assert np.array_equal(got, exp.astype(got.dtype)), f
