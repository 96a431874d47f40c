"""GPU parity of the downstream far-field layouts: dense level-major M2L
buffers and displacement classes, against the oracle's restatements.

Given the same far list, the HIP outputs must equal the oracle's bit for
bit. Against the oracle's own lists (whose within-row order is an
implementation detail, as in the reference) the comparison is per slot or
per class as sets, and class keys, offsets and displacements bit-exact."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from tests.test_gpu_parity import _cfg, _run_both, CASES
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

LAYOUT_CASES = ["clustered_32k_octree", "getting_started_radix", "plummer4k_engblom_octree"]


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _np_lists(inter):
    return {k: parity.np_(getattr(inter, k)) for k in
            ("counts", "offsets", "sources", "targets", "level_offsets")}


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _sorted_slots(src, disp):
    order = np.argsort(src, axis=-1, kind="stable")
    return (np.take_along_axis(src, order, -1),
            np.take_along_axis(disp, order[..., None], -2))


@pytest.mark.parametrize("name", LAYOUT_CASES)
def test_dense_interactions_match_oracle(ygg, name):
    from yggdrax_amd import dense_interactions as di
    (ref, rgeom, rint, _), (tree, _, geom, inter, _), _ = _run_both(ygg, name)
    d = di.densify_interactions(tree, geom, inter)
    torch.cuda.synchronize()
    got = {f: parity.np_(getattr(d, f)) for f in
           ("m2l_sources", "m2l_mask", "m2l_counts", "m2l_displacements")}
    same = oracle.densify_interactions(ref, rgeom, _np_lists(inter))
    for f, v in got.items():
        assert v.shape == same[f].shape, f
        assert np.array_equal(_bits(v), _bits(same[f].astype(v.dtype))), f
    assert d.sparse_interactions is inter
    lm = ygg.geometry_to_level_major(tree, geom)
    assert torch.equal(d.geometry.node_indices, lm.node_indices)
    own = oracle.densify_interactions(ref, rgeom, rint)
    assert np.array_equal(got["m2l_counts"], own["m2l_counts"])
    assert np.array_equal(got["m2l_mask"], own["m2l_mask"])
    gs, gd = _sorted_slots(got["m2l_sources"], got["m2l_displacements"])
    os_, od = _sorted_slots(own["m2l_sources"], own["m2l_displacements"])
    assert np.array_equal(gs, os_)
    assert np.array_equal(_bits(gd), _bits(od))


def _grouped_np(g):
    return {f: parity.np_(getattr(g, f)) for f in g._fields}


@pytest.mark.parametrize("name", LAYOUT_CASES)
def test_grouped_interactions_match_oracle(ygg, name):
    from yggdrax_amd import grouped_interactions as gi
    _, _, _, theta, mac, cfg = CASES[name]
    (ref, rgeom, rint, _), (tree, _, geom, inter, _), _ = _run_both(ygg, name)
    g = _grouped_np(gi.build_grouped_interactions(tree, geom, inter))
    lists = _np_lists(inter)
    same = oracle.grouped_interactions_from_pairs(ref, rgeom, lists["sources"], lists["targets"],
                                                  lists["level_offsets"])
    print("classes", same["class_keys"].shape[0], "pairs", lists["sources"].size, flush=True)
    for f, v in g.items():
        assert v.shape == same[f].shape, f
        assert np.array_equal(_bits(v), _bits(same[f].astype(v.dtype))), f
    own = oracle.grouped_interactions_from_pairs(ref, rgeom, rint["sources"], rint["targets"],
                                                 rint["level_offsets"])
    for f in ("class_keys", "class_offsets", "class_displacements", "level_offsets",
              "level_nodes"):
        assert np.array_equal(_bits(g[f]), _bits(own[f].astype(g[f].dtype))), f
    # per-class member sets: (class, source, target) triples
    n = g["class_sources"].size
    cls = np.repeat(np.arange(g["class_offsets"].size - 1), np.diff(g["class_offsets"]))
    mine = np.sort(cls * (1 << 40) + g["class_sources"] * (1 << 20) + g["class_targets"])
    exp = np.sort(cls * (1 << 40) + own["class_sources"] * (1 << 20) + own["class_targets"])
    assert n < (1 << 20) and np.array_equal(mine, exp)
    # through the public entry point
    out = ygg.build_interactions_and_neighbors(tree, geom, theta=theta, mac_type=mac,
                                               traversal_config=_cfg(cfg), return_grouped=True)
    assert len(out) == 3
    og = _grouped_np(out[2])
    assert np.array_equal(og["class_keys"], g["class_keys"])
    assert np.array_equal(og["class_offsets"], g["class_offsets"])
    out4 = ygg.build_interactions_and_neighbors(tree, geom, theta=theta, mac_type=mac,
                                                traversal_config=_cfg(cfg), return_result=True,
                                                return_compact_far_pairs=True,
                                                return_grouped=True)
    assert len(out4) == 5
    with pytest.raises(ValueError, match="requires return_interactions=True"):
        ygg.build_interactions_and_neighbors(tree, geom, theta=theta, return_grouped=True,
                                             return_interactions=False)


def test_grouped_wide_and_constant_keys(ygg):
    """Arbitrary node pairs (deep levels, far apart: indices beyond int32 and
    keys wider than 64 bits, sorted in two words) and one repeated pair (a
    single class, no sort) against the oracle."""
    from yggdrax_amd import grouped_interactions as gi
    pos, mass = inputs.clustered(1 << 14, 9)
    ref = oracle.build_tree(pos, mass, leaf_size=4)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=4, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    T = int(ref["parent"].shape[0])
    rng = np.random.default_rng(5)
    exps = []
    for src, tgt in ((rng.integers(0, T, 200003), rng.integers(0, T, 200003)),
                     (np.full(777, 3), np.full(777, T - 1))):
        print("pairs", src.size, flush=True)
        g = _grouped_np(gi.build_grouped_interactions_from_pairs(
            tree, geom, torch.from_numpy(src).cuda(), torch.from_numpy(tgt).cuda()))
        exp = oracle.grouped_interactions_from_pairs(ref, rgeom, src, tgt)
        for f, v in g.items():
            assert np.array_equal(_bits(v), _bits(exp[f].astype(v.dtype))), f
        exps.append(exp)
    k = exps[0]["class_keys"].astype(np.int64)
    bits = sum(int(k[:, i].max() - k[:, i].min()).bit_length() for i in range(5))
    assert bits > 64 and (k == -(1 << 31)).any()
    assert exps[1]["class_keys"].shape[0] == 1


def test_layouts_of_a_single_leaf_tree(ygg):
    """One leaf, no far pairs: dense buffers of one slot and one invalid
    entry; grouped buffers empty."""
    from yggdrax_amd import dense_interactions as di
    from yggdrax_amd import grouped_interactions as gi
    pos, mass = inputs.uniform_cube(10, 1)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, _ = ygg.build_interactions_and_neighbors(tree, geom)
    d = di.densify_interactions(tree, geom, inter)
    assert parity.np_(d.m2l_sources).tolist() == [[[-1]]]
    assert parity.np_(d.m2l_counts).tolist() == [[0]]
    assert not parity.np_(d.m2l_mask).any()
    g = gi.build_grouped_interactions(tree, geom, inter)
    assert tuple(g.class_keys.shape) == (0, 5) and parity.np_(g.class_offsets).tolist() == [0]
    assert g.class_ids.numel() == 0
    b = di.build_dense_interactions(tree, geom)
    assert torch.equal(b.m2l_counts, d.m2l_counts)


def test_layouts_reject_lists_outside_the_tree(ygg):
    """A far list naming a node outside the tree, or a CSR row reaching past
    its sources, is refused with an error (the kernels bounds-check every
    read and flag it) instead of being densified or grouped."""
    from yggdrax_amd import dense_interactions as di
    from yggdrax_amd import grouped_interactions as gi
    (ref, rgeom, rint, _), (tree, _, geom, inter, _), _ = _run_both(ygg, "getting_started_radix")
    T = tree.num_nodes
    bad_src = inter.sources.clone()
    bad_src[0] = T + 5
    with pytest.raises(ValueError, match="not a CSR list"):
        di.densify_interactions(tree, geom, inter._replace(sources=bad_src))
    short = inter._replace(sources=inter.sources[:-1])
    with pytest.raises(ValueError, match="not a CSR list"):
        di.densify_interactions(tree, geom, short)
    with pytest.raises(IndexError, match="out of range"):
        gi.build_grouped_interactions_from_pairs(tree, geom, bad_src, inter.targets)
    neg = inter.targets.clone()
    neg[-1] = -1
    with pytest.raises(IndexError, match="out of range"):
        gi.build_grouped_interactions_from_pairs(tree, geom, inter.sources, neg)
    # the untouched list still works after the refusals
    g = gi.build_grouped_interactions(tree, geom, inter)
    assert int(g.class_offsets[-1]) == inter.sources.numel()
