"""Comparison helpers: GPU results (torch) against the CPU oracle (numpy)."""

from __future__ import annotations

import numpy as np
import torch


def np_(x):
    if isinstance(x, torch.Tensor):
        if x.dtype == torch.uint64:
            return x.view(torch.int64).cpu().numpy().view(np.uint64)
        return x.cpu().numpy()
    return np.asarray(x)


RADIX_FIELDS = ("parent", "left_child", "right_child", "left_is_leaf", "right_is_leaf",
                "particle_indices", "morton_codes", "node_ranges", "node_level",
                "level_offsets", "nodes_by_level", "leaf_codes")
OCT_FIELDS = ("oct_parent", "oct_children", "oct_child_counts", "oct_child_mask",
              "oct_valid_mask", "oct_node_codes", "oct_node_depths", "oct_node_ranges",
              "oct_nodes_by_level", "oct_level_offsets", "oct_leaf_mask", "oct_leaf_nodes",
              "radix_node_to_oct", "radix_leaf_to_oct")


def assert_topology_equal(tree, ref, octree=False):
    topo = tree.topology
    for f in RADIX_FIELDS + (OCT_FIELDS if octree else ()):
        got = np_(getattr(topo, f))
        exp = np.asarray(ref[f])
        assert got.shape == exp.shape, (f, got.shape, exp.shape)
        if got.dtype == np.bool_ or exp.dtype == np.bool_:
            got, exp = got.astype(bool), exp.astype(bool)
        assert np.array_equal(got, exp.astype(got.dtype)), f"{f} differs"
    assert int(topo.num_levels) == int(ref["num_levels"])
    if octree:
        assert int(topo.oct_num_levels) == int(ref["oct_num_levels"])
    assert np.array_equal(np_(topo.bounds_min).view(np.uint32),
                          np.asarray(ref["bounds_min"], np.float32).view(np.uint32))
    assert np.array_equal(np_(topo.bounds_max).view(np.uint32),
                          np.asarray(ref["bounds_max"], np.float32).view(np.uint32))


def assert_geometry_close(geom, ref, exact=True):
    for f in ("center", "half_extent", "radius", "max_extent"):
        got = np_(getattr(geom, f))
        exp = np.asarray(ref[f], np.float32)
        np.testing.assert_allclose(got, exp, rtol=1e-6, atol=0, err_msg=f)
        if exact:
            assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"{f} not bit-exact"


def _rows(offsets, counts, values):
    return [np.sort(values[o:o + c]) for o, c in zip(offsets, counts)]


def assert_interactions_equal(inter, ref):
    counts = np_(inter.counts)
    assert np.array_equal(counts, ref["counts"]), "far counts differ"
    assert np.array_equal(np_(inter.offsets), ref["offsets"]), "far offsets differ"
    assert np.array_equal(np_(inter.level_offsets), ref["level_offsets"]), "level offsets differ"
    assert np.array_equal(np_(inter.targets), ref["targets"]), "far targets differ"
    assert np.array_equal(np_(inter.target_levels), ref["target_levels"]), "target levels differ"
    src = np_(inter.sources)
    off = ref["offsets"][:-1]
    for node, (a, b) in enumerate(zip(_rows(off, counts, src), _rows(off, counts, ref["sources"]))):
        assert np.array_equal(a, b), f"far row {node} differs"


def assert_neighbors_equal(nbrs, ref, dense=True):
    counts = np_(nbrs.counts)
    assert np.array_equal(counts, ref["counts"]), "near counts differ"
    assert np.array_equal(np_(nbrs.offsets), ref["offsets"]), "near offsets differ"
    assert np.array_equal(np_(nbrs.leaf_indices), ref["leaf_indices"])
    assert np.array_equal(np_(nbrs.particle_order_leaf_indices), ref["particle_order_leaf_indices"])
    assert np.array_equal(np_(nbrs.particle_order_to_native_leaf),
                          ref["particle_order_to_native_leaf"])
    got = np_(nbrs.neighbors)
    off = ref["offsets"][:-1]
    for r, (a, b) in enumerate(zip(_rows(off, counts, got), _rows(off, counts, ref["neighbors"]))):
        assert np.array_equal(a, b), f"near row {r} differs"
    if not dense:
        return
    nlp = np_(nbrs.neighbor_leaf_positions)
    assert nlp.shape == ref["neighbor_leaf_positions"].shape
    assert np.array_equal(np.sort(nlp, axis=1), np.sort(ref["neighbor_leaf_positions"], axis=1))
