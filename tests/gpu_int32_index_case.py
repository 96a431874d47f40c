"""Runs in its own process with YGGDRAX_INDEX_PRECISION=int32 (the index
dtype is fixed at import): the pipeline, the downstream layouts and the
batched path with int32 index outputs, against the oracle. Started by
tests/test_gpu_int32_index.py; prints INT32 OK at the end."""

from __future__ import annotations

import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import oracle  # noqa: E402
from tests import parity  # noqa: E402
from yggdrax_amd import inputs  # noqa: E402


def main():
    import yggdrax_amd as ygg
    from yggdrax_amd import batch, dense_interactions as di, grouped_interactions as gi
    assert ygg.INDEX_DTYPE == torch.int32, ygg.INDEX_DTYPE
    cfg = ygg.DualTreeTraversalConfig(8192, 256, 2048, 2048)
    for name, (pos, mass), octree, leaf, theta in (
            ("c1_uniform512_octree", inputs.uniform_cube(512, 0), True, 16, 0.6),
            ("plummer4k_radix", inputs.plummer_sphere(4096, 5, a=0.3), False, 8, 0.5)):
        print("[int32] case", name, flush=True)
        ob = oracle.build_octree if octree else oracle.build_tree
        ref = ob(pos, mass, leaf_size=leaf)
        rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
        rint, rnb, _ = oracle.build_interactions_and_neighbors(
            ref, rgeom, theta=theta, mac_type="dehnen", traversal_config=cfg)
        gb = ygg.build_octree if octree else ygg.build_tree
        tree, ps, _, inv = gb(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                              leaf_size=leaf, return_reordered=True)
        assert tree.parent.dtype == torch.int32 and inv.dtype == torch.int32
        parity.assert_topology_equal(tree, ref, octree=octree)
        geom = ygg.compute_tree_geometry(tree, ps)
        parity.assert_geometry_close(geom, rgeom)
        inter, nbrs, res, compact, grouped = ygg.build_interactions_and_neighbors(
            tree, geom, theta=theta, mac_type="dehnen", traversal_config=cfg,
            return_result=True, return_compact_far_pairs=True, return_grouped=True)
        for t in (inter.sources, inter.offsets, nbrs.neighbors, nbrs.offsets, compact.sources,
                  grouped.class_offsets, grouped.class_ids):
            assert t.dtype == torch.int32, t.dtype
        parity.assert_interactions_equal(inter, rint)
        parity.assert_neighbors_equal(nbrs, rnb)
        lists = {k: parity.np_(getattr(inter, k)) for k in
                 ("counts", "offsets", "sources", "targets", "level_offsets")}
        d = di.densify_interactions(tree, geom, inter)
        exp = oracle.densify_interactions(ref, rgeom, lists)
        for f in ("m2l_sources", "m2l_mask", "m2l_counts", "m2l_displacements"):
            got = parity.np_(getattr(d, f))
            assert np.array_equal(got, exp[f].astype(got.dtype)), f
        assert d.m2l_sources.dtype == torch.int32
        g = gi.build_grouped_interactions(tree, geom, inter)
        exp = oracle.grouped_interactions_from_pairs(ref, rgeom, lists["sources"],
                                                     lists["targets"], lists["level_offsets"])
        for f in g._fields:
            got = parity.np_(getattr(g, f))
            assert np.array_equal(got, np.asarray(exp[f]).astype(got.dtype)), f
        lm = ygg.geometry_to_level_major(tree, geom)
        assert lm.node_indices.dtype == torch.int32
        assert np.array_equal(parity.np_(lm.node_indices),
                              oracle.geometry_to_level_major(ref, rgeom)["node_indices"])
    print("[int32] batched", flush=True)
    rng = np.random.default_rng(2)
    sizes = [300, 1, 2000, 77]
    pos = [inputs.plummer_sphere(n, 10 + k, a=float(rng.uniform(0.1, 0.4)))[0]
           for k, n in enumerate(sizes)]
    mass = [np.full(n, 1.0 / n, np.float32) for n in sizes]
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    bt = batch.build_octree_batched(torch.from_numpy(np.concatenate(pos)).cuda(),
                                    torch.from_numpy(np.concatenate(mass)).cuda(), offsets,
                                    leaf_size=16)
    bg = batch.compute_tree_geometry_batched(bt)
    bl = batch.build_interactions_and_neighbors_batched(bt, bg, 0.6, mac_type="dehnen")
    for k in range(len(sizes)):
        ref = oracle.build_octree(pos[k], mass[k], leaf_size=16)
        rg = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
        ri, rn, _ = oracle.build_interactions_and_neighbors(ref, rg, theta=0.6,
                                                            mac_type="dehnen")
        parity.assert_topology_equal(bt.system(k), ref, octree=True)
        inter, nbrs = bl.system(k)
        assert inter.sources.dtype == torch.int32
        parity.assert_interactions_equal(inter, ri)
        parity.assert_neighbors_equal(nbrs, rn, dense=False)
    print("INT32 OK", flush=True)


if __name__ == "__main__":
    main()
