"""Batched mode on the GPU: every system of a batch equals the oracle's
single-system result (topology, octree metadata, geometry, far and near
lists), including one-leaf systems and a malformed system (identical points)
that takes the per-system fallback."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


def _systems(sizes, seed):
    rng = np.random.default_rng(seed)
    pos, mass = [], []
    for k, n in enumerate(sizes):
        p, m = inputs.plummer_sphere(int(n), seed * 1000 + k, a=float(rng.uniform(0.05, 0.5)))
        pos.append(p)
        mass.append(m)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    return pos, mass, offsets


@pytest.mark.parametrize("octree", [True, False])
def test_batch_matches_oracle_per_system(octree):
    from yggdrax_amd import batch
    sizes = [1, 2, 33, 100, 513, 2000, 4097, 1, 64, 700]
    pos, mass, offsets = _systems(sizes, 3)
    pos[6] = np.full_like(pos[6], 0.25)  # identical points: malformed tree, fallback path
    P = torch.from_numpy(np.concatenate(pos)).cuda()
    M = torch.from_numpy(np.concatenate(mass)).cuda()
    leaf, theta = 16, 0.6
    build = batch.build_octree_batched if octree else batch.build_tree_batched
    bt = build(P, M, offsets, leaf_size=leaf)
    assert 6 in bt.fallback
    geom = batch.compute_tree_geometry_batched(bt)
    lists = batch.build_interactions_and_neighbors_batched(bt, geom, theta, mac_type="dehnen")
    ob = oracle.build_octree if octree else oracle.build_tree
    for k in range(len(sizes)):
        print("system", k, "n", sizes[k], flush=True)
        ref = ob(pos[k], mass[k], leaf_size=leaf)
        t = bt.system(k)
        parity.assert_topology_equal(t, ref, octree=octree)
        p0, p1 = offsets[k], offsets[k + 1]
        assert np.array_equal(parity.np_(bt.positions_sorted[p0:p1]), ref["positions_sorted"])
        rg = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
        parity.assert_geometry_close(geom.system(k), rg)
        ri, rn, _ = oracle.build_interactions_and_neighbors(ref, rg, theta=theta,
                                                            mac_type="dehnen")
        inter, nbrs = lists.system(k)
        parity.assert_interactions_equal(inter, ri)
        parity.assert_neighbors_equal(nbrs, rn, dense=False)


def test_batch_matches_single_system_path_c4_like():
    """64 Plummer systems with log-uniform sizes: each system's lists equal the
    single-system GPU path's (itself pinned to the oracle)."""
    import yggdrax_amd as ygg
    from yggdrax_amd import batch
    rng = np.random.default_rng(4)
    sizes = np.round(np.exp(rng.uniform(np.log(256), np.log(16384), 64))).astype(np.int64)
    pos, mass, offsets = _systems(sizes, 4)
    P = torch.from_numpy(np.concatenate(pos)).cuda()
    M = torch.from_numpy(np.concatenate(mass)).cuda()
    bt = batch.build_octree_batched(P, M, offsets, leaf_size=32)
    geom = batch.compute_tree_geometry_batched(bt)
    lists = batch.build_interactions_and_neighbors_batched(bt, geom, 0.6, mac_type="dehnen")
    for k in range(0, 64, 7):
        t, ps, _, _ = ygg.build_octree(torch.from_numpy(pos[k]).cuda(),
                                       torch.from_numpy(mass[k]).cuda(), leaf_size=32,
                                       return_reordered=True)
        g = ygg.compute_tree_geometry(t, ps)
        si, sn = ygg.build_interactions_and_neighbors(t, g, theta=0.6, mac_type="dehnen",
                                                      neighbor_positions=False)
        bi, bn = lists.system(k)
        for a, b in ((si.counts, bi.counts), (si.offsets, bi.offsets), (si.targets, bi.targets),
                     (si.level_offsets, bi.level_offsets), (si.target_levels, bi.target_levels),
                     (sn.counts, bn.counts), (sn.offsets, bn.offsets),
                     (sn.leaf_indices, bn.leaf_indices)):
            assert torch.equal(a, b)
        for (o, c, x1, x2) in ((si.offsets, si.counts, si.sources, bi.sources),
                               (sn.offsets, sn.counts, sn.neighbors, bn.neighbors)):
            off, cnt = parity.np_(o), parity.np_(c)
            order = np.argsort(off[: len(cnt)], kind="stable")
            rows = np.repeat(order, cnt[order]).astype(np.int64)
            k1 = np.sort(rows << 32 | parity.np_(x1).astype(np.int64))
            k2 = np.sort(rows << 32 | parity.np_(x2).astype(np.int64))
            assert np.array_equal(k1, k2)


def _assert_batches_equal(a, b, systems):
    for k in systems:
        ai, an = a.system(k)
        bi, bn = b.system(k)
        for x, y in ((ai.counts, bi.counts), (ai.offsets, bi.offsets), (ai.targets, bi.targets),
                     (ai.level_offsets, bi.level_offsets), (an.counts, bn.counts),
                     (an.offsets, bn.offsets)):
            assert torch.equal(x, y), k
        for (o, c, x1, x2) in ((ai.offsets, ai.counts, ai.sources, bi.sources),
                               (an.offsets, an.counts, an.neighbors, bn.neighbors)):
            off, cnt = parity.np_(o), parity.np_(c)
            order = np.argsort(off[: len(cnt)], kind="stable")
            rows = np.repeat(order, cnt[order]).astype(np.int64)
            k1 = np.sort(rows << 32 | parity.np_(x1).astype(np.int64))
            k2 = np.sort(rows << 32 | parity.np_(x2).astype(np.int64))
            assert np.array_equal(k1, k2), k


@pytest.mark.parametrize("records", [True, False], ids=["records", "entries"])
def test_batched_graph_replays_match_eager(monkeypatch, records):
    """The HIP-graph capture of the batched pipeline, replayed several times
    on new positions of the same layout (eager work and synchronisation in
    between), gives every replay the eager pipeline's lists; with the walk
    writing pair records (the default) and writing entries."""
    from yggdrax_amd import batch
    from yggdrax_amd import interactions as I
    monkeypatch.setattr(I, "PAIR_RECORDS", records)
    print(f"[batch] graph replays records={records}", flush=True)
    rng = np.random.default_rng(6)
    sizes = np.round(np.exp(rng.uniform(np.log(256), np.log(8192), 48))).astype(np.int64)
    pos, mass, offsets = _systems(sizes, 6)
    P = torch.from_numpy(np.concatenate(pos)).cuda()
    M = torch.from_numpy(np.concatenate(mass)).cuda()
    g = batch.BatchedGraph(P, M, offsets, leaf_size=32, theta=0.6, mac_type="dehnen",
                           slack=1.3)
    assert g.pair_records == records
    for step in range(3):
        Pn = P + torch.from_numpy(rng.normal(0, 1e-3, tuple(P.shape)).astype(np.float32)).cuda()
        bt, geom, lists = g.replay(Pn, M)
        print("replay", step, "fell back", g.last_replay_fell_back, flush=True)
        assert not g.last_replay_fell_back
        eb = batch.build_octree_batched(Pn, M, offsets, leaf_size=32)
        eg = batch.compute_tree_geometry_batched(eb)
        el = batch.build_interactions_and_neighbors_batched(eb, eg, 0.6, mac_type="dehnen")
        torch.cuda.synchronize()
        assert torch.equal(bt.sorted_codes, eb.sorted_codes)
        assert torch.equal(geom.radius, eg.radius)
        _assert_batches_equal(lists, el, range(0, len(sizes), 5))
