"""Fallback paths that a capacity can force from the host, every kernel
exiting normally on its own bounds checks: an undersized far or near hint
(single system and batch) grows its stream during the walk, an undersized
queue hint reruns the walk without the hint, and a HIP-graph replay whose
captured capacities or generation count are too small returns the eager
pipeline's result. Outputs never depend on the capacity."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

TINY_HINT = {"queue": 4, "far": 1, "near": 1}


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _assert_lists_equal(a, b):
    ai, an = a
    bi, bn = b
    for x, y in ((ai.counts, bi.counts), (ai.offsets, bi.offsets), (ai.targets, bi.targets),
                 (ai.level_offsets, bi.level_offsets), (ai.target_levels, bi.target_levels),
                 (an.counts, bn.counts), (an.offsets, bn.offsets),
                 (an.leaf_indices, bn.leaf_indices)):
        assert torch.equal(x, y)
    for (o, c, x1, x2) in ((ai.offsets, ai.counts, ai.sources, bi.sources),
                           (an.offsets, an.counts, an.neighbors, bn.neighbors)):
        off, cnt = parity.np_(o), parity.np_(c)
        order = np.argsort(off[: len(cnt)], kind="stable")
        rows = np.repeat(order, cnt[order]).astype(np.int64)
        k1 = np.sort(rows << 32 | parity.np_(x1).astype(np.int64))
        k2 = np.sort(rows << 32 | parity.np_(x2).astype(np.int64))
        assert np.array_equal(k1, k2)


@pytest.mark.parametrize("hint,outcome", [
    (TINY_HINT, "hint_overflows"),
    ({"queue": 1 << 20, "far": 1, "near": 1 << 20}, "hint_growths"),
    ({"queue": 1 << 20, "far": 1 << 22, "near": 1}, "hint_growths"),
    ({"queue": 8, "far": 1 << 22, "near": 1 << 22}, "hint_overflows")])
def test_undersized_hint_falls_back_single_system(ygg, hint, outcome):
    from yggdrax_amd.interactions import FLAT_STATS
    print(f"[fallback] single hint={hint}", flush=True)
    pos, mass = inputs.clustered(1 << 15, 7)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                      leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    kw = dict(theta=0.5, mac_type="dehnen")
    plain = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    before = dict(FLAT_STATS)
    hinted = ygg.build_interactions_and_neighbors(tree, geom, capacity_hint=hint, **kw)
    torch.cuda.synchronize()
    other = "hint_growths" if outcome == "hint_overflows" else "hint_overflows"
    assert FLAT_STATS[outcome] == before[outcome] + 1
    assert FLAT_STATS[other] == before[other]
    if outcome == "hint_growths":
        assert FLAT_STATS["stream_growths"] > before["stream_growths"]
        assert FLAT_STATS["plain_walks"] == before["plain_walks"]
    else:  # the walk reran without the hint
        assert FLAT_STATS["plain_walks"] == before["plain_walks"] + 1
    _assert_lists_equal(hinted, plain)
    parity.assert_neighbors_equal(hinted[1], {k: parity.np_(getattr(plain[1], k)) for k in
                                              plain[1]._fields}, dense=True)


def _batch(num, seed, nmax=4096):
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(64), np.log(nmax), num))).astype(np.int64)
    pos, mass = [], []
    for k, n in enumerate(sizes):
        p, m = inputs.plummer_sphere(int(n), seed * 1000 + k, a=float(rng.uniform(0.05, 0.5)))
        pos.append(p)
        mass.append(m)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    return (torch.from_numpy(np.concatenate(pos)).cuda(),
            torch.from_numpy(np.concatenate(mass)).cuda(), offsets)


@pytest.mark.parametrize("hint,outcome", [
    (TINY_HINT, "hint_overflows"),
    ({"queue": 1 << 22, "far": 1, "near": 1}, "hint_growths")])
def test_undersized_hint_falls_back_batched(ygg, hint, outcome):
    """A queue hint below the batch's root count is raised to it and then
    overflows, so the batch reruns without the hint; one-entry far and near
    hints with room in the queue grow during the walk."""
    from yggdrax_amd import batch
    from yggdrax_amd.interactions import FLAT_STATS
    print(f"[fallback] batched hint={hint}", flush=True)
    P, M, offsets = _batch(24, 11)
    bt = batch.build_octree_batched(P, M, offsets, leaf_size=32)
    geom = batch.compute_tree_geometry_batched(bt)
    plain = batch.build_interactions_and_neighbors_batched(bt, geom, 0.6, mac_type="dehnen")
    h0 = FLAT_STATS[outcome]
    hinted = batch.build_interactions_and_neighbors_batched(bt, geom, 0.6, mac_type="dehnen",
                                                            capacity_hint=hint)
    torch.cuda.synchronize()
    assert FLAT_STATS[outcome] == h0 + 1
    for k in range(24):
        _assert_lists_equal(hinted.system(k), plain.system(k))


@pytest.mark.parametrize("override", [dict(capacities={"far": 4096, "near": 4096,
                                                       "queue": 1 << 16}),
                                      dict(capacities={"far": 1 << 24, "near": 1 << 24,
                                                       "queue": 64}),
                                      dict(generations=3)])
def test_graph_replay_with_small_capacities_falls_back(ygg, override):
    from yggdrax_amd import batch
    print(f"[fallback] graph {override}", flush=True)
    P, M, offsets = _batch(32, 12)
    g = batch.BatchedGraph(P, M, offsets, leaf_size=32, theta=0.6, mac_type="dehnen",
                           **override)
    rng = np.random.default_rng(1)
    Pn = P + torch.from_numpy(rng.normal(0, 1e-3, tuple(P.shape)).astype(np.float32)).cuda()
    bt, geom, lists = g.replay(Pn, M)
    assert g.last_replay_fell_back
    eb = batch.build_octree_batched(Pn, M, offsets, leaf_size=32)
    eg = batch.compute_tree_geometry_batched(eb)
    el = batch.build_interactions_and_neighbors_batched(eb, eg, 0.6, mac_type="dehnen")
    torch.cuda.synchronize()
    assert torch.equal(bt.sorted_codes, eb.sorted_codes)
    assert torch.equal(geom.radius, eg.radius)
    for k in range(32):
        _assert_lists_equal(lists.system(k), el.system(k))


def test_graph_without_masses_refuses_masses(ygg):
    from yggdrax_amd import batch
    P, _, offsets = _batch(8, 13, nmax=1024)
    g = batch.BatchedGraph(P, None, offsets, leaf_size=32, theta=0.6, mac_type="dehnen")
    g.replay(P)
    with pytest.raises(ValueError, match="without masses"):
        g.replay(P, torch.ones(P.shape[0], device=P.device))
