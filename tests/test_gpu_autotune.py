"""The autotuner in the measured path: every tuned plan is the argmin of
the candidate timings it recorded, the tuned pipeline's lists equal the
default plans' lists, and a HIP graph captured with tuning uses the plans."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuner():
    from yggdrax_amd import autotune
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    autotune.clear()
    yield autotune
    autotune.enable(False)
    autotune.clear()


def _lists_equal(a, b):
    ai, an = a
    bi, bn = b
    for x, y in ((ai.counts, bi.counts), (ai.offsets, bi.offsets), (ai.targets, bi.targets),
                 (ai.level_offsets, bi.level_offsets), (an.counts, bn.counts),
                 (an.offsets, bn.offsets)):
        assert torch.equal(x, y)
    for (o, c, x1, x2) in ((ai.offsets, ai.counts, ai.sources, bi.sources),
                           (an.offsets, an.counts, an.neighbors, bn.neighbors)):
        off, cnt = parity.np_(o), parity.np_(c)
        order = np.argsort(off[: len(cnt)], kind="stable")
        rows = np.repeat(order, cnt[order]).astype(np.int64)
        assert np.array_equal(np.sort(rows << 32 | parity.np_(x1).astype(np.int64)),
                              np.sort(rows << 32 | parity.np_(x2).astype(np.int64)))


@pytest.mark.parametrize("row_lists", ["bucket", "sort"])
def test_tuned_plans_are_the_argmin_and_keep_the_lists(tuner, monkeypatch, row_lists):
    """The walk's grid and tile are tuned in both row-list modes; the sort's
    digit width only where the lists come from the stable sort."""
    import yggdrax_amd as ygg
    from yggdrax_amd import interactions
    monkeypatch.setattr(interactions, "ROW_LISTS", row_lists)
    pos, mass = inputs.clustered(1 << 18, 21)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                      leaf_size=32, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    kw = dict(theta=0.5, mac_type="dehnen", neighbor_positions=False)
    default = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    tuner.enable(True)
    tuned = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    tuner.enable(False)
    again = ygg.build_interactions_and_neighbors(tree, geom, **kw)   # plans, no tuning
    torch.cuda.synchronize()
    plans, times = tuner.plans(), tuner.timings()
    print("[autotune] plans", plans, flush=True)
    kinds = {k.split("/")[0] + "/" + k.split("/")[1] for k in plans}
    assert kinds >= {"walk/grid", "walk/tile"}
    assert ("sort/k4" in kinds) == (row_lists == "sort")
    for key, plan in plans.items():
        assert key in times and len(times[key]) >= 3
        assert plan == min(times[key], key=times[key].get)
        assert all(t > 0 for t in times[key].values())
    _lists_equal(tuned, default)
    _lists_equal(again, default)


def test_graph_captured_with_tuning_uses_the_plans(tuner):
    from yggdrax_amd import batch
    rng = np.random.default_rng(5)
    sizes = np.round(np.exp(rng.uniform(np.log(256), np.log(8192), 40))).astype(np.int64)
    pos, mass = [], []
    for k, n in enumerate(sizes):
        p, m = inputs.plummer_sphere(int(n), 500 + k, a=float(rng.uniform(0.05, 0.5)))
        pos.append(p)
        mass.append(m)
    off = np.concatenate([[0], np.cumsum(sizes)])
    P = torch.from_numpy(np.concatenate(pos)).cuda()
    M = torch.from_numpy(np.concatenate(mass)).cuda()
    g = batch.BatchedGraph(P, M, off, leaf_size=32, theta=0.6, mac_type="dehnen", tune=True)
    T = g.lay.num_nodes
    c = T.bit_length()
    assert g.plans[f"walk/grid/n{c}"] == g.walk_grid
    assert g.plans[f"walk/tile/n{c}"] == g.walk_tile
    assert not tuner.enabled()
    bt, geom, lists = g.replay(P + 1e-4, M)
    assert not g.last_replay_fell_back
    eb = batch.build_octree_batched(P + 1e-4, M, off, leaf_size=32)
    eg = batch.compute_tree_geometry_batched(eb)
    el = batch.build_interactions_and_neighbors_batched(eb, eg, 0.6, mac_type="dehnen")
    torch.cuda.synchronize()
    for k in range(0, 40, 3):
        _lists_equal(lists.system(k), el.system(k))
