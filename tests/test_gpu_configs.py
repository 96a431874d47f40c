"""Full-size configurations against the oracle, field by field, through the
digest fixtures the oracle wrote on the CPU (tests/golden/make_fullsize.py):

* 2^20 Plummer points, build_tree leaf 32, dehnen theta 0.5;
* 1024 Plummer systems in one batch (build_octree leaf 32, dehnen theta 0.6),
  every system; its HIP-graph replay on perturbed positions, against the
  oracle run here on a sample of systems;
* the learned-capacity 256-system batch (same parameters), every system.

The 2^24 clustered workload (and its learned-capacity run) is in
test_gpu_fullsize.py."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from tests.golden import fullsize as F
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _check(got, fixture, k):
    bad = F.compare(got, fixture, k)
    assert not bad, f"system {k}: digests differ for {bad}"


def test_c2_plummer_2p20_matches_oracle(ygg):
    print("[configs] c2 plummer 2^20", flush=True)
    fx = F.load(os.path.join(GOLDEN, "c2_plummer_2p20.npz"))
    pos, mass = inputs.plummer_sphere(1 << 20, 2)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=32, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.5, mac_type="dehnen", neighbor_positions=False, return_result=True)
    torch.cuda.synchronize()
    dec = (res.accept_decisions, res.near_decisions, res.refine_decisions)
    _check(F.gpu_system(tree, geom, inter, nbrs, octree=False, decisions=dec), fx, 0)


def _batch_digests(ygg, bt, geom, lists, num):
    out = []
    for k in range(num):
        t = bt.system(k)
        inter, nbrs = lists.system(k)
        out.append(F.gpu_system(t, geom.system(k), inter, nbrs, octree=True))
    return out


def test_c4_batch_1024_matches_oracle_every_system(ygg):
    from yggdrax_amd import batch
    print("[configs] c4 1024 systems", flush=True)
    fx = F.load(os.path.join(GOLDEN, "c4_plummer_batch_1024.npz"))
    pos, mass, off = inputs.plummer_batch(1024, 4)
    P, M = torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda()
    bt = batch.build_octree_batched(P, M, off, leaf_size=32)
    geom = batch.compute_tree_geometry_batched(bt)
    lists = batch.build_interactions_and_neighbors_batched(bt, geom, 0.6, mac_type="dehnen")
    torch.cuda.synchronize()
    for k, got in enumerate(_batch_digests(ygg, bt, geom, lists, 1024)):
        _check(got, fx, k)


def test_c4_graph_replay_on_new_positions_matches_oracle(ygg):
    """The captured pipeline replayed on perturbed positions (capacities from
    the capture run): sampled systems equal the oracle on the new positions,
    and every system equals the eager batched pipeline."""
    from yggdrax_amd import batch
    from tests.golden.fullsize import oracle_system
    print("[configs] c4 graph replay", flush=True)
    pos, mass, off = inputs.plummer_batch(1024, 4)
    P, M = torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda()
    g = batch.BatchedGraph(P, M, off, leaf_size=32, theta=0.6, mac_type="dehnen", slack=1.2)
    rng = np.random.default_rng(44)
    sample = sorted(set(rng.integers(0, 1024, 12).tolist()) | {0, 1023})
    for step in range(2):
        noise = rng.normal(0, 1e-4, pos.shape).astype(np.float32)
        new = pos + noise
        bt, geom, lists = g.replay(torch.from_numpy(new).cuda(), M)
        torch.cuda.synchronize()
        print(f"[configs] replay {step} fell back {g.last_replay_fell_back}", flush=True)
        assert not g.last_replay_fell_back
        got = _batch_digests(ygg, bt, geom, lists, 1024)
        eb = batch.build_octree_batched(torch.from_numpy(new).cuda(), M, off, leaf_size=32)
        eg = batch.compute_tree_geometry_batched(eb)
        el = batch.build_interactions_and_neighbors_batched(eb, eg, 0.6, mac_type="dehnen")
        eager = _batch_digests(ygg, eb, eg, el, 1024)
        for k in range(1024):
            assert got[k] == eager[k], k
        for k in sample:
            a, b = int(off[k]), int(off[k + 1])
            exp = oracle_system(new[a:b], mass[a:b], octree=True, leaf=32, theta=0.6)
            bad = [f for f in exp if f in got[k] and int(got[k][f]) != int(exp[f])]
            assert not bad, (step, k, bad)


def test_c5_learned_batch_256_matches_oracle_every_system(ygg):
    from yggdrax_amd import batch, learned
    from yggdrax_amd.interactions import FLAT_STATS
    print("[configs] c5 learned 256 systems", flush=True)
    fx = F.load(os.path.join(GOLDEN, "c5_plummer_batch_256.npz"))
    model = learned.CapacityModel.load()
    pos, mass, off = inputs.plummer_batch(256, 5)
    P, M = torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda()
    bt = batch.build_octree_batched(P, M, off, leaf_size=32)
    geom = batch.compute_tree_geometry_batched(bt)
    hint = model.capacity_hint_batched(bt, geom)
    before = dict(FLAT_STATS)
    lists = batch.build_interactions_and_neighbors_batched(bt, geom, 0.6, mac_type="dehnen",
                                                           capacity_hint=hint)
    torch.cuda.synchronize()
    assert FLAT_STATS["hinted_walks"] == before["hinted_walks"] + 1
    print(f"[configs] c5 hint {hint} overflows "
          f"{FLAT_STATS['hint_overflows'] - before['hint_overflows']} growths "
          f"{FLAT_STATS['hint_growths'] - before['hint_growths']}", flush=True)
    for k in ("count_walks", "plain_walks", "hint_overflows", "hint_growths"):
        assert FLAT_STATS[k] == before[k], f"the learned hint did not hold ({k})"
    for k, got in enumerate(_batch_digests(ygg, bt, geom, lists, 256)):
        _check(got, fx, k)
