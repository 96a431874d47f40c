"""Stream sizing of the walk. The default, one resumable emission walk, is
run from its default capacities and from starting capacities small enough
that the walk pauses and grows its far stream, its near stream or both many
times; the reference's count pass (a counting walk, then the emission walk
at the exact totals) is run beside them. Each runs with the walk writing
pair records (one per accepted or near pair, the default) and writing the
two entries of each pair. Every variant gives the oracle's lists, single
system and batch, and only the resumable walks grow."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

CASES = {
    "clustered_32k_octree_leaf16": (lambda: inputs.clustered(1 << 15, 3), "octree", 16, 0.5,
                                    "dehnen"),
    "plummer_16k_radix_leaf8": (lambda: inputs.plummer_sphere(1 << 14, 2), "radix", 8, 0.6,
                                "bh"),
}

# starting (far, near) capacities; None = the default ones
STARTS = {"default": None, "both_tiny": (1, 1), "far_tiny": (64, 1 << 24),
          "near_tiny": (1 << 24, 64), "count_pass": "count"}


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


@pytest.fixture(scope="module")
def refs():
    out = {}
    for name, (gen, kind, leaf, theta, mac) in CASES.items():
        pos, mass = gen()
        build = oracle.build_octree if kind == "octree" else oracle.build_tree
        ref = build(pos, mass, leaf_size=leaf)
        rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
        rint, rnb, _ = oracle.build_interactions_and_neighbors(ref, rgeom, theta=theta,
                                                                mac_type=mac)
        out[name] = (pos, mass, rint, rnb)
    return out


def _sizing(monkeypatch, start):
    from yggdrax_amd import interactions as I
    s = STARTS[start]
    if s == "count":
        monkeypatch.setattr(I, "STREAM_SIZING", "count")
    else:
        monkeypatch.setattr(I, "STREAM_SIZING", "grow")
        if s is not None:
            monkeypatch.setattr(I, "_start_capacities", lambda T, L: s)


@pytest.mark.parametrize("records", [True, False], ids=["records", "entries"])
@pytest.mark.parametrize("start", list(STARTS))
@pytest.mark.parametrize("name", list(CASES))
def test_stream_sizing_matches_oracle(ygg, refs, monkeypatch, name, start, records):
    from yggdrax_amd import interactions as I
    _, kind, leaf, theta, mac = CASES[name]
    pos, mass, rint, rnb = refs[name]
    print(f"[grow] case={name} start={start} records={records}", flush=True)
    monkeypatch.setattr(I, "PAIR_RECORDS", records)
    gbuild = ygg.build_octree if kind == "octree" else ygg.build_tree
    tree, ps, _, _ = gbuild(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                            leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    _sizing(monkeypatch, start)
    before = dict(I.FLAT_STATS)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=theta, mac_type=mac)
    torch.cuda.synchronize()
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)
    walk = dict(I.LAST_WALK)
    assert walk["stream_bytes_per_entry"] == (4 if records else 8)
    growths = I.FLAT_STATS["stream_growths"] - before["stream_growths"]
    counts = I.FLAT_STATS["count_walks"] - before["count_walks"]
    print(f"[grow] {walk} growths={growths} count_walks={counts}", flush=True)
    assert walk["far_total"] == int(rint["counts"].sum())
    assert walk["near_total"] == int(rnb["counts"].sum())
    if start == "count_pass":
        assert walk["sizing"] == "count" and counts == 1 and growths == 0
        assert walk["far_capacity"] == walk["far_total"]
    else:
        assert walk["sizing"] == "grow" and counts == 0
        assert walk["stream_growths"] == growths
        if start == "default":
            assert growths == 0
        else:
            assert growths >= 2
        # the capacity covers the entries written and the worst case of the
        # generation that last grew it
        assert walk["far_capacity"] >= walk["far_total"]
        assert walk["near_capacity"] >= walk["near_total"]


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


def test_batched_stream_sizings_agree(ygg, monkeypatch):
    """The forest walk with tiny starting capacities (many growths), the
    default ones and the count pass: every system's lists are the same."""
    from yggdrax_amd import batch
    from yggdrax_amd import interactions as I
    print("[grow] batched 40 systems", flush=True)
    P, M, offsets = _batch(40, 21)
    bt = batch.build_octree_batched(P, M, offsets, leaf_size=16)
    geom = batch.compute_tree_geometry_batched(bt)
    outs = {}
    for start in ("default", "both_tiny", "count_pass", "entries"):
        with monkeypatch.context() as m:
            _sizing(m, "default" if start == "entries" else start)
            # pair records unless the walk is asked for entries
            m.setattr(I, "PAIR_RECORDS", start != "entries")
            g0 = I.FLAT_STATS["stream_growths"]
            outs[start] = batch.build_interactions_and_neighbors_batched(bt, geom, 0.6,
                                                                         mac_type="dehnen")
            torch.cuda.synchronize()
            grew = I.FLAT_STATS["stream_growths"] - g0
            assert (grew >= 2) if start == "both_tiny" else grew == 0
    for k in range(40):
        ref = outs["count_pass"].system(k)
        for start in ("default", "both_tiny", "entries"):
            got = outs[start].system(k)
            for a, b in zip(ref, got):
                for f in ("counts", "offsets", "targets", "level_offsets", "leaf_indices"):
                    if hasattr(a, f):
                        assert torch.equal(getattr(a, f), getattr(b, f)), (k, start, f)
            for (a, b, f) in ((ref[0], got[0], "sources"), (ref[1], got[1], "neighbors")):
                off, cnt = parity.np_(a.offsets), parity.np_(a.counts)
                order = np.argsort(off[: len(cnt)], kind="stable")
                rows = np.repeat(order, cnt[order]).astype(np.int64)
                ka = np.sort(rows << 32 | parity.np_(getattr(a, f)).astype(np.int64))
                kb = np.sort(rows << 32 | parity.np_(getattr(b, f)).astype(np.int64))
                assert np.array_equal(ka, kb), (k, start, f)


def test_descriptor_walk_grows_its_tag_stream(ygg, monkeypatch):
    """A pair-policy descriptor with tags (opening ratio buckets) from tiny
    starting capacities: the far, tag and near streams grow together and the
    (target, source, tag) multiset and near lists equal the count pass's."""
    from yggdrax_amd import interactions as I
    print("[grow] descriptor tags 16k", flush=True)
    desc = ygg.PairPolicyDescriptor(accept="mac", near="leaf_pairs", tag="opening_ratio",
                                    tag_thresholds=(0.01, 0.09), tag_values=(2, 4, 6))
    pos, mass = inputs.plummer_sphere(1 << 14, 9)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    outs = {}
    for start in ("both_tiny", "count_pass"):
        with monkeypatch.context() as m:
            _sizing(m, start)
            g0 = I.FLAT_STATS["stream_growths"]
            outs[start] = ygg.build_interactions_and_neighbors(
                tree, geom, theta=0.6, mac_type="dehnen", pair_policy=desc, return_result=True)
            torch.cuda.synchronize()
            grew = I.FLAT_STATS["stream_growths"] - g0
            assert (grew >= 2) if start == "both_tiny" else grew == 0
    (gi, gn, gres), (ci, cn, cres) = outs["both_tiny"], outs["count_pass"]
    assert torch.equal(gi.counts, ci.counts) and torch.equal(gi.offsets, ci.offsets)
    assert torch.equal(gn.counts, cn.counts) and torch.equal(gn.offsets, cn.offsets)
    T = tree.num_nodes

    def far_key(r):
        return torch.sort((r.interaction_targets.long() * T + r.interaction_sources.long()) * 16
                          + r.interaction_tags.long())[0]
    assert torch.equal(far_key(gres), far_key(cres))
    assert bool((gres.interaction_tags >= 2).all())
    L = gn.counts.numel()
    rows = torch.repeat_interleave(torch.arange(L, device=gn.counts.device), gn.counts)
    assert torch.equal(torch.sort(rows * T + gn.neighbors.long())[0],
                       torch.sort(rows * T + cn.neighbors.long())[0])
    assert (gres.accept_decisions, gres.near_decisions, gres.refine_decisions) == (
        cres.accept_decisions, cres.near_decisions, cres.refine_decisions)
