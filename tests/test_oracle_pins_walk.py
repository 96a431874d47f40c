"""CPU pins of the oracle's retry driver and raw walk result: the event
sequence follows from the capacity ladders and the walk's own statistics,
and the raw result converts back to the interaction lists by the
reference's own consumer (_result_to_interactions, 3561-3637)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import oracle as O
from tests import oracle_checks as C
from yggdrax_amd import inputs


class _Cfg:
    def __init__(self, q, pb, k, kn):
        self.max_pair_queue, self.process_block = q, pb
        self.max_interactions_per_node, self.max_neighbors_per_leaf = k, kn


def _system(kind, gen, leaf, theta, mac):
    pos, mass = gen()
    t = (oracle.build_octree if kind == "octree" else oracle.build_tree)(pos, mass,
                                                                        leaf_size=leaf)
    return t, oracle.compute_tree_geometry(t, t["positions_sorted"]), theta, mac


RETRY = {
    "queue": (("octree", lambda: inputs.plummer_sphere(4096, 5, a=0.3), 8, 0.5, "dehnen"),
              _Cfg(64, 4, 2048, 2048)),
    "neighbor": (("radix", lambda: inputs.uniform_cube(4096, 11), 8, 0.6, "bh"),
                 _Cfg(1 << 16, 256, 2048, 2)),
    "queue_and_neighbor": (("octree", lambda: inputs.clustered(1 << 13, 3), 8, 0.5, "engblom"),
                           _Cfg(64, 4, 2048, 3)),
}


@pytest.mark.parametrize("name", sorted(RETRY))
def test_retry_events_follow_the_ladders(name):
    """Expected events from the ladders and an independent plain-Python
    generation walk: each attempt stops after the first generation that
    raised a flag, the queue flag is checked first, then far (raise), then
    near (next neighbour capacity); success is reported only after a retry;
    attempts count from 1 (3917-3958, 4306-4405). The stored pair counts of
    every event are checked too."""
    sysargs, cfg = RETRY[name]
    t, g, theta, mac = _system(*sysargs)
    T = t["parent"].shape[0]
    NI = t["num_internal_nodes"]
    L = T - NI
    gens = C.bfs_generations(t, g, theta, mac)
    qc = O._queue_candidates_bounded(cfg.max_pair_queue, cfg.process_block)
    qc = list(dict.fromkeys(qc + [c for c in O._auto_pair_queue_candidates(T, NI) if c > qc[-1]]))
    nbr = O._neighbor_capacity_candidates(cfg.max_neighbors_per_leaf, L)
    K = cfg.max_interactions_per_node
    exp, attempt, done = [], 0, False
    for Q in qc:
        for Kn in nbr:
            attempt += 1
            (qo, fo, no), f, n = C.simulate_attempt(gens, Q, K, Kn)
            assert not fo or qo
            if qo:
                exp.append(("queue_overflow", attempt, Q, K, f, n))
                break
            if no:
                exp.append(("neighbor_overflow", attempt, Q, K, f, n))
                continue
            if attempt > 1:
                exp.append(("success", attempt, Q, K, f, n))
            done = True
            break
        if done:
            break
    ev = []
    oracle.build_interactions_and_neighbors(t, g, theta=theta, mac_type=mac, traversal_config=cfg,
                                            events=ev)
    assert ev == exp
    statuses = {e[0] for e in ev}
    if "queue" in name:
        assert "queue_overflow" in statuses
    if "neighbor" in name:
        assert "neighbor_overflow" in statuses
    # a second call with the cache starts at the successful queue capacity
    cache = {}
    oracle.build_interactions_and_neighbors(t, g, theta=theta, mac_type=mac,
                                            traversal_config=cfg, queue_cache=cache)
    ev2 = []
    oracle.build_interactions_and_neighbors(t, g, theta=theta, mac_type=mac,
                                            traversal_config=cfg, events=ev2, queue_cache=cache)
    assert list(cache.values()) == [exp[-1][2]]
    assert not any(e[0] == "queue_overflow" for e in ev2)


def _reference_result_to_interactions(res, t):
    """_result_to_interactions (3561-3637) on a raw result: rank nodes by
    nodes_by_level and stable-sort the first far_pair_count entries by the
    rank of their target."""
    T = t["parent"].shape[0]
    F = int(res["far_pair_count"])
    nbl = np.asarray(t["nodes_by_level"], np.int64)
    rank = np.zeros(T, np.int64)
    rank[nbl] = np.arange(T)
    src = res["interaction_sources"][:F]
    tgt = res["interaction_targets"][:F]
    order = np.argsort(rank[tgt], kind="stable")
    return src[order], tgt[order]


@pytest.mark.parametrize("path", ["fixed", "count"])
def test_raw_result_layout_and_round_trip(path):
    t, g, theta, mac = _system("octree", lambda: inputs.plummer_sphere(3000, 4, a=0.4), 8,
                               0.5, "dehnen")
    cfg = _Cfg(1 << 16, 256, 512, 256) if path == "fixed" else None
    inter, nbrs, raw = oracle.build_interactions_and_neighbors(
        t, g, theta=theta, mac_type=mac, traversal_config=cfg)
    assert raw["path"] == path
    res = O.to_result(raw, t)
    T = t["parent"].shape[0]
    L = T - t["num_internal_nodes"]
    F, M = int(inter["counts"].sum()), int(nbrs["counts"].sum())
    src, tgt = _reference_result_to_interactions(res, t)
    assert np.array_equal(tgt, inter["targets"])
    for k in range(T):
        o, c = inter["offsets"][k], inter["counts"][k]
        assert np.array_equal(np.sort(src[o:o + c]), np.sort(inter["sources"][o:o + c]))
    assert np.array_equal(res["interaction_offsets"],
                          np.concatenate([[0], np.cumsum(inter["counts"])]))
    assert np.array_equal(res["neighbor_offsets"], nbrs["offsets"])
    assert res["accept_decisions"] * 2 == F and res["near_decisions"] * 2 == M
    if path == "fixed":
        _, K, Kn = raw["attempt_caps"]
        assert K == 512 and Kn in O._neighbor_capacity_candidates(256, L)
        assert res["interaction_sources"].shape == (T * K,)
        assert res["neighbor_indices"].shape == (L * Kn,)
        assert np.all(res["interaction_sources"][F:] == -1)
        assert np.all(res["neighbor_indices"][M:] == -1)
        # B9: the node-order offsets do not index the level-major rows
        off = res["interaction_offsets"]
        rows_by_offsets = [np.sort(res["interaction_targets"][off[k]:off[k + 1]])
                           for k in range(T) if inter["counts"][k]]
        assert any(np.any(r != r[0]) or r[0] != k for r, k in
                   zip(rows_by_offsets, [k for k in range(T) if inter["counts"][k]]))
    else:
        assert res["interaction_sources"].shape == (F,)
        assert res["neighbor_indices"].shape == (M,)
        assert np.all(np.diff(res["interaction_targets"]) >= 0)
    assert np.all(res["interaction_tags"] == -1)


def check_target_blocks(counts, nlp, k, rows, src, valid, boff):
    """Invariants of near-field target blocks with block size k: row r has
    ceil(count_r / k) blocks, consecutive in block order; every block but the
    row's last is full; the valid slots of a row's blocks, read in order,
    are exactly the row's neighbour leaf positions; invalid slots hold 0."""
    L = len(counts)
    per = [-(-int(c) // k) for c in counts]
    assert list(boff) == [0] + list(np.cumsum(per))
    assert len(rows) == boff[-1] == src.shape[0] == valid.shape[0]
    for r in range(L):
        blk = range(int(boff[r]), int(boff[r + 1]))
        assert all(int(rows[b]) == r for b in blk)
        got = [int(src[b, j]) for b in blk for j in range(k) if valid[b, j]]
        assert got == [int(v) for v in nlp[r, :counts[r]]]
        for q, b in enumerate(blk):
            n_valid = int(valid[b].sum())
            assert n_valid == (k if q < len(blk) - 1 else int(counts[r]) - k * (len(blk) - 1))
    assert np.all(src[~valid] == 0)


@pytest.mark.parametrize("k", [1, 3, 8, 64])
def test_target_blocks_invariants(k):
    t, g, theta, mac = _system("octree", lambda: inputs.plummer_sphere(3000, 4, a=0.4), 8,
                               0.5, "dehnen")
    _, nbrs, raw = oracle.build_interactions_and_neighbors(t, g, theta=theta, mac_type=mac)
    nb = O.to_neighbors(raw, t, target_block_size=k)
    assert nb["target_block_size"] == k
    check_target_blocks(nb["counts"], nb["neighbor_leaf_positions"], k,
                        nb["target_block_leaf_ids"], nb["target_block_source_leaf_ids"],
                        nb["target_block_valid_mask"], nb["target_block_offsets"])
    off = O.to_neighbors(raw, t, target_block_size=0)
    assert off["target_block_leaf_ids"].shape == (0,)
    assert off["target_block_offsets"].shape == (t["parent"].shape[0]
                                                - t["num_internal_nodes"] + 1,)
