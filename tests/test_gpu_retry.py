"""The reference's fixed-capacity method on the GPU: the attempt-by-attempt
walk with per-row slots (run whenever a retry_logger is given, or when a
queue and a far overflow meet), its retry-event sequence, its errors, and
the raw DualTreeWalkResult of return_result=True, all against the oracle's
retry driver (_interactions_impl.py:3840-4440, 4606-4627)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import oracle as O
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


class _OCfg:
    def __init__(self, c):
        (self.max_pair_queue, self.process_block, self.max_interactions_per_node,
         self.max_neighbors_per_leaf) = c


CASES = {
    # name: (generator, tree kind, leaf, theta, mac, capacity keywords)
    "c1_cfg": (lambda: inputs.uniform_cube(512, 0), "octree", 16, 0.6, "dehnen",
               dict(traversal_config=(8192, 256, 2048, 2048))),
    "getting_started_cfg": (lambda: inputs.uniform_cube(256, 7), "radix", 16, 0.8, "dehnen",
                            dict(traversal_config=(8192, 256, 2048, 2048))),
    "plummer16k_explicit": (lambda: inputs.plummer_sphere(1 << 14, 2), "radix", 32, 0.5,
                            "dehnen", dict(max_interactions_per_node=4096,
                                           max_pair_queue=1 << 20)),
    "queue_retries": (lambda: inputs.plummer_sphere(4096, 5, a=0.3), "octree", 8, 0.5, "dehnen",
                      dict(traversal_config=(64, 4, 2048, 2048))),
    "neighbor_retries": (lambda: inputs.uniform_cube(4096, 11), "radix", 8, 0.6, "bh",
                         dict(traversal_config=(1 << 16, 256, 2048, 2))),
    "queue_and_neighbor_retries": (lambda: inputs.clustered(1 << 13, 3), "octree", 8, 0.5,
                                   "engblom", dict(traversal_config=(64, 4, 2048, 3))),
    "count_path": (lambda: inputs.uniform_cube(512, 0), "octree", 16, 0.6, "dehnen", dict()),
}

ERROR_CASES = {
    "far_overflow": (lambda: inputs.uniform_cube(512, 0), "octree", 16, 0.6, "dehnen",
                     dict(traversal_config=(8192, 256, 1, 2048))),
    "queue_error_user": (lambda: inputs.uniform_cube(512, 0), "radix", 8, 0.6, "dehnen",
                         dict(max_pair_queue=4, max_interactions_per_node=64)),
    "queue_and_far_overflow": (lambda: inputs.plummer_sphere(4096, 5, a=0.3), "octree", 8, 0.5,
                               "dehnen", dict(traversal_config=(64, 4, 8, 2048))),
}


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _kw(ygg, kw, gpu):
    out = dict(kw)
    if "traversal_config" in kw:
        c = kw["traversal_config"]
        out["traversal_config"] = ygg.DualTreeTraversalConfig(*c) if gpu else _OCfg(c)
    return out


def _setup(ygg, case):
    gen, kind, leaf, theta, mac, kw = case
    pos, mass = gen()
    build = oracle.build_octree if kind == "octree" else oracle.build_tree
    ref = build(pos, mass, leaf_size=leaf)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    gbuild = ygg.build_octree if kind == "octree" else ygg.build_tree
    tree, ps, _, _ = gbuild(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                            leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    return ref, rgeom, tree, geom, theta, mac, kw


def _gpu_events(evs):
    return [(e.status, e.attempt, e.queue_capacity, e.interaction_capacity, e.far_pair_count,
             e.near_pair_count) for e in evs]


def _runs(x):
    x = np.asarray(x)
    return np.concatenate([[0], np.cumsum(x[1:] != x[:-1])]) if x.size else x


def _assert_far_layout_equal(res, exp):
    """Targets are laid out identically (row structure); within a row the
    sources compare as sets (the reference's in-row order is unspecified)."""
    tg, te = parity.np_(res.interaction_targets), exp["interaction_targets"]
    assert tg.shape == te.shape and np.array_equal(tg, te), "raw far targets differ"
    sg, se = parity.np_(res.interaction_sources), exp["interaction_sources"]
    assert sg.shape == se.shape
    r = _runs(te)
    assert np.array_equal(sg[np.lexsort((sg, r))], se[np.lexsort((se, r))]), "raw far rows differ"
    assert np.array_equal(parity.np_(res.interaction_tags), exp["interaction_tags"])


def _assert_near_layout_equal(res, exp):
    nb, ne = parity.np_(res.neighbor_indices), exp["neighbor_indices"]
    assert nb.shape == ne.shape
    off = exp["neighbor_offsets"]
    for r in range(len(off) - 1):
        assert np.array_equal(np.sort(nb[off[r]:off[r + 1]]), np.sort(ne[off[r]:off[r + 1]]))
    assert np.all(nb[off[-1]:] == -1) and np.all(ne[off[-1]:] == -1)


def _assert_result_equal(res, exp):
    for f in ("interaction_offsets", "interaction_counts", "neighbor_offsets",
              "neighbor_counts", "leaf_indices"):
        assert np.array_equal(parity.np_(getattr(res, f)), exp[f]), f
    for f in ("far_pair_count", "near_pair_count", "queue_overflow", "far_overflow",
              "near_overflow", "accept_decisions", "near_decisions", "refine_decisions"):
        assert int(getattr(res, f)) == int(exp[f]), f
    _assert_far_layout_equal(res, exp)
    _assert_near_layout_equal(res, exp)


@pytest.mark.parametrize("name", list(CASES))
def test_retry_logger_path_matches_oracle(ygg, name):
    """Events (status, attempt, capacities, stored pair counts), lists and the
    raw result of the attempt-by-attempt walk against the oracle; a second
    call reuses the cached queue capacity exactly as the oracle's cache does."""
    from yggdrax_amd import interactions as gi
    print(f"[retry] case={name}", flush=True)
    ref, rgeom, tree, geom, theta, mac, kw = _setup(ygg, CASES[name])
    gi._DUAL_TREE_QUEUE_CACHE.clear()
    cache = {}
    for call in range(2):
        oev, gev = [], []
        rint, rnb, raw = oracle.build_interactions_and_neighbors(
            ref, rgeom, theta=theta, mac_type=mac, events=oev, queue_cache=cache,
            **_kw(ygg, kw, False))
        inter, nbrs, res = ygg.build_interactions_and_neighbors(
            tree, geom, theta=theta, mac_type=mac, retry_logger=gev.append, return_result=True,
            **_kw(ygg, kw, True))
        torch.cuda.synchronize()
        print(f"[retry] case={name} call={call} events={oev}", flush=True)
        assert _gpu_events(gev) == oev
        parity.assert_interactions_equal(inter, rint)
        parity.assert_neighbors_equal(nbrs, rnb)
        _assert_result_equal(res, O.to_result(raw, ref))


@pytest.mark.parametrize("name", list(ERROR_CASES))
@pytest.mark.parametrize("logger", [True, False])
def test_capacity_errors_match_oracle(ygg, name, logger):
    """The same error, and with a logger the same events before it, on the
    attempt-by-attempt path (with a logger) and the flat path with the
    ladder emulation (without one; a queue and a far overflow together go
    to the attempt-by-attempt walk there too)."""
    from yggdrax_amd import interactions as gi
    print(f"[retry-error] case={name} logger={logger}", flush=True)
    ref, rgeom, tree, geom, theta, mac, kw = _setup(ygg, ERROR_CASES[name])
    gi._DUAL_TREE_QUEUE_CACHE.clear()
    oev, gev = [], []
    with pytest.raises(RuntimeError) as oerr:
        oracle.build_interactions_and_neighbors(ref, rgeom, theta=theta, mac_type=mac,
                                                events=oev, **_kw(ygg, kw, False))
    with pytest.raises(RuntimeError) as gerr:
        ygg.build_interactions_and_neighbors(tree, geom, theta=theta, mac_type=mac,
                                             retry_logger=gev.append if logger else None,
                                             **_kw(ygg, kw, True))
    torch.cuda.synchronize()
    assert str(gerr.value) == str(oerr.value)
    if logger:
        assert _gpu_events(gev) == oev


@pytest.mark.parametrize("name", list(CASES))
def test_flat_path_raw_result_matches_oracle(ygg, name):
    """Without a logger the flat walk plus the ladder emulation runs; its
    raw result takes the successful attempt's capacities from the emulation
    and must equal the oracle's attempt-by-attempt result."""
    from yggdrax_amd import interactions as gi
    print(f"[flat-result] case={name}", flush=True)
    ref, rgeom, tree, geom, theta, mac, kw = _setup(ygg, CASES[name])
    gi._DUAL_TREE_QUEUE_CACHE.clear()
    rint, rnb, raw = oracle.build_interactions_and_neighbors(
        ref, rgeom, theta=theta, mac_type=mac, **_kw(ygg, kw, False))
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=theta, mac_type=mac, return_result=True, **_kw(ygg, kw, True))
    torch.cuda.synchronize()
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)
    _assert_result_equal(res, O.to_result(raw, ref))
