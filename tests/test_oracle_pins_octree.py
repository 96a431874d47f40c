"""CPU pins of the oracle's octree ranges, radix-to-octree map, octree
levels and the walk's split tie-break (no GPU)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests import oracle_checks as C
from yggdrax_amd import inputs


def _cases():
    rng = np.random.default_rng(11)
    return {
        "uniform_700_leaf8": (rng.uniform(-1, 1, (700, 3)).astype(np.float32), 8),
        "plummer_3000_leaf16": (inputs.plummer_sphere(3000, 7, a=0.2)[0], 16),
        "uniform_2_leaf1": (rng.uniform(-1, 1, (2, 3)).astype(np.float32), 1),
        "uniform_9_leaf9": (rng.uniform(-1, 1, (9, 3)).astype(np.float32), 9),
        "clustered_4096_leaf4": (inputs.clustered(4096, 5, clusters=6)[0], 4),
        "b3_root_clamp": (C.b3_root_clamp_case()[0], 1),
        "mapping_8_leaf2": (C.mapping_case(), 2),
    }


CASES = _cases()


def _sorted_resolved(t):
    codes = [c for _, c in C.resolved_cells(t)]
    return codes == sorted(codes)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oct_node_ranges_brute_force(name):
    """Sorted resolved codes: exact ranges by enumeration. Unsorted ones
    (B4): the transition-point property of the bisection."""
    pos, leaf = CASES[name]
    C.check_oct_node_ranges(oracle.build_octree(pos, None, leaf_size=leaf))


@pytest.mark.parametrize("name", sorted(CASES))
def test_radix_node_to_oct_brute_force(name):
    pos, leaf = CASES[name]
    C.check_radix_node_to_oct(oracle.build_octree(pos, None, leaf_size=leaf))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oct_levels_by_direct_histogram(name):
    pos, leaf = CASES[name]
    C.check_oct_levels(oracle.build_octree(pos, None, leaf_size=leaf))


def test_b3_root_range_is_clamped():
    """octree.py:292-294 known answer: the root range stops at the last
    lower-z leaf."""
    pos, exp = C.b3_root_clamp_case()
    t = oracle.build_octree(pos, None, leaf_size=1)
    assert all(int(c) >> 62 == (1 if i >= 4 else 0) for i, c in enumerate(t["morton_codes"]))
    assert int(t["oct_node_depths"][0]) == 0 and int(t["oct_node_codes"][0]) == 0
    assert tuple(int(v) for v in t["oct_node_ranges"][0]) == exp


def test_octree_radix_node_mapping_covers_all_nodes():
    """tests/unit/test_octree_topology.py:65-93 on the oracle, plus the exact
    map: every radix node lands on its own cell, which exists as a slot."""
    t = oracle.build_octree(C.mapping_case(), None, leaf_size=2)
    T = t["parent"].shape[0]
    NI = t["num_internal_nodes"]
    n_oct = int(np.count_nonzero(t["oct_valid_mask"]))
    r2o = t["radix_node_to_oct"]
    assert r2o.shape == (T,)
    assert t["radix_leaf_to_oct"].shape == (T - NI,)
    assert np.all(r2o >= 0) and np.all(r2o < n_oct)
    assert np.all(t["radix_leaf_to_oct"] < n_oct)
    slots = C.valid_slot_cells(t)
    sc = t["morton_codes"]
    for i in range(NI):
        a, b = t["node_ranges"][i]
        d = C.common_depth(sc[a], sc[b])
        assert slots[int(r2o[i])] == (d, C.prefix(sc[a], d))


def test_walk_split_tie_break_known_answer():
    C.run_tie_break_on_oracle()


def test_unsorted_resolved_codes_occur():
    """The unsorted-resolved-codes case (B4) exists on real inputs; the
    reference's ranges there follow jnp.searchsorted on unsorted input and
    stay unpinned. This keeps that statement honest: at least one of the
    sampled inputs is unsorted, and the sorted ones are checked above."""
    flags = []
    for seed in range(40):
        pos = np.random.default_rng(seed).normal(size=(300, 3)).astype(np.float32) ** 3
        flags.append(_sorted_resolved(oracle.build_octree(pos, None, leaf_size=7)))
    assert not all(flags)
