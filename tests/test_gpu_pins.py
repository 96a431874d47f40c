"""The brute-force octree checks and the hand-derived known answers, applied
directly to the HIP pipeline's outputs (not only through equality with the
oracle)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests import oracle_checks as C
from tests import parity
from tests.test_oracle_pins_octree import CASES

pytestmark = pytest.mark.gpu

TOPO_FIELDS = ("num_internal_nodes", "morton_codes", "node_ranges", "parent", "left_child",
               "right_child", "oct_valid_mask", "oct_node_depths", "oct_node_codes",
               "oct_node_ranges", "radix_node_to_oct", "radix_leaf_to_oct", "oct_level_offsets",
               "oct_num_levels")


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _as_dict(tree):
    topo = tree.topology
    out = {}
    for f in TOPO_FIELDS:
        v = getattr(topo, f)
        out[f] = int(v) if not isinstance(v, torch.Tensor) else parity.np_(v)
    return out


def _octree(ygg, pos, leaf):
    return ygg.build_octree(torch.from_numpy(pos).cuda(), None, leaf_size=leaf)


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_octree_fields_brute_force(ygg, name):
    print(f"[pins] case={name}", flush=True)
    pos, leaf = CASES[name]
    t = _as_dict(_octree(ygg, pos, leaf))
    C.check_oct_node_ranges(t)
    C.check_radix_node_to_oct(t)
    C.check_oct_levels(t)


def test_gpu_b3_root_range_is_clamped(ygg):
    pos, exp = C.b3_root_clamp_case()
    t = _as_dict(_octree(ygg, pos, 1))
    assert int(t["oct_node_depths"][0]) == 0
    assert tuple(int(v) for v in t["oct_node_ranges"][0]) == exp


def test_gpu_octree_radix_node_mapping_covers_all_nodes(ygg):
    """tests/unit/test_octree_topology.py:65-93."""
    tree = _octree(ygg, C.mapping_case(), 2)
    r2o = parity.np_(tree.topology.radix_node_to_oct)
    assert r2o.shape == (tree.num_nodes,)
    assert parity.np_(tree.topology.radix_leaf_to_oct).shape == (tree.num_leaves,)
    assert np.all(r2o >= 0) and np.all(r2o < tree.oct_num_nodes)
    assert np.all(parity.np_(tree.topology.radix_leaf_to_oct) < tree.oct_num_nodes)


def test_gpu_walk_split_tie_break_known_answer(ygg):
    pos, mass, theta, _ = C.tie_break_case()
    tree = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                          leaf_size=1)
    ps = torch.from_numpy(pos).cuda()[tree.particle_indices]
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=theta,
                                                       mac_type="dehnen")
    assert [int(v) for v in parity.np_(tree.topology.left_child)] == [1, 3, 5]
    off, cnt = parity.np_(inter.offsets), parity.np_(inter.counts)
    src = parity.np_(inter.sources)
    lists = {k: sorted(int(s) for s in src[off[k]:off[k] + cnt[k]]) for k in range(7)}
    C.check_tie_break(lists, int(parity.np_(nbrs.counts).sum()), cnt)
