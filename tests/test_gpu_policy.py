"""pair_policy parity (tests/unit/test_traversal_policy.py:61-233 policies):
HIP policy walk against the oracle's policy walk, including directed tags."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import oracle as O
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

ACCEPT, NEAR, REFINE = 0, 1, 2


def _where(c, a, b):
    if isinstance(c, torch.Tensor):
        a = torch.as_tensor(a, device=c.device)
        b = torch.as_tensor(b, device=c.device)
        return torch.where(c, a, b)
    return np.where(c, a, b)


def distance_bucket_policy(state, *, valid_pairs, mac_ok, different_nodes, target_leaf,
                           source_leaf, same_node, target_nodes, source_nodes, center_target,
                           center_source, dist_sq, extent_target, extent_source):
    accept = valid_pairs & different_nodes & (dist_sq >= state["far_sq"])
    near = valid_pairs & different_nodes & (~accept) & target_leaf & source_leaf
    actions = _where(accept, ACCEPT, _where(near, NEAR, REFINE))
    tags = _where(dist_sq >= state["tag_split_sq"], 2, 1)
    return actions, _where(accept, tags, -1)


def target_parity_policy(state, *, valid_pairs, mac_ok, different_nodes, target_leaf,
                         source_leaf, same_node, target_nodes, source_nodes, center_target,
                         center_source, dist_sq, extent_target, extent_source):
    accept = valid_pairs & different_nodes
    actions = _where(accept, ACCEPT, REFINE)
    return actions, _where(accept, target_nodes % 2, -1)


def mac_policy(state, *, valid_pairs, mac_ok, different_nodes, target_leaf, source_leaf,
               same_node, target_nodes, source_nodes, center_target, center_source, dist_sq,
               extent_target, extent_source):
    near = valid_pairs & (~mac_ok) & target_leaf & source_leaf & different_nodes
    return _where(mac_ok, ACCEPT, _where(near, NEAR, REFINE)), _where(mac_ok, 7, -1)


@pytest.mark.parametrize("policy,state,n,leaf", [
    (distance_bucket_policy, {"far_sq": 0.36, "tag_split_sq": 1.0}, 256, 8),
    (target_parity_policy, None, 32, 8),
    (mac_policy, None, 2048, 16),
])
def test_policy_walk_parity(policy, state, n, leaf):
    import yggdrax_amd as ygg
    pos, mass = inputs.uniform_cube(n, 23)
    ref = oracle.build_tree(pos, mass, leaf_size=leaf)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    far, near, _ = O.walk_with_policy(ref, rgeom, 0.6, "dehnen", 1.0, policy, state)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen", pair_policy=policy, policy_state=state,
        return_result=True)
    torch.cuda.synchronize()
    src = parity.np_(inter.sources)
    tags = parity.np_(res.interaction_tags)
    off = parity.np_(inter.offsets)
    cnt = parity.np_(inter.counts)
    for node in range(len(far)):
        got = sorted(zip(src[off[node]:off[node] + cnt[node]].tolist(),
                         tags[off[node]:off[node] + cnt[node]].tolist()))
        assert got == sorted(far[node]), f"node {node}"
    noff = parity.np_(nbrs.offsets)
    nb = parity.np_(nbrs.neighbors)
    for r in range(len(near)):
        assert sorted(nb[noff[r]:noff[r + 1]].tolist()) == sorted(near[r])
    if policy is target_parity_policy:
        t = parity.np_(inter.targets)
        assert np.all(np.isin(tags, [0, 1])) and np.array_equal(tags, t % 2)
    if policy is mac_policy:  # identical to the built-in MAC walk
        plain, pn = ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen")
        assert np.array_equal(parity.np_(plain.counts), cnt)
        assert np.array_equal(parity.np_(pn.counts), parity.np_(nbrs.counts))
