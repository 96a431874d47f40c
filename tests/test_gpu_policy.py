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


def opening_ratio_policy(state, *, valid_pairs, mac_ok, different_nodes, target_leaf,
                         source_leaf, same_node, target_nodes, source_nodes, center_target,
                         center_source, dist_sq, extent_target, extent_source):
    """Built-in MAC acceptance with a tag from the opening ratio
    (e_t + e_s)^2 / d^2 in three bands (an adaptive-order choice)."""
    near = valid_pairs & (~mac_ok) & target_leaf & source_leaf & different_nodes
    s = extent_target + extent_source
    if isinstance(dist_sq, torch.Tensor):
        ratio = (s * s) / dist_sq
    else:
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = (s * s).astype(np.float32) / dist_sq
    tag = _where(ratio >= np.float32(0.09), 6, _where(ratio >= np.float32(0.01), 4, 2))
    return _where(mac_ok, ACCEPT, _where(near, NEAR, REFINE)), _where(mac_ok, tag, -1)


def _descriptors():
    from yggdrax_amd import PairPolicyDescriptor as D
    return [
        (D(accept="distance", accept_dist_sq=0.36, near="leaf_pairs", tag="dist_sq",
           tag_thresholds=(1.0,), tag_values=(1, 2)),
         distance_bucket_policy, {"far_sq": np.float32(0.36), "tag_split_sq": np.float32(1.0)},
         256, 8),
        (D(accept="different", near="never", tag="target_mod", tag_modulus=2),
         target_parity_policy, None, 32, 8),
        (D(accept="mac", near="leaf_pairs", tag="opening_ratio", tag_thresholds=(0.01, 0.09),
           tag_values=(2, 4, 6)), opening_ratio_policy, None, 2048, 16),
    ]


@pytest.mark.parametrize("case", range(3))
def test_policy_descriptor_matches_callable_and_oracle(case):
    """A policy given as a descriptor (evaluated by the HIP kernel) gives the
    same per-node far sets with tags and near sets as the oracle's policy
    walk with the equivalent callable, and the same lists as the host
    callable path."""
    import yggdrax_amd as ygg
    desc, policy, state, n, leaf = _descriptors()[case]
    pos, mass = inputs.uniform_cube(n, 23)
    ref = oracle.build_tree(pos, mass, leaf_size=leaf)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    far, near, _ = O.walk_with_policy(ref, rgeom, 0.6, "dehnen", 1.0, policy, state)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen", pair_policy=desc, return_result=True)
    torch.cuda.synchronize()
    src, tags = parity.np_(inter.sources), parity.np_(res.interaction_tags)
    off, cnt = parity.np_(inter.offsets), parity.np_(inter.counts)
    print("descriptor", case, "far", int(cnt.sum()), flush=True)
    for node in range(len(far)):
        got = sorted(zip(src[off[node]:off[node] + cnt[node]].tolist(),
                         tags[off[node]:off[node] + cnt[node]].tolist()))
        assert got == sorted(far[node]), f"node {node}"
    noff, nb = parity.np_(nbrs.offsets), parity.np_(nbrs.neighbors)
    for r in range(len(near)):
        assert sorted(nb[noff[r]:noff[r + 1]].tolist()) == sorted(near[r])
    hi, hn, hres = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen", pair_policy=policy, policy_state=state,
        return_result=True)
    assert torch.equal(hi.counts, inter.counts) and torch.equal(hi.offsets, inter.offsets)
    assert torch.equal(hn.counts, nbrs.counts)
    assert sorted(zip(parity.np_(hi.targets).tolist(), parity.np_(hi.sources).tolist(),
                      parity.np_(hres.interaction_tags).tolist())) == sorted(
        zip(parity.np_(inter.targets).tolist(), src.tolist(), tags.tolist()))
