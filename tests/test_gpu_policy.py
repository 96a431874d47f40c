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
    # count path: the raw result holds node-order rows at node-order offsets
    src = parity.np_(res.interaction_sources)
    tags = parity.np_(res.interaction_tags)
    off = parity.np_(res.interaction_offsets)
    cnt = parity.np_(res.interaction_counts)
    assert np.array_equal(cnt, parity.np_(inter.counts))
    for node in range(len(far)):
        got = sorted(zip(src[off[node]:off[node] + cnt[node]].tolist(),
                         tags[off[node]:off[node] + cnt[node]].tolist()))
        assert got == sorted(far[node]), f"node {node}"
    noff = parity.np_(nbrs.offsets)
    nb = parity.np_(nbrs.neighbors)
    for r in range(len(near)):
        assert sorted(nb[noff[r]:noff[r + 1]].tolist()) == sorted(near[r])
    if policy is target_parity_policy:
        t = parity.np_(res.interaction_targets)
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
    src, tags = parity.np_(res.interaction_sources), parity.np_(res.interaction_tags)
    off, cnt = parity.np_(res.interaction_offsets), parity.np_(res.interaction_counts)
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
    assert sorted(zip(parity.np_(hres.interaction_targets).tolist(),
                      parity.np_(hres.interaction_sources).tolist(),
                      parity.np_(hres.interaction_tags).tolist())) == sorted(
        zip(parity.np_(res.interaction_targets).tolist(), src.tolist(), tags.tolist()))


def test_descriptor_default_rules_equal_plain_walk():
    """The descriptor with the built-in rules (MAC accept, near leaf pairs,
    no tag) runs the flat walk with its rule fields set and gives the plain
    walk's lists exactly, every tag -1 (65536 clustered points, leaf 16)."""
    import yggdrax_amd as ygg
    pos, mass = inputs.clustered(1 << 16, 5)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(),
                                      torch.from_numpy(mass).cuda(), leaf_size=16,
                                      return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    desc = ygg.PairPolicyDescriptor(accept="mac", near="leaf_pairs", tag="none")
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.5, mac_type="dehnen", pair_policy=desc, return_result=True)
    plain, pn, pres = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.5, mac_type="dehnen", return_result=True)
    torch.cuda.synchronize()
    print("default descriptor far", inter.sources.numel(), flush=True)
    assert torch.equal(inter.counts, plain.counts) and torch.equal(inter.offsets, plain.offsets)
    assert torch.equal(inter.level_offsets, plain.level_offsets)
    assert torch.equal(nbrs.counts, pn.counts) and torch.equal(nbrs.offsets, pn.offsets)
    T = tree.num_nodes
    key = lambda i: torch.sort(i.targets.long() * T + i.sources.long())[0]  # noqa: E731
    assert torch.equal(key(inter), key(plain))
    L = nbrs.counts.numel()
    rows = torch.repeat_interleave(torch.arange(L, device=nbrs.counts.device), nbrs.counts)
    nkey = lambda n: torch.sort(rows * T + n.neighbors.long())[0]  # noqa: E731
    assert torch.equal(nkey(nbrs), nkey(pn))
    assert bool(torch.all(res.interaction_tags == -1))
    assert (res.accept_decisions, res.near_decisions, res.refine_decisions) == (
        pres.accept_decisions, pres.near_decisions, pres.refine_decisions)


@pytest.mark.parametrize("case", [0, 2])
def test_descriptor_flat_walk_matches_callable_at_16k(case):
    """Descriptor policies in the flat walk against the generation-by-generation
    host callable path at 16384 Plummer points (leaf 16): identical far
    (target, source, tag) multisets and near lists."""
    import yggdrax_amd as ygg
    desc, policy, state, _, _ = _descriptors()[case]
    pos, mass = inputs.plummer_sphere(1 << 14, 9)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen", pair_policy=desc, return_result=True)
    hi, hn, hres = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen", pair_policy=policy, policy_state=state,
        return_result=True)
    torch.cuda.synchronize()
    print("descriptor 16k", case, "far", inter.sources.numel(), flush=True)
    assert torch.equal(hi.counts, inter.counts) and torch.equal(hi.offsets, inter.offsets)
    assert torch.equal(hn.counts, nbrs.counts)
    T = tree.num_nodes
    k1 = ((res.interaction_targets.long() * T + res.interaction_sources.long()) * 16
          + res.interaction_tags.long())
    k2 = ((hres.interaction_targets.long() * T + hres.interaction_sources.long()) * 16
          + hres.interaction_tags.long())
    assert torch.equal(torch.sort(k1)[0], torch.sort(k2)[0])
    L = nbrs.counts.numel()
    rows = torch.repeat_interleave(torch.arange(L, device=nbrs.counts.device), nbrs.counts)
    assert torch.equal(torch.sort(rows * T + nbrs.neighbors.long())[0],
                       torch.sort(rows * T + hn.neighbors.long())[0])
    assert (res.accept_decisions, res.near_decisions) == (hres.accept_decisions,
                                                          hres.near_decisions)


def all_near_policy(state, *, valid_pairs, mac_ok, different_nodes, target_leaf, source_leaf,
                    same_node, target_nodes, source_nodes, center_target, center_source,
                    dist_sq, extent_target, extent_source):
    """Nothing accepted: every pair of different leaves is near."""
    near = valid_pairs & target_leaf & source_leaf & different_nodes
    return _where(near, NEAR, REFINE), _where(near, -1, -1)


def test_descriptor_never_accept_gives_all_leaf_pairs_near():
    """accept="never" with near leaf pairs: no far entries, and every leaf's
    near row holds every other leaf (oracle policy walk with the same rule,
    256 uniform points, leaf 8)."""
    import yggdrax_amd as ygg
    pos, mass = inputs.uniform_cube(256, 29)
    ref = oracle.build_tree(pos, mass, leaf_size=8)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    far, near, _ = O.walk_with_policy(ref, rgeom, 0.6, "dehnen", 1.0, all_near_policy, None)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=8, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    desc = ygg.PairPolicyDescriptor(accept="never", near="leaf_pairs")
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=0.6,
                                                       mac_type="dehnen", pair_policy=desc)
    torch.cuda.synchronize()
    L = nbrs.counts.numel()
    print("never-accept near", nbrs.neighbors.numel(), "leaves", L, flush=True)
    assert inter.sources.numel() == 0 and all(len(r) == 0 for r in far)
    assert nbrs.neighbors.numel() == L * (L - 1)
    noff, nb = parity.np_(nbrs.offsets), parity.np_(nbrs.neighbors)
    for r in range(len(near)):
        assert sorted(nb[noff[r]:noff[r + 1]].tolist()) == sorted(near[r])


def test_descriptor_never_accept_never_near_is_empty():
    import yggdrax_amd as ygg
    pos, mass = inputs.uniform_cube(512, 31)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                    leaf_size=8, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    desc = ygg.PairPolicyDescriptor(accept="never", near="never")
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.6, mac_type="dehnen", pair_policy=desc, return_result=True)
    torch.cuda.synchronize()
    assert inter.sources.numel() == 0 and nbrs.neighbors.numel() == 0
    assert int(inter.counts.sum()) == 0 and int(nbrs.counts.sum()) == 0
    assert res.accept_decisions == 0 and res.near_decisions == 0 and res.refine_decisions > 0
