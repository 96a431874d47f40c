"""The public accessors, the stage-timing callback and the traversal
diagnostics of the drop-in surface, on the GPU against the oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    import yggdrax_amd as ygg
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    pos, mass = inputs.plummer_sphere(4096, 8, a=0.3)
    ref = oracle.build_octree(pos, mass, leaf_size=16)
    rg = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    rint, rnb, _ = oracle.build_interactions_and_neighbors(ref, rg, theta=0.6, mac_type="dehnen")
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                      leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    return ygg, ref, rint, rnb, tree, geom


def test_interactions_for_node_and_neighbors_for_leaf(system):
    """_interactions_impl.py:6147-6170 over every node and every leaf."""
    ygg, ref, rint, rnb, tree, geom = system
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen")
    T = tree.num_nodes
    NI = T - tree.num_leaves
    for k in range(T):
        got = np.sort(parity.np_(ygg.interactions_for_node(inter, k)))
        o, c = rint["offsets"][k], rint["counts"][k]
        assert np.array_equal(got, np.sort(rint["sources"][o:o + c])), k
    for leaf in range(NI, T):
        got = np.sort(parity.np_(ygg.neighbors_for_leaf(nbrs, leaf)))
        r = int(np.nonzero(rnb["leaf_indices"] == leaf)[0][0])
        exp = np.sort(rnb["neighbors"][rnb["offsets"][r]:rnb["offsets"][r + 1]])
        assert np.array_equal(got, exp), leaf
    with pytest.raises(ValueError, match="not present"):
        ygg.neighbors_for_leaf(nbrs, 0)


def test_timing_callback_reports_both_stages(system):
    ygg, ref, rint, rnb, tree, geom = system
    stages = []
    compact, nbrs = ygg.build_compact_far_pairs_and_leaf_neighbor_lists(
        tree, geom, theta=0.6, mac_type="dehnen",
        timing_callback=lambda name, sec: stages.append((name, sec)))
    assert [n for n, _ in stages] == ["dual_split_shared_count",
                                      "dual_split_shared_combined_fill"]
    assert all(sec >= 0.0 for _, sec in stages)
    assert int(compact.far_pair_count) == int(rint["counts"].sum())


def test_traversal_diagnostics_lines(system, capsys, monkeypatch):
    ygg, ref, rint, rnb, tree, geom = system
    from yggdrax_amd import interactions as I
    monkeypatch.setattr(I, "_TRAVERSAL_DIAGNOSTICS", True)
    ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen")
    out = capsys.readouterr().out
    assert "[yggdrax.traversal] walk start" in out
    assert "[yggdrax.traversal] count pass" in out
    assert f"total_far_pairs={int(rint['counts'].sum())}" in out
    monkeypatch.setattr(I, "_TRAVERSAL_DIAGNOSTICS", False)
    ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen")
    assert "[yggdrax.traversal]" not in capsys.readouterr().out
