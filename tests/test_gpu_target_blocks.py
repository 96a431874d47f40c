"""Near-field target blocks (YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE > 0,
_interactions_impl.py:3689-3770) on the GPU against the oracle: the block
layout (offsets, rows, validity) exactly, the neighbours of each row as a
set (their order within a row is unspecified), and the block invariants."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from tests.test_oracle_pins_walk import check_target_blocks
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 4, 32])
@pytest.mark.parametrize("kind", ["octree", "radix"])
def test_target_blocks_match_oracle(monkeypatch, k, kind):
    import yggdrax_amd as ygg
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    monkeypatch.setenv("YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE", str(k))
    print(f"[target-blocks] kind={kind} k={k}", flush=True)
    pos, mass = inputs.plummer_sphere(8192, 6, a=0.3)
    build = oracle.build_octree if kind == "octree" else oracle.build_tree
    ref = build(pos, mass, leaf_size=16)
    rg = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    _, rnb, _ = oracle.build_interactions_and_neighbors(ref, rg, theta=0.6, mac_type="dehnen")
    gb = ygg.build_octree if kind == "octree" else ygg.build_tree
    tree, ps, _, _ = gb(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                        leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    for kw in (dict(), dict(retry_logger=lambda e: None, max_interactions_per_node=4096)):
        _, nb = ygg.build_interactions_and_neighbors(tree, geom, theta=0.6, mac_type="dehnen",
                                                     **kw)
        torch.cuda.synchronize()
        assert nb.target_block_size == k
        for f in ("target_block_offsets", "target_block_leaf_ids", "target_block_valid_mask"):
            assert np.array_equal(parity.np_(getattr(nb, f)), rnb[f]), f
        src = parity.np_(nb.target_block_source_leaf_ids)
        valid = parity.np_(nb.target_block_valid_mask)
        boff = parity.np_(nb.target_block_offsets)
        for r in range(len(boff) - 1):
            a, b = int(boff[r]), int(boff[r + 1])
            got = np.sort(src[a:b][valid[a:b]])
            exp = np.sort(rnb["target_block_source_leaf_ids"][a:b][rnb[
                "target_block_valid_mask"][a:b]])
            assert np.array_equal(got, exp), r
        check_target_blocks(parity.np_(nb.counts), parity.np_(nb.neighbor_leaf_positions), k,
                            parity.np_(nb.target_block_leaf_ids), src, valid, boff)
