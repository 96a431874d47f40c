"""Launch shapes of the flat walk from the previous walk's generation
profile (ygg_walk_flat's gen_pairs_hint): a profile only sizes the grids
and the first batch of launches, so the lists stay those of the oracle
with no profile, with the true profile, and with a profile that is far
too small (every generation then runs on the 256-workgroup floor) or too
short (generations past it launch from the bound alone)."""

from __future__ import annotations

import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import yggdrax_amd as ygg
    pos, mass = inputs.clustered(1 << 16, 11)
    ref = oracle.build_octree(pos, mass, leaf_size=16)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    rint, rnb, _ = oracle.build_interactions_and_neighbors(ref, rgeom, theta=0.5,
                                                            mac_type="dehnen")
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                      leaf_size=16, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    return ygg, tree, geom, rint, rnb


@pytest.mark.parametrize("profile", ["none", "true", "tiny", "short"])
def test_walk_profile_keeps_the_lists(case, monkeypatch, profile):
    from yggdrax_amd import interactions as I
    ygg, tree, geom, rint, rnb = case
    print(f"[walk profile] {profile}", flush=True)
    monkeypatch.setattr(I, "_GEN_PROFILES", type(I._GEN_PROFILES)())
    kw = dict(theta=0.5, mac_type="dehnen")
    if profile != "none":
        ygg.build_interactions_and_neighbors(tree, geom, **kw)  # records the profile
        torch.cuda.synchronize()
        assert len(I._GEN_PROFILES) == 1
        (key, (prof, arr)), = I._GEN_PROFILES.items()
        assert prof[-1] == 0 and max(prof) > 10_000 and list(arr[:len(prof)]) == prof
        if profile == "tiny":
            for k in range(len(prof)):
                arr[k] = 1
        elif profile == "short":
            for k in range(len(prof) // 3, len(prof)):
                arr[k] = 0xFFFFFFFF
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    torch.cuda.synchronize()
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)
