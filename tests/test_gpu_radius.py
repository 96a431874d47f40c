"""compute_tree_geometry(particle_radius=...) on the GPU against the oracle,
bit for bit (segment-tree path and the level-synchronous path of malformed
trees), and the reference's own particle_radius checks on GPU outputs."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from tests import test_oracle_pins_radius as R
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _gpu_geom_fn(ygg):
    def fn(t, ps, r):
        pos = torch.from_numpy(np.ascontiguousarray(R.sample_tree()[1]))
        tree = ygg.build_tree(torch.from_numpy(_sample_positions()).cuda(),
                              torch.ones(8, device="cuda"), R.BOUNDS, leaf_size=R.LEAF)
        assert np.array_equal(parity.np_(tree.particle_indices), t["particle_indices"])
        g = ygg.compute_tree_geometry(tree, pos.cuda(), max_leaf_size=R.LEAF,
                                      particle_radius=None if r is None else
                                      torch.from_numpy(r).cuda())
        return {f: parity.np_(getattr(g, f)) for f in g._fields}
    return fn


def _sample_positions():
    return np.random.default_rng(3).uniform(-1.0, 1.0, size=(8, 3)).astype(np.float32)


@pytest.mark.parametrize("check", [R.check_zero_radius, R.check_balls_inside_boxes,
                                   R.check_inflation_reaches_internal_nodes,
                                   R.check_invariants])
def test_reference_particle_radius_checks_on_gpu(ygg, check):
    check(_gpu_geom_fn(ygg))


CASES = {
    "plummer_16k_radix": (lambda: inputs.plummer_sphere(1 << 14, 2), "radix", 32),
    "clustered_32k_octree": (lambda: inputs.clustered(1 << 15, 3), "octree", 64),
    "uniform_1000_leaf1": (lambda: inputs.uniform_cube(1000, 5), "radix", 1),
    "identical_points": (lambda: (np.full((300, 3), 0.25, np.float32),
                                  np.ones(300, np.float32)), "radix", 4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_radius_geometry_matches_oracle(ygg, name):
    print(f"[radius] case={name}", flush=True)
    gen, kind, leaf = CASES[name]
    pos, mass = gen()
    build = oracle.build_octree if kind == "octree" else oracle.build_tree
    ref = build(pos, mass, leaf_size=leaf)
    r = np.random.default_rng(7).uniform(0.0, 0.05, pos.shape[0]).astype(np.float32)
    rg = oracle.compute_tree_geometry(ref, ref["positions_sorted"], particle_radius=r)
    gb = ygg.build_octree if kind == "octree" else ygg.build_tree
    tree, ps, _, _ = gb(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                        leaf_size=leaf, return_reordered=True)
    g = ygg.compute_tree_geometry(tree, ps, particle_radius=torch.from_numpy(r).cuda())
    torch.cuda.synchronize()
    parity.assert_geometry_close(g, rg)
    with pytest.raises(ValueError, match="particle_radius"):
        ygg.compute_tree_geometry(tree, ps, particle_radius=torch.zeros(3, device="cuda"))
