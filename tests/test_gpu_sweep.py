"""Seeded sweep over the pipeline's parameters: generator, size, leaf size,
theta, MAC and tree kind drawn at random (fixed seed), each compared with
the oracle end to end (topology, geometry, far and near lists)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu


def _cases(count=9, seed=2026):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(dict(gen=("uniform", "plummer", "clustered")[i % 3],
                        n=int(rng.choice([777, 2048, 5000, 12000])),
                        leaf=int(rng.choice([1, 4, 8, 16, 32, 64])),
                        theta=float(rng.choice([0.3, 0.5, 0.6, 0.8, 1.0])),
                        mac=("bh", "dehnen", "engblom")[int(rng.integers(0, 3))],
                        octree=bool(rng.integers(0, 2)),
                        seed=int(rng.integers(0, 1000))))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: "{gen}-{n}-leaf{leaf}-th{theta}-{mac}-{t}".format(
    t="oct" if c["octree"] else "radix", **c))
def test_pipeline_matches_oracle(case):
    import yggdrax_amd as ygg
    gen, n, seed = case["gen"], case["n"], case["seed"]
    if gen == "uniform":
        pos, mass = inputs.uniform_cube(n, seed)
    elif gen == "plummer":
        pos, mass = inputs.plummer_sphere(n, seed)
    else:
        pos, mass = inputs.clustered(n, seed)
    pos = np.ascontiguousarray(pos, np.float32)
    mass = np.ascontiguousarray(mass, np.float32)
    kw = dict(theta=case["theta"], mac_type=case["mac"], max_neighbors_per_leaf=1 << 20)
    ref = (oracle.build_octree if case["octree"] else oracle.build_tree)(
        pos, mass, leaf_size=case["leaf"])
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    rint, rnb, _ = oracle.build_interactions_and_neighbors(ref, rgeom, **kw)
    build = ygg.build_octree if case["octree"] else ygg.build_tree
    tree, ps, _, _ = build(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                           leaf_size=case["leaf"], return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    torch.cuda.synchronize()
    print("sweep", case, "far", inter.sources.numel(), "near", nbrs.neighbors.numel(), flush=True)
    parity.assert_topology_equal(tree, ref, octree=case["octree"])
    parity.assert_geometry_close(geom, rgeom)
    parity.assert_interactions_equal(inter, rint)
    parity.assert_neighbors_equal(nbrs, rnb)
