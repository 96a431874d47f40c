"""Writes tests/golden/c1_uniform512.npz: the oracle's complete result for
the reference's CPU-runnable configuration (512 points uniform in
[-1,1]^3 from jax.random.split(PRNGKey(0)), masses in [0.5, 1.5];
build_octree leaf 16; compute_tree_geometry; build_interactions_and_neighbors
theta 0.6 dehnen with DualTreeTraversalConfig(8192, 256, 2048, 2048)).

    python tests/golden/make_golden.py

The oracle restates the reference (oracle/ygg_oracle.cpp, oracle/oracle.py);
this fixture freezes its output so the HIP path and later oracle changes are
both checked against the same bytes. Far-list rows are stored sorted (the
reference's within-row order comes from an unstable argsort).
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from yggdrax_amd import inputs  # noqa: E402

from types import SimpleNamespace  # noqa: E402

CONFIG = SimpleNamespace(max_pair_queue=8192, process_block=256, max_interactions_per_node=2048,
                         max_neighbors_per_leaf=2048)


def sorted_rows(offsets, counts, values):
    out = np.array(values, copy=True)
    for k in range(len(counts)):
        a = int(offsets[k])
        out[a:a + int(counts[k])] = np.sort(out[a:a + int(counts[k])])
    return out


def compute():
    pos, mass = inputs.uniform_cube(512, 0)
    topo = oracle.build_octree(pos, mass, leaf_size=16)
    geom = oracle.compute_tree_geometry(topo, topo["positions_sorted"])
    inter, nbrs, _ = oracle.build_interactions_and_neighbors(
        topo, geom, theta=0.6, mac_type="dehnen", traversal_config=CONFIG)
    out = {"positions": pos, "masses": mass}
    for k, v in topo.items():
        if isinstance(v, np.ndarray):
            out["topo_" + k] = v
        elif isinstance(v, (int, np.integer, bool, np.bool_)):
            out["topo_" + k] = np.asarray(v)
    for k in ("center", "half_extent", "radius", "max_extent"):
        out["geom_" + k] = np.asarray(geom[k])
    for k, v in inter.items():
        out["far_" + k] = np.asarray(v)
    out["far_sources"] = sorted_rows(inter["offsets"], inter["counts"], inter["sources"])
    for k, v in nbrs.items():
        if isinstance(v, np.ndarray):
            out["near_" + k] = v
    out["near_neighbors"] = sorted_rows(nbrs["offsets"], nbrs["counts"], nbrs["neighbors"])
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "c1_uniform512.npz"), **compute())
    print("wrote", os.path.join(HERE, "c1_uniform512.npz"))
