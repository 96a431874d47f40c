"""CPU pins of compute_tree_geometry(particle_radius=...) on the oracle
(_geometry_impl.py:107-112): the reference's own checks
(tests/unit/test_tree_geometry_particle_radius.py) restated on the oracle,
plus every node box equal to the direct min/max over its particles'
[p - r, p + r] cubes."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

LEAF = 1
BOUNDS = (np.full(3, -2.0, np.float32), np.full(3, 2.0, np.float32))


def sample_tree(n=8, seed=3):
    """tests/unit/test_tree_geometry_particle_radius.py:28-36: one particle
    per leaf, fixed bounds [-2, 2]^3."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    t = oracle.build_tree(pos, np.ones(n, np.float32), BOUNDS, leaf_size=LEAF)
    return t, t["positions_sorted"]


def check_zero_radius(geom_fn):
    t, ps = sample_tree()
    point = geom_fn(t, ps, None)
    zero = geom_fn(t, ps, np.zeros(ps.shape[0], np.float32))
    for f in ("center", "half_extent", "radius", "max_extent"):
        assert np.array_equal(np.asarray(point[f]).view(np.uint32),
                              np.asarray(zero[f]).view(np.uint32)), f


def check_balls_inside_boxes(geom_fn):
    t, ps = sample_tree()
    r = np.linspace(0.05, 0.4, ps.shape[0]).astype(np.float32)
    g = geom_fn(t, ps, r)
    c, h = np.asarray(g["center"]), np.asarray(g["half_extent"])
    for node, (lo, hi) in enumerate(t["node_ranges"]):
        for p in range(lo, hi + 1):
            assert np.all(ps[p] - r[p] >= c[node] - h[node] - 1e-5)
            assert np.all(ps[p] + r[p] <= c[node] + h[node] + 1e-5)


def check_inflation_reaches_internal_nodes(geom_fn):
    t, ps = sample_tree()
    pad = 0.25
    point = geom_fn(t, ps, None)
    grown = geom_fn(t, ps, np.full(ps.shape[0], pad, np.float32))
    NI = t["num_internal_nodes"]
    assert NI > 0
    grew = np.asarray(grown["half_extent"])[:NI] - np.asarray(point["half_extent"])[:NI]
    assert np.all(grew >= -1e-6) and np.max(grew) > 0.5 * pad
    # a uniform pad grows every half extent by the pad (floor not binding)
    np.testing.assert_allclose(grew, pad, rtol=0, atol=1e-6)


def check_invariants(geom_fn):
    t, ps = sample_tree()
    g = geom_fn(t, ps, np.full(ps.shape[0], 0.3, np.float32))
    h = np.asarray(g["half_extent"])
    np.testing.assert_allclose(np.asarray(g["radius"]), np.linalg.norm(h, axis=1), rtol=1e-6)
    np.testing.assert_allclose(np.asarray(g["max_extent"]), h.max(axis=1), rtol=1e-6)


def _oracle_geom(t, ps, r):
    return oracle.compute_tree_geometry(t, ps, max_leaf_size=LEAF, particle_radius=r)


@pytest.mark.parametrize("check", [check_zero_radius, check_balls_inside_boxes,
                                   check_inflation_reaches_internal_nodes, check_invariants])
def test_reference_particle_radius_checks_on_oracle(check):
    check(_oracle_geom)


@pytest.mark.parametrize("n,leaf", [(1000, 8), (333, 1), (64, 64)])
def test_radius_geometry_is_direct_min_max_of_cubes(n, leaf):
    rng = np.random.default_rng(n)
    pos = rng.normal(size=(n, 3)).astype(np.float32)
    t = oracle.build_tree(pos, None, leaf_size=leaf)
    ps = t["positions_sorted"]
    r = rng.uniform(0.0, 0.2, n).astype(np.float32)
    g = oracle.compute_tree_geometry(t, ps, particle_radius=r)
    lo_all, hi_all = ps - r[:, None], ps + r[:, None]
    for k in range(t["parent"].shape[0]):
        a, b = t["node_ranges"][k]
        mn, mx = lo_all[a:b + 1].min(0), hi_all[a:b + 1].max(0)
        np.testing.assert_array_equal(g["center"][k], np.float32(0.5) * (mn + mx))
        np.testing.assert_array_equal(
            g["half_extent"][k], np.maximum(np.float32(0.5) * (mx - mn), np.float32(1e-6)))
