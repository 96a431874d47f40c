"""GPU per-node mass moments against the oracle.

The oracle differences extended-precision prefix sums, i.e. it gives the
node sums of the reference's per-particle terms almost exactly; the GPU
accumulates each node sum in float64 and rounds once. Both are within one
rounding of the exact sum, so the bar is one unit in the last place of the
output dtype (centres: the division adds one more rounding). The
reference's own float32 prefix difference is less accurate than either and
its rounding order is not fixed by its code, so bit parity with it is not
claimed."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from tests.test_gpu_parity import _run_both
from yggdrax_amd import inputs
from yggdrax_amd import tree_moments as tm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ygg():
    import yggdrax_amd
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return yggdrax_amd


def _within_ulps(got, exp, ulps):
    got = np.asarray(got, np.float64)
    e = np.asarray(exp)
    tol = ulps * np.spacing(np.abs(e)).astype(np.float64)
    return np.abs(got - e.astype(np.float64)) <= tol


def _check(mm, ref_mm, label):
    m_ok = _within_ulps(parity.np_(mm.mass), ref_mm["mass"], 1)
    c = parity.np_(mm.center_of_mass)
    rc = ref_mm["center_of_mass"]
    # a centre is moment / mass: one rounding each in the moment, the mass
    # and the quotient
    c_ok = _within_ulps(c, rc, 3) | (np.abs(c - rc) <= 4 * np.spacing(
        np.abs(rc).max(axis=1, keepdims=True)))
    exact = float(np.mean(parity.np_(mm.mass) == ref_mm["mass"]))
    print(label, "mass exactly equal", exact, flush=True)
    assert parity.np_(mm.mass).dtype == ref_mm["mass"].dtype
    assert c.dtype == rc.dtype
    assert m_ok.all(), np.flatnonzero(~m_ok)[:5]
    assert c_ok.all(), np.argwhere(~c_ok)[:5]


@pytest.mark.parametrize("name", ["clustered_32k_octree", "plummer_16k_radix",
                                  "getting_started_radix"])
def test_mass_moments_match_oracle(ygg, name):
    (ref, rgeom, _, _), (tree, ps, geom, _, _), _ = _run_both(ygg, name)
    ms = tree.masses_sorted
    mm = tm.compute_tree_mass_moments(tree, ps, ms)
    exp = oracle.compute_tree_mass_moments(ref, ref["positions_sorted"], ref["masses_sorted"])
    _check(mm, exp, name)
    # float64 inputs, and mixed dtypes
    for pdt, mdt in ((torch.float64, torch.float64), (torch.float32, torch.float64),
                     (torch.float64, torch.float32)):
        mm = tm.compute_tree_mass_moments(tree, ps.to(pdt), ms.to(mdt))
        exp = oracle.compute_tree_mass_moments(
            ref, ref["positions_sorted"].astype(parity.np_(ps.to(pdt)[:1]).dtype),
            ref["masses_sorted"].astype(parity.np_(ms.to(mdt)[:1]).dtype))
        if pdt == torch.float32 and mdt == torch.float32:
            continue
        # float64 node sums: the oracle's extended-precision prefix sums carry
        # an error up to n * 2^-64 times the prefix magnitude
        got_m = parity.np_(mm.mass).astype(np.float64)
        tol = 1e-12 * float(np.abs(exp["mass"]).max())
        assert np.all(np.abs(got_m - exp["mass"]) <= tol), (pdt, mdt)
        assert parity.np_(mm.center_of_mass).dtype == np.float64


def test_mass_moments_leaf_size_one_sample(ygg):
    """tests/unit/test_tree_moments.py:18-40, 69-83, 135-160 on the GPU: the
    three-particle sample tree with leaf size 1."""
    pos = torch.tensor([[-0.5, -0.5, -0.5], [0.1, 0.0, 0.3], [0.6, 0.4, -0.2]],
                       dtype=torch.float32).cuda()
    mass = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float32).cuda()
    bounds = (torch.full((3,), -1.0).cuda(), torch.ones(3).cuda())
    tree, ps, ms, _ = ygg.build_tree(pos, mass, bounds, return_reordered=True, leaf_size=1)
    mm = tm.compute_tree_mass_moments(tree, ps, ms)
    ni = int(tree.num_internal_nodes)
    assert mm.mass.shape == (tree.num_nodes,) and mm.center_of_mass.shape == (tree.num_nodes, 3)
    assert torch.equal(mm.mass[ni:], ms)
    assert torch.allclose(mm.center_of_mass[ni:], ps, rtol=1e-6, atol=0)
    root = int(torch.nonzero(tree.parent == -1)[0])
    assert float(mm.mass[root]) == 6.0
    want = (mass[:, None].double() * pos.double()).sum(0) / 6.0
    assert torch.allclose(mm.center_of_mass[root].double(), want, rtol=1e-6, atol=1e-7)


def test_mass_moments_zero_mass_and_errors(ygg):
    """Zero-mass nodes get a zero centre; mismatched inputs raise the
    reference's errors."""
    pos, mass = inputs.uniform_cube(1000, 5)
    mass[:500] = 0.0
    tree, ps, ms, _ = ygg.build_tree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                     leaf_size=8, return_reordered=True)
    mm = tm.compute_tree_mass_moments(tree, ps, ms)
    zero = mm.mass == 0
    assert bool(zero.any())
    assert torch.all(mm.center_of_mass[zero] == 0)
    with pytest.raises(ValueError, match="positions_sorted must match tree.num_particles"):
        tm.compute_tree_mass_moments(tree, ps[:-1], ms)
    with pytest.raises(ValueError, match="masses_sorted must match tree.num_particles"):
        tm.compute_tree_mass_moments(tree, ps, ms[:-1])
    with pytest.raises(ValueError, match=r"positions must have shape \(N, 3\)"):
        tm.compute_tree_mass_moments(tree, ps[:, :2], ms)


def test_mass_moments_large_root_total(ygg):
    """2^22 clustered points: the root holds the total mass and the
    mass-weighted mean (float64 reductions) within one float32 rounding."""
    pos, mass = inputs.clustered(1 << 22, 3)
    tree, ps, ms, _ = ygg.build_octree(torch.from_numpy(pos).cuda(),
                                       torch.from_numpy(mass).cuda(), leaf_size=64,
                                       return_reordered=True)
    mm = tm.compute_tree_mass_moments(tree, ps, ms)
    root = int(torch.nonzero(tree.parent == -1)[0])
    m64 = ms.double()
    total = float(m64.sum())
    assert abs(float(mm.mass[root]) - total) <= float(np.spacing(np.float32(total)))
    w = (ms[:, None] * ps).double().sum(0)  # float32 terms, float64 sum
    com = (w / float(mm.mass[root])).cpu().numpy()
    got = parity.np_(mm.center_of_mass[root]).astype(np.float64)
    assert np.all(np.abs(got - com) <= 4 * np.spacing(np.abs(com).astype(np.float32)))
