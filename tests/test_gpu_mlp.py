"""Capacity MLP: HIP inference and training kernels against the numpy
reference (oracle/mlp_ref.py, itself pinned by finite differences), and the
learned capacity path against the plain path (identical lists)."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from oracle import mlp_ref
from tests import parity

pytestmark = pytest.mark.gpu


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (B, 8)).astype(np.float32)
    t = rng.normal(0, 2, (B, 3)).astype(np.float32)
    m = (rng.uniform(size=(B, 3)) > 0.25).astype(np.float32)
    return x, t, m


def test_mlp_forward_matches_reference():
    from yggdrax_amd import learned
    p = learned.init_params(3)
    x, _, _ = _batch(1000, 1)
    model = learned.CapacityModel(p)
    y = model.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = mlp_ref.forward(p, x)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("B", [256, 4096])
def test_mlp_gradients_match_reference(B):
    from yggdrax_amd import learned
    p = learned.init_params(5)
    x, t, m = _batch(B, 2)
    tr = learned.Trainer(p, B)
    g = tr.gradients(*(torch.from_numpy(a).cuda() for a in (x, t, m))).cpu().numpy()
    loss_ref, g_ref = mlp_ref.loss_and_grad(p, x, t, m)
    scale = np.abs(g_ref).max()
    np.testing.assert_allclose(g, g_ref, rtol=1e-3, atol=1e-5 * scale)
    assert abs(float(tr.loss) - loss_ref) <= 1e-4 * abs(loss_ref)


def test_mlp_training_reduces_loss():
    from yggdrax_amd import learned
    rng = np.random.default_rng(0)
    x = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    w = rng.normal(0, 1, (8, 3)).astype(np.float32)
    y = np.log2(1 + np.abs(x @ w) * 4).astype(np.float32)
    m = np.ones_like(y)
    xt, yt, mt = (torch.from_numpy(a).cuda() for a in (x, y, m))
    p0 = learned.init_params(0)
    loss0, _ = mlp_ref.loss_and_grad(p0, x[:4096], y[:4096], m[:4096])
    p = learned.train(xt, yt, mt, steps=600, batch=1024, lr=3e-3, log=None)
    loss1, _ = mlp_ref.loss_and_grad(p, x[:4096], y[:4096], m[:4096])
    assert loss1 < 0.2 * loss0


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "yggdrax_amd", "data", "capacity_mlp.npz")),
    reason="trained capacity model not present")
def test_learned_capacity_path_matches_plain_path():
    """Held-out clustered system: the one-pass walk sized by the trained
    network returns the same lists as the plain count-then-fill walk, and
    the network's hint suffices (no fallback)."""
    import yggdrax_amd as ygg
    from yggdrax_amd import inputs, learned
    model = learned.CapacityModel.load()
    pos, mass = inputs.clustered(1 << 18, 4242)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                      leaf_size=64, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    kw = dict(theta=0.5, mac_type="dehnen", neighbor_positions=False, strict_int32=False)
    plain_i, plain_n = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    hint = model.capacity_hint(tree, geom)
    assert hint["far"] >= plain_i.sources.numel() and hint["near"] >= plain_n.neighbors.numel()
    li, ln = ygg.build_interactions_and_neighbors(tree, geom, capacity_hint=hint, **kw)
    for a, b in ((plain_i.counts, li.counts), (plain_i.offsets, li.offsets),
                 (plain_i.targets, li.targets), (plain_n.counts, ln.counts),
                 (plain_n.offsets, ln.offsets)):
        assert torch.equal(a, b)
    off = parity.np_(li.offsets)
    cnt = parity.np_(li.counts)
    s1, s2 = parity.np_(plain_i.sources), parity.np_(li.sources)
    for k in range(len(cnt)):
        assert np.array_equal(np.sort(s1[off[k]:off[k] + cnt[k]]), np.sort(s2[off[k]:off[k] + cnt[k]]))
