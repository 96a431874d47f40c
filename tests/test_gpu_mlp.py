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
    t = rng.poisson(2.0, (B, 3)).astype(np.float32)
    m = (rng.uniform(size=(B, 3)) > 0.25).astype(np.float32)
    return x, t, m


@pytest.mark.parametrize("B", [1000, 3 * 1024 * 16 + 5])
def test_mlp_forward_matches_reference(B):
    """MFMA inference against the float32 reference. 1000 samples end in a
    partial 16-sample tile; 49157 samples give each of the 1024 wavefronts
    about three tiles (the tile loop and its feature prefetch)."""
    from yggdrax_amd import learned
    p = learned.init_params(3)
    x, _, _ = _batch(B, 1)
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
    y = rng.poisson(np.exp2(np.clip(x @ w, -4, 6))).astype(np.float32)
    m = np.ones_like(y)
    xt, yt, mt = (torch.from_numpy(a).cuda() for a in (x, y, m))
    p, scale = learned.train(xt, yt, mt, steps=600, batch=1024, lr=3e-3, log=None)
    ys = y[:4096] / scale
    # loss of the constant-rate model (the best a model without features can do)
    p0 = np.zeros_like(p)
    p0[-3:] = np.log2(ys.mean(0))
    loss0, _ = mlp_ref.loss_and_grad(p0, x[:4096], ys, m[:4096])
    loss1, _ = mlp_ref.loss_and_grad(p, x[:4096], ys, m[:4096])
    # Poisson NLL is offset by the targets' own entropy term; compare the
    # excess over the saturated model
    sat = float((ys * (1 - np.log(np.maximum(ys, 1e-30)))).sum(1).mean())
    assert (loss1 - sat) < 0.3 * (loss0 - sat)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "yggdrax_amd", "data", "capacity_mlp.npz")),
    reason="trained capacity model not present")
def test_learned_capacity_path_matches_plain_path():
    """Held-out clustered system: the one-pass walk sized by the trained
    network returns the same lists as the plain walk, with no counting walk,
    no rerun and no stream growth, and sizes its streams well below the
    reference's fixed per-node capacities."""
    import time
    import yggdrax_amd as ygg
    from yggdrax_amd import inputs, interactions, learned
    model = learned.CapacityModel.load()
    pos, mass = inputs.clustered(1 << 18, 4242)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(), torch.from_numpy(mass).cuda(),
                                      leaf_size=64, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    kw = dict(theta=0.5, mac_type="dehnen", neighbor_positions=False, strict_int32=False)
    st = interactions.FLAT_STATS
    p0 = st["plain_walks"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_i, plain_n = ygg.build_interactions_and_neighbors(tree, geom, **kw)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    assert st["plain_walks"] == p0 + 1
    before = dict(st)
    t0 = time.perf_counter()
    hint = model.capacity_hint(tree, geom)
    li, ln = ygg.build_interactions_and_neighbors(tree, geom, capacity_hint=hint, **kw)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    print(f"plain {t_plain * 1e3:.2f} ms, learned {t_learned * 1e3:.2f} ms, hint {hint}, "
          f"true far {plain_i.sources.numel()} near {plain_n.neighbors.numel()}")
    for k in ("count_walks", "plain_walks", "hint_overflows", "hint_growths"):
        assert st[k] == before[k], f"the learned hint did not hold ({k})"
    nf, nn = plain_i.sources.numel(), plain_n.neighbors.numel()
    assert nf <= hint["far"] <= 3 * nf + (1 << 20)
    assert nn <= hint["near"] <= 3 * nn + (1 << 20)
    # the reference's fixed-capacity buffers: max_neighbors_per_leaf (2048)
    # slots per leaf and its smallest far capacity per node
    fixed = 2048 * ln.counts.numel() + 64 * tree.num_nodes
    assert hint["far"] + hint["near"] < fixed
    for a, b in ((plain_i.counts, li.counts), (plain_i.offsets, li.offsets),
                 (plain_i.targets, li.targets), (plain_n.counts, ln.counts),
                 (plain_n.offsets, ln.offsets)):
        assert torch.equal(a, b)
    for (o, c, x1, x2) in ((li.offsets, li.counts, plain_i.sources, li.sources),
                           (ln.offsets, ln.counts, plain_n.neighbors, ln.neighbors)):
        off, cnt = parity.np_(o), parity.np_(c)
        s1, s2 = parity.np_(x1), parity.np_(x2)
        # entry e belongs to the row whose [offset, offset + count) holds it
        order = np.argsort(off[: len(cnt)], kind="stable")
        rows = np.repeat(order, cnt[order]).astype(np.int64)
        assert len(rows) == len(s1) == len(s2)
        k1 = np.sort(rows << 32 | s1.astype(np.int64))
        k2 = np.sort(rows << 32 | s2.astype(np.int64))
        assert np.array_equal(k1, k2)
