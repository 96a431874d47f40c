"""The benchmark workload at full size (2^24 clustered points, build_octree
leaf 64, dehnen theta 0.5): every output field against the oracle's digest
fixture (tests/golden/c3_clustered_2p24.npz, written on the CPU by
tests/golden/make_fullsize.py), the learned-capacity run against the same
fixture, and properties that hold whatever the oracle says:

* the CSR lists tile their entry arrays exactly;
* far and near lists are symmetric (the walk emits both directions);
* sampled far entries satisfy the MAC and sampled near entries are pairs of
  different leaves that fail it (float32, in the kernel's operation order);
* every sampled ordered pair of different leaves (a, b) is covered exactly
  once: by one far entry y in row x with x an ancestor-or-self of a and y
  of b, or by b in a's near row (the dual-tree walk partitions the leaf
  product).
"""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from tests import parity
from tests.golden import fullsize as FS
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "c3_clustered_2p24.npz")


def _ancestors(parent, node):
    out = [node]
    while parent[out[-1]] >= 0:
        out.append(int(parent[out[-1]]))
    return out


@pytest.fixture(scope="module")
def c3():
    import yggdrax_amd as ygg
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    print("[fullsize] c3 clustered 2^24", flush=True)
    pos, mass = inputs.clustered(1 << 24, 3)
    tree, ps, _, _ = ygg.build_octree(torch.from_numpy(pos).cuda(),
                                      torch.from_numpy(mass).cuda(), leaf_size=64,
                                      return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs, res = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.5, mac_type="dehnen", neighbor_positions=False,
        strict_int32=False, return_result=True)
    torch.cuda.synchronize()
    return tree, geom, inter, nbrs, res


def test_clustered_2p24_every_field_matches_oracle(c3):
    tree, geom, inter, nbrs, res = c3
    fx = FS.load(FIXTURE)
    dec = (res.accept_decisions, res.near_decisions, res.refine_decisions)
    got = FS.gpu_system(tree, geom, inter, nbrs, octree=True, decisions=dec)
    print("[fullsize] far", got["far_total"], "near", got["near_total"], flush=True)
    bad = FS.compare(got, fx, 0)
    assert not bad, bad


def test_clustered_2p24_learned_capacity_matches_oracle(c3):
    """The one-pass walk sized by the capacity model: no counting walk, no
    stream growth, and every list field equal to the oracle's."""
    import yggdrax_amd as ygg
    from yggdrax_amd import learned
    from yggdrax_amd.interactions import FLAT_STATS
    tree, geom, _, _, _ = c3
    model = learned.CapacityModel.load()
    hint = model.capacity_hint(tree, geom)
    before = dict(FLAT_STATS)
    inter, nbrs = ygg.build_interactions_and_neighbors(
        tree, geom, theta=0.5, mac_type="dehnen", neighbor_positions=False,
        strict_int32=False, capacity_hint=hint)
    torch.cuda.synchronize()
    for k in ("count_walks", "plain_walks", "hint_overflows", "hint_growths"):
        assert FLAT_STATS[k] == before[k], f"the learned hint did not hold ({k})"
    got = FS.gpu_system(tree, geom, inter, nbrs, octree=True)
    bad = FS.compare(got, FS.load(FIXTURE), 0)
    assert not bad, bad


def test_clustered_2p24_list_properties(c3):
    tree, geom, inter, nbrs, _ = c3
    theta = 0.5
    T = tree.num_nodes
    L = tree.num_leaves
    NI = T - L
    F = inter.sources.numel()
    N = nbrs.neighbors.numel()
    print("far", F, "near", N, flush=True)

    # CSR tiling: far rows are placed level-major, near rows in native order
    cnt = inter.counts.long()
    off = inter.offsets[:T].long()
    nz = cnt > 0  # empty rows share their offset with the next row
    order = torch.argsort(off[nz])
    so, sc = off[nz][order], cnt[nz][order]
    assert int(so[0]) == 0 and int(so[-1] + sc[-1]) == F
    assert torch.equal(so[1:], so[:-1] + sc[:-1])
    assert bool(torch.all((off[~nz] >= 0) & (off[~nz] <= F)))
    ncnt, noff = nbrs.counts.long(), nbrs.offsets.long()
    assert int(noff[0]) == 0 and int(noff[-1]) == N
    assert torch.equal(noff[1:] - noff[:-1], ncnt)

    # symmetry
    tgt, src = inter.targets.long(), inter.sources.long()
    a = torch.sort(tgt * T + src)[0]
    b = torch.sort(src * T + tgt)[0]
    assert torch.equal(a, b)
    del a, b
    leaf_of_row = nbrs.leaf_indices.long()
    rows = torch.repeat_interleave(torch.arange(L, device=ncnt.device), ncnt)
    nt = leaf_of_row[rows]
    nb = nbrs.neighbors.long()
    assert bool(torch.all(nb >= NI)) and bool(torch.all(nb != nt))
    a = torch.sort(nt * T + nb)[0]
    b = torch.sort(nb * T + nt)[0]
    assert torch.equal(a, b)
    del a, b

    # MAC on samples, float32 in the kernel's order
    rng = np.random.default_rng(11)
    c = parity.np_(geom.center)
    e = parity.np_(geom.radius)  # dehnen, radius scale 1
    th = np.float32(theta)
    th2 = np.float32(th * th)

    def mac_ok(t, s):
        dx = c[t, 0] - c[s, 0]
        dy = c[t, 1] - c[s, 1]
        dz = c[t, 2] - c[s, 2]
        d2 = dx * dx
        d2 = d2 + dy * dy
        d2 = d2 + dz * dz
        sm = e[t] + e[s]
        return (t != s) & (d2 > 0) & (sm * sm <= th2 * d2)

    fi = rng.integers(0, F, 200000)
    ft, fs = parity.np_(tgt[fi]), parity.np_(src[fi])
    assert mac_ok(ft, fs).all()
    ni = rng.integers(0, N, 200000)
    nt_s, nb_s = parity.np_(nt[ni]), parity.np_(nb[ni])
    assert not mac_ok(nt_s, nb_s).any()

    # partition of the leaf product, on samples
    parent = parity.np_(tree.parent)
    srcs = parity.np_(src)
    offs, cnts = parity.np_(off), parity.np_(cnt)
    row_of_leaf = np.empty(T, np.int64)
    row_of_leaf[parity.np_(leaf_of_row)] = np.arange(L)
    nbh = parity.np_(nb)
    noffs = parity.np_(noff)
    pairs = [tuple(rng.integers(NI, T, 2)) for _ in range(300)]
    pairs += [(int(nt_s[k]), int(nb_s[k])) for k in range(150)]
    pairs += [(int(ft[k]), int(fs[k])) for k in range(150) if ft[k] >= NI and fs[k] >= NI]
    checked = 0
    for la, lb in pairs:
        la, lb = int(la), int(lb)
        if la == lb:
            continue
        # a far entry (x, y) covers (a, b) when x is an ancestor-or-self of a
        # and y of b
        anc_b = set(_ancestors(parent, lb))
        far_cover = 0
        for x in _ancestors(parent, la):
            row = srcs[offs[x]:offs[x] + cnts[x]]
            far_cover += int(np.isin(row, list(anc_b)).sum())
        r = row_of_leaf[la]
        near_cover = int((nbh[noffs[r]:noffs[r + 1]] == lb).sum())
        assert far_cover + near_cover == 1, (la, lb, far_cover, near_cover)
        checked += 1
    print("leaf pairs checked", checked, flush=True)
    assert checked >= 400
