"""Learned stream capacities for the dual-tree walk.

A small MLP (8 -> 64 -> 64 -> 64 -> 3, ReLU, bfloat16 weights, float32
accumulation; HIP kernels in csrc/ygg_mlp.hip) predicts for every node the
log2 of a Poisson rate: the expected far-list count, near-list count
(leaves) and number of pairs with this node as target in the largest
wavefront, each over a per-output scale. Per-node features: level,
particle count, max half extent, sphere radius, parent far count, sibling
particle count, density relative to the root, leaf flag (log-scaled). The
parent far count is known exactly when training (the walk produces it); at
inference a first pass predicts every node's far count with that feature
zeroed and a second pass feeds the parent's prediction.

Trained on the Poisson negative log-likelihood, the rates are unbiased
estimates of the counts, so their sums estimate the far/near stream sizes
and the pair-queue size; a calibration factor, a margin and a small
per-node floor (which covers small trees, whose far totals are tiny and
erratic) turn them into capacities. The walk then runs once instead of
counting first. The walk detects every overflow exactly and falls back to
counting, so lists never depend on the model.
"""

from __future__ import annotations

import ctypes
import math
import os
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib
from .interactions import _WalkContext, build_interactions_and_neighbors
from .tree import device_topology

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "capacity_mlp.npz")
N_IN, N_HID, N_OUT = 8, 64, 3


def num_params() -> int:
    return int(_lib.load().ygg_mlp_num_params())


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the kernel's parameter layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for nin, nout in ((N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)):
        lim = math.sqrt(6.0 / nin)
        parts.append(rng.uniform(-lim, lim, (nout, nin)).astype(np.float32).ravel())
        parts.append(np.zeros(nout, np.float32))
    p = np.concatenate(parts)
    assert p.size == num_params()
    return p


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 -> float32 with round-to-nearest-even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def features(dt, geom, parent_far: Optional[torch.Tensor] = None) -> torch.Tensor:
    lib = _lib.load()
    dev = geom.center.device
    x = torch.empty((dt.num_nodes, N_IN), dtype=torch.float32, device=dev)
    left = dt.left if dt.left.numel() else None
    right = dt.right if dt.right.numel() else None
    _lib.check(lib.ygg_mlp_features(
        dt.num_nodes, dt.node_level.data_ptr(), dt.node_ranges.data_ptr(), dt.parent.data_ptr(),
        _lib.ptr(left), _lib.ptr(right), geom.half_extent.data_ptr(), geom.radius.data_ptr(),
        geom.max_extent.data_ptr(), _lib.ptr(parent_far), x.data_ptr(), _lib.stream_ptr(dev)),
        "mlp features")
    return x


# Per-node capacity floors added to the calibrated sums: far entries per
# node, near entries per leaf, queue pairs per node.
FLOORS = np.array([8.0, 64.0, 8.0], np.float32)


@dataclass
class CapacityModel:
    params: np.ndarray                      # float32 master parameters
    scale: np.ndarray = field(default_factory=lambda: np.ones(3, np.float32))
    calib: np.ndarray = field(default_factory=lambda: np.ones(3, np.float32))
    margin: np.ndarray = field(default_factory=lambda: np.full(3, 0.1, np.float32))
    floor: np.ndarray = field(default_factory=lambda: FLOORS.copy())
    meta: dict = field(default_factory=dict)
    _dev: dict = field(default_factory=dict)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "CapacityModel":
        z = np.load(path, allow_pickle=False)
        meta = {k[5:]: str(z[k]) for k in z.files if k.startswith("meta_")}
        return cls(z["params"].astype(np.float32), z["scale"].astype(np.float32),
                   z["calib"].astype(np.float32), z["margin"].astype(np.float32),
                   z["floor"].astype(np.float32), meta)

    def save(self, path: str = DEFAULT_WEIGHTS) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        extra = {f"meta_{k}": np.array(str(v)) for k, v in self.meta.items()}
        np.savez(path, params=self.params, scale=self.scale, calib=self.calib,
                 margin=self.margin, floor=self.floor, **extra)

    def device_weights(self, device) -> torch.Tensor:
        key = str(device)
        w = self._dev.get(key)
        if w is None:
            p = torch.from_numpy(self.params).to(device)
            w = torch.empty(p.numel(), dtype=torch.int16, device=device)
            _lib.check(_lib.load().ygg_mlp_params_to_bf16(p.data_ptr(), w.data_ptr(),
                                                          _lib.stream_ptr(device)), "bf16")
            self._dev[key] = w
        return w

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty((x.shape[0], N_OUT), dtype=torch.float32, device=x.device)
        w = self.device_weights(x.device)
        _lib.check(_lib.load().ygg_mlp_forward(w.data_ptr(), x.data_ptr(), x.shape[0],
                                               y.data_ptr(), _lib.stream_ptr(x.device)),
                   "mlp forward")
        return y

    def predict_nodes(self, tree, geom) -> torch.Tensor:
        """Two-pass per-node prediction (log2 rate / scale per output)."""
        lib = _lib.load()
        dt = device_topology(tree)
        dev = geom.center.device
        y0 = self.forward(features(dt, geom, None))
        pf = torch.empty(dt.num_nodes, dtype=torch.float32, device=dev)
        _lib.check(lib.ygg_mlp_far_counts(y0.data_ptr(), dt.num_nodes, float(self.scale[0]),
                                          pf.data_ptr(), _lib.stream_ptr(dev)), "far counts")
        return self.forward(features(dt, geom, pf))

    def predict_totals(self, tree, geom) -> np.ndarray:
        """Uncalibrated sums of the predicted rates (far, near, queue)."""
        dt = device_topology(tree)
        y = self.predict_nodes(tree, geom)
        tot = torch.empty(3, dtype=torch.float64, device=y.device)
        s = [float(v) for v in self.scale]
        _lib.check(_lib.load().ygg_mlp_totals(y.data_ptr(), dt.num_nodes, s[0], s[1], s[2],
                                              tot.data_ptr(), _lib.stream_ptr(y.device)),
                   "totals")
        return tot.cpu().numpy()

    @staticmethod
    def _units(tree_or_nodes) -> np.ndarray:
        T = int(getattr(tree_or_nodes, "num_nodes", tree_or_nodes))
        return np.array([T, (T + 1) // 2, T], np.float64)

    def capacities(self, tree, totals: np.ndarray) -> np.ndarray:
        return totals * self.calib * (1.0 + self.margin) + self.floor * self._units(tree)

    def capacity_hint(self, tree, geom) -> dict:
        """Far / near stream and queue capacities for a one-pass walk."""
        c = self.capacities(tree, self.predict_totals(tree, geom))
        return {"far": int(c[0]) + 1024, "near": int(c[1]) + 1024, "queue": int(c[2]) + 1024}

    def predict_totals_batched(self, trees, geom) -> np.ndarray:
        """Uncalibrated predicted totals over every system of a batch
        (two-pass inference with each system's own root)."""
        lib = _lib.load()
        lay = trees.layout
        dev = geom.center.device
        st = _lib.stream_ptr(dev)
        T = lay.num_nodes
        r = trees.rad

        def feats(pf):
            x = torch.empty((T, N_IN), dtype=torch.float32, device=dev)
            _lib.check(lib.ygg_batch_mlp_features(
                lay.ptr(), r["node_level"].data_ptr(), r["node_ranges"].data_ptr(),
                r["parent"].data_ptr(), r["left_child"].data_ptr(), r["right_child"].data_ptr(),
                geom.half_extent.data_ptr(), geom.radius.data_ptr(), geom.max_extent.data_ptr(),
                _lib.ptr(pf), x.data_ptr(), st), "batch features")
            return x

        y0 = self.forward(feats(None))
        pf = torch.empty(T, dtype=torch.float32, device=dev)
        _lib.check(lib.ygg_mlp_far_counts(y0.data_ptr(), T, float(self.scale[0]), pf.data_ptr(),
                                          st), "far counts")
        y = self.forward(feats(pf))
        tot = torch.empty(3, dtype=torch.float64, device=dev)
        sc = [float(v) for v in self.scale]
        _lib.check(lib.ygg_batch_mlp_totals(lay.ptr(), y.data_ptr(), sc[0], sc[1], sc[2],
                                            tot.data_ptr(), st), "batch totals")
        return tot.cpu().numpy()

    def capacity_hint_batched(self, trees, geom) -> dict:
        """Far / near stream and queue capacities for one forest walk over a
        batch (sums over its systems; floors per node of the whole batch)."""
        lay = trees.layout
        units = np.array([lay.num_nodes, lay.num_leaves, lay.num_nodes], np.float64)
        t = self.predict_totals_batched(trees, geom)
        c = t * self.calib * (1.0 + self.margin) + self.floor * units
        return {"far": int(c[0]) + 1024, "near": int(c[1]) + 1024,
                "queue": int(c[2]) + 1024 + lay.num_systems}


# ---- training data from the walk itself ------------------------------------

def node_targets(tree, geom, theta, mac_type="dehnen"):
    """Per-node target counts (far-list entries, near-list entries, pairs
    with the node as target in the peak wavefront) and masks, from the walk."""
    dt = device_topology(tree)
    dev = geom.center.device
    inter, nbrs = build_interactions_and_neighbors(tree, geom, theta=theta, mac_type=mac_type,
                                                   neighbor_positions=False, strict_int32=False)
    far = inter.counts.to(torch.float32)
    near = torch.zeros(dt.num_nodes, dtype=torch.float32, device=dev)
    near[dt.leaf_indices.long()] = nbrs.counts.to(torch.float32)
    ctx = _WalkContext(dt, geom, theta, mac_type, 1.0)
    T, NI = dt.num_nodes, dt.num_internal
    qbig = max(T * 128, 4096)
    hist = torch.zeros(T, dtype=torch.int32, device=dev)
    if NI > 0:
        tot = ctx.run_flat(emit=False, queue_cap=qbig)
        g_peak = int(np.argmax(tot["gen_sizes"]))
        ctx.run_flat(emit=False, queue_cap=qbig, hist_gen=g_peak, node_hist=hist)
        peak = tot["peak"]
    else:
        peak = 1
    y = torch.stack([far, near, hist.to(torch.float32)], dim=1).contiguous()
    mask = torch.ones_like(y)
    mask[:NI, 1] = 0.0
    totals = np.array([float(inter.counts.sum()), float(nbrs.counts.sum()), float(peak)])
    return y, mask, far, totals


def collect(systems, device="cuda", keep_trees=False):
    """systems: iterable of (name, builder, make, leaf, theta), make() ->
    (positions, masses). Returns features (with exact parent far counts),
    targets, masks and per-system records (true totals; the tree and
    geometry too when keep_trees)."""
    import yggdrax_amd as ygg
    xs, ys, ms, recs = [], [], [], []
    for name, builder, make, leaf, theta in systems:
        pos, mass = make()
        build = ygg.build_octree if builder == "octree" else ygg.build_tree
        tree, ps, _, _ = build(torch.from_numpy(pos).to(device), torch.from_numpy(mass).to(device),
                               leaf_size=leaf, return_reordered=True)
        geom = ygg.compute_tree_geometry(tree, ps)
        y, m, far, totals = node_targets(tree, geom, theta)
        dt = device_topology(tree)
        xs.append(features(dt, geom, far))
        ys.append(y)
        ms.append(m)
        rec = dict(name=name, totals=totals, num_nodes=int(tree.num_nodes))
        if keep_trees:
            rec.update(tree=tree, geom=geom)
        recs.append(rec)
    return torch.cat(xs), torch.cat(ys), torch.cat(ms), recs


def training_systems(seeds=range(1000, 1200)):
    """Seeded systems from the three benchmark generators with each
    generator's tree parameters: uniform cube (octree, leaf 16, theta 0.6,
    2^9..2^16 points), Plummer sphere (radix, leaf 32, theta 0.5,
    2^13..2^20) and clustered (octree, leaf 64, theta 0.5, 2^15..2^24);
    sizes log-uniform. Positions are generated lazily."""
    from . import inputs
    out = []
    for k, seed in enumerate(seeds):
        kind = k % 3
        rng = np.random.default_rng(seed)
        if kind == 0:
            n = int(round(2 ** rng.uniform(9, 16)))
            out.append((f"uniform{n}_{seed}", "octree",
                        lambda n=n, s=seed: inputs.uniform_cube(n, s), 16, 0.6))
        elif kind == 1:
            n = int(round(2 ** rng.uniform(13, 20)))
            out.append((f"plummer{n}_{seed}", "radix",
                        lambda n=n, s=seed: inputs.plummer_sphere(n, s), 32, 0.5))
        else:
            n = int(round(2 ** rng.uniform(15, 24)))
            out.append((f"clustered{n}_{seed}", "octree",
                        lambda n=n, s=seed: inputs.clustered(n, s), 64, 0.5))
    return out


class Trainer:
    """Adam on the masked Poisson NLL with the HIP training kernels."""

    def __init__(self, params: np.ndarray, batch: int, device="cuda"):
        self.lib = _lib.load()
        self.dev = torch.device(device)
        self.B = int(batch)
        self.p = torch.from_numpy(params.copy()).to(self.dev)
        self.pbf = torch.empty(self.p.numel(), dtype=torch.int16, device=self.dev)
        _lib.check(self.lib.ygg_mlp_params_to_bf16(self.p.data_ptr(), self.pbf.data_ptr(),
                                                   self._st()), "bf16")
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.g = torch.zeros_like(self.p)
        self.ws = torch.empty(int(self.lib.ygg_mlp_train_workspace_floats(self.B)),
                              dtype=torch.float32, device=self.dev)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self.step_no = 0

    def _st(self):
        return _lib.stream_ptr(self.dev)

    def step(self, x, y, m, lr: float) -> None:
        assert x.shape[0] == self.B
        self.step_no += 1
        _lib.check(self.lib.ygg_mlp_train_step(
            self.p.data_ptr(), self.pbf.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
            self.step_no, ctypes.c_float(lr).value, x.data_ptr(), y.data_ptr(), m.data_ptr(),
            self.B, self.g.data_ptr(), self.ws.data_ptr(), self.loss.data_ptr(), self._st()),
            "mlp train step")

    def gradients(self, x, y, m) -> torch.Tensor:
        """Loss gradient on one batch without updating (lr = 0)."""
        _lib.check(self.lib.ygg_mlp_train_step(
            self.p.data_ptr(), self.pbf.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), 1,
            0.0, x.data_ptr(), y.data_ptr(), m.data_ptr(), self.B, self.g.data_ptr(),
            self.ws.data_ptr(), self.loss.data_ptr(), self._st()), "mlp gradients")
        return self.g.clone()


def target_scale(y, m) -> np.ndarray:
    """Per-output masked mean count (the unit each output's rate is in)."""
    s = (y * m).sum(0) / m.sum(0).clamp(min=1.0)
    return np.maximum(s.cpu().numpy().astype(np.float32), 1e-3)


def train(x, y, m, *, steps=4000, batch=4096, lr=3e-3, seed=0, log=print):
    """Trains on features x, target counts y and masks m; returns (params,
    scale). Half of every batch has the parent far-count feature zeroed so
    one network serves both inference passes."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    scale = target_scale(y, m)
    ys = (y / torch.from_numpy(scale).to(y.device)).contiguous()
    tr = Trainer(init_params(seed), batch, x.device)
    n = x.shape[0]
    t0 = time.time()
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(x.device)
        xb = x[idx].clone()
        xb[: batch // 2, 4] = 0.0
        rate = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
        tr.step(xb.contiguous(), ys[idx].contiguous(), m[idx].contiguous(), rate)
        if log and (it % 500 == 0 or it == steps - 1):
            log(f"[train] step {it} loss {float(tr.loss):.5f} ({time.time() - t0:.1f} s)")
    return tr.p.cpu().numpy(), scale


def calibrate(model: CapacityModel, recs, quantile=0.98) -> None:
    """calib = median(true / predicted total) per output; margin so that the
    `quantile` of calibration systems fit once the per-node floors are
    added."""
    ratios, need = [], []
    for r in recs:
        pred = np.maximum(model.predict_totals(r["tree"], r["geom"]), 1.0)
        true = np.asarray(r["totals"], np.float64)
        ratios.append(true / pred)
        need.append((true - model.floor * model._units(r["tree"]), pred))
    med = np.median(np.array(ratios), axis=0)
    req = np.array([(nd / (med * pr)) - 1.0 for nd, pr in need])
    model.calib = med.astype(np.float32)
    model.margin = np.maximum(np.quantile(req, quantile, axis=0), 0.05).astype(np.float32)


__all__ = ["CapacityModel", "Trainer", "bf16_round", "calibrate", "collect", "features",
           "init_params", "node_targets", "target_scale", "train", "training_systems"]
