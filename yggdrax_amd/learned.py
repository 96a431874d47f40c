"""Learned stream capacities for the dual-tree walk.

A small MLP (8 -> 64 -> 64 -> 64 -> 3, ReLU, bfloat16 weights, float32
accumulation; HIP kernels in csrc/ygg_mlp.hip) predicts for every node
log2(1 + far-list count), log2(1 + near-list count) (leaves) and
log2(1 + number of pairs with this node as target in the largest
wavefront). Per-node features: level, particle count, max half extent,
sphere radius, parent far count, sibling particle count, density relative
to the root, leaf flag (log-scaled). The parent far count is known exactly
when training (the walk produces it); at inference a first pass predicts
every node's far count with that feature zeroed and a second pass feeds
the parent's prediction.

Summed and calibrated, the predictions give the far/near stream sizes and
the pair-queue size, so the walk runs once instead of counting first. The
walk detects every overflow exactly and falls back to counting, so lists
never depend on the model.
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


@dataclass
class CapacityModel:
    params: np.ndarray                      # float32 master parameters
    calib: np.ndarray = field(default_factory=lambda: np.ones(3, np.float32))
    margin: np.ndarray = field(default_factory=lambda: np.full(3, 0.1, np.float32))
    meta: dict = field(default_factory=dict)
    _dev: dict = field(default_factory=dict)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "CapacityModel":
        z = np.load(path, allow_pickle=False)
        meta = {k[5:]: str(z[k]) for k in z.files if k.startswith("meta_")}
        return cls(z["params"].astype(np.float32), z["calib"].astype(np.float32),
                   z["margin"].astype(np.float32), meta)

    def save(self, path: str = DEFAULT_WEIGHTS) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        extra = {f"meta_{k}": np.array(str(v)) for k, v in self.meta.items()}
        np.savez(path, params=self.params, calib=self.calib, margin=self.margin, **extra)

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
        """Two-pass per-node prediction (log2(1 + count) per output)."""
        lib = _lib.load()
        dt = device_topology(tree)
        dev = geom.center.device
        y0 = self.forward(features(dt, geom, None))
        pf = torch.empty(dt.num_nodes, dtype=torch.float32, device=dev)
        _lib.check(lib.ygg_mlp_far_counts(y0.data_ptr(), dt.num_nodes, pf.data_ptr(),
                                          _lib.stream_ptr(dev)), "far of")
        return self.forward(features(dt, geom, pf))

    def predict_totals(self, tree, geom) -> np.ndarray:
        dt = device_topology(tree)
        y = self.predict_nodes(tree, geom)
        tot = torch.empty(3, dtype=torch.float64, device=y.device)
        _lib.check(_lib.load().ygg_mlp_totals(y.data_ptr(), dt.num_nodes, tot.data_ptr(),
                                              _lib.stream_ptr(y.device)), "totals")
        return tot.cpu().numpy()

    def capacity_hint(self, tree, geom) -> dict:
        """Far / near stream and queue capacities for a one-pass walk."""
        t = self.predict_totals(tree, geom) * self.calib * (1.0 + self.margin)
        return {"far": int(t[0]) + 1024, "near": int(t[1]) + 1024, "queue": int(t[2]) + 1024}


# ---- training data from the walk itself ------------------------------------

def node_targets(tree, geom, theta, mac_type="dehnen"):
    """Per-node targets (log2(1 + far), log2(1 + near), log2(1 + peak-wavefront
    pairs as target)) and masks, from the exact walk."""
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
    y = torch.stack([torch.log2(1 + far), torch.log2(1 + near),
                     torch.log2(1 + hist.to(torch.float32))], dim=1).contiguous()
    mask = torch.ones_like(y)
    mask[:NI, 1] = 0.0
    totals = np.array([float(inter.counts.sum()), float(nbrs.counts.sum()), float(peak)])
    return y, mask, far, totals


def collect(systems, device="cuda"):
    """systems: iterable of (name, builder, pos, mass, leaf, theta). Returns
    features (with exact parent far counts), targets, masks and per-system
    true totals."""
    import yggdrax_amd as ygg
    xs, ys, ms, recs = [], [], [], []
    for name, builder, pos, mass, leaf, theta in systems:
        build = ygg.build_octree if builder == "octree" else ygg.build_tree
        tree, ps, _, _ = build(torch.from_numpy(pos).to(device), torch.from_numpy(mass).to(device),
                               leaf_size=leaf, return_reordered=True)
        geom = ygg.compute_tree_geometry(tree, ps)
        y, m, far, totals = node_targets(tree, geom, theta)
        dt = device_topology(tree)
        x = features(dt, geom, far)
        xs.append(x)
        ys.append(y)
        ms.append(m)
        recs.append(dict(name=name, tree=tree, geom=geom, totals=totals))
    return torch.cat(xs), torch.cat(ys), torch.cat(ms), recs


def training_systems(seeds=range(1000, 1200)):
    """200 seeded systems from the three benchmark generators (uniform cube,
    Plummer sphere, clustered), at sizes that keep collection fast."""
    from . import inputs
    out = []
    for k, seed in enumerate(seeds):
        kind = k % 3
        rng = np.random.default_rng(seed)
        if kind == 0:
            n = int(rng.choice([512, 2048, 8192]))
            pos, mass = inputs.uniform_cube(n, seed)
            out.append((f"uniform{n}_{seed}", "octree", pos, mass, 16, 0.6))
        elif kind == 1:
            n = int(2 ** rng.integers(13, 18))
            pos, mass = inputs.plummer_sphere(n, seed)
            out.append((f"plummer{n}_{seed}", "radix", pos, mass, 32, 0.5))
        else:
            n = int(2 ** rng.integers(15, 20))
            pos, mass = inputs.clustered(n, seed)
            out.append((f"clustered{n}_{seed}", "octree", pos, mass, 64, 0.5))
    return out


class Trainer:
    """Adam on the masked MSE with the HIP training kernels."""

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


def train(x, y, m, *, steps=4000, batch=4096, lr=3e-3, seed=0, log=print):
    """Trains on (x, y, m). Half of every batch has the parent far-count
    feature zeroed so one network serves both inference passes."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    tr = Trainer(init_params(seed), batch, x.device)
    n = x.shape[0]
    t0 = time.time()
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(x.device)
        xb = x[idx].clone()
        xb[: batch // 2, 4] = 0.0
        rate = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
        tr.step(xb.contiguous(), y[idx].contiguous(), m[idx].contiguous(), rate)
        if log and (it % 500 == 0 or it == steps - 1):
            log(f"[train] step {it} loss {float(tr.loss):.5f} ({time.time() - t0:.1f} s)")
    return tr.p.cpu().numpy()


def calibrate(model: CapacityModel, recs, quantile=0.98) -> None:
    """calib = median(true / predicted total) per output; margin so that the
    `quantile` of calibration systems fit."""
    ratios = []
    for r in recs:
        pred = model.predict_totals(r["tree"], r["geom"])
        ratios.append(np.asarray(r["totals"], np.float64) / np.maximum(pred, 1.0))
    ratios = np.array(ratios)
    med = np.median(ratios, axis=0)
    q = np.quantile(ratios / med, quantile, axis=0)
    model.calib = med.astype(np.float32)
    model.margin = np.maximum(q - 1.0, 0.02).astype(np.float32)


__all__ = ["CapacityModel", "Trainer", "bf16_round", "calibrate", "collect", "features",
           "init_params", "node_targets", "train", "training_systems"]
