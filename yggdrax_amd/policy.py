"""Custom pair policies (PairPolicy, yggdrax/_interactions_impl.py:324, 731-810).

A policy is a Python callable ``policy(policy_state, **kw) -> (actions, tags)``
over arrays of one walk generation (torch tensors on the device here, with
the reference's keyword names). It is evaluated in both directions; a pair
is accepted only when both directions accept and is near only when both
say near, otherwise it is refined. Accepted pairs carry the directed tags
(forward tag on the target's row, reverse tag on the source's row), exposed
in DualTreeWalkResult.interaction_tags.

A policy can also be given as data, a ``PairPolicyDescriptor``: the same
contract for the policies it can describe (accept by the built-in MAC, by
distance or always; near for leaf pairs or never; tags by distance or
opening-ratio buckets or target id modulo k), evaluated for both directions
by a HIP kernel instead of on the host.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import torch

from . import _lib
from .dtypes import INDEX_DTYPE
from .interactions import (MAX_GENERATIONS, _FlatResult, _WalkContext, _emulate_ladder,
                           _flat_outputs, _ladder, _to_result, DEFAULT_PAIR_QUEUE_MULTIPLIER)
from .tree import device_topology, resolve_tree_topology

ACTION_ACCEPT, ACTION_NEAR, ACTION_REFINE = 0, 1, 2

_ACCEPT = {"mac": 0, "distance": 1, "different": 2, "never": 3}
_NEAR = {"leaf_pairs": 0, "never": 1}
_TAG = {"none": 0, "dist_sq": 1, "opening_ratio": 2, "target_mod": 3}
MAX_TAG_THRESHOLDS = 8


@dataclass(frozen=True)
class PairPolicyDescriptor:
    """A pair policy as data, evaluated by the walk's kernels in both
    directions (the PairPolicy contract, _interactions_impl.py:731-810).

    accept: "mac" (the built-in MAC of the call: mac_ok), "distance"
      (different nodes with dist_sq >= accept_dist_sq), "different" (any two
      different nodes) or "never".
    near: "leaf_pairs" (two different leaves that are not accepted, the
      default rule of _default_pair_actions, 694-711) or "never".
    tag (accepted pairs only, -1 otherwise): "none"; "dist_sq" or
      "opening_ratio" ((extent_target + extent_source)^2 / dist_sq), giving
      tag_values[k] with k the number of tag_thresholds (ascending, at most
      8) <= the metric; or "target_mod" (target node % tag_modulus).
    All comparisons are in float32 on the same dist_sq and extents a policy
    callable receives.
    """

    accept: str = "mac"
    accept_dist_sq: float = 0.0
    near: str = "leaf_pairs"
    tag: str = "none"
    tag_thresholds: Sequence[float] = ()
    tag_values: Sequence[int] = (-1,)
    tag_modulus: int = 1

    def __post_init__(self):
        if self.accept not in _ACCEPT:
            raise ValueError(f"accept must be one of {sorted(_ACCEPT)}")
        if self.near not in _NEAR:
            raise ValueError(f"near must be one of {sorted(_NEAR)}")
        if self.tag not in _TAG:
            raise ValueError(f"tag must be one of {sorted(_TAG)}")
        th = [float(x) for x in self.tag_thresholds]
        if len(th) > MAX_TAG_THRESHOLDS:
            raise ValueError(f"at most {MAX_TAG_THRESHOLDS} tag thresholds")
        if any(b < a for a, b in zip(th, th[1:])):
            raise ValueError("tag_thresholds must be ascending")
        if self.tag in ("dist_sq", "opening_ratio") and len(self.tag_values) != len(th) + 1:
            raise ValueError("tag_values needs one more entry than tag_thresholds")
        if self.tag == "target_mod" and int(self.tag_modulus) < 1:
            raise ValueError("tag_modulus must be >= 1")

    def to_c(self) -> "_lib.PolicyDesc":
        d = _lib.PolicyDesc()
        d.accept, d.near, d.tag_metric = _ACCEPT[self.accept], _NEAR[self.near], _TAG[self.tag]
        th = [float(x) for x in self.tag_thresholds]
        d.num_thresholds = len(th)
        d.accept_dist_sq = float(self.accept_dist_sq)
        d.tag_modulus = int(self.tag_modulus)
        for j, v in enumerate(th):
            d.thresholds[j] = v
        vals = list(self.tag_values) if self.tag in ("dist_sq", "opening_ratio") else [-1]
        for j, v in enumerate(vals):
            d.tag_values[j] = int(v)
        return d

def _lib_policy():
    return _lib.load()


def _policy_walk(ctx: _WalkContext, dt, pair_policy, policy_state, qcap, fcap, ncap):
    lib = _lib_policy()
    dev, st = ctx.dev, ctx.stream
    wf = [torch.empty(qcap, dtype=torch.int64, device=dev) for _ in range(2)]
    wf[0][0] = 0  # root pair (0, 0)
    far_k = torch.empty(fcap, dtype=torch.int32, device=dev)
    far_v = torch.empty(fcap, dtype=torch.int32, device=dev)
    far_t = torch.empty(fcap, dtype=torch.int32, device=dev)
    near_k = torch.empty(ncap, dtype=torch.int32, device=dev)
    near_v = torch.empty(ncap, dtype=torch.int32, device=dev)
    counters = torch.zeros(2, dtype=torch.int64, device=dev)
    flags = torch.zeros(4, dtype=torch.int32, device=dev)
    out_size = torch.zeros(1, dtype=torch.int32, device=dev)
    desc = pair_policy.to_c() if isinstance(pair_policy, PairPolicyDescriptor) else None
    size, gens, cur = 1, [1], 0
    while size > 0:
        if len(gens) > MAX_GENERATIONS:
            raise RuntimeError("pair-policy walk exceeded the generation limit")
        n = size
        if desc is not None:
            fa, ra, ftag, rtag = (torch.empty(n, dtype=torch.int32, device=dev)
                                  for _ in range(4))
            _lib.check(lib.ygg_policy_descriptor_actions(
                ctx.nodes.data_ptr(), wf[cur].data_ptr(), n, dt.num_internal, ctx.mac,
                ctx.theta_sq, ctx.theta_root, ctypes.byref(desc), fa.data_ptr(),
                ra.data_ptr(), ftag.data_ptr(), rtag.data_ptr(), st), "policy descriptor")
        else:
            fa, ra, ftag, rtag = _host_policy_actions(ctx, dt, pair_policy, policy_state,
                                                      wf[cur], n)
        out_size.zero_()
        _lib.check(lib.ygg_policy_apply(
            ctx.nodes.data_ptr(), wf[cur].data_ptr(), n, dt.num_internal, fa.data_ptr(),
            ra.data_ptr(), ftag.data_ptr(), rtag.data_ptr(), wf[1 - cur].data_ptr(), qcap,
            out_size.data_ptr(), far_k.data_ptr(), far_v.data_ptr(), far_t.data_ptr(), fcap,
            near_k.data_ptr(), near_v.data_ptr(), ncap, counters.data_ptr(), flags.data_ptr(),
            st), "policy apply")
        size = int(out_size)
        gens.append(size)
        cur = 1 - cur
        if int(flags.max()) != 0:
            return None
    c = [int(v) for v in counters.cpu()]
    return dict(far_total=c[0], near_total=c[1], gen_sizes=gens, peak=max(gens),
                far_keys=far_k, far_vals=far_v, far_tags=far_t, near_keys=near_k,
                near_vals=near_v, decisions=(c[0] // 2, c[1] // 2, sum(gens) - (c[0] + c[1]) // 2))


def _descriptor_walk(ctx: _WalkContext, dt, desc: "PairPolicyDescriptor"):
    """A descriptor policy evaluated inside the flat walk's own kernels: the
    emission walk writes far entries with their tags and near entries in one
    pass per generation, into streams sized as for the plain walk (one
    resumable walk that grows them, or a counting walk first)."""
    from . import interactions as I
    T, NI = dt.num_nodes, dt.num_internal
    c = desc.to_c()
    qbig = max(T * DEFAULT_PAIR_QUEUE_MULTIPLIER, I._auto_pair_queue_candidates(T, NI)[-1])
    em = None
    if I.STREAM_SIZING == "grow":
        f0, n0 = I._start_capacities(T, dt.num_leaves)
        em = ctx.run_flat(emit=True, queue_cap=qbig, far_cap=f0, near_cap=n0, policy=c,
                          grow=True)
        if em["queue_overflow"]:
            raise RuntimeError("Count-pass traversal wavefront overflowed its conservative "
                               "buffer.")
    else:
        tot = ctx.run_flat(emit=False, queue_cap=qbig, policy=c)
        if tot["queue_overflow"]:
            raise RuntimeError("Count-pass traversal wavefront overflowed its conservative "
                               "buffer.")
        em = ctx.run_flat(emit=True, queue_cap=tot["peak"], far_cap=tot["far_total"],
                          near_cap=tot["near_total"], policy=c)
    if em["queue_overflow"] or em["far_overflow"] or em["near_overflow"]:
        raise RuntimeError("pair-policy emission walk overflowed its streams")
    return dict(far_total=em["far_total"], near_total=em["near_total"],
                gen_sizes=em["gen_sizes"], peak=em["peak"], far_keys=em["far_keys"],
                far_vals=em["far_vals"], far_tags=em["far_tags"], near_keys=em["near_keys"],
                near_vals=em["near_vals"], decisions=em["decisions"])


def _host_policy_actions(ctx: _WalkContext, dt, pair_policy, policy_state, pairs, n):
    """One generation's geometry from the device, then the policy callable
    in both directions on the host."""
    lib = _lib_policy()
    dev, st = ctx.dev, ctx.stream
    idx = INDEX_DTYPE
    f32 = dict(dtype=torch.float32, device=dev)
    t = torch.empty(n, dtype=torch.int32, device=dev)
    s = torch.empty(n, dtype=torch.int32, device=dev)
    ct = torch.empty((n, 3), **f32)
    cs = torch.empty((n, 3), **f32)
    d2 = torch.empty(n, **f32)
    et = torch.empty(n, **f32)
    es = torch.empty(n, **f32)
    mok = torch.empty(n, dtype=torch.uint8, device=dev)
    tl = torch.empty(n, dtype=torch.uint8, device=dev)
    sl = torch.empty(n, dtype=torch.uint8, device=dev)
    _lib.check(lib.ygg_policy_geometry(
        ctx.nodes.data_ptr(), pairs.data_ptr(), n, dt.num_internal, ctx.mac,
        ctx.theta_sq, ctx.theta_root, t.data_ptr(), s.data_ptr(), ct.data_ptr(),
        cs.data_ptr(), d2.data_ptr(), et.data_ptr(), es.data_ptr(), mok.data_ptr(),
        tl.data_ptr(), sl.data_ptr(), st), "policy geometry")
    valid = torch.ones(n, dtype=torch.bool, device=dev)
    ti, si = t.to(idx), s.to(idx)
    same = ti == si
    common = dict(valid_pairs=valid, mac_ok=mok.bool(), different_nodes=~same,
                  same_node=same, dist_sq=d2)
    fa, ftag = pair_policy(policy_state, target_leaf=tl.bool(), source_leaf=sl.bool(),
                           target_nodes=ti, source_nodes=si, center_target=ct,
                           center_source=cs, extent_target=et, extent_source=es, **common)
    ra, rtag = pair_policy(policy_state, target_leaf=sl.bool(), source_leaf=tl.bool(),
                           target_nodes=si, source_nodes=ti, center_target=cs,
                           center_source=ct, extent_target=es, extent_source=et, **common)
    conv = lambda a: torch.as_tensor(a, device=dev).to(torch.int32).reshape(n).contiguous()  # noqa
    return conv(fa), conv(ra), conv(ftag), conv(rtag)


def build_with_pair_policy(tree, geometry, theta, mac_type, dehnen_radius_scale, pair_policy,
                           policy_state, *, max_interactions_per_node=None,
                           max_neighbors_per_leaf=2048, max_pair_queue=None,
                           traversal_config=None, return_result=False):
    """build_interactions_and_neighbors with a custom pair policy."""
    from .interactions import _empty_outputs, _sort_far_and_near, _to_neighbors

    dt = device_topology(tree)
    topo = resolve_tree_topology(tree)
    if float(dehnen_radius_scale) <= 0.0:
        raise ValueError("dehnen_radius_scale must be > 0")
    if dt.num_internal == 0:
        inter, nbrs = _empty_outputs(dt, topo), _to_neighbors(None, dt)
        return (inter, nbrs, _to_result(None, dt, inter, nbrs)) if return_result else (inter, nbrs)
    ctx = _WalkContext(dt, geometry, theta, mac_type, dehnen_radius_scale)
    lad = _ladder(dt, mac_type, max_interactions_per_node=max_interactions_per_node,
                  max_neighbors_per_leaf=max_neighbors_per_leaf, max_pair_queue=max_pair_queue,
                  traversal_config=traversal_config)
    T, L = dt.num_nodes, dt.num_leaves
    if isinstance(pair_policy, PairPolicyDescriptor):
        res = _descriptor_walk(ctx, dt, pair_policy)
    else:
        qcap = T * DEFAULT_PAIR_QUEUE_MULTIPLIER
        fcap = ncap = 16 * T + 1024
        res = None
        for _ in range(8):
            res = _policy_walk(ctx, dt, pair_policy, policy_state, qcap, fcap, ncap)
            if res is not None:
                break
            fcap, ncap = fcap * 4, ncap * 4
        if res is None:
            raise RuntimeError("pair-policy walk did not fit its buffers")
    nf, nn = res["far_total"], res["near_total"]
    # far entries sort with their stream positions, which then gather the
    # sources and tags
    iota = torch.arange(max(1, nf), dtype=torch.int32, device=ctx.dev)
    (fk, perm, far_off), (_, nv, near_off) = _sort_far_and_near(
        res["far_keys"], iota, nf, T, res["near_keys"], res["near_vals"], nn, L)
    pl = perm[:nf].long()
    fv = res["far_vals"][pl]
    tags = res["far_tags"][pl]
    mx = torch.stack([torch.diff(far_off).max(), torch.diff(near_off).max()]).cpu()
    stats = dict(peak=res["peak"], gen_sizes=res["gen_sizes"], decisions=res["decisions"],
                 max_far=int(mx[0]), max_near=int(mx[1]), far_total=nf, near_total=nn)
    if not lad.use_count:
        caps = _emulate_ladder(lad, res["peak"], stats["max_far"], stats["max_near"])
        if caps is None:
            raise RuntimeError("overflow order of this capacity configuration needs the "
                               "per-generation policy walk with capacities")
        stats["queue_capacity"] = caps[0]
        stats["attempt_caps"] = caps
    fr = _FlatResult(fk, fv, nv, far_off, near_off, stats)
    inter, nbrs = _flat_outputs(fr, dt, topo)
    if return_result:
        return inter, nbrs, _to_result(fr, dt, inter, nbrs, tags=tags)
    return inter, nbrs


__all__ = ["ACTION_ACCEPT", "ACTION_NEAR", "ACTION_REFINE", "MAX_TAG_THRESHOLDS",
           "PairPolicyDescriptor", "build_with_pair_policy"]
