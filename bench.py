"""Benchmark: build + traversal throughput of the yggdrax tree pipeline on one MI355X.

Workload (one "step"): 2^24 clustered float32 points (64 Plummer clusters,
scale radii 10^U(-3,-1), Dirichlet(0.5) weights, 10 % uniform background,
lognormal masses; PCG64 seed 3), resident in HBM; build_octree(leaf_size=64)
-> compute_tree_geometry -> build_interactions_and_neighbors(theta=0.5,
mac_type="dehnen") with no capacities (count pass + exact-length fill), all
outputs in the reference's layout (int64 indices). Metric: Mpts/s =
N / wall time per step.

Prints ONE JSON line. The roofline object is for the dominant kernel (the
walk's generation kernel, emission pass); its algorithmic bytes use the
per-pair model 2 x 28 B node records + 8 B queue read per evaluated pair,
8 B per pushed child pair and 4 B per emitted list entry, with the measured
HBM traffic of the same kernel beside it. `build_roofline` prices sort +
build + geometry at 287 compulsory bytes per point against HBM peak. The `learned` object compares
the walk with learned capacities (one emission walk) against the plain
count-then-fill walk on the same tree and checks the lists are identical.
The `layouts` object times the downstream far-field layouts (level-major
geometry, dense M2L buffers, displacement classes) on the 2^20-point
Plummer workload, the `moments` object the per-node mass moments of the
benchmark's own tree. The CPU baseline is the oracle (C++) on the same
2^24-point workload (build, geometry and the walk with every list entry), on
one thread and on every host thread available (OpenMP), with the CPU model.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpts/s build+traversal, 2^24 pts, 1 GPU; learned/batched vs plain speedup"


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# The reference's count-pass path rejects this workload (a background leaf
# next to a dense cluster has more than (2^31-1)/L near neighbours, which its
# dense (L x capacity) buffers cannot index), so the benchmark call lifts that
# guard and skips the dense neighbour-position matrix; the CSR lists are
# complete and exact.
WALK_KW = dict(mac_type="dehnen", neighbor_positions=False, strict_int32=False)


def run_pipeline(ygg, pos, mass, leaf, theta):
    tree, ps, _, _ = ygg.build_octree(pos, mass, leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=theta, **WALK_KW)
    return tree, geom, inter, nbrs


def stage_times(ygg, torch, pos, mass, leaf, theta):
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    tree, ps, _, _ = ygg.build_octree(pos, mass, leaf_size=leaf, return_reordered=True)
    ev[1].record()
    geom = ygg.compute_tree_geometry(tree, ps)
    ev[2].record()
    inter, nbrs = ygg.build_interactions_and_neighbors(tree, geom, theta=theta, **WALK_KW)
    ev[3].record()
    torch.cuda.synchronize()
    return dict(build_ms=ev[0].elapsed_time(ev[1]), geometry_ms=ev[1].elapsed_time(ev[2]),
                traversal_ms=ev[2].elapsed_time(ev[3]),
                far_entries=int(inter.sources.numel()), near_entries=int(nbrs.neighbors.numel()),
                num_nodes=int(tree.num_nodes), num_levels=int(tree.num_levels))


def measured_traffic(log2n, leaf, theta):
    """HBM bytes per emission-walk launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy), for the
    workload it was measured on; None otherwise."""
    if (log2n, leaf, theta) != (24, 64, 0.5):
        return None
    path = os.path.join(ROOT, "profiles", "pmc_hbm_traffic_clustered_2p24.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch_k_walk_flat_emit:"):
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


def walk_roofline(ygg, torch, pos, mass, leaf, theta):
    """Roofline of the walk's generation kernel (emission pass), timed with
    HIP events recorded on the launch stream around every generation."""
    from yggdrax_amd import interactions as I
    I.WALK_TIMING["enabled"] = True
    I.WALK_TIMING["log"].clear()
    try:
        run_pipeline(ygg, pos, mass, leaf, theta)
        torch.cuda.synchronize()
    finally:
        I.WALK_TIMING["enabled"] = False
    recs = [r for r in I.WALK_TIMING["log"] if r.get("emit")]
    rec = recs[-1]
    acc, near, ref = rec["decisions"]
    pairs = acc + near + ref
    pushes = sum(rec["gen_sizes"][1:])
    entries = rec["far_total"] + rec["near_total"]
    total_bytes = (pairs * PAIR_BYTES + pushes * PUSH_BYTES + entries * ENTRY_BYTES)
    gen_ms = [m for m, g in zip(rec["gen_ms"], rec["gen_sizes"]) if g > 0]
    total_ms = sum(gen_ms)
    launches = len(gen_ms)
    avg_ms = total_ms / max(1, launches)
    avg_bytes = total_bytes / max(1, launches)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    count_ms = sum(sum(r["gen_ms"]) for r in I.WALK_TIMING["log"] if not r.get("emit"))
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="k_walk_flat<true> (emission pass)", launches=launches,
                avg_launch_ms=round(avg_ms, 5), bytes_per_launch=round(avg_bytes),
                pairs=pairs, pairs_per_s=round(pairs / (total_ms * 1e-3)),
                bytes_per_pair=round(total_bytes / max(1, pairs), 2),
                emit_walk_ms=round(total_ms, 4), count_walk_ms=round(count_ms, 4))


# Algorithmic bytes of the walk: per evaluated pair two 28 B node records
# (centre 12, extent 4, children 8, leaf row 4) and the 8 B queue read; 8 B
# per pushed child pair; 4 B per emitted directed list entry (int32).
PAIR_BYTES, PUSH_BYTES, ENTRY_BYTES = 2 * 28 + 8, 8, 4


# Compulsory bytes per point of sort + build + geometry at radix digit width
# d: the sort's P x (8 + 4) x 2 + 8 with P = ceil(63 / d) passes, plus 87 for
# the rest (bounds 12, encode 20, reorder 40, Karras + levels ~0.6 and
# octree metadata ~1 per point at leaf 64, geometry ~13; 287 in all at
# d = 8). The Morton sort uses 9-bit digits (7 passes): 263 B/pt.
MORTON_DIGIT_BITS = 9


def build_bytes_per_point(digit_bits: int = MORTON_DIGIT_BITS) -> float:
    passes = -(-63 // digit_bits)
    return float(passes * (8 + 4) * 2 + 8 + 87)


BUILD_BYTES_PER_POINT = build_bytes_per_point()


def build_roofline(n, stages):
    ms = stages["build_ms"] + stages["geometry_ms"]
    achieved = BUILD_BYTES_PER_POINT * n / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), bytes_per_point=BUILD_BYTES_PER_POINT,
                morton_digit_bits=MORTON_DIGIT_BITS,
                stages="build_octree + compute_tree_geometry", ms=round(ms, 4))


def _same_lists(torch, a, b, num_rows):
    """Both CSR lists hold the same per-row sets (row order of entries is
    unspecified): compare counts/offsets exactly and the (row, value) pairs
    after a sort."""
    if not (torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])):
        return False
    # entry e belongs to the row whose [offset, offset + count) holds it
    order = torch.argsort(a[0][:num_rows], stable=True)
    rows = torch.repeat_interleave(order, a[1][order].long())
    ka = torch.sort(rows * (1 << 32) + a[2].long())[0]
    kb = torch.sort(rows * (1 << 32) + b[2].long())[0]
    return bool(torch.equal(ka, kb))


def learned_leg(ygg, torch, pos, mass, leaf, theta, steps):
    """Walk time of the three stream sizings on the same tree: the plain
    call (one resumable emission walk from the default capacities), the
    reference's count pass (counting walk, then the emission walk at exact
    sizes) and the learned path (capacity MLP inference, then one emission
    walk at the predicted sizes), with a check that all produce the same
    lists, and the stream memory each one allocated."""
    from yggdrax_amd import interactions as I
    from yggdrax_amd import learned
    model = learned.CapacityModel.load()
    tree, ps, _, _ = ygg.build_octree(pos, mass, leaf_size=leaf, return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)

    def plain():
        return ygg.build_interactions_and_neighbors(tree, geom, theta=theta, **WALK_KW)

    def counted():
        old, I.STREAM_SIZING = I.STREAM_SIZING, "count"
        try:
            return plain()
        finally:
            I.STREAM_SIZING = old

    def with_model():
        hint = model.capacity_hint(tree, geom)
        return ygg.build_interactions_and_neighbors(tree, geom, theta=theta, capacity_hint=hint,
                                                    **WALK_KW), hint

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3, out

    t_plain, (pi, pn) = timed(plain)
    plain_walk = dict(I.LAST_WALK)
    t_count, (ci, cn) = timed(counted)
    runs0 = I.FLAT_STATS["hint_overflows"] + I.FLAT_STATS["hint_growths"]
    t_learn, ((li, ln), hint) = timed(with_model)
    fallbacks = I.FLAT_STATS["hint_overflows"] + I.FLAT_STATS["hint_growths"] - runs0
    T = int(tree.num_nodes)
    L = int(pn.counts.numel())

    def same(a, b):
        return (_same_lists(torch, (a[0].offsets, a[0].counts, a[0].sources),
                            (b[0].offsets, b[0].counts, b[0].sources), T)
                and torch.equal(a[0].targets, b[0].targets)
                and _same_lists(torch, (a[1].offsets, a[1].counts, a[1].neighbors),
                                (b[1].offsets, b[1].counts, b[1].neighbors), L))
    nf, nn = int(pi.sources.numel()), int(pn.neighbors.numel())
    # slot memory the fixed-capacity method would need on this tree: one row
    # of the largest count per node (far) and per leaf (near), int32 entries
    max_far, max_near = int(pi.counts.max()), int(pn.counts.max())
    # far and near stream bytes: a 32-bit key and value per entry, or per
    # pair record (two entries) when the walk writes records
    bpe = plain_walk["stream_bytes_per_entry"]
    return dict(plain_walk_ms=round(t_plain, 3), count_pass_walk_ms=round(t_count, 3),
                learned_walk_ms=round(t_learn, 3),
                max_far_per_node=max_far, max_near_per_leaf=max_near,
                fixed_capacity_slot_bytes=4 * (T * max_far + L * max_near),
                walk_speedup=round(t_plain / t_learn, 3),
                walk_speedup_vs_count_pass=round(t_count / t_learn, 3),
                identical_lists=same((pi, pn), (li, ln)) and same((pi, pn), (ci, cn)),
                fallback_walks=fallbacks, hint=hint, true_far=nf, true_near=nn,
                far_overalloc=round(hint["far"] / max(1, nf), 3),
                near_overalloc=round(hint["near"] / max(1, nn), 3),
                plain_stream_bytes=plain_walk["stream_bytes"],
                plain_stream_growths=plain_walk["stream_growths"],
                learned_stream_bytes=bpe * (hint["far"] + hint["near"]),
                count_pass_stream_bytes=bpe * (nf + nn))


_MIX_C = 0x9E3779B97F4A7C15 - (1 << 64)
_MIX_1 = 0xBF58476D1CE4E5B9 - (1 << 64)
_MIX_2 = 0x94D049BB133111EB - (1 << 64)


def _mix(z):
    """splitmix64 finaliser on int64 bits (logical shifts by masking)."""
    z = z + _MIX_C
    z = (z ^ ((z >> 30) & ((1 << 34) - 1))) * _MIX_1
    z = (z ^ ((z >> 27) & ((1 << 37) - 1))) * _MIX_2
    return z ^ ((z >> 31) & ((1 << 33) - 1))


def _entry_hash(torch, system, row, value, salt):
    return _mix(_mix((system * (1 << 32) + row) * _MIX_C + salt) ^ value.long())


def _batch_digests(torch, lists):
    """Per-system digests [S, 4] of a BatchedLists: far (target, source)
    entries, near (leaf row, neighbour) entries, far counts, near counts;
    order-free within rows. Two results are equal system k exactly when row
    k of their digests is (up to 64-bit hash collisions)."""
    lay = lists.layout
    dev = lists.sources.device
    S = lay.num_systems
    out = torch.zeros((S, 4), dtype=torch.int64, device=dev)
    node_sys = torch.repeat_interleave(torch.arange(S, device=dev),
                                       torch.from_numpy(np.diff(lay.node_off)).to(dev))
    leaf_sys = torch.repeat_interleave(torch.arange(S, device=dev),
                                       torch.from_numpy(np.diff(lay.leaf_off)).to(dev))
    node_loc = torch.arange(node_sys.numel(), device=dev) - torch.from_numpy(
        lay.node_off[:-1]).to(dev)[node_sys]
    leaf_loc = torch.arange(leaf_sys.numel(), device=dev) - torch.from_numpy(
        lay.leaf_off[:-1]).to(dev)[leaf_sys]
    # valid entries lead the (possibly capacity-sized) entry arrays, system
    # after system
    fcnt, ncnt = lists.far_counts.long(), lists.near_counts.long()
    nf, nn = int(fcnt.sum()), int(ncnt.sum())
    fb = torch.from_numpy(np.asarray(lists.far_base, np.int64)).to(dev)
    e = torch.arange(nf, device=dev)
    esys = torch.searchsorted(fb, e, right=True) - 1
    out[:, 0].index_add_(0, esys, _entry_hash(torch, esys, lists.targets[:nf].long(),
                                              lists.sources[:nf], 1))
    nrow = torch.repeat_interleave(leaf_loc, ncnt)
    nsys = torch.repeat_interleave(leaf_sys, ncnt)
    out[:, 1].index_add_(0, nsys, _entry_hash(torch, nsys, nrow, lists.neighbors[:nn], 2))
    out[:, 2].index_add_(0, node_sys, _entry_hash(torch, node_sys, node_loc,
                                                  lists.far_counts, 3))
    out[:, 3].index_add_(0, leaf_sys, _entry_hash(torch, leaf_sys, leaf_loc,
                                                  lists.near_counts, 4))
    return out


def _single_digests(torch, outs):
    """The same digests for single-system results (inter, nbrs), one row per
    system."""
    rows = []
    for k, (inter, nbrs) in enumerate(outs):
        dev = inter.counts.device
        kk = torch.full((1,), k, dtype=torch.int64, device=dev)
        T, L = inter.counts.numel(), nbrs.counts.numel()
        nrow = torch.repeat_interleave(torch.arange(L, device=dev), nbrs.counts.long())
        rows.append(torch.stack([
            _entry_hash(torch, kk, inter.targets.long(), inter.sources, 1).sum(),
            _entry_hash(torch, kk, nrow, nbrs.neighbors, 2).sum(),
            _entry_hash(torch, kk, torch.arange(T, device=dev), inter.counts, 3).sum(),
            _entry_hash(torch, kk, torch.arange(L, device=dev), nbrs.counts, 4).sum()]))
    return torch.stack(rows)


def _mismatches(torch, a, b):
    return int((a != b).any(dim=1).sum())


def batched_leg(torch, steps, warmup):
    """Batched workload: 1024 Plummer systems, N_i log-uniform in [256, 65536]
    (PCG64 seed 4), build_octree leaf 32, dehnen theta 0.6. Times, on the same
    systems: the plain batched pipeline (one resumable emission walk over
    the forest from the default stream capacities), the same with the
    reference's count pass (counting walk, then the emission walk at exact
    sizes), the learned batched pipeline (capacity model, then one emission
    walk) -- the equal-batching comparisons of the learned capacities --,
    its HIP-graph capture replayed on new positions every step
    with stream capacities from the capacity model, one single-system call
    per system, and the reference's fixed-capacity method (per-row slots of
    2048 far and 2048 near entries, its automatic pair-queue ladder, retry
    on overflow) one call per system. Every system's lists are compared
    across all of them (per-system digests)."""
    from yggdrax_amd import batch, inputs
    import yggdrax_amd as ygg
    pos_np, mass_np, offsets = inputs.plummer_batch(1024, 4)
    P = torch.from_numpy(pos_np).cuda()
    M = torch.from_numpy(mass_np).cuda()
    S = len(offsets) - 1
    kw = dict(mac_type="dehnen")

    def run_batched(p=P):
        bt = batch.build_octree_batched(p, M, offsets, leaf_size=32)
        g = batch.compute_tree_geometry_batched(bt)
        return bt, batch.build_interactions_and_neighbors_batched(bt, g, 0.6, **kw)

    from yggdrax_amd import learned
    from yggdrax_amd import interactions as I
    model = learned.CapacityModel.load()

    def run_learned():
        bt = batch.build_octree_batched(P, M, offsets, leaf_size=32)
        g = batch.compute_tree_geometry_batched(bt)
        hint = model.capacity_hint_batched(bt, g)
        return bt, batch.build_interactions_and_neighbors_batched(bt, g, 0.6, capacity_hint=hint,
                                                                  **kw), hint

    def run_one(k, **walk):
        p0, p1 = int(offsets[k]), int(offsets[k + 1])
        t, ps, _, _ = ygg.build_octree(P[p0:p1], M[p0:p1], leaf_size=32, return_reordered=True)
        g = ygg.compute_tree_geometry(t, ps)
        return ygg.build_interactions_and_neighbors(t, g, theta=0.6, neighbor_positions=False,
                                                    **kw, **walk)

    # the reference's fixed-capacity method: per-row slots of the given
    # capacities and its retry ladder on overflow (a retry logger selects
    # the attempt-by-attempt walk), one call per system
    fixed = dict(max_interactions_per_node=2048, max_neighbors_per_leaf=2048,
                 retry_logger=lambda event: None)

    for _ in range(max(1, warmup)):
        run_batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bt, lists = run_batched()
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / steps * 1e3
    plain_walk = dict(I.LAST_WALK)
    sizing, I.STREAM_SIZING = I.STREAM_SIZING, "count"
    try:
        run_batched()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            _, clists = run_batched()
        torch.cuda.synchronize()
        t_c = (time.perf_counter() - t0) / steps * 1e3
    finally:
        I.STREAM_SIZING = sizing
    run_learned()
    torch.cuda.synchronize()
    c0 = I.FLAT_STATS["hint_overflows"] + I.FLAT_STATS["hint_growths"]
    t0 = time.perf_counter()
    for _ in range(steps):
        _, llists, hint = run_learned()
    torch.cuda.synchronize()
    t_l = (time.perf_counter() - t0) / steps * 1e3
    learned_fallbacks = I.FLAT_STATS["hint_overflows"] + I.FLAT_STATS["hint_growths"] - c0
    # the learned batched pipeline captured once as a HIP graph (stream
    # capacities from the capacity model), replayed on new positions: a
    # small random displacement of every point per step, inputs in HBM
    from yggdrax_amd import autotune
    graph = batch.BatchedGraph(P, M, offsets, leaf_size=32, theta=0.6, model=model,
                               tune=autotune.enabled(), **kw)
    rng = np.random.default_rng(404)
    moved = [P + torch.from_numpy(rng.normal(0.0, 1e-4, pos_np.shape).astype(np.float32)).cuda()
             for _ in range(steps)]
    graph.replay(moved[0])
    torch.cuda.synchronize()
    fell_back = 0
    t0 = time.perf_counter()
    for k in range(steps):
        _, _, glists = graph.replay(moved[k])
        fell_back += int(graph.last_replay_fell_back)
    torch.cuda.synchronize()
    t_g = (time.perf_counter() - t0) / steps * 1e3
    g_dig = _batch_digests(torch, glists)
    _, moved_lists = run_batched(moved[-1])
    moved_dig = _batch_digests(torch, moved_lists)
    # one call per system meets many tree sizes: no tuning inside its timing
    tuning = autotune.enabled()
    autotune.enable(False)
    run_one(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [run_one(k) for k in range(S)]
    torch.cuda.synchronize()
    t_s = (time.perf_counter() - t0) * 1e3
    run_one(0, **fixed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fouts = [run_one(k, **fixed) for k in range(S)]
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) * 1e3
    autotune.enable(tuning)
    b_dig = _batch_digests(torch, lists)
    l_dig = _batch_digests(torch, llists)
    s_dig = _single_digests(torch, outs)
    f_dig = _single_digests(torch, fouts)
    c_dig = _batch_digests(torch, clists)
    mism = dict(learned_vs_plain_batched=_mismatches(torch, l_dig, b_dig),
                count_pass_vs_plain_batched=_mismatches(torch, c_dig, b_dig),
                graph_replay_vs_eager_on_new_positions=_mismatches(torch, g_dig, moved_dig),
                per_system_vs_batched=_mismatches(torch, s_dig, b_dig),
                fixed_capacity_vs_per_system=_mismatches(torch, f_dig, s_dig))
    n = int(offsets[-1])
    return dict(workload="1024 Plummer systems, N_i log-uniform [256, 65536] (seed 4), "
                         "build_octree leaf 32, dehnen theta 0.6",
                points=n, batched_ms=round(t_b, 3), batched_mpts_s=round(n / t_b / 1e3, 3),
                learned_batched_ms=round(t_l, 3),
                learned_batched_mpts_s=round(n / t_l / 1e3, 3),
                count_pass_batched_ms=round(t_c, 3),
                learned_fallback_walks=learned_fallbacks,
                learned_speedup_vs_batched_plain=round(t_b / t_l, 3),
                learned_speedup_vs_batched_count_pass=round(t_c / t_l, 3),
                plain_stream_bytes=plain_walk["stream_bytes"],
                plain_stream_growths=plain_walk["stream_growths"],
                learned_stream_bytes=plain_walk["stream_bytes_per_entry"] * (hint["far"]
                                                                             + hint["near"]),
                graph_replay_ms=round(t_g, 3), graph_replay_mpts_s=round(n / t_g / 1e3, 3),
                graph_replay_positions="new positions every step (N(0, 1e-4) displacement)",
                graph_capacities=dict(graph.caps, generations=graph.generations,
                                      source="capacity model"),
                graph_replays_fell_back=fell_back,
                learned_far_overalloc=round(hint["far"] / max(1, lists.stats["far_total"]), 3),
                learned_near_overalloc=round(hint["near"] / max(1, lists.stats["near_total"]), 3),
                per_system_loop_ms=round(t_s, 3), speedup_vs_per_system=round(t_s / t_b, 2),
                learned_speedup_vs_per_system=round(t_s / t_l, 2),
                graph_speedup_vs_per_system=round(t_s / t_g, 2),
                fixed_capacity_per_system_ms=round(t_f, 3),
                learned_speedup_vs_fixed_capacity=round(t_f / t_l, 2),
                far_entries=int(lists.stats["far_total"]),
                near_entries=int(lists.stats["near_total"]), fallback_systems=len(bt.fallback),
                systems_compared=S, list_mismatches=mism,
                all_systems_identical=not any(mism.values()))


def moments_leg(ygg, torch, pos, mass, leaf, reps=5):
    """Per-node mass moments (compute_tree_mass_moments) of the benchmark's
    own tree, timed over `reps` calls. Algorithmic bytes: 16 B per particle
    read (position and mass), 8 B per node range read, 16 B per node written
    (mass and centre)."""
    from yggdrax_amd import tree_moments as tm
    tree, ps, ms, _ = ygg.build_octree(pos, mass, leaf_size=leaf, return_reordered=True)
    tm.compute_tree_mass_moments(tree, ps, ms)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        tm.compute_tree_mass_moments(tree, ps, ms)
    b.record()
    b.synchronize()
    ms_call = a.elapsed_time(b) / reps
    n, t = int(ps.shape[0]), int(tree.num_nodes)
    nbytes = 16 * n + 24 * t
    return dict(ms=round(ms_call, 4), points=n, nodes=t, algorithmic_bytes=nbytes,
                gbs=round(nbytes / ms_call / 1e6, 1),
                frac_of_hbm_peak=round(nbytes / ms_call / 1e6 / HBM_PEAK_GBS, 3))


def layouts_leg(ygg, torch, reps=3):
    """Downstream far-field layouts on the 2^20-point Plummer workload
    (build_tree leaf 32, dehnen theta 0.5): level-major geometry, dense M2L
    buffers and displacement classes, each timed over `reps` calls on lists
    already in HBM. The dense kernel is write-bound: its algorithmic bytes
    are 21 B per (slot, entry) (8 B source, 12 B displacement, 1 B mask),
    8 B per slot count and 8 B per far source read. Returns the figures and
    the inputs the CPU baseline of the grouping reuses."""
    from yggdrax_amd import dense_interactions as di, grouped_interactions as gi, inputs
    pos_np, mass_np = inputs.plummer_sphere(1 << 20, 2)
    tree, ps, _, _ = ygg.build_tree(torch.from_numpy(pos_np).cuda(),
                                    torch.from_numpy(mass_np).cuda(), leaf_size=32,
                                    return_reordered=True)
    geom = ygg.compute_tree_geometry(tree, ps)
    inter, _ = ygg.build_interactions_and_neighbors(tree, geom, theta=0.5, **WALK_KW)
    torch.cuda.synchronize()

    def timed(fn):
        out = fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            out = fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps, out

    t_lm, lm = timed(lambda: ygg.geometry_to_level_major(tree, geom))
    t_d, d = timed(lambda: di.densify_interactions(tree, geom, inter))
    slots = int(d.m2l_counts.numel())
    k = int(d.m2l_sources.shape[2])
    far = int(inter.sources.numel())
    dense_bytes = slots * k * 21 + slots * 8 + far * 8
    del d
    torch.cuda.empty_cache()
    t_g, g = timed(lambda: gi.build_grouped_interactions(tree, geom, inter))
    classes = int(g.class_keys.shape[0])
    del g
    figures = dict(
        workload="plummer 2^20 points, build_tree leaf 32, dehnen theta 0.5", far_pairs=far,
        level_major_ms=round(t_lm, 4), level_major_shape=list(lm.node_indices.shape),
        dense_ms=round(t_d, 3), dense_shape=[*lm.node_indices.shape, k],
        dense_algorithmic_bytes=dense_bytes, dense_gbs=round(dense_bytes / t_d / 1e6, 1),
        dense_frac_of_hbm_peak=round(dense_bytes / t_d / 1e6 / HBM_PEAK_GBS, 3),
        grouped_ms=round(t_g, 3), grouped_classes=classes,
        grouped_gpairs_s=round(far / t_g / 1e6, 3))
    return figures, (pos_np, mass_np, tree, geom, inter)


def layouts_cpu_baseline(torch, state, cpu_pairs=1 << 21):
    """The oracle's NumPy grouping (the reference's own host algorithm) timed
    on the first `cpu_pairs` far pairs of the layouts workload, with a check
    that the GPU grouping of the same pairs is identical."""
    import oracle
    from yggdrax_amd import grouped_interactions as gi
    pos_np, mass_np, tree, geom, inter = state
    m = min(cpu_pairs, int(inter.sources.numel()))
    src = inter.sources[:m].cpu().numpy()
    tgt = inter.targets[:m].cpu().numpy()
    ref = oracle.build_tree(pos_np, mass_np, leaf_size=32)
    rgeom = oracle.compute_tree_geometry(ref, ref["positions_sorted"])
    t0 = time.perf_counter()
    exp = oracle.grouped_interactions_from_pairs(ref, rgeom, src, tgt)
    t_cpu = time.perf_counter() - t0
    gs = gi.build_grouped_interactions_from_pairs(tree, geom, inter.sources[:m],
                                                  inter.targets[:m])
    same = True
    for f in gs._fields:
        got = getattr(gs, f).cpu().numpy()
        same &= bool(np.array_equal(got, np.asarray(exp[f]).astype(got.dtype)))
    return dict(value=round(m / t_cpu / 1e6, 4), unit="Mpairs/s", cores=1, kind="port",
                sample=f"oracle NumPy grouping (the reference's host algorithm) of the first "
                       f"{m} far pairs, {t_cpu:.2f} s", gpu_identical=same)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(int(env), n) if env and env.isdigit() else n)


def cpu_baseline(sample_n, leaf, theta, pos=None, mass=None):
    """The oracle (C++) on the same clustered workload: build_octree,
    compute_tree_geometry and the walk with every list entry, on one thread
    and on every host thread this process may use (OpenMP; the outputs do
    not depend on the thread count)."""
    import oracle
    from oracle import oracle as O
    from yggdrax_amd import inputs
    if pos is None or pos.shape[0] != sample_n:
        pos, mass = inputs.clustered(sample_n, 3)

    def run(threads):
        O.lib.ygo_set_threads(threads)
        t0 = time.perf_counter()
        topo = oracle.build_octree(pos, mass, leaf_size=leaf)
        geom = oracle.compute_tree_geometry(topo, topo["positions_sorted"])
        cap = max(topo["parent"].shape[0] * 128, 1 << 16)
        w = O.walk(topo, geom, theta, "dehnen", queue_cap=cap, far_cap=1 << 62,
                   near_cap=1 << 62)
        dt = time.perf_counter() - t0
        return dt, int(w["far_counts"].sum()) + int(w["near_counts"].sum())

    threads = _host_threads()
    d1, e1 = run(1)
    dn, en = run(threads)
    O.lib.ygo_set_threads(threads)
    model = _cpu_model()
    what = (f"oracle (C++) build_octree + geometry + walk with every list entry on {sample_n} "
            f"points of the same clustered workload (seed 3); {model}")
    return dict(value=round(sample_n / dn / 1e6, 4), unit="Mpts/s", cores=threads, kind="port",
                sample=f"{what}, {threads} threads, {dn:.1f} s", cpu_model=model,
                single_thread=dict(value=round(sample_n / d1 / 1e6, 4), unit="Mpts/s", cores=1,
                                   seconds=round(d1, 2)),
                entries=en, same_entries_1_and_n_threads=bool(e1 == en))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--cpu-sample-log2n", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-layouts", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (scales the HBM "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        _log("copy probe: 3 copies of 1 GiB")
        return
    import yggdrax_amd as ygg
    from yggdrax_amd import autotune, inputs

    # plans (row-sort digit widths, walk grid and queue tile) are tuned in
    # the untimed warm-up calls and applied in every timed one
    autotune.enable(not args.no_autotune)
    n = 1 << args.log2n
    _log(f"generating clustered input n={n}")
    pos_np, mass_np = inputs.clustered(n, 3)
    pos = torch.from_numpy(pos_np).cuda()
    mass = torch.from_numpy(mass_np).cuda()
    torch.cuda.synchronize()
    for w in range(args.warmup):
        run_pipeline(ygg, pos, mass, args.leaf, args.theta)
        torch.cuda.synchronize()
        _log(f"warmup {w} done")
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for s in range(args.steps):
        run_pipeline(ygg, pos, mass, args.leaf, args.theta)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(end) / args.steps
    _log(f"timed {args.steps} steps: {ms:.3f} ms/step (wall {wall / args.steps * 1e3:.3f} ms)")
    stages = stage_times(ygg, torch, pos, mass, args.leaf, args.theta)
    _log(f"stages {stages}")
    roof = walk_roofline(ygg, torch, pos, mass, args.leaf, args.theta)
    roof["traffic"] = measured_traffic(args.log2n, args.leaf, args.theta)
    if roof["traffic"]:
        hbm = roof["traffic"] / (roof["avg_launch_ms"] * 1e-3) / 1e9
        roof["hbm_measured_gbs"] = round(hbm, 1)
        roof["hbm_measured_frac"] = round(hbm / HBM_PEAK_GBS, 4)
    _log(f"roofline {roof}")
    build_roof = build_roofline(n, stages)
    _log(f"build roofline {build_roof}")
    lrn = None
    if not args.no_learned:
        lrn = learned_leg(ygg, torch, pos, mass, args.leaf, args.theta, args.steps)
        _log(f"learned {lrn}")
    bat = None
    if not args.no_batched:
        bat = batched_leg(torch, args.steps, args.warmup)
        _log(f"batched {bat}")
    mom = moments_leg(ygg, torch, pos, mass, args.leaf)
    _log(f"moments {mom}")
    lay = lay_state = None
    if not args.no_layouts:
        lay, lay_state = layouts_leg(ygg, torch)
        _log(f"layouts {lay}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(1 << args.cpu_sample_log2n, args.leaf, args.theta, pos_np, mass_np)
        _log(f"cpu baseline {cpu}")
        if lay is not None:
            lay["grouped_cpu_baseline"] = layouts_cpu_baseline(torch, lay_state)
            _log(f"layouts cpu baseline {lay['grouped_cpu_baseline']}")
    value = n / (ms * 1e-3) / 1e6
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"clustered 2^{args.log2n} points (64 Plummer clusters + 10% "
                               f"uniform background, seed 3), build_octree leaf {args.leaf}, "
                               f"dehnen theta {args.theta}, one resumable emission walk from the "
                               f"default stream capacities, int64 outputs",
                   "parallelism": "single GPU"},
        "stages": stages, "learned": lrn, "batched": bat, "layouts": lay, "moments": mom,
        "roofline": roof,
        "build_roofline": build_roof,
        "cpu_baseline": cpu,
        "autotune": {"enabled": not args.no_autotune, "plans": autotune.plans()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
