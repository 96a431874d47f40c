"""Digest schema of the full-size configurations.

One digest per output field of a system: every topology array (radix and
octree), the four geometry arrays (float32 bit patterns), the far lists
(counts, level offsets, and the (target node, source) entry set) and the near
lists (counts, leaf rows, and the (leaf row, neighbour) entry set). Entry
sets are order-free, as the reference's within-row order is unspecified.
The same names come from the oracle's outputs (numpy) and from the HIP
path's outputs (torch), so a fixture written from the oracle checks the GPU
field by field.
"""

from __future__ import annotations

import numpy as np

from tests import digest as D
from tests.parity import OCT_FIELDS, RADIX_FIELDS

GEOM_FIELDS = ("center", "half_extent", "radius", "max_extent")


def _salt(name: str) -> int:
    return int.from_bytes(name.encode()[:8].ljust(8, b"\0"), "little") ^ len(name)


# ---- oracle (numpy) -------------------------------------------------------

def oracle_system(pos, mass, *, octree: bool, leaf: int, theta: float, mac: str = "dehnen"):
    """Digests of the oracle's complete result for one system."""
    import oracle
    from oracle import oracle as O
    topo = (oracle.build_octree if octree else oracle.build_tree)(pos, mass, leaf_size=leaf)
    geom = oracle.compute_tree_geometry(topo, topo["positions_sorted"])
    T = topo["parent"].shape[0]
    NI = topo["num_internal_nodes"]
    L = T - NI
    out = {}
    for f in RADIX_FIELDS + (OCT_FIELDS if octree else ()):
        out["topo_" + f] = D.array_digest_np(topo[f], _salt(f))
    out["num_levels"] = int(topo["num_levels"])
    if octree:
        out["oct_num_levels"] = int(topo["oct_num_levels"])
    for f in GEOM_FIELDS:
        out["geom_" + f] = D.array_digest_np(np.asarray(geom[f], np.float32), _salt(f))
    if NI > 0:
        raw = O.walk(topo, geom, theta, mac, queue_cap=max(T * 128, 1 << 16),
                     far_cap=1 << 62, near_cap=1 << 62)
        assert not (raw["queue_overflow"] or raw["far_overflow"] or raw["near_overflow"])
        fc, nc = raw["far_counts"], raw["near_counts"]
        fe, ne = raw["far_entries"], raw["near_entries"]
        decisions = (raw["accept_decisions"], raw["near_decisions"], raw["refine_decisions"])
        leaf_indices = raw["leaf_indices"]
    else:
        fc, nc = np.zeros(T, np.int64), np.zeros(L, np.int64)
        fe = ne = np.zeros(0, np.int64)
        decisions = (0, 0, 0)
        leaf_indices = O.leaf_ordering(topo)[0]
    nbl = np.asarray(topo["nodes_by_level"], np.int64)
    obl = np.concatenate([[0], np.cumsum(fc[nbl])]).astype(np.int64)
    lvl = np.asarray(topo["level_offsets"], np.int64)[: int(topo["num_levels"]) + 1]
    out["far_counts"] = D.array_digest_np(fc, _salt("far_counts"))
    out["far_level_offsets"] = D.array_digest_np(obl[lvl] if fe.size else
                                                 np.zeros(lvl.size, np.int64),
                                                 _salt("far_level_offsets"))
    out["far_entries"] = D.digest_np(np.repeat(np.arange(T, dtype=np.int64), fc), fe,
                                     _salt("far_entries"))
    out["far_total"] = int(fc.sum())
    out["near_counts"] = D.array_digest_np(nc, _salt("near_counts"))
    out["near_entries"] = D.digest_np(np.repeat(np.arange(L, dtype=np.int64), nc), ne,
                                      _salt("near_entries"))
    out["near_total"] = int(nc.sum())
    out["leaf_indices"] = D.array_digest_np(leaf_indices, _salt("leaf_indices"))
    out["decisions"] = D.array_digest_np(np.asarray(decisions, np.int64), _salt("decisions"))
    return out


# ---- HIP path (torch) -----------------------------------------------------

def gpu_system(tree, geom, inter, nbrs, *, octree: bool, decisions=None):
    """The same digests from the HIP path's outputs (local node ids)."""
    import torch
    topo = tree.topology
    out = {}
    for f in RADIX_FIELDS + (OCT_FIELDS if octree else ()):
        out["topo_" + f] = D.array_digest_t(getattr(topo, f), _salt(f))
    out["num_levels"] = int(topo.num_levels)
    if octree:
        out["oct_num_levels"] = int(topo.oct_num_levels)
    for f in GEOM_FIELDS:
        out["geom_" + f] = D.array_digest_t(getattr(geom, f), _salt(f))
    dev = inter.counts.device
    T = int(inter.counts.numel())
    L = int(nbrs.counts.numel())
    out["far_counts"] = D.array_digest_t(inter.counts, _salt("far_counts"))
    out["far_level_offsets"] = D.array_digest_t(inter.level_offsets, _salt("far_level_offsets"))
    out["far_entries"] = D.digest_t(inter.targets, inter.sources, _salt("far_entries"))
    out["far_total"] = int(inter.sources.numel())
    out["near_counts"] = D.array_digest_t(nbrs.counts, _salt("near_counts"))
    rows = torch.repeat_interleave(torch.arange(L, device=dev), nbrs.counts.long())
    out["near_entries"] = D.digest_t(rows, nbrs.neighbors, _salt("near_entries"))
    out["near_total"] = int(nbrs.neighbors.numel())
    out["leaf_indices"] = D.array_digest_t(nbrs.leaf_indices, _salt("leaf_indices"))
    if decisions is not None:
        out["decisions"] = D.array_digest_np(np.asarray(decisions, np.int64), _salt("decisions"))
    assert T == int(tree.num_nodes)
    return out


def save(path, rows: list):
    """rows: one digest dict per system -> npz of uint64 arrays [S]."""
    keys = sorted(rows[0])
    np.savez_compressed(path, **{k: np.array([r[k] for r in rows], dtype=np.uint64)
                                 for k in keys})


def load(path) -> dict:
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def compare(got: dict, fixture: dict, index: int) -> list:
    """Names of the fields whose digest differs from the fixture's row."""
    return [k for k in sorted(got) if k in fixture and int(got[k]) != int(fixture[k][index])]
