"""The bucketing row lists (ygg_far_lists_sort / ygg_near_lists_sort with
digit_bits 0): entries grouped by row with no stable sort. Values are node
indices (below 2 * rows for the near lists, below rows for the far lists),
packed with the low digit where they fit. Against numpy:
row counts and offsets exactly, each row's values as a multiset (the order
inside a row is unspecified, as the walk's emission order is), and nothing
written past the valid entries. Cases cover one bucket (row keys of at most
11 bits), every split width up to 22 bits, keys at or above the row count
(unused capacity, listed nowhere), a device-side entry count with garbage
behind it, no entries, all entries in one row (one LDS counter takes every
atomic) and one bucket holding most entries (many tiles per bucket). The
far targets and levels equal the stable sort's."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from yggdrax_amd import _lib
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return _lib.load()


def _near(lib, keys, vals, n, count, L, bits, out_i64=1, leaf_nodes=None, nend=None):
    """Near lists of n entries, or of n pair records of node ids with
    `nend`-bit keys whose rows follow `leaf_nodes` (2n entries)."""
    from yggdrax_amd import _lib
    dev = keys.device
    nend = nend or max(1, (L - 1).bit_length())
    dt = torch.int64 if out_i64 else torch.int32
    ne = 2 * n if leaf_nodes is not None else n
    out = dict(neighbors=torch.full((max(1, ne),), -7, dtype=dt, device=dev),
               counts=torch.empty(L, dtype=dt, device=dev),
               offsets=torch.empty(L + 1, dtype=dt, device=dev),
               rows=torch.empty(L + 1, dtype=torch.int64, device=dev),
               scratch=torch.empty(max(1, ne), dtype=torch.int32, device=dev))
    ws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(ne, nend, bits), dev)
    _lib.check(lib.ygg_near_lists_sort(
        keys.data_ptr(), vals.data_ptr(), n, _lib.ptr(count), L, nend, bits, None,
        _lib.ptr(leaf_nodes), out_i64,
        out["scratch"].data_ptr(), out["rows"].data_ptr(), out["neighbors"].data_ptr(),
        out["counts"].data_ptr(), out["offsets"].data_ptr(), ws.data_ptr(), ws.numel(),
        _lib.stream_ptr(dev)), "near lists")
    return out


def _far(lib, keys, vals, n, count, T, nbl, lvl, bits, records=False):
    from yggdrax_amd import _lib
    dev = keys.device
    fend = max(1, (T - 1).bit_length())
    ne = 2 * n if records else n
    out = dict(sources=torch.full((max(1, ne),), -7, dtype=torch.int64, device=dev),
               targets=torch.full((max(1, ne),), -7, dtype=torch.int64, device=dev),
               levels=torch.full((max(1, ne),), -7, dtype=torch.int64, device=dev),
               counts=torch.empty(T, dtype=torch.int64, device=dev),
               node_off=torch.empty(T + 1, dtype=torch.int64, device=dev),
               counts32=torch.empty(T, dtype=torch.int32, device=dev),
               lm=torch.empty(T + 1, dtype=torch.int64, device=dev),
               scratch=torch.empty(max(1, ne), dtype=torch.int32, device=dev))
    ws = _lib.workspace(lib.ygg_far_lists_sort_workspace_bytes(ne, fend, bits), dev)
    _lib.check(lib.ygg_far_lists_sort(
        keys.data_ptr(), vals.data_ptr(), n, _lib.ptr(count), T, fend, bits, nbl.data_ptr(),
        lvl.data_ptr(), None, int(records), 1, out["scratch"].data_ptr(),
        out["lm"].data_ptr(),
        out["sources"].data_ptr(), out["targets"].data_ptr(), out["levels"].data_ptr(),
        out["counts"].data_ptr(), out["node_off"].data_ptr(), out["counts32"].data_ptr(),
        ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev)), "far lists")
    return out


def _expected(keys: np.ndarray, vals: np.ndarray, rows: int):
    """Row counts and the values ordered by (row, value) of keys < rows."""
    keep = keys < rows
    k, v = keys[keep].astype(np.int64), vals[keep].astype(np.int64)
    counts = np.bincount(k, minlength=rows)
    order = np.lexsort((v, k))
    return counts, v[order]


def _row_sorted(values: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """values with each row's segment sorted (rows laid out by counts)."""
    rows = np.repeat(np.arange(len(counts)), counts)
    return values[np.lexsort((values, rows))]


def _keys(rng, kind: str, n: int, rows: int) -> np.ndarray:
    if kind == "uniform":
        return rng.integers(0, rows, n)
    if kind == "one_row":
        return np.full(n, rows // 3)
    if kind == "skewed":  # most entries in one bucket, runs of equal rows
        k = rng.integers(0, rows, n)
        hot = rng.random(n) < 0.8
        k[hot] = rng.integers(0, min(rows, 300), int(hot.sum()))
        return np.sort(k[: n // 2]).tolist() + k[n // 2:].tolist()
    if kind == "spill":  # 5 % unused capacity (keys >= rows)
        k = rng.integers(0, rows, n)
        k[rng.random(n) < 0.05] = rows + 7
        return k
    raise ValueError(kind)


CASES = [  # (rows, n, key kind)
    (1000, 200_000, "uniform"),          # one bucket, 10-bit keys
    (2048, 100_000, "spill"),            # one bucket, 11-bit keys
    (5000, 300_000, "uniform"),          # 13 bits: 6 + 7
    (5000, 300_000, "one_row"),
    (200_000, 1_000_000, "skewed"),      # 18 bits: 9 + 9
    (300_000, 2_000_000, "spill"),       # 19 bits: 9 + 10
    (1 << 20, 3_000_000, "uniform"),     # 20 bits: 10 + 10, values packed with the low digit
    (1_500_000, 2_000_000, "spill"),     # 21 bits: 10 + 11, values too wide to pack
    ((1 << 22) - 5, 2_500_000, "uniform"),  # 22 bits: 11 + 11, values too wide to pack
    (70_000, 0, "uniform"),              # no entries
]


@pytest.mark.parametrize("rows,n,kind", CASES)
def test_near_buckets_match_numpy(lib, rows, n, kind):
    print(f"[row buckets] near rows={rows} n={n} {kind}", flush=True)
    rng = np.random.default_rng(rows + n)
    keys = np.asarray(_keys(rng, kind, n, rows), dtype=np.int64).astype(np.int32)
    # near values are node indices below 2 * rows, down to the last one
    vals = rng.integers(0, 2 * rows, n).astype(np.int32)
    if n:
        vals[n // 2] = 2 * rows - 1
    dev = torch.device("cuda")
    kt, vt = torch.from_numpy(keys).to(dev), torch.from_numpy(vals).to(dev)
    for out_i64 in (1, 0):
        got = _near(lib, kt, vt, n, None, rows, 0, out_i64)
        torch.cuda.synchronize()
        counts, vals_sorted = _expected(keys, vals, rows)
        valid = int(counts.sum())
        assert np.array_equal(got["counts"].cpu().numpy(), counts)
        off = np.concatenate([[0], np.cumsum(counts)])
        assert np.array_equal(got["offsets"].cpu().numpy(), off)
        assert np.array_equal(got["rows"].cpu().numpy(), off)
        nb = got["neighbors"].cpu().numpy().astype(np.int64)
        assert np.array_equal(_row_sorted(nb[:valid], counts), vals_sorted)
        assert bool((nb[valid:n] == -7).all()), "written past the valid entries"


@pytest.mark.parametrize("rows,n,kind", [(5000, 300_000, "uniform"),
                                         (300_000, 2_000_000, "skewed")])
def test_far_buckets_match_the_stable_sort(lib, rows, n, kind):
    print(f"[row buckets] far rows={rows} n={n} {kind}", flush=True)
    rng = np.random.default_rng(rows * 3 + n)
    keys = np.asarray(_keys(rng, kind, n, rows), dtype=np.int64).astype(np.int32)
    vals = rng.integers(0, rows, n).astype(np.int32)  # source nodes
    vals[0] = rows - 1
    dev = torch.device("cuda")
    kt, vt = torch.from_numpy(keys).to(dev), torch.from_numpy(vals).to(dev)
    nbl = torch.from_numpy(rng.permutation(rows).astype(np.int32)).to(dev)
    lvl = torch.from_numpy(rng.integers(0, 30, rows).astype(np.int32)).to(dev)
    got = _far(lib, kt, vt, n, None, rows, nbl, lvl, 0)
    ref = _far(lib, kt, vt, n, None, rows, nbl, lvl, 10)
    torch.cuda.synchronize()
    for k in ("counts", "node_off", "counts32", "lm", "targets", "levels"):
        assert torch.equal(got[k], ref[k]), k
    counts = np.bincount(keys.astype(np.int64), minlength=rows)
    src = got["sources"].cpu().numpy()
    assert np.array_equal(_row_sorted(src, counts), _row_sorted(ref["sources"].cpu().numpy(),
                                                                counts))


@pytest.mark.parametrize("cap,valid", [(300_000, 123_457), (1 << 20, (1 << 20) - 1),
                                       (50_000, 0)])
def test_buckets_stop_at_the_device_count(lib, cap, valid):
    print(f"[row buckets] device count cap={cap} valid={valid}", flush=True)
    rng = np.random.default_rng(cap + 3 * valid)
    L = 3000 if cap < (1 << 20) else 150_000
    dev = torch.device("cuda")
    keys = rng.integers(0, L, cap).astype(np.int32)
    vals = rng.integers(0, 2 * L, cap).astype(np.int32)
    kt, vt = torch.from_numpy(keys).to(dev), torch.from_numpy(vals).to(dev)
    count = torch.tensor([valid], dtype=torch.int64, device=dev)
    got = _near(lib, kt, vt, cap, count, L, 0)
    torch.cuda.synchronize()
    counts, vals_sorted = _expected(keys[:valid], vals[:valid], L)
    assert np.array_equal(got["counts"].cpu().numpy(), counts)
    nb = got["neighbors"].cpu().numpy()
    assert np.array_equal(_row_sorted(nb[:valid], counts), vals_sorted)
    assert bool((nb[valid:] == -7).all()), "written past the device count"


def test_buckets_repeat_with_a_smaller_input(lib):
    """Two calls in a row on one workspace, the second smaller: the look-back
    flags, tile counter and bucket tables are reset by each call."""
    print("[row buckets] repeat", flush=True)
    rng = np.random.default_rng(5)
    L = 100_000
    dev = torch.device("cuda")
    from yggdrax_amd import _lib
    big = 1_500_000
    keys = rng.integers(0, L, big).astype(np.int32)
    vals = rng.integers(0, 2 * L, big).astype(np.int32)
    kt, vt = torch.from_numpy(keys).to(dev), torch.from_numpy(vals).to(dev)
    nend = (L - 1).bit_length()
    ws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(big, nend, 0), dev)
    for n in (big, 400_000):
        nb = torch.full((big,), -7, dtype=torch.int64, device=dev)
        counts = torch.empty(L, dtype=torch.int64, device=dev)
        offs = torch.empty(L + 1, dtype=torch.int64, device=dev)
        rows = torch.empty(L + 1, dtype=torch.int64, device=dev)
        scratch = torch.empty(big, dtype=torch.int32, device=dev)
        _lib.check(lib.ygg_near_lists_sort(
            kt.data_ptr(), vt.data_ptr(), n, None, L, nend, 0, None, None, 1, scratch.data_ptr(),
            rows.data_ptr(), nb.data_ptr(), counts.data_ptr(), offs.data_ptr(), ws.data_ptr(),
            ws.numel(), _lib.stream_ptr(dev)), "near lists")
        torch.cuda.synchronize()
        c, v = _expected(keys[:n], vals[:n], L)
        assert np.array_equal(counts.cpu().numpy(), c)
        assert np.array_equal(_row_sorted(nb.cpu().numpy()[:n], c), v)


def test_row_sub_makes_values_local(lib):
    """row_sub[r] is taken off every value of row r (and off the far target
    of row r): the batched forest's system-local ids. The stable sort does
    not take it."""
    from yggdrax_amd import _lib
    print("[row buckets] row_sub", flush=True)
    rng = np.random.default_rng(17)
    dev = torch.device("cuda")
    rows, n = 200_000, 1_500_000
    keys = rng.integers(0, rows, n).astype(np.int32)
    sub = rng.integers(0, 1000, rows).astype(np.int64)
    vals = (sub[keys] + rng.integers(0, 2 * rows - 1000, n)).astype(np.int32)
    kt, vt = torch.from_numpy(keys).to(dev), torch.from_numpy(vals).to(dev)
    st = torch.from_numpy(sub).to(dev)
    nend = (rows - 1).bit_length()
    out = torch.full((n,), -7, dtype=torch.int64, device=dev)
    cnt = torch.empty(rows, dtype=torch.int64, device=dev)
    off = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    roff = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    scr = torch.empty(n, dtype=torch.int32, device=dev)
    ws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(n, nend, 0), dev)
    _lib.check(lib.ygg_near_lists_sort(
        kt.data_ptr(), vt.data_ptr(), n, None, rows, nend, 0, st.data_ptr(), None, 1, scr.data_ptr(),
        roff.data_ptr(), out.data_ptr(), cnt.data_ptr(), off.data_ptr(), ws.data_ptr(),
        ws.numel(), _lib.stream_ptr(dev)), "near lists")
    torch.cuda.synchronize()
    counts, expect = _expected(keys, (vals - sub[keys]).astype(np.int64), rows)
    assert np.array_equal(cnt.cpu().numpy(), counts)
    assert np.array_equal(_row_sorted(out.cpu().numpy(), counts), expect)
    # far: sources (node ids below T) and targets both local
    T = rows
    vt = torch.from_numpy((sub[keys] + rng.integers(0, T - 1000, n)).astype(np.int32)).to(dev)
    nbl = torch.from_numpy(rng.permutation(T).astype(np.int32)).to(dev)
    lvl = torch.from_numpy(rng.integers(0, 30, T).astype(np.int32)).to(dev)
    got = _far(lib, kt, vt, n, None, T, nbl, lvl, 0)
    ref_t = got["targets"].clone()
    src = torch.full((n,), -7, dtype=torch.int64, device=dev)
    tgt = torch.full((n,), -7, dtype=torch.int64, device=dev)
    levels = torch.full((n,), -7, dtype=torch.int64, device=dev)
    c64 = torch.empty(T, dtype=torch.int64, device=dev)
    noff = torch.empty(T + 1, dtype=torch.int64, device=dev)
    c32 = torch.empty(T, dtype=torch.int32, device=dev)
    lm = torch.empty(T + 1, dtype=torch.int64, device=dev)
    ws = _lib.workspace(lib.ygg_far_lists_sort_workspace_bytes(n, nend, 0), dev)
    _lib.check(lib.ygg_far_lists_sort(
        kt.data_ptr(), vt.data_ptr(), n, None, T, nend, 0, nbl.data_ptr(), lvl.data_ptr(),
        st.data_ptr(), 0, 1, scr.data_ptr(), lm.data_ptr(), src.data_ptr(), tgt.data_ptr(),
        levels.data_ptr(), c64.data_ptr(), noff.data_ptr(), c32.data_ptr(), ws.data_ptr(),
        ws.numel(), _lib.stream_ptr(dev)), "far lists")
    torch.cuda.synchronize()
    assert torch.equal(lm, got["lm"]) and torch.equal(levels, got["levels"])
    per_rank = (lm[1:] - lm[:-1])  # entries lie in level-major rank order
    rows_of = torch.repeat_interleave(torch.arange(T, device=dev), per_rank)
    assert torch.equal(tgt, ref_t - st[rows_of])
    pr = per_rank.cpu().numpy()
    assert np.array_equal(_row_sorted(src.cpu().numpy(), pr),
                          _row_sorted((got["sources"] - st[rows_of]).cpu().numpy(), pr))
    # the stable sort refuses it
    ws9 = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(n, nend, 9), dev)
    rc = lib.ygg_near_lists_sort(
        kt.data_ptr(), vt.data_ptr(), n, None, rows, nend, 9, st.data_ptr(), None, 1, scr.data_ptr(),
        roff.data_ptr(), out.data_ptr(), cnt.data_ptr(), off.data_ptr(), ws9.data_ptr(),
        ws9.numel(), _lib.stream_ptr(dev))
    assert rc != 0



def _pair_tree(rng, L: int):
    """Node ids of a tree of L leaves (T = 2L - 1 nodes): the node of each
    leaf row (a permutation of the last L ids) and each node's leaf row
    (-1 for internal nodes)."""
    T = 2 * L - 1
    leaf_nodes = (L - 1 + rng.permutation(L)).astype(np.int32)
    row_of = np.full(T, -1, dtype=np.int64)
    row_of[leaf_nodes] = np.arange(L)
    return T, leaf_nodes, row_of


def _records(rng, kind: str, n: int, L: int, leaf_nodes: np.ndarray, dead_key: int):
    """n near pair records (t, s) of leaf node ids; "skewed" crowds one
    range of leaf rows, "spill" leaves 5 % as unused capacity (both keys at
    dead_key, beyond the key range)."""
    pa = np.asarray(_keys(rng, "skewed" if kind == "skewed" else "uniform", n, L),
                    dtype=np.int64)
    pb = rng.integers(0, L, n)
    a, b = leaf_nodes[pa].astype(np.int64), leaf_nodes[pb].astype(np.int64)
    if kind == "spill":
        dead = rng.random(n) < 0.05
        a[dead] = dead_key
        b[dead] = dead_key
    return a.astype(np.uint32).view(np.int32), b.astype(np.uint32).view(np.int32)


PAIR_CASES = [  # (leaves, records, kind)
    (500, 150_000, "uniform"),           # 999 nodes, 10-bit keys: the split into one bucket
    (1024, 60_000, "spill"),             # 2047 nodes, 11 bits, unused capacity
    (3000, 200_000, "uniform"),          # 13 bits: 6 + 7
    (100_000, 800_000, "skewed"),        # 18 bits: 9 + 9
    (150_000, 1_000_000, "spill"),       # 19 bits: 9 + 10
    (1_000_000, 1_000_000, "uniform"),   # 21 bits: 10 + 11, values just packed
    (1_500_000, 1_000_000, "uniform"),   # 22 bits: 11 + 11, values too wide to pack
    (35_000, 0, "uniform"),              # no records
]


@pytest.mark.parametrize("L,n,kind", PAIR_CASES)
def test_near_pair_records_match_their_entries(lib, L, n, kind):
    """Near pair records (t, s) of leaf node ids list s in t's row and t in
    s's row, rows in leaf order: counts, offsets and every row's
    neighbours equal those of the two entries each record stands for."""
    print(f"[row buckets] near pair records L={L} n={n} {kind}", flush=True)
    rng = np.random.default_rng(L + 7 * n)
    T, leaf_nodes, row_of = _pair_tree(rng, L)
    nend = max(1, (T - 1).bit_length())
    a, b = _records(rng, kind, n, L, leaf_nodes, (1 << nend) + 5)
    dev = torch.device("cuda")
    at, bt = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    lt = torch.from_numpy(leaf_nodes).to(dev)
    live = a.view(np.uint32) < (1 << nend)
    a64, b64 = a[live].astype(np.int64), b[live].astype(np.int64)
    ke = np.concatenate([row_of[a64], row_of[b64]])
    ve = np.concatenate([b64, a64])
    counts, vals_sorted = _expected(ke, ve, L)
    for out_i64 in (1, 0):
        got = _near(lib, at, bt, n, None, L, 0, out_i64, leaf_nodes=lt, nend=nend)
        torch.cuda.synchronize()
        valid = int(counts.sum())
        assert valid == 2 * int(live.sum())
        assert np.array_equal(got["counts"].cpu().numpy(), counts)
        off = np.concatenate([[0], np.cumsum(counts)])
        assert np.array_equal(got["rows"].cpu().numpy(), off)
        assert np.array_equal(got["offsets"].cpu().numpy(), off)
        nb = got["neighbors"].cpu().numpy().astype(np.int64)
        assert np.array_equal(_row_sorted(nb[:valid], counts), vals_sorted)
        assert bool((nb[valid:2 * n] == -7).all()), "written past the valid entries"


@pytest.mark.parametrize("T,n", [(5000, 200_000), (300_000, 1_000_000)])
def test_far_pair_records_match_their_entries(lib, T, n):
    """Far pair records (t, s) of node ids give the lists of their entries
    (level-major rank of t, s) and (rank of s, t): rows in level-major
    order through nodes_by_level."""
    print(f"[row buckets] far pair records T={T} n={n}", flush=True)
    rng = np.random.default_rng(T * 5 + n)
    a = np.asarray(_keys(rng, "skewed", n, T), dtype=np.int64).astype(np.int32)
    b = rng.integers(0, T, n).astype(np.int32)
    nbl_np = rng.permutation(T).astype(np.int32)  # level-major rank -> node
    rank = np.empty(T, dtype=np.int64)
    rank[nbl_np] = np.arange(T)
    ke = np.concatenate([rank[a], rank[b]])
    ve = np.concatenate([b, a]).astype(np.int64)
    dev = torch.device("cuda")
    nbl = torch.from_numpy(nbl_np).to(dev)
    lvl = torch.from_numpy(rng.integers(0, 30, T).astype(np.int32)).to(dev)
    got = _far(lib, torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), n, None, T,
               nbl, lvl, 0, records=True)
    kt = torch.from_numpy(ke.astype(np.int32)).to(dev)
    vt = torch.from_numpy(ve.astype(np.int32)).to(dev)
    ref = _far(lib, kt, vt, 2 * n, None, T, nbl, lvl, 10)
    torch.cuda.synchronize()
    for k in ("counts", "node_off", "counts32", "lm", "targets", "levels"):
        assert torch.equal(got[k], ref[k]), k
    counts = np.bincount(ke, minlength=T)
    assert np.array_equal(_row_sorted(got["sources"].cpu().numpy(), counts),
                          _row_sorted(ref["sources"].cpu().numpy(), counts))


def test_pair_records_stop_at_the_device_count(lib):
    """A device record count below the capacity: only those records are
    listed, nothing is written past their entries."""
    print("[row buckets] pair records device count", flush=True)
    rng = np.random.default_rng(11)
    L, cap, valid = 150_000, 600_000, 234_567
    T, leaf_nodes, row_of = _pair_tree(rng, L)
    nend = (T - 1).bit_length()
    a, b = _records(rng, "uniform", cap, L, leaf_nodes, 0)
    dev = torch.device("cuda")
    count = torch.tensor([valid], dtype=torch.int64, device=dev)
    got = _near(lib, torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), cap, count, L,
                0, leaf_nodes=torch.from_numpy(leaf_nodes).to(dev), nend=nend)
    torch.cuda.synchronize()
    a64, b64 = a[:valid].astype(np.int64), b[:valid].astype(np.int64)
    counts, vals_sorted = _expected(np.concatenate([row_of[a64], row_of[b64]]),
                                    np.concatenate([b64, a64]), L)
    assert np.array_equal(got["counts"].cpu().numpy(), counts)
    nb = got["neighbors"].cpu().numpy()
    assert np.array_equal(_row_sorted(nb[:2 * valid], counts), vals_sorted)
    assert bool((nb[2 * valid:] == -7).all()), "written past the device count"


def test_pair_records_need_the_bucketing(lib):
    """The stable sort cannot list pair records: it refuses them."""
    from yggdrax_amd import _lib
    dev = torch.device("cuda")
    L, n = 5000, 1000
    kt = torch.zeros(n, dtype=torch.int32, device=dev)
    lt = torch.arange(L, dtype=torch.int32, device=dev)
    nend = (2 * L - 2).bit_length()
    ws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(2 * n, nend, 9), dev)
    out = torch.empty(2 * n, dtype=torch.int64, device=dev)
    scr = torch.empty(2 * n, dtype=torch.int32, device=dev)
    roff = torch.empty(L + 1, dtype=torch.int64, device=dev)
    cnt = torch.empty(L, dtype=torch.int64, device=dev)
    off = torch.empty(L + 1, dtype=torch.int64, device=dev)
    rc = lib.ygg_near_lists_sort(
        kt.data_ptr(), kt.data_ptr(), n, None, L, nend, 9, None, lt.data_ptr(), 1,
        scr.data_ptr(), roff.data_ptr(), out.data_ptr(), cnt.data_ptr(), off.data_ptr(),
        ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev))
    assert rc != 0
