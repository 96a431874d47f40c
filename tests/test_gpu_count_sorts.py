"""The fused row-key sorts with a device-side entry count (the HIP-graph
path's streams are allocated at capacity and filled up to the walk's
totals): with garbage past the count, the far and near lists, row offsets
and counts equal those of the same sort run on exactly the valid prefix,
and nothing past the count is written."""

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


def _far(lib, keys, vals, n, count, T, nbl, lvl):
    from yggdrax_amd import _lib
    dev = keys.device
    fend = max(1, (T - 1).bit_length())
    out = dict(sources=torch.full((n,), -7, dtype=torch.int64, device=dev),
               targets=torch.full((n,), -7, dtype=torch.int64, device=dev),
               levels=torch.full((n,), -7, dtype=torch.int64, device=dev),
               counts=torch.empty(T, dtype=torch.int64, device=dev),
               node_off=torch.empty(T + 1, dtype=torch.int64, device=dev),
               counts32=torch.empty(T, dtype=torch.int32, device=dev),
               lm=torch.empty(T + 1, dtype=torch.int64, device=dev),
               sorted=torch.empty(max(1, n), dtype=torch.int32, device=dev))
    ws = _lib.workspace(lib.ygg_far_lists_sort_workspace_bytes(n, fend, 10), dev)
    _lib.check(lib.ygg_far_lists_sort(
        keys.data_ptr(), vals.data_ptr(), n, _lib.ptr(count), T, fend, 10, nbl.data_ptr(),
        lvl.data_ptr(), None, 0, 1, out["sorted"].data_ptr(), out["lm"].data_ptr(),
        out["sources"].data_ptr(), out["targets"].data_ptr(), out["levels"].data_ptr(),
        out["counts"].data_ptr(), out["node_off"].data_ptr(), out["counts32"].data_ptr(),
        ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev)), "far lists sort")
    return out


def _near(lib, keys, vals, n, count, L):
    from yggdrax_amd import _lib
    dev = keys.device
    nend = max(1, (L - 1).bit_length())
    out = dict(neighbors=torch.full((n,), -7, dtype=torch.int64, device=dev),
               counts=torch.empty(L, dtype=torch.int64, device=dev),
               offsets=torch.empty(L + 1, dtype=torch.int64, device=dev),
               rows=torch.empty(L + 1, dtype=torch.int64, device=dev),
               sorted=torch.empty(max(1, n), dtype=torch.int32, device=dev))
    ws = _lib.workspace(lib.ygg_near_lists_sort_workspace_bytes(n, nend, 9), dev)
    _lib.check(lib.ygg_near_lists_sort(
        keys.data_ptr(), vals.data_ptr(), n, _lib.ptr(count), L, nend, 9, None, None, 1,
        out["sorted"].data_ptr(), out["rows"].data_ptr(), out["neighbors"].data_ptr(),
        out["counts"].data_ptr(), out["offsets"].data_ptr(), ws.data_ptr(), ws.numel(),
        _lib.stream_ptr(dev)), "near lists sort")
    return out


@pytest.mark.parametrize("cap,valid", [(300_000, 123_457), (70_000, 70_000), (50_000, 0),
                                       (1 << 20, (1 << 20) - 1)])
def test_fused_sorts_stop_at_the_device_count(lib, cap, valid):
    print(f"[count sorts] cap={cap} valid={valid}", flush=True)
    rng = np.random.default_rng(cap + valid)
    T, L = 5000, 3000
    dev = torch.device("cuda")
    # valid keys are row keys; the tail holds garbage keys inside the row range
    fk = torch.from_numpy(rng.integers(0, T, cap).astype(np.int32)).to(dev)
    fv = torch.from_numpy(rng.integers(0, T, cap).astype(np.int32)).to(dev)
    nk = torch.from_numpy(rng.integers(0, L, cap).astype(np.int32)).to(dev)
    nv = torch.from_numpy(rng.integers(0, 2 * L, cap).astype(np.int32)).to(dev)
    nbl = torch.from_numpy(rng.permutation(T).astype(np.int32)).to(dev)
    lvl = torch.from_numpy(rng.integers(0, 30, T).astype(np.int32)).to(dev)
    count = torch.tensor([valid], dtype=torch.int64, device=dev)
    got_f = _far(lib, fk, fv, cap, count, T, nbl, lvl)
    got_n = _near(lib, nk, nv, cap, count, L)
    ref_f = _far(lib, fk[:valid].contiguous(), fv[:valid].contiguous(), valid, None, T, nbl, lvl)
    ref_n = _near(lib, nk[:valid].contiguous(), nv[:valid].contiguous(), valid, None, L)
    torch.cuda.synchronize()
    for k in ("counts", "node_off", "counts32", "lm"):
        assert torch.equal(got_f[k], ref_f[k]), k
    for k in ("sources", "targets", "levels"):
        assert torch.equal(got_f[k][:valid], ref_f[k][:valid]), k
        assert bool((got_f[k][valid:] == -7).all()), f"{k} written past the count"
    for k in ("counts", "offsets", "rows"):
        assert torch.equal(got_n[k], ref_n[k]), k
    assert torch.equal(got_n["neighbors"][:valid], ref_n["neighbors"][:valid])
    assert bool((got_n["neighbors"][valid:] == -7).all()), "neighbours written past the count"
    # and the prefix sort is the stable sort by row key
    if valid:
        order = np.argsort(fk[:valid].cpu().numpy(), kind="stable")
        assert np.array_equal(got_f["sources"][:valid].cpu().numpy(),
                              fv[:valid].cpu().numpy()[order].astype(np.int64))
        assert int(got_f["lm"][T]) == valid and int(got_n["rows"][L]) == valid
