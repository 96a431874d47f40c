"""ygg_morton_sort_noalloc (workspace in a file-scope device array) through
the C ABI: two calls in a row, the second with a smaller input, each equal
to ygg_morton_sort with a caller workspace and to the oracle's codes and
stable order; sizes beyond the maximum are refused before any launch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import oracle
from tests import parity
from yggdrax_amd import inputs

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 20   # YGG_NOALLOC_MAX_POINTS
TOO_LARGE = 5          # YGG_ERR_TOO_LARGE


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from yggdrax_amd import _lib
    return _lib.load()


def _inputs(n, seed):
    pos, _ = inputs.clustered(n, seed, clusters=8)
    lo, hi = oracle.infer_bounds(pos)
    return pos, np.concatenate([lo, hi]).astype(np.float32)


def _both(lib, pos, bounds):
    from yggdrax_amd import _lib
    n = pos.shape[0]
    P = torch.from_numpy(pos).cuda()
    B = torch.from_numpy(bounds).cuda()
    st = _lib.stream_ptr(P.device)
    c1 = torch.empty(n, dtype=torch.int64, device=P.device)
    p1 = torch.empty(n, dtype=torch.int32, device=P.device)
    c2 = torch.empty_like(c1)
    p2 = torch.empty_like(p1)
    ws = _lib.workspace(lib.ygg_morton_sort_workspace_bytes(n), P.device)
    _lib.check(lib.ygg_morton_sort(P.data_ptr(), n, B.data_ptr(), c1.data_ptr(), p1.data_ptr(),
                                   ws.data_ptr(), ws.numel(), st), "morton_sort")
    _lib.check(lib.ygg_morton_sort_noalloc(P.data_ptr(), n, B.data_ptr(), c2.data_ptr(),
                                           p2.data_ptr(), st), "morton_sort_noalloc")
    torch.cuda.synchronize()
    return c1, p1, c2, p2


def test_noalloc_matches_workspace_entry_on_two_calls(lib):
    for n, seed in ((MAX_POINTS, 31), (70001, 32)):
        print(f"[noalloc] n={n} seed={seed}", flush=True)
        pos, bounds = _inputs(n, seed)
        c1, p1, c2, p2 = _both(lib, pos, bounds)
        assert torch.equal(c1, c2) and torch.equal(p1, p2)
        codes = oracle.morton_encode(pos, (bounds[:3], bounds[3:]))
        perm = np.argsort(codes, kind="stable")
        assert np.array_equal(parity.np_(p2).astype(np.int64), perm)
        assert np.array_equal(parity.np_(c2).view(np.uint64), codes[perm])


def test_noalloc_refuses_sizes_beyond_the_maximum(lib):
    from yggdrax_amd import _lib
    dev = torch.device("cuda")
    rc = lib.ygg_morton_sort_noalloc(ctypes.c_void_p(16), MAX_POINTS + 1, ctypes.c_void_p(16),
                                     ctypes.c_void_p(16), ctypes.c_void_p(16),
                                     _lib.stream_ptr(dev))
    assert rc == TOO_LARGE
    with pytest.raises(RuntimeError, match="larger than the fixed maximum"):
        _lib.check(rc, "morton_sort_noalloc")
    assert lib.ygg_morton_sort_noalloc(None, 0, None, None, None, _lib.stream_ptr(dev)) == 1
