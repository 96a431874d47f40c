"""Autotuner for the pipeline's launch parameters.

Plans map (parameter, key width, key range, size class) to a choice. With
tuning enabled (``YGGDRAX_AUTOTUNE=1`` or :func:`enable`) the first call in
a size class (bit length of the element count) times every candidate on
the device and keeps the fastest; otherwise the library defaults apply.
Plans can be saved to and loaded from a JSON file (``YGGDRAX_AUTOTUNE_CACHE``).

Tuned parameters:
  * radix digit width (8..11 bits) of the row-key sorts that turn the
    walk's emission streams into CSR lists (when the lists come from the
    stable sort; the bucketing path has no digit width to tune), and of
    the Morton-code sort;
  * the flat walk's grid size (persistent blocks per generation launch:
    the analogue of the reference's process_block) and its queue tile
    (pairs per thread of one chunk, i.e. 1024-4096-pair chunks, each
    reserving its outputs with one atomic per stream).
"""

from __future__ import annotations

import json
import os
import threading
from typing import Callable, Dict, Sequence

import torch

from . import _lib

_LOCK = threading.Lock()
_PLANS: Dict[str, int] = {}
# measured milliseconds per candidate of every plan tuned in this process
_TIMINGS: Dict[str, Dict[int, float]] = {}
_ENABLED = os.environ.get("YGGDRAX_AUTOTUNE", "0") not in ("", "0", "false", "False")
_CACHE = os.environ.get("YGGDRAX_AUTOTUNE_CACHE", "")

SORT_DIGIT_BITS = (8, 9, 10, 11)
WALK_GRID_BLOCKS = (1024, 2048, 4096, 8192)
WALK_TILES = (4, 8, 16)


def enable(flag: bool = True) -> None:
    global _ENABLED
    _ENABLED = bool(flag)


def enabled() -> bool:
    return _ENABLED


def plans() -> Dict[str, int]:
    with _LOCK:
        return dict(_PLANS)


def timings() -> Dict[str, Dict[int, float]]:
    """Candidate timings (ms) of the plans tuned in this process."""
    with _LOCK:
        return {k: dict(v) for k, v in _TIMINGS.items()}


def clear() -> None:
    """Forget every plan and timing (the defaults apply until retuned)."""
    with _LOCK:
        _PLANS.clear()
        _TIMINGS.clear()


def save(path: str) -> None:
    with _LOCK, open(path, "w") as f:
        json.dump(_PLANS, f, indent=1, sort_keys=True)


def load(path: str) -> None:
    with open(path) as f:
        data = json.load(f)
    with _LOCK:
        _PLANS.update({str(k): int(v) for k, v in data.items()})


if _CACHE and os.path.exists(_CACHE):
    load(_CACHE)


def _time_ms(fn: Callable[[], None], reps: int = 3) -> float:
    fn()  # warm-up (allocations, code load)
    st = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(reps):
        fn()
    b.record(st)
    b.synchronize()
    return a.elapsed_time(b) / reps


def choose(key: str, candidates: Sequence[int], timer: Callable[[int], float],
           default: int) -> int:
    """The plan for `key`: cached, else (when tuning) the fastest candidate
    by `timer`, else `default`."""
    with _LOCK:
        if key in _PLANS:
            return _PLANS[key]
    if not _ENABLED:
        return default
    measured = {int(c): float(timer(c)) for c in candidates}
    best = min(measured, key=measured.get)
    with _LOCK:
        _PLANS[key] = int(best)
        _TIMINGS[key] = measured
        if _CACHE:
            with open(_CACHE, "w") as f:
                json.dump(_PLANS, f, indent=1, sort_keys=True)
    return int(best)


def _size_class(n: int) -> int:
    return max(1, int(n)).bit_length()


def sort_digit_bits(key_bytes: int, range_bits: int, n: int) -> int:
    """Digit width for a stable radix sort of n (key, u32) pairs whose keys
    span range_bits bits."""
    lib = _lib.load()
    default = int(lib.ygg_sort_default_digit_bits(key_bytes, 0, range_bits))
    key = f"sort/k{key_bytes}/r{range_bits}/n{_size_class(n)}"

    def timer(bits: int) -> float:
        dev = torch.device("cuda", torch.cuda.current_device())
        g = torch.Generator(device=dev).manual_seed(1)
        m = 1 << _size_class(n) - 1
        hi = 1 << min(range_bits, 62)
        dt = torch.int64 if key_bytes == 8 else torch.int32
        keys = torch.randint(0, min(hi, 1 << (8 * key_bytes - 1)), (m,), generator=g,
                             device=dev, dtype=dt)
        ko = torch.empty_like(keys)
        vo = torch.empty(m, dtype=torch.int32, device=dev)
        ws = _lib.workspace(lib.ygg_sort_workspace_bytes(m, key_bytes, 0, range_bits, bits), dev)
        fn_c = lib.ygg_sort_pairs_u64 if key_bytes == 8 else lib.ygg_sort_pairs_u32
        st = _lib.stream_ptr(dev)

        def run():
            _lib.check(fn_c(keys.data_ptr(), None, ko.data_ptr(), vo.data_ptr(), m, 0,
                            range_bits, bits, ws.data_ptr(), ws.numel(), st), "autotune sort")
        return _time_ms(run)

    return choose(key, SORT_DIGIT_BITS, timer, default)


def walk_plan(num_nodes: int, grid_default: int, tile_default: int) -> tuple:
    """(grid blocks, queue tile) of the flat walk for trees of this size:
    the tuned plans where they exist, else the defaults (no tuning here)."""
    c = _size_class(num_nodes)
    with _LOCK:
        return (_PLANS.get(f"walk/grid/n{c}", grid_default),
                _PLANS.get(f"walk/tile/n{c}", tile_default))


def walk_tile(num_nodes: int, run_with: Callable[[int], None], default: int) -> int:
    """Queue tile of the flat walk (pairs per thread of a 256-thread chunk);
    `run_with(per)` runs one walk with that tile."""
    key = f"walk/tile/n{_size_class(num_nodes)}"
    return choose(key, WALK_TILES, lambda t: _time_ms(lambda: run_with(t), reps=1), default)


def walk_grid_blocks(num_nodes: int, run_with: Callable[[int], None], default: int) -> int:
    """Grid size of the flat walk's generation kernel; `run_with(blocks)`
    runs one walk with that grid (used only while tuning)."""
    key = f"walk/grid/n{_size_class(num_nodes)}"
    return choose(key, WALK_GRID_BLOCKS, lambda b: _time_ms(lambda: run_with(b), reps=1), default)


__all__ = ["choose", "clear", "enable", "enabled", "load", "plans", "save", "sort_digit_bits",
           "timings", "walk_grid_blocks", "walk_plan", "walk_tile"]
