"""Host logic of the walk's stream sizing (CPU only): the starting far and
near capacities of a walk without a hint are the reference's default
per-row capacities summed over the rows (512 per node, 2048 per leaf;
_interactions_impl.py:47-48), scaled down together to the byte limit."""

from __future__ import annotations

from yggdrax_amd import interactions as I


def test_start_capacities_are_the_default_row_capacities():
    assert I._DEFAULT_MAX_INTERACTIONS == 512 and I._DEFAULT_MAX_NEIGHBORS == 2048
    T, L = 524287, 262144  # the 2^24-point, leaf-64 octree
    far, near = I._start_capacities(T, L)
    assert (far, near) == (512 * T, 2048 * L)
    assert 8 * (far + near) <= I.STREAM_START_MAX_BYTES
    assert I._start_capacities(1, 1) == (512, 2048)
    assert I._start_capacities(0, 0) == (512, 2048)


def test_start_capacities_respect_the_byte_limit():
    T, L = 1 << 25, 1 << 24
    far, near = I._start_capacities(T, L)
    limit = I.STREAM_START_MAX_BYTES // 8
    assert far + near <= limit
    assert far + near >= limit - 2
    # scaled together: the far / near ratio of the default capacities stays
    assert abs(far / near - (512 * T) / (2048 * L)) < 1e-6


def test_default_sizing_is_the_resumable_walk():
    assert I.STREAM_SIZING in ("grow", "count")
    import os
    if "YGGDRAX_STREAM_SIZING" not in os.environ:
        assert I.STREAM_SIZING == "grow"
