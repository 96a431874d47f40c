"""Index dtype policy (yggdrax/dtypes.py:12-40).

INDEX_DTYPE is torch.int64 unless YGGDRAX_INDEX_PRECISION (or
JACCPOT_INDEX_PRECISION) selects int32. Device kernels compute with int32
indices; public outputs are widened to INDEX_DTYPE.
"""

from __future__ import annotations

import os

import torch


def _resolve_index_dtype() -> torch.dtype:
    raw = os.environ.get("YGGDRAX_INDEX_PRECISION")
    if raw is None:
        raw = os.environ.get("JACCPOT_INDEX_PRECISION", "int64")
    raw = str(raw).strip().lower()
    if raw in ("int32", "i32", "32"):
        return torch.int32
    return torch.int64


INDEX_DTYPE: torch.dtype = _resolve_index_dtype()


def as_index(x, device=None) -> torch.Tensor:
    return torch.as_tensor(x, dtype=INDEX_DTYPE, device=device)


__all__ = ["INDEX_DTYPE", "as_index"]
