"""64-bit multiset digests of (key, value) pairs, computed identically with
numpy (for the oracle's outputs) and torch (for device outputs).

digest(keys, values, salt) = sum over j of mix(mix(keys[j] * C + salt) ^ values[j])
modulo 2^64, where mix is the splitmix64 finaliser. With keys = positions it
digests an array exactly (up to collisions of a 64-bit hash); with keys =
row ids it digests the set of (row, value) entries of a CSR list, which does
not depend on the order of entries within or across rows. Values are taken
as 64-bit integers (float32 arrays by their bit patterns).
"""

from __future__ import annotations

import numpy as np

_C = 0x9E3779B97F4A7C15
_M1 = 0xBF58476D1CE4E5B9
_M2 = 0x94D049BB133111EB
_MASK = (1 << 64) - 1


def _s64(v: int) -> int:
    """An unsigned 64-bit constant as the signed value with the same bits."""
    return v - (1 << 64) if v >= (1 << 63) else v


# ---- numpy ----------------------------------------------------------------

def _mix_np(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(_C)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(_M1)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(_M2)
    return z ^ (z >> np.uint64(31))


def _as_u64_np(a) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype == np.float32:
        a = a.view(np.uint32)
    elif a.dtype == np.float64:
        a = a.view(np.uint64)
    elif a.dtype == np.bool_:
        a = a.astype(np.uint8)
    return a.astype(np.int64).view(np.uint64) if a.dtype != np.uint64 else a


def digest_np(keys, values, salt: int = 0, chunk: int = 1 << 24) -> int:
    k = _as_u64_np(keys).reshape(-1)
    v = _as_u64_np(values).reshape(-1)
    assert k.shape == v.shape
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for a in range(0, k.size, chunk):
            kk = k[a:a + chunk] * np.uint64(_C) + np.uint64(salt & _MASK)
            total += _mix_np(_mix_np(kk) ^ v[a:a + chunk]).sum(dtype=np.uint64)
    return int(total)


def array_digest_np(a, salt: int = 0) -> int:
    a = np.asarray(a).reshape(-1)
    return digest_np(np.arange(a.size, dtype=np.int64), a, salt)


# ---- torch ----------------------------------------------------------------

def _srl(z, k: int):
    """Logical right shift of int64 bits."""
    return (z >> k) & ((1 << (64 - k)) - 1)


def _mix_t(z):
    z = z + _s64(_C)
    z = (z ^ _srl(z, 30)) * _s64(_M1)
    z = (z ^ _srl(z, 27)) * _s64(_M2)
    return z ^ _srl(z, 31)


def _as_i64_t(x):
    import torch
    if x.dtype == torch.float32:
        x = x.contiguous().view(torch.int32)
        return x.to(torch.int64) & 0xFFFFFFFF
    if x.dtype == torch.uint64:
        return x.contiguous().view(torch.int64)
    if x.dtype == torch.bool:
        x = x.to(torch.uint8)
    return x.to(torch.int64)


def digest_t(keys, values, salt: int = 0, chunk: int = 1 << 26) -> int:
    import torch
    k = _as_i64_t(keys).reshape(-1)
    v = _as_i64_t(values).reshape(-1)
    assert k.shape == v.shape
    total = torch.zeros((), dtype=torch.int64, device=k.device)
    for a in range(0, k.numel(), chunk):
        kk = k[a:a + chunk] * _s64(_C) + _s64(salt & _MASK)
        total = total + _mix_t(_mix_t(kk) ^ v[a:a + chunk]).sum()
    return int(total.item()) & _MASK


def array_digest_t(x, salt: int = 0) -> int:
    import torch
    x = x.reshape(-1)
    return digest_t(torch.arange(x.numel(), dtype=torch.int64, device=x.device), x, salt)
