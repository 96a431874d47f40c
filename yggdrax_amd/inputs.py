"""Seeded synthetic particle systems for the tree pipeline.

* ``jax_uniform``: a numpy restatement of ``jax.random.PRNGKey`` /
  ``jax.random.split`` / ``jax.random.uniform`` (threefry2x32-20, Random123),
  so the reference's own examples can be reproduced without JAX.
* ``uniform_cube`` (512 points in [-1,1]^3, jax PRNGKey 0),
  ``plummer_sphere`` (2^20 points), ``clustered`` (2^24 points: 64 Plummer
  clusters plus 10 % uniform background) and ``plummer_batch`` (1024
  systems of log-uniform size) generate the benchmark workloads.
"""

from __future__ import annotations

import numpy as np

_ROT = (13, 15, 26, 6, 17, 29, 16, 24)


def _rotl(x: np.ndarray, r: int) -> np.ndarray:
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def threefry2x32(k0, k1, x0, x1):
    """Threefry-2x32 with 20 rounds (Salmon et al., SC'11, Random123)."""
    with np.errstate(over="ignore"):
        k0 = np.uint32(k0)
        k1 = np.uint32(k1)
        x0 = np.asarray(x0, np.uint32).copy()
        x1 = np.asarray(x1, np.uint32).copy()
        ks = (k0, k1, np.uint32(k0 ^ k1 ^ np.uint32(0x1BD11BDA)))
        x0 += ks[0]
        x1 += ks[1]
        for i in range(5):
            for r in (_ROT[0:4] if i % 2 == 0 else _ROT[4:8]):
                x0 += x1
                x1 = _rotl(x1, r)
                x1 ^= x0
            x0 += ks[(i + 1) % 3]
            x1 += ks[(i + 2) % 3] + np.uint32(i + 1)
    return x0, x1


def prng_key(seed: int) -> tuple[int, int]:
    """jax.random.PRNGKey(seed) for the threefry implementation."""
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return (seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF


def _iota_hi_lo(size: int):
    idx = np.arange(size, dtype=np.uint64)
    return (idx >> np.uint64(32)).astype(np.uint32), (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def _threefry_counts(key, counts: np.ndarray) -> np.ndarray:
    """jax's threefry_2x32(keypair, count): split the flat counts in halves."""
    c = counts.ravel()
    odd = c.size % 2
    if odd:
        c = np.concatenate([c, np.zeros(1, np.uint32)])
    h = c.size // 2
    o0, o1 = threefry2x32(key[0], key[1], c[:h], c[h:])
    out = np.concatenate([o0, o1])
    return out[:-1] if odd else out


def split(key, num: int = 2, partitionable: bool = True):
    """jax.random.split: fold-like counters when partitionable, else iota."""
    if partitionable:
        hi, lo = _iota_hi_lo(num)
        b1, b2 = threefry2x32(key[0], key[1], hi, lo)
        return [(int(b1[i]), int(b2[i])) for i in range(num)]
    out = _threefry_counts(key, np.arange(2 * num, dtype=np.uint32)).reshape(num, 2)
    return [(int(out[i, 0]), int(out[i, 1])) for i in range(num)]


def random_bits32(key, shape, partitionable: bool = True) -> np.ndarray:
    size = int(np.prod(shape))
    if partitionable:
        hi, lo = _iota_hi_lo(size)
        b1, b2 = threefry2x32(key[0], key[1], hi, lo)
        bits = b1 ^ b2
    else:
        bits = _threefry_counts(key, np.arange(size, dtype=np.uint32))
    return bits.reshape(shape)


def jax_uniform(key, shape, minval=0.0, maxval=1.0, partitionable: bool = True):
    """jax.random.uniform for float32: mantissa-randomised floats in [1,2) - 1."""
    bits = random_bits32(key, shape, partitionable)
    fb = (bits >> np.uint32(9)) | np.uint32(0x3F800000)
    floats = fb.view(np.float32) - np.float32(1.0)
    lo = np.float32(minval)
    hi = np.float32(maxval)
    return np.maximum(lo, floats * (hi - lo) + lo).astype(np.float32)


def uniform_cube(n: int = 512, seed: int = 0, partitionable: bool = True):
    """Points uniform in [-1,1]^3 and masses uniform in [0.5,1.5] from
    jax.random.split(PRNGKey(seed)) -> (pos_key, mass_key)."""
    kp, km = split(prng_key(seed), 2, partitionable)
    pos = jax_uniform(kp, (n, 3), -1.0, 1.0, partitionable)
    mass = jax_uniform(km, (n,), 0.5, 1.5, partitionable)
    return pos, mass


def _plummer_radii(rng: np.random.Generator, n: int, a: float, rmax: float) -> np.ndarray:
    out = np.empty(n, np.float64)
    filled = 0
    while filled < n:
        m = n - filled
        u = rng.random(m)
        u = np.clip(u, 1e-300, 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[np.isfinite(r) & (r <= rmax)]
        out[filled:filled + r.size] = r
        filled += r.size
    return out


def _isotropic(rng: np.random.Generator, r: np.ndarray) -> np.ndarray:
    n = r.shape[0]
    cz = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    return np.stack([r * sz * np.cos(phi), r * sz * np.sin(phi), r * cz], axis=1)


def plummer_sphere(n: int = 1 << 20, seed: int = 2, a: float = 1.0, rmax: float = 20.0):
    """Plummer sphere (scale radius a, truncated at rmax), equal masses 1/N."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = _isotropic(rng, _plummer_radii(rng, n, a, rmax)).astype(np.float32)
    mass = np.full(n, 1.0 / n, np.float32)
    return pos, mass


def clustered(n: int = 1 << 24, seed: int = 3, clusters: int = 64,
              background: float = 0.10):
    """64 Plummer clusters (a_k = 10^U(-3,-1), centres U[-1,1]^3, weights
    Dirichlet(0.5), truncated at 20 a_k) plus a uniform background in
    [-1,1]^3; masses lognormal(0, 0.5)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = 10.0 ** rng.uniform(-3.0, -1.0, clusters)
    centres = rng.uniform(-1.0, 1.0, (clusters, 3))
    w = rng.dirichlet(np.full(clusters, 0.5))
    n_bg = int(round(background * n))
    counts = rng.multinomial(n - n_bg, w)
    parts = []
    for k in range(clusters):
        if counts[k] == 0:
            continue
        r = _plummer_radii(rng, int(counts[k]), a[k], 20.0 * a[k])
        parts.append(_isotropic(rng, r) + centres[k])
    parts.append(rng.uniform(-1.0, 1.0, (n_bg, 3)))
    pos = np.concatenate(parts, axis=0)
    pos = pos[rng.permutation(n)].astype(np.float32)
    mass = rng.lognormal(0.0, 0.5, n).astype(np.float32)
    return pos, mass


def plummer_batch(num_systems: int = 1024, seed: int = 4, n_min: int = 256,
                  n_max: int = 65536):
    """Independent Plummer spheres with N_i = round(exp(U[ln n_min, ln n_max]))
    and scale radius a_i ~ U[0.05, 0.5]. Returns (positions, masses, offsets)
    with system i occupying rows offsets[i]:offsets[i+1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), num_systems))).astype(np.int64)
    radii = rng.uniform(0.05, 0.5, num_systems)
    pos = []
    for i in range(num_systems):
        r = _plummer_radii(rng, int(sizes[i]), radii[i], 20.0 * radii[i])
        pos.append(_isotropic(rng, r))
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    p = np.concatenate(pos).astype(np.float32)
    m = np.repeat((1.0 / sizes).astype(np.float32), sizes)
    return p, m, offsets
