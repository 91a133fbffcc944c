"""numpy restatement of the JAX default PRNG path used by the reference's
``make_initial_state`` (test infrastructure, part of the oracle).

The reference draws its random initial q' with ``jax.random.PRNGKey``,
``jax.random.split`` and ``jax.random.normal`` (nhqg/solver.py:1351-1357).
JAX is a third-party dependency absent from the reference tree
(requirements.txt:13 pins ``jax>=0.10.0``); its published algorithm is restated
here:

* Threefry-2x32 with 20 rounds (Salmon et al., "Parallel random numbers: as
  easy as 1, 2, 3", SC'11; the Random123 reference constants), pinned below
  against the Random123 known-answer vectors.
* ``PRNGKey(seed)``: key = (seed >> 32, seed & 0xffffffff) as uint32.
* ``split``: two variants exist in JAX.  ``partitionable`` (the default since
  JAX 0.5) hashes the counter pair (0, i) for each new key i; ``original``
  hashes ``iota(2*num)`` split in halves.
* ``random_bits`` (64-bit): ``partitionable`` hashes the counter pair
  (hi(i), lo(i)) of the flat index i and forms (b1 << 32) | b2; ``original``
  hashes ``iota(2*size)`` in halves and joins the two halves as hi|lo words.
* ``uniform`` float64: the top 52 bits become the mantissa of a number in
  [1, 2), minus 1, scaled to [lo, hi) with lo = nextafter(-1, 0), hi = 1.
* ``normal``: sqrt(2) * erfinv(u).  XLA's erfinv and scipy's agree to a few
  ulp; the fixture comparison is made at 1e-12 relative, far above that.

Which split/bits variant produced the reference's golden fixture is decided
by the fixture test itself (tests/test_oracle_golden.py).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erfinv

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))
_PARITY = np.uint32(0x1BD11BDA)


def _rotl(x, r):
    r = np.uint32(r)
    return (x << r) | (x >> np.uint32(32 - int(r)))


def threefry2x32(key, x0, x1):
    """Threefry-2x32-20 block function on uint32 arrays x0, x1 with key (k0, k1)."""
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    ks = (k0, k1, np.uint32(k0 ^ k1 ^ _PARITY))
    a = np.asarray(x0, dtype=np.uint32).copy()
    b = np.asarray(x1, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        a = a + ks[0]
        b = b + ks[1]
        for blk in range(5):
            for r in _ROT[blk % 2]:
                a = a + b
                b = _rotl(b, r)
                b = b ^ a
            a = a + ks[(blk + 1) % 3]
            b = b + ks[(blk + 2) % 3] + np.uint32(blk + 1)
    return a, b


def prng_key(seed: int):
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return (np.uint32(s >> 32), np.uint32(s & 0xFFFFFFFF))


def split(key, num: int = 2, partitionable: bool = True):
    if partitionable:
        lo = np.arange(num, dtype=np.uint32)
        hi = np.zeros(num, dtype=np.uint32)
        b1, b2 = threefry2x32(key, hi, lo)
        return [(b1[i], b2[i]) for i in range(num)]
    counts = np.arange(2 * num, dtype=np.uint32)
    b1, b2 = threefry2x32(key, counts[:num], counts[num:])
    flat = np.concatenate([b1, b2])
    return [(flat[2 * i], flat[2 * i + 1]) for i in range(num)]


def random_bits64(key, size: int, partitionable: bool = True) -> np.ndarray:
    if partitionable:
        idx = np.arange(size, dtype=np.uint64)
        hi = (idx >> np.uint64(32)).astype(np.uint32)
        lo = (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        b1, b2 = threefry2x32(key, hi, lo)
        return (b1.astype(np.uint64) << np.uint64(32)) | b2.astype(np.uint64)
    count = 2 * size  # 64 bits per element = two 32-bit words
    if count >= 0xFFFFFFFF:
        raise NotImplementedError("restatement covers fewer than 2**31 draws")
    iota = np.arange(count, dtype=np.uint32)
    half = count // 2
    b1, b2 = threefry2x32(key, iota[:half], iota[half:])
    words = np.concatenate([b1, b2])
    hi = words[:size].astype(np.uint64)
    lo = words[size:].astype(np.uint64)
    return (hi << np.uint64(32)) | lo


def uniform64(key, shape, minval, maxval, partitionable: bool = True) -> np.ndarray:
    size = int(np.prod(shape))
    bits = random_bits64(key, size, partitionable)
    mant = (bits >> np.uint64(12)) | np.uint64(0x3FF0000000000000)
    u = mant.view(np.float64) - 1.0
    out = u * (maxval - minval) + minval
    return np.maximum(minval, out).reshape(shape)


def normal64(key, shape, partitionable: bool = True) -> np.ndarray:
    lo = np.nextafter(-1.0, 0.0)
    u = uniform64(key, shape, lo, 1.0, partitionable)
    return np.sqrt(2.0) * erfinv(u)
