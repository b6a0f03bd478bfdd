"""Oracle: jax.random-style Threefry-2x32 stream for the initial tensors.

Test infrastructure only (see oracle/__init__.py).  Restates the generator the
reference draws its initial tensors from (src/tenax/core/tensor.py:655-660:
``fold_in(key, i)`` per sorted valid block, ``normal(subkey, shape)``;
src/tenax/algorithms/dmrg.py:1484: ``PRNGKey(seed + i)``), i.e. jax.random's
published construction (jax is not installed here, so its output is not
available to compare against):

* Threefry-2x32, 20 rounds (Salmon et al., SC'11), pinned by the Random123
  known-answer vectors in tests/test_oracle_random.py;
* PRNGKey(seed) = (seed >> 32, seed & 0xffffffff); fold_in(k, d) = threefry(k, (0, d));
* 64-bit words: layout 1 ("partitionable", jax's current default) word i =
  threefry(k, (i >> 32, i & 0xffffffff)); layout 0 ("original") word i =
  (y0 << 32) | y1 with (y0, y1) = threefry(k, (i, n + i)) for n words;
* uniform: top 52 bits as the mantissa of [1, 2), minus 1, mapped to
  [nextafter(-1, 0), 1); normal = sqrt(2) * erfinv(u) (scipy's erfinv).
"""

from __future__ import annotations

import numpy as np

LAYOUT = 1
_M32 = np.uint64(0xFFFFFFFF)


def _rotl(v, r):
    return ((v << np.uint32(r)) | (v >> np.uint32(32 - r))).astype(np.uint32)


def threefry2x32(k0, k1, c0, c1):
    """Threefry-2x32 on uint32 arrays (broadcast), 20 rounds."""
    with np.errstate(over="ignore"):
        k0, k1 = np.uint32(k0), np.uint32(k1)
        ks = (k0, k1, np.uint32(k0 ^ k1 ^ np.uint32(0x1BD11BDA)))
        x0 = (np.asarray(c0, np.uint32) + ks[0]).astype(np.uint32)
        x1 = (np.asarray(c1, np.uint32) + ks[1]).astype(np.uint32)
        rot = ((13, 15, 26, 6), (17, 29, 16, 24))
        for g in range(5):
            for r in rot[g & 1]:
                x0 = (x0 + x1).astype(np.uint32)
                x1 = _rotl(x1, r)
                x1 = (x1 ^ x0).astype(np.uint32)
            x0 = (x0 + ks[(g + 1) % 3]).astype(np.uint32)
            x1 = (x1 + ks[(g + 2) % 3] + np.uint32(g + 1)).astype(np.uint32)
    return x0, x1


def prng_key(seed: int):
    s = int(seed)
    return (s >> 32) & 0xFFFFFFFF, s & 0xFFFFFFFF


def fold_in(key, data: int):
    a, b = threefry2x32(key[0], key[1], np.uint32(0), np.uint32(int(data) & 0xFFFFFFFF))
    return int(a), int(b)


def key_from_words(words):
    """PRNGKey(words[0]) folded with every later word."""
    key = prng_key(words[0])
    for w in words[1:]:
        key = fold_in(key, w)
    return key


def random_words(key, n: int, layout: int = LAYOUT) -> np.ndarray:
    i = np.arange(n, dtype=np.uint64)
    if layout == 1:
        a, b = threefry2x32(key[0], key[1], (i >> np.uint64(32)).astype(np.uint32), (i & _M32).astype(np.uint32))
    else:
        a, b = threefry2x32(key[0], key[1], i.astype(np.uint32), (i + np.uint64(n)).astype(np.uint32))
    return (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)


def uniform_pm1(words: np.ndarray) -> np.ndarray:
    f = ((words >> np.uint64(12)) | np.uint64(0x3FF0000000000000)).view(np.float64) - 1.0
    lo = np.nextafter(-1.0, 0.0)
    return np.maximum(lo, f * (1.0 - lo) + lo)


def normal(key, shape, layout: int = LAYOUT) -> np.ndarray:
    from scipy.special import erfinv

    n = int(np.prod(shape)) if len(shape) else 1
    u = uniform_pm1(random_words(key, n, layout))
    return (np.sqrt(2.0) * erfinv(u)).reshape(shape)
