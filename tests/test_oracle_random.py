"""Initial-state generator: Threefry-2x32 pinned by the Random123 known-answer
vectors (also checked by jax's own test suite), the product's host Threefry
(libtnx, no GPU needed) against the oracle's, and the normal draw's
distribution.  jax itself is not installed, so the final normals of the
reference stay unpinned beyond the published construction."""

import numpy as np
import pytest

from oracle import random as orand

KAT = [((0, 0), (0, 0), (0x6B200159, 0x99BA4EFE)),
       ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
       ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0))]


@pytest.mark.parametrize("key,ctr,want", KAT)
def test_oracle_threefry_known_answers(key, ctr, want):
    a, b = orand.threefry2x32(key[0], key[1], np.uint32(ctr[0]), np.uint32(ctr[1]))
    assert (int(a), int(b)) == want


@pytest.mark.parametrize("key,ctr,want", KAT)
def test_library_threefry_known_answers(key, ctr, want):
    import ctypes as C

    from tenax_amd import _lib

    lib = _lib.load(require_gpu=False)
    out = (C.c_uint32 * 2)()
    _lib.check(lib.tnx_threefry2x32(key[0], key[1], ctr[0], ctr[1], out))
    assert (int(out[0]), int(out[1])) == want


def test_key_derivation_matches_library():
    from tenax_amd.core.tensor import PRNGKey, fold_in

    for words in [(0,), (2,), (2, 5), (3, 0, 7), (-1,), (1 << 33, 4)]:
        k = PRNGKey(words[0])
        for w in words[1:]:
            k = fold_in(k, w)
        assert k.key == orand.key_from_words(words), words
    assert orand.prng_key(-1) == (0xFFFFFFFF, 0xFFFFFFFF)


@pytest.mark.parametrize("layout", [0, 1])
def test_normal_distribution(layout):
    x = orand.normal(orand.key_from_words((7,)), (200000,), layout)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1.0) < 0.01
    from scipy.stats import kstest

    assert kstest(x, "norm").pvalue > 1e-3
    # the uniform never reaches -1 and stays below 1
    u = orand.uniform_pm1(orand.random_words(orand.key_from_words((7,)), 1000, layout))
    assert u.min() > -1.0 and u.max() < 1.0


def test_layouts_differ_and_are_prefix_stable():
    k = orand.key_from_words((1, 2))
    a1, a0 = orand.random_words(k, 64, 1), orand.random_words(k, 64, 0)
    assert not np.array_equal(a1, a0)
    # partitionable words do not depend on the draw size; original ones do
    assert np.array_equal(orand.random_words(k, 16, 1), a1[:16])
    assert not np.array_equal(orand.random_words(k, 16, 0), a0[:16])
