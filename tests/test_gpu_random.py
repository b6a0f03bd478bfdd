"""Device Threefry stream against the oracle: identical 64-bit words for both
jax layouts, normals within a few ulp (the device erfinv is Giles'
approximation refined by Newton steps; the oracle uses scipy's erfinv), and
SymmetricTensor.random_normal block for block."""

import numpy as np
import pytest
import torch

from oracle import core as oc
from oracle import random as orand
from tests.helpers import from_oracle, to_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("words", [(0,), (3,), (2, 17), (12345, 1, 2)])
def test_device_words_match_oracle(gpu, layout, words):
    from tenax_amd import _lib

    k = orand.key_from_words(words)
    n = 100003
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    lib = _lib.load()
    _lib.check(lib.tnx_random_bits(k[0], k[1], n, layout, _lib.dptr(out), _lib.stream_handle()))
    got = out.cpu().numpy().view(np.uint64)
    assert np.array_equal(got, orand.random_words(k, n, layout))


@pytest.mark.parametrize("layout", [0, 1])
def test_device_normals_match_oracle(gpu, layout, monkeypatch):
    from tenax_amd.core import tensor as T

    monkeypatch.setattr(T, "RANDOM_LAYOUT", layout)
    for words in [(0,), (3,), (5, 9)]:
        got = T.normal(T.PRNGKey(words), (300007,))
        want = orand.normal(orand.key_from_words(words), (300007,), layout)
        err = np.abs(got - want)
        assert np.all(err <= 8 * np.finfo(float).eps * np.maximum(1.0, np.abs(want))), (words, err.max())


def test_random_normal_matches_oracle(gpu):
    from tenax_amd import FlowDirection, SymmetricTensor, TensorIndex, U1Symmetry

    sym = U1Symmetry()
    ch = np.array([-1, 0, 1, 0, 2], np.int32)
    idx = (TensorIndex(sym, ch, FlowDirection.IN, "a"), TensorIndex(sym, ch, FlowDirection.IN, "b"),
           TensorIndex(sym, ch, FlowDirection.OUT, "c"))
    t = SymmetricTensor.random_normal(idx, 4, stddev=0.5)
    U1 = oc.U1()
    o = oc.SymT.random_normal(tuple(oc.Leg(U1, ch, f, l) for l, f in (("a", 1), ("b", 1), ("c", -1))), (4,),
                              stddev=0.5)
    po = to_oracle(t)
    assert set(po.blocks) == set(o.blocks)
    for k in o.blocks:
        np.testing.assert_allclose(po.blocks[k], o.blocks[k], rtol=0, atol=1e-14)
    assert from_oracle is not None
