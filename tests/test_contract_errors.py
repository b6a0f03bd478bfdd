"""contract()'s argument errors, with the reference's exception types and
message texts (src/tenax/contraction/contractor.py:86-91, 100-104, 130-136,
521-522 and 565-569).  Raised on the host before any device work, so these
run without a GPU."""

import re

import numpy as np
import pytest
import torch

from tenax_amd import DenseTensor, FlowDirection, SymmetricTensor, TensorIndex, U1Symmetry, contract
from tenax_amd.contraction.contractor import contract_with_subscripts

U1 = U1Symmetry()


def _dense(*labels, dim=2):
    """A dense tensor held in host memory (DenseTensor() itself moves data to
    the GPU; the checks under test run before any device work)."""
    idx = tuple(TensorIndex(U1, np.zeros(dim, np.int32), FlowDirection.IN, l) for l in labels)
    t = DenseTensor.__new__(DenseTensor)
    t._data = torch.zeros((dim,) * len(labels), dtype=torch.float64)
    t._indices = idx
    return t


def _symmetric(*labels):
    """A one-block symmetric tensor held in host memory (no device work is done)."""
    idx = tuple(TensorIndex(U1, np.zeros(1, np.int32), FlowDirection.IN, l) for l in labels)
    r = len(labels)
    return SymmetricTensor._wrap(idx, np.zeros((1, r), np.int32), np.ones((1, r), np.int32), np.zeros(1, np.int64),
                                 torch.zeros(1, dtype=torch.float64))


def test_no_tensors():
    with pytest.raises(ValueError, match=re.escape("contract() requires at least one tensor")):
        contract()


def test_label_used_three_times():
    with pytest.raises(ValueError, match=re.escape(
            "Label 'k' appears 3 times across tensors. Labels must appear at most 2 times "
            "(one per tensor to contract).")):
        contract(_dense("i", "k"), _dense("k", "j"), _dense("k", "l"))


def test_more_than_52_labels():
    labels = [f"x{i}" for i in range(53)]
    ts = [_dense(l, dim=1) for l in labels]
    with pytest.raises(ValueError, match=re.escape(
            "Too many unique labels (53) for einsum encoding. Maximum supported is 52 (a-z + A-Z).")):
        contract(*ts)


def test_exactly_52_labels_is_accepted_by_the_encoder():
    from tenax_amd.contraction.contractor import _labels_to_subscripts

    ts = [_dense(f"x{i}", dim=1) for i in range(52)]
    subs, out = _labels_to_subscripts(ts, None)
    assert len(out) == 52 and subs.count(",") == 51


def test_output_label_not_free():
    with pytest.raises(ValueError, match=re.escape(
            "output_labels contains 'k' which is not a free label. Free labels are: ['i', 'j']")):
        contract(_dense("i", "k"), _dense("k", "j"), output_labels=("i", "k"))


def test_mixed_tensor_types():
    with pytest.raises(TypeError, match=re.escape(
            "Cannot mix DenseTensor and SymmetricTensor in a single contraction. "
            "Got types: ['DenseTensor', 'SymmetricTensor']. Convert all tensors to the same type first.")):
        contract(_dense("i", "k"), _symmetric("k", "j"))


def test_mixed_types_through_subscripts():
    A, B = _symmetric("i", "k"), _dense("k", "j")
    with pytest.raises(TypeError, match="Cannot mix DenseTensor and SymmetricTensor"):
        contract_with_subscripts([A, B], "ik,kj->ij", (A.indices[0], B.indices[1]))


def test_label_subscripts_sorted_by_str():
    """contractor.py:97-108: letters follow the labels sorted by str(); free
    labels keep their order of first appearance."""
    from tenax_amd.contraction.contractor import _labels_to_subscripts

    subs, out = _labels_to_subscripts([_dense("b", 10), _dense(10, "a")], None)
    # sorted by str: "10" < "a" < "b" -> 10:a, a:b, b:c
    assert subs == "ca,ab->cb"
    assert [ix.label for ix in out] == ["b", "a"]


def test_fused_output_order_run_count():
    """The GEMM writes a requested output order itself only when the A free
    legs (rows) and the B free legs (columns) each fall into at most three
    runs of legs that stay adjacent in that order."""
    from tenax_amd.contraction.contractor import _fused_order_fits

    assert _fused_order_fits((0, 1, 2, 3), 2)          # natural order: one run a side
    assert _fused_order_fits((3, 2, 1, 0), 2)          # two legs a side, reversed: two runs each
    assert _fused_order_fits((2, 0, 3, 1, 4, 5), 3)    # rows 0,1,2 at 1,3,0 -> 3 runs
    assert not _fused_order_fits((3, 2, 4, 1, 0), 4)   # rows a,b,c,d reversed: 4 runs
    assert _fused_order_fits((4, 0, 1, 2, 3), 4)       # rows contiguous, columns first
