"""Host-side helpers of the packed block layout (no GPU): the block range index
that gathers / scatters whole blocks between packed layouts (used by the
two-site problem's theta scatter and eigenvector extraction, key reordering,
union-of-keys addition and inner products), checked against its definition as
the concatenation of each block's element range."""

import numpy as np
import pytest
import torch

from tenax_amd.core.tensor import SymmetricTensor, _block_ranges_index


def _naive(offsets, shapes, order):
    sizes = shapes.prod(axis=1).astype(np.int64)
    parts = [np.arange(offsets[i], offsets[i] + sizes[i], dtype=np.int64) for i in order]
    return np.concatenate(parts) if parts else np.zeros(0, np.int64)


@pytest.mark.parametrize("seed", range(25))
def test_block_ranges_index_matches_definition(seed):
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(1, 20))
    shapes = rng.integers(0, 6, size=(nb, int(rng.integers(1, 5)))).astype(np.int32)  # zero-size blocks included
    offsets, total = SymmetricTensor._layout(shapes)
    order = list(rng.permutation(nb)[: int(rng.integers(0, nb + 1))])
    got = _block_ranges_index(offsets, shapes, order, "cpu")
    assert got.dtype == torch.int64
    np.testing.assert_array_equal(got.numpy(), _naive(offsets, shapes, order))
    if got.numel():
        assert int(got.max()) < total


def test_block_ranges_index_empty_and_repeated():
    shapes = np.array([[2, 3], [1, 4], [0, 5]], np.int32)
    offsets, _ = SymmetricTensor._layout(shapes)
    assert _block_ranges_index(offsets, shapes, [], "cpu").numel() == 0
    np.testing.assert_array_equal(_block_ranges_index(offsets, shapes, [2], "cpu").numpy(), np.zeros(0))
    np.testing.assert_array_equal(_block_ranges_index(offsets, shapes, [1, 1, 0], "cpu").numpy(),
                                  _naive(offsets, shapes, [1, 1, 0]))
