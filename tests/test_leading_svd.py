"""The leading-triplet factorisation used for HOTRG isometries
(tenax_amd.linalg._leading_isometries: block subspace iteration with shifted
Cholesky QR and a Jordan-Wielandt Rayleigh-Ritz step), checked on the host
against LAPACK's SVD (numpy.linalg.svd) of the same matrices: singular values
to eps * s_0, orthonormal factors, the SVD's (u, v) pairing, and the
fallbacks for small and rank-deficient matrices.  torch runs these on the CPU."""

import numpy as np
import pytest
import torch

from tenax_amd import linalg as L


def _matrix(rng, R, C, decay, rank=None):
    n = min(R, C)
    Q1, _ = np.linalg.qr(rng.standard_normal((R, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((C, n)))
    s = np.exp(-np.arange(n) / decay)
    s[7:9] = s[7]  # a degenerate pair inside the kept set
    if rank is not None:
        s[rank:] = 0.0
    return (Q1 * s) @ Q2.T


def _check(M, U, s, Vh, k):
    se = np.linalg.svd(M, compute_uv=False)
    n = min(k, min(M.shape))
    U, s, Vh = U.numpy(), s.numpy(), Vh.numpy()
    assert s.shape == (n,) and U.shape == (M.shape[0], n) and Vh.shape == (n, M.shape[1])
    assert np.all(np.diff(s) <= 0)
    assert np.max(np.abs(s - se[:n])) <= 1e-14 * se[0]
    # U^T M V^T = diag(s): factors orthonormal and paired like an SVD's
    assert np.max(np.abs(U.T @ M @ Vh.T - np.diag(s))) <= 1e-13 * se[0]
    assert np.max(np.abs(U.T @ U - np.eye(n))) <= 1e-9
    assert np.max(np.abs(Vh @ Vh.T - np.eye(n))) <= 1e-9


@pytest.mark.parametrize("shape,decay", [((400, 400), 6.0), ((420, 380), 12.0), ((300, 512), 25.0)])
def test_leading_triplets_match_lapack(shape, decay):
    rng = np.random.default_rng(sum(shape))
    k = 24
    Ms = [_matrix(rng, *shape, decay) for _ in range(3)]
    before = dict(L.LEADING_STATS)
    out = L._leading_isometries([torch.tensor(M) for M in Ms], k)
    assert L.LEADING_STATS["fallback"] == before["fallback"]
    assert L.LEADING_STATS["iterations"] > before["iterations"]
    for M, (U, s, Vh) in zip(Ms, out):
        _check(M, U, s, Vh, k)


def test_small_matrices_take_the_full_factorisation():
    rng = np.random.default_rng(5)
    k = 24
    Ms = [_matrix(rng, 60, 50, 5.0), _matrix(rng, 30, 90, 5.0)]  # short side < 4k
    before = L.LEADING_STATS["iterations"]
    out = L._leading_isometries([torch.tensor(M) for M in Ms], k)
    assert L.LEADING_STATS["iterations"] == before
    for M, (U, s, Vh) in zip(Ms, out):
        se = np.linalg.svd(M, compute_uv=False)
        n = min(k, min(M.shape))
        assert s.shape == (n,)
        assert np.max(np.abs(s.numpy() - se[:n])) <= 1e-14 * se[0]


@pytest.mark.parametrize("kind", ["exact_rank", "spectrum_below_eps"])
def test_rank_deficient_blocks_are_completed(kind):
    """A block of 2k vectors that the matrix cannot fill (rank below 2k, or a
    spectrum that falls under eps * s_0 within 2k values) breaks plain
    Cholesky QR; the iteration then completes the null directions and still
    returns the leading triplets, without the full factorisation, for that
    matrix and the healthy one in its group."""
    rng = np.random.default_rng(9)
    k = 16
    good = _matrix(rng, 256, 256, 8.0)
    if kind == "exact_rank":
        low = _matrix(rng, 256, 256, 8.0, rank=20)
    else:
        Q1, _ = np.linalg.qr(rng.standard_normal((256, 256)))
        Q2, _ = np.linalg.qr(rng.standard_normal((256, 256)))
        low = (Q1 * 10.0 ** (-np.arange(256) / 1.5)) @ Q2.T  # s_24 ~ 1e-16 s_0
    before = L.LEADING_STATS["fallback"]
    out = L._leading_isometries([torch.tensor(good), torch.tensor(low)], k)
    assert L.LEADING_STATS["fallback"] == before
    _check(good, *out[0], k)
    U, s, Vh = (x.numpy() for x in out[1])
    se = np.linalg.svd(low, compute_uv=False)
    assert np.max(np.abs(s - se[:k])) <= 1e-14 * se[0]
    # vectors of s_i resolve to eps * s_0 / s_i (here s_16 ~ 1e-10 s_0), as
    # for a Jordan-Wielandt eigh of the whole matrix
    assert np.max(np.abs(U.T @ U - np.eye(k))) <= 1e-9 and np.max(np.abs(Vh @ Vh.T - np.eye(k))) <= 1e-9
    assert np.max(np.abs(U.T @ low @ Vh.T - np.diag(s))) <= 1e-13 * se[0]


@pytest.mark.gpu
def test_truncated_svd_leading_mode_equals_full_mode(gpu):
    """truncated_svd_symmetric_many(leading=True) keeps the same singular
    values and spans as the full Jordan-Wielandt path on a Z2 tensor (on the
    GPU: the sector matrices are assembled and unpacked by the device copy
    kernel)."""
    from tenax_amd.core.index import FlowDirection, TensorIndex
    from tenax_amd.core.symmetry import ZnSymmetry
    from tenax_amd.core.tensor import SymmetricTensor

    rng = np.random.default_rng(3)
    Z2 = ZnSymmetry(2)
    ch = np.array([0] * 12 + [1] * 12, np.int32)
    idx = (TensorIndex(Z2, ch, FlowDirection.IN, "a"), TensorIndex(Z2, ch, FlowDirection.IN, "b"),
           TensorIndex(Z2, ch, FlowDirection.OUT, "c"), TensorIndex(Z2, ch, FlowDirection.OUT, "d"))
    keys, shapes = [], []
    for q in np.ndindex(2, 2, 2, 2):
        if (q[0] + q[1] - q[2] - q[3]) % 2 == 0:
            keys.append(q)
            shapes.append((12, 12, 12, 12))
    keys = np.array(keys, np.int32)
    shapes = np.array(shapes, np.int32)
    offs, tot = SymmetricTensor._layout(shapes)
    # a decaying spectrum: blocks scaled by a smooth profile
    data = torch.tensor(rng.standard_normal(tot) * np.exp(-np.arange(tot) / (tot / 8)), device="cuda")
    t = SymmetricTensor._wrap(idx, keys, shapes, offs, data)
    full = L.truncated_svd_symmetric_many([t], ["a", "b"], ["c", "d"], "x", 20, absorb="isometries")[0]
    lead = L.truncated_svd_symmetric_many([t], ["a", "b"], ["c", "d"], "x", 20, absorb="isometries",
                                          leading=True)[0]
    assert torch.allclose(lead[1], full[1], rtol=0, atol=1e-13 * float(full[1][0]))
    assert np.array_equal(lead[0].keys, full[0].keys) and np.array_equal(lead[0].shapes, full[0].shapes)
    # same kept subspace per sector (bond charge q): the rows of all U blocks
    # of the sector stacked; |U_lead^T U_full| has unit singular values
    for q in np.unique(lead[0].keys[:, -1]):
        rows = [i for i in range(lead[0].n_blocks) if lead[0].keys[i, -1] == q]
        a = torch.cat([lead[0].block(i).reshape(-1, int(lead[0].shapes[i, -1])) for i in rows])
        b = torch.cat([full[0].block(i).reshape(-1, int(full[0].shapes[i, -1])) for i in rows])
        sv = torch.linalg.svdvals(a.T @ b)
        assert torch.allclose(sv, torch.ones_like(sv), atol=1e-10)


def test_nearly_equal_shapes_iterate_in_one_padded_group(monkeypatch):
    """Sector matrices whose shapes differ by a few rows or columns (instances
    whose charge splits differ) iterate as one zero-padded batch; each still
    gets its own leading triplets, unpadded, to LAPACK's accuracy.  A shape
    too far from the others keeps a group of its own."""
    rng = np.random.default_rng(17)
    k = 24
    shapes = [(400, 400), (396, 404), (390, 388), (402, 398), (250, 260)]
    Ms = [_matrix(rng, *sh, 10.0) for sh in shapes]
    groups = []
    real_orth = L._orth

    def spy(X):
        groups.append(X.shape)
        return real_orth(X)

    monkeypatch.setattr(L, "_orth", spy)
    before = dict(L.LEADING_STATS)
    out = L._leading_isometries([torch.tensor(M) for M in Ms], k)
    assert L.LEADING_STATS["fallback"] == before["fallback"]
    # the first _orth of each group orthonormalises its random start block
    starts = [g for g in groups if g[2] == 2 * k]
    assert (4, 404, 2 * k) in starts and (1, 260, 2 * k) in starts
    for M, (U, s, Vh) in zip(Ms, out):
        _check(M, U, s, Vh, k)
