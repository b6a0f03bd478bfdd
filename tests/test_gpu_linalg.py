"""GPU block-sparse truncated SVD against the FP64 oracle (linalg.py:75-333).

Both decompositions are checked: rocSOLVER SVD per sector (``absorb=None``)
and the batched Gram-eigh factorisation used by the DMRG split (``absorb`` =
"right"/"left").  Checked per case: the kept singular values and s_full, the
new bond's charge order (the global stable sort), the truncated reconstruction
U s Vh (gauge-free), and that the isometric factor is orthonormal.
"""

import numpy as np
import pytest
import torch

from oracle import contract as ocn
from oracle import dmrg as od
from oracle import linalg as ola
from tests.helpers import from_oracle, to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def theta10():
    mpo = od.build_mpo_heisenberg(10)
    mps = od.build_random_symmetric_mps(10, bond_dim=4, seed=5)
    _, _, mps, _, _ = od.dmrg(mpo, mps, od.DMRGConfig(max_bond_dim=24, num_sweeps=1, lanczos_max_iter=8))
    return ocn.contract(mps[4], mps[5])


def _dense_product(U, s, Vh, absorb):
    Ud, Vd = to_oracle(U).todense(), to_oracle(Vh).todense()
    if absorb is None:
        Ud = Ud * s
    return np.tensordot(Ud, Vd, axes=([Ud.ndim - 1], [0]))


def _isometry(U, Vh, absorb):
    if absorb == "left":
        V = to_oracle(Vh).todense()
        m = V.reshape(V.shape[0], -1).T
    else:
        Ud = to_oracle(U).todense()
        m = Ud.reshape(-1, Ud.shape[-1])
    return float(np.abs(m.T @ m - np.eye(m.shape[1])).max())


@pytest.fixture(params=["direct", "gram"])
def factor_path(request, monkeypatch):
    """Run the DMRG split through the direct per-sector SVD (small problems)
    and through the batched Gram-eigh path (forced with SMALL_SECTOR = 0)."""
    import tenax_amd.linalg as la

    if request.param == "gram":
        monkeypatch.setattr(la, "SMALL_SECTOR", 0)
    return request.param


@pytest.mark.parametrize("absorb", [None, "right", "left"])
@pytest.mark.parametrize("chi", [5, 16, 40])
def test_truncated_svd_matches_oracle(gpu, theta10, absorb, chi, factor_path):
    from tenax_amd.linalg import truncated_svd_symmetric

    labs = list(theta10.labels())
    Uo, so, Vo, sfo, _ = ola.truncated_svd_symmetric(theta10, labs[:2], labs[2:], "bond", chi)
    U, s, Vh, sf = truncated_svd_symmetric(from_oracle(theta10), labs[:2], labs[2:], "bond", chi,
                                           absorb=absorb)
    s, sf = s.cpu().numpy(), sf.cpu().numpy()
    smax = float(sfo[0])
    assert len(s) == len(so)
    assert np.abs(s - so).max() < 1e-12 * smax
    assert np.abs(sf - sfo).max() < 1e-12 * smax
    if chi < len(sfo):
        # bond charge order of the kept values (global stable sort by -s)
        assert np.array_equal(U.indices[-1].charges, Uo.legs[-1].charges)
    ref = np.tensordot(Uo.todense() * so, Vo.todense(), axes=([2], [0]))
    got = _dense_product(U, s, Vh, absorb)
    assert np.abs(got - ref).max() < 1e-11 * smax
    assert _isometry(U, Vh, absorb) < 1e-13


def test_truncation_error_rule(gpu, theta10):
    from tenax_amd.linalg import truncated_svd_symmetric

    labs = list(theta10.labels())
    for err in (1e-1, 1e-3):
        _, so, _, _, _ = ola.truncated_svd_symmetric(theta10, labs[:2], labs[2:], "bond", 100, err)
        for absorb in (None, "right"):
            _, s, _, _ = truncated_svd_symmetric(from_oracle(theta10), labs[:2], labs[2:], "bond", 100, err,
                                                 absorb=absorb)
            assert len(s) == len(so)


def test_gram_factorisation_at_chi256(gpu, monkeypatch):
    """Large sectors (up to a few hundred rows): Gram-eigh vs rocSOLVER SVD."""
    import tenax_amd.linalg as la

    monkeypatch.setattr(la, "SMALL_SECTOR", 0)
    from tenax_amd.contraction import contract
    from tenax_amd.linalg import truncated_svd_symmetric
    from tenax_amd.workloads import synthetic_dmrg_mps

    mps = synthetic_dmrg_mps(24, 256, seed=7)
    th = contract(mps.get_tensor(11), mps.get_tensor(12))
    labs = list(th.labels())
    U0, s0, V0, sf0 = truncated_svd_symmetric(th, labs[:2], labs[2:], "bond", 256)
    ref = _dense_product(U0, s0.cpu().numpy(), V0, None)
    smax = float(sf0[0])
    for absorb in ("right", "left"):
        U, s, Vh, sf = truncated_svd_symmetric(th, labs[:2], labs[2:], "bond", 256, absorb=absorb)
        assert torch.allclose(sf, sf0, atol=1e-12 * smax, rtol=0)
        assert np.array_equal(U.indices[-1].charges, U0.indices[-1].charges)
        got = _dense_product(U, s.cpu().numpy(), Vh, absorb)
        assert np.abs(got - ref).max() < 1e-11 * smax
        assert _isometry(U, Vh, absorb) < 1e-13


@pytest.mark.parametrize("absorb", [None, "isometries"])
def test_non_leading_left_legs_follow_reference_reshape(gpu, theta10, absorb):
    """Left legs not leading: each block enters its sector matrix reshaped in
    stored leg order (linalg.py:181), like the oracle."""
    from tenax_amd.linalg import truncated_svd_symmetric

    labs = list(theta10.labels())
    left, right = [labs[0], labs[2]], [labs[1], labs[3]]
    Uo, so, Vo, sfo, _ = ola.truncated_svd_symmetric(theta10, left, right, "bond", 12)
    U, s, Vh, sf = truncated_svd_symmetric(from_oracle(theta10), left, right, "bond", 12, absorb=absorb)
    smax = float(sfo[0])
    assert np.abs(s.cpu().numpy() - so).max() < 1e-12 * smax
    assert np.array_equal(U.indices[-1].charges, Uo.legs[-1].charges)
    ref = np.tensordot(Uo.todense() * so, Vo.todense(), axes=([2], [0]))
    got = _dense_product(U, s.cpu().numpy(), Vh, None)
    assert np.abs(got - ref).max() < 1e-11 * smax


def test_copy2d_batched_rectangles(gpu):
    """tnx_copy2d_execute: many strided rectangles (1x1 up to chunk-spanning
    ones, scaled) in one launch, against torch slicing; nothing outside the
    destination rectangles is touched."""
    from tenax_amd import _lib

    rng = np.random.default_rng(5)
    src = torch.tensor(rng.standard_normal(200_000), device="cuda")
    dst = torch.full((300_000,), 7.0, dtype=torch.float64, device="cuda")
    ref = dst.clone()
    copies, at = [], 0
    for rows, cols in [(1, 1), (3, 5), (17, 300), (100, 41), (1, 5000), (64, 64), (7, 1)]:
        sld = cols + int(rng.integers(0, 9))
        so = int(rng.integers(0, src.numel() - rows * sld))
        dld = cols + int(rng.integers(0, 5))
        scale = float(rng.choice([1.0, -1.0, 0.25]))
        copies.append((src, so, sld, dst, at, dld, rows, cols, scale))
        for r in range(rows):
            ref[at + r * dld: at + r * dld + cols] = scale * src[so + r * sld: so + r * sld + cols]
        at += rows * dld + 3
    _lib.copy2d(copies, dst.device)
    assert torch.equal(dst, ref)


@pytest.mark.parametrize("absorb", [None, "right", "left"])
def test_tied_values_across_sectors_and_keep_all(gpu, absorb, factor_path):
    """Two sectors holding the same block have tied singular values; the
    global sort is stable (linalg.py:211), so the earlier sector's value comes
    first and a cut inside a tie keeps it, and the new bond's charges follow.
    A bound above the rank keeps every value.  Checked against the oracle on
    the direct and the Gram path."""
    from oracle import core as oc
    from tenax_amd.linalg import truncated_svd_symmetric

    U1 = oc.U1()
    rng = np.random.default_rng(12)
    rows = oc.Leg(U1, np.array([0] * 4 + [1] * 4 + [2] * 3), oc.IN, "r")
    cols = oc.Leg(U1, np.array([0] * 5 + [1] * 5 + [2] * 4), oc.OUT, "c")
    m0 = rng.standard_normal((4, 5))
    T = oc.SymT({(0, 0): m0, (1, 1): m0.copy(), (2, 2): 0.01 * rng.standard_normal((3, 4))}, (rows, cols))
    Td = from_oracle(T)
    for chi in (1, 3, 5, 9, 100):
        Uo, so, Vo, sfo, _ = ola.truncated_svd_symmetric(T, ["r"], ["c"], "bond", chi)
        U, s, Vh, sf = truncated_svd_symmetric(Td, ["r"], ["c"], "bond", chi, absorb=absorb)
        s = s.cpu().numpy()
        assert len(s) == len(so) == min(chi, 11), chi
        assert np.abs(s - so).max() < 1e-12 * so[0], chi
        assert np.array_equal(U.indices[-1].charges, Uo.legs[-1].charges), (chi, U.indices[-1].charges)
        ref = np.tensordot(Uo.todense() * so, Vo.todense(), axes=([1], [0]))
        got = _dense_product(U, s, Vh, absorb)
        assert np.abs(got - ref).max() < 1e-11 * so[0], chi
        assert _isometry(U, Vh, absorb) < 1e-13, chi
    # the ties themselves: sector 0 and sector 1 give bitwise-equal spectra
    _, s, _, _ = truncated_svd_symmetric(Td, ["r"], ["c"], "bond", 100, absorb=absorb)
    s = s.cpu().numpy()
    assert np.array_equal(s[0:8:2], s[1:8:2])
