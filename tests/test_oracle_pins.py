"""Independent pins of the oracle's truncation rule and DMRG anchors (CPU).

* Sparse exact diagonalisation (built from spin operators, not the MPO) fixes
  the ground energies the DMRG oracle must reach, including the L=20 anchor
  of the committed C1 trajectory fixture.
* The truncated block SVD (reference linalg.py:211-235) must keep exactly the
  k largest singular values of the densified matrix (Eckart-Young: the
  reconstruction error equals the tail of the dense spectrum), in
  non-increasing order, and the truncation-error rule must keep the fewest
  values whose discarded tail stays within the tolerance.
"""

import json
import os

import numpy as np
import pytest

from oracle import contract as ocn
from oracle import dmrg as od
from oracle import linalg as ol

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_oracle.json")


@pytest.mark.parametrize("L,two_sz,anchor", [(4, 0, -1.616025403784438), (6, 0, -2.493577133887925),
                                             (4, 2, -0.957106781186548)])
def test_sector_ed_matches_dense_ed(L, two_sz, anchor):
    """The sparse sector ED agrees with the dense Kronecker Hamiltonian restricted to the sector."""
    H = od.heisenberg_ed_matrix(L)
    # heisenberg_ed_matrix: site j is bit L-1-j of the basis index, bit 0 = spin up (Sz = +1/2)
    sz2 = np.array([sum(1 if not (s >> (L - 1 - j)) & 1 else -1 for j in range(L)) for s in range(2 ** L)])
    sel = np.nonzero(sz2 == two_sz)[0]
    e_dense = np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0]
    assert abs(e_dense - anchor) < 1e-12
    assert abs(od.heisenberg_ed_sector_energy(L, two_sz) - anchor) < 1e-12


def test_c1_fixture_reaches_l20_ed_anchor():
    """Both committed oracle trajectories (with and without re-orthogonalisation)
    end within 1e-10 of the independently computed L=20, 2Sz=0 ground energy."""
    gold = json.load(open(GOLDEN))
    e_ed = od.heisenberg_ed_sector_energy(20, 0)
    assert abs(e_ed - gold["ed_anchor_2sz0"]) < 1e-12
    for run in ("reorth", "no_reorth"):
        r = gold[run]
        assert r["converged"]
        assert abs(r["energy"] - e_ed) < 1e-10, (run, r["energy"] - e_ed)
        # variational: every sweep energy lies above the ground energy, and decreases
        es = np.array(r["energies_per_sweep"])
        assert np.all(es >= e_ed - 1e-12)
        assert np.all(np.diff(es) <= 1e-12)
        # the bond multiplicities sum to the bond dimension, capped at chi=64
        for i, b in enumerate(r["bond_multiplicities"]):
            assert sum(b.values()) == min(2 ** (i + 1), 2 ** (19 - i), 64)


def _theta(seed=3):
    mps = od.build_random_symmetric_mps(8, bond_dim=12, seed=seed)
    return ocn.contract(mps[3], mps[4])


@pytest.mark.parametrize("k", [1, 5, 17, 40])
def test_truncated_svd_keeps_top_k_of_dense_spectrum(k):
    th = _theta()
    left, right = ["v2_3", "p3"], ["p4", "v4_5"]
    U, s, Vh, s_full, _ = ol.truncated_svd_symmetric(th, left, right, "bond", max_singular_values=k)
    dense = th.todense()
    m = dense.reshape(dense.shape[0] * dense.shape[1], -1)
    sig = np.linalg.svd(m, compute_uv=False)
    k_eff = min(k, len(s_full))
    assert len(s) == k_eff
    assert np.all(np.diff(s) <= 0.0)  # non-increasing
    np.testing.assert_allclose(s, sig[:k_eff], rtol=0, atol=1e-12 * sig[0])
    # Eckart-Young: the kept factors reconstruct the best rank-k approximation
    rec = np.einsum("apk,k,kqb->apqb", U.todense(), s, Vh.todense())
    err = np.linalg.norm(rec - dense)
    assert abs(err - np.sqrt(np.sum(sig[k_eff:] ** 2))) < 1e-10 * sig[0]
    # the full spectrum is the dense spectrum (structural zeros aside)
    np.testing.assert_allclose(s_full, sig[: len(s_full)], atol=1e-12 * sig[0])


@pytest.mark.parametrize("eps", [1e-1, 3e-2, 1e-3])
def test_truncation_error_rule_keeps_fewest_within_tolerance(eps):
    th = _theta(seed=5)
    U, s, Vh, s_full, _ = ol.truncated_svd_symmetric(th, ["v2_3", "p3"], ["p4", "v4_5"], "bond",
                                                     max_truncation_err=eps)
    dense = th.todense()
    sig = np.linalg.svd(dense.reshape(dense.shape[0] * dense.shape[1], -1), compute_uv=False)
    tot = np.sum(sig ** 2)
    tails = np.concatenate([np.cumsum((sig ** 2)[::-1])[::-1], [0.0]])  # tails[n] = sum_{j>=n} s_j^2
    n_ref = max(1, int(np.argmax(tails <= eps ** 2 * tot)))
    assert len(s) == n_ref
    assert np.sqrt(tails[n_ref] / tot) <= eps
    assert np.sqrt(tails[n_ref - 1] / tot) > eps or n_ref == 1


@pytest.mark.parametrize("site", [0, 3, 6])
def test_factored_matvec_equals_per_combination_matvec(site):
    """The oracle's stage-by-stage matvec (used at chi=2048) reaches the same
    blocks as the per-combination restatement of dmrg.py:1048-1151 with the
    same values to rounding, and agrees with the dense einsum."""
    mpo = od.build_mpo_heisenberg(8)
    mps = od.build_random_symmetric_mps(8, bond_dim=4, seed=3)
    _, _, mps, _, _ = od.dmrg(mpo, mps, od.DMRGConfig(max_bond_dim=12, num_sweeps=1, lanczos_max_iter=8))
    oL = od.trivial_left_env()
    for k in range(site):
        oL = od.update_left_env(oL, mps[k], mpo[k])
    oR = od.trivial_right_env()
    for k in range(7, site + 1, -1):
        oR = od.update_right_env(oR, mps[k], mpo[k])
    th = ocn.contract(mps[site], mps[site + 1])
    if th.ndim == 3:
        th = od.pad(th, not th.labels()[0].startswith("v"))
    a = od.matvec(oL, th, mpo[site], mpo[site + 1], oR)
    b = od.FactoredMatvec(oL, mpo[site], mpo[site + 1], oR)(th)
    assert sorted(a.blocks) == sorted(b.blocks)
    scale = max(float(np.abs(v).max()) for v in a.blocks.values())
    for k, v in a.blocks.items():
        np.testing.assert_allclose(b.blocks[k], v, rtol=0, atol=1e-14 * scale)
    dense = np.einsum("abc,apqd,bpse,eqtf,dfg->cstg", oL.todense(), th.todense(), mpo[site].todense(),
                      mpo[site + 1].todense(), oR.todense())
    np.testing.assert_allclose(b.todense(), dense, rtol=0, atol=1e-13 * scale)
