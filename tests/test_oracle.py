"""CPU tests pinning the oracle against closed forms, brute force and ED anchors."""

import itertools

import numpy as np
import pytest

from oracle import contract as ocn
from oracle import core as oc
from oracle import dmrg as od
from oracle import linalg as ol

U1 = oc.U1()


def brute_valid(legs, target=0):
    sym = legs[0].sym
    out = []
    for key in itertools.product(*[sorted(set(l.charges.tolist())) for l in legs]):
        v = sum(l.flow * q for l, q in zip(legs, key))
        if sym.n is not None:
            v %= sym.n
        if v == target:
            out.append(key)
    return sorted(out)


@pytest.mark.parametrize("sym", [oc.U1(), oc.Zn(2), oc.Zn(3)])
@pytest.mark.parametrize("rank", [2, 3, 4])
def test_valid_blocks_brute_force(sym, rank):
    rng = np.random.default_rng(rank)
    legs = []
    for i in range(rank):
        ch = rng.integers(-2, 3, size=5) if sym.n is None else rng.integers(0, sym.n, size=4)
        legs.append(oc.Leg(sym, ch, oc.IN if i % 2 == 0 else oc.OUT))
    assert sorted(oc.compute_valid_blocks(legs)) == brute_valid(legs)


def test_valid_blocks_target_u1():
    legs = [oc.Leg(U1, [-1, 0, 1], oc.IN), oc.Leg(U1, [1, -1], oc.IN), oc.Leg(U1, [-2, 0, 2], oc.OUT)]
    assert sorted(oc.compute_valid_blocks(legs, target=1)) == brute_valid(legs, 1)


def _rand_sym(legs, seed):
    return oc.SymT.random_normal(legs, (seed,))


def test_contract_symmetric_matches_dense_einsum():
    a = oc.Leg(U1, [-1, 0, 0, 1, 2], oc.IN, "a")
    p = oc.Leg(U1, [1, -1], oc.IN, "p")
    b = oc.Leg(U1, [-1, 0, 1, 1, 2, 3], oc.OUT, "b")
    c = oc.Leg(U1, [0, 1, 2, 2, 3, 4], oc.OUT, "c")
    A = _rand_sym((a, p, b), 1)
    B = _rand_sym((oc.Leg(U1, b.charges, oc.IN, "b"), oc.Leg(U1, [1, -1], oc.IN, "q"), c), 2)
    C, combos = ocn.contract_symmetric([A, B], "apb,bqc->apqc")
    ref = np.einsum("apb,bqc->apqc", A.todense(), B.todense())
    np.testing.assert_allclose(C.todense(), ref, atol=1e-12)
    assert combos  # nonempty


def test_blockwise_env_update_matches_dense():
    mpo = od.build_mpo_heisenberg(4)
    mps = od.build_random_symmetric_mps(4, bond_dim=4, seed=42)
    L0 = od.trivial_left_env()
    L1 = od.update_left_env(L0, mps[0], mpo[0])
    A = od.pad(mps[0], True)
    ref = np.einsum("abc,apd,bpxe,cxf->def", L0.todense(), A.todense(), mpo[0].todense(), A.todense())
    np.testing.assert_allclose(L1.todense(), ref, atol=1e-12)


def test_mpo_bond_charges_known_answer():
    # tests/test_auto_mpo.py:628-656: Heisenberg bond charges [0,0,2,-2,0]
    mpo = od.build_mpo_heisenberg(6)
    for i in range(1, 5):
        np.testing.assert_array_equal(mpo[i].legs[0].charges, [0, 0, 2, -2, 0])


def test_mpo_contracts_to_ed_hamiltonian():
    L = 4
    mpo = od.build_mpo_heisenberg(L)
    W = [m.todense() for m in mpo]
    H = W[0]
    for w in W[1:]:
        H = np.einsum("a...b,bpqc->a...pqc", H, w)
    # legs: (1, t0, b0, t1, b1, ..., 1) -> matrix rows = tops, cols = bots
    H = H.reshape([2] * (2 * L))
    perm = list(range(0, 2 * L, 2)) + list(range(1, 2 * L, 2))
    Hm = H.transpose(perm).reshape(2 ** L, 2 ** L)
    np.testing.assert_allclose(Hm, od.heisenberg_ed_matrix(L), atol=1e-14)


def test_truncated_svd_reconstructs():
    mps = od.build_random_symmetric_mps(6, bond_dim=6, seed=3)
    th = ocn.contract(mps[2], mps[3])
    U, s, Vh, s_full, _ = ol.truncated_svd_symmetric(th, ["v1_2", "p2"], ["p3", "v3_4"], "bond")
    dense = th.todense()
    rec = np.einsum("apk,k,kqb->apqb", U.todense(), s, Vh.todense())
    np.testing.assert_allclose(rec, dense, atol=1e-12)
    ref_s = np.linalg.svd(dense.reshape(dense.shape[0] * 2, -1), compute_uv=False)
    np.testing.assert_allclose(np.sort(s_full)[::-1], ref_s[: len(s_full)], atol=1e-12)


@pytest.mark.parametrize("L,tgt,anchor", [(4, 0, -1.616025403784438), (6, 0, -2.493577133887925),
                                          (4, 2, -0.957106781186548)])
def test_oracle_dmrg_matches_ed_anchor(L, tgt, anchor):
    mpo = od.build_mpo_heisenberg(L)
    mps = od.build_random_symmetric_mps(L, bond_dim=4, seed=7, target_charge=tgt)
    cfg = od.DMRGConfig(max_bond_dim=16, num_sweeps=6, lanczos_max_iter=30)
    e, *_ = od.dmrg(mpo, mps, cfg)
    assert abs(e - anchor) < 1e-10


def test_ed_matrix_anchor_independent():
    H = od.heisenberg_ed_matrix(4)
    assert abs(np.linalg.eigvalsh(H)[0] - (-1.616025403784438)) < 1e-12


def test_parity_error():
    with pytest.raises(ValueError, match="parity"):
        od.build_random_symmetric_mps(3, bond_dim=4, target_charge=0)
