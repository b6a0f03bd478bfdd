"""Oracle fermionic sign rules (tensor.py:32-50, 853-856; contractor.py:206-285,
462-470; linalg.py:171-186) pinned independently of their restatement.

* Koszul sign: the reference's own known answers (test_fermionic.py:339-369).
* Inversion pairs: against an explicit bubble sort of the graded legs into the
  target order, counting swaps of two odd legs.
* Transpose: round trip, composition, bosonic = plain transpose.
* Contraction: covariance -- contracting a fermionically transposed operand, or
  asking for another output order, gives the fermionic transpose of the same
  result (the Koszul sign is a cocycle) -- plus conservation and the bosonic
  limit (no signs at all when every charge is even).
* SVD: reconstruction of fermionic tensors (test_fermionic.py:497-575).
"""

import itertools

import numpy as np
import pytest

from oracle import contract as ocn
from oracle import core as oc
from oracle import linalg as ol

FP = oc.FermionParity()
FU1 = oc.FermionicU1()


@pytest.mark.parametrize("par,perm,want", [
    ([1, 1, 1], (0, 1, 2), 1), ([1, 1], (1, 0), -1), ([0, 1], (1, 0), 1), ([0, 0], (1, 0), 1),
    ([1, 0, 1], (2, 0, 1), -1), ([1, 1, 1], (2, 1, 0), -1), ([1, 1, 1, 1], (1, 0, 3, 2), 1), ([], (), 1)])
def test_koszul_known_answers(par, perm, want):
    assert oc.koszul_sign(par, perm) == want


def _bubble_sign(input_subs, output, charge_of):
    """Sign of carrying the concatenated legs to [output, contracted pairs] by
    adjacent swaps, -1 per swap of two odd legs."""
    seq = [(c, k) for s in input_subs for k, c in enumerate(s)]
    chars = [c for c, _ in seq]
    cnt = {c: chars.count(c) for c in chars}
    target = list(output) + list(dict.fromkeys(c for c in chars if cnt[c] == 2))
    occ = {}
    items = []
    for c in chars:
        o = occ.get(c, 0)
        occ[c] = o + 1
        items.append((2 * target.index(c) + o, charge_of[c] % 2))
    sign = 1
    a = list(items)
    for i in range(len(a)):
        for j in range(len(a) - 1 - i):
            if a[j][0] > a[j + 1][0]:
                if a[j][1] and a[j + 1][1]:
                    sign = -sign
                a[j], a[j + 1] = a[j + 1], a[j]
    return sign


@pytest.mark.parametrize("seed", range(20))
def test_inversion_pairs_match_bubble_sort(seed):
    rng = np.random.default_rng(seed)
    letters = "abcdefg"
    nA, nB, nc = rng.integers(1, 4), rng.integers(1, 4), None
    nc = int(rng.integers(0, min(nA, nB) + 1))
    cont = list(letters[:nc])
    fa = list(letters[nc:nc + nA - nc]) if nA > nc else []
    fb = list(letters[nc + len(fa):nc + len(fa) + nB - nc]) if nB > nc else []
    sa = list(rng.permutation(cont + fa))
    sb = list(rng.permutation(cont + fb))
    out = "".join(rng.permutation(fa + fb))
    sa, sb = "".join(sa), "".join(sb)
    pairs = ocn.inversion_pairs([sa, sb], out)
    for _ in range(8):
        charge_of = {c: int(rng.integers(-2, 3)) for c in set(sa + sb)}
        s = 1
        for x, y in pairs:
            if charge_of[x] % 2 and charge_of[y] % 2:
                s = -s
        assert s == _bubble_sign([sa, sb], out, charge_of), (sa, sb, out, charge_of)


def _legs(sym, spec, charges):
    return tuple(oc.Leg(sym, charges, f, l) for l, f in spec)


def _rand(sym, spec, charges, seed):
    return oc.SymT.random_normal(_legs(sym, spec, charges), (seed, 0))


CH = {"fp": np.array([0, 1, 1, 0, 1]), "fu1": np.array([-1, 0, 1, 1, 2, -2])}
SYMS = {"fp": FP, "fu1": FU1}


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_transpose_round_trip_and_composition(name):
    sym, ch = SYMS[name], CH[name]
    t = _rand(sym, [("a", 1), ("b", 1), ("c", -1), ("d", -1)], ch, 3)
    for perm in itertools.permutations(range(4)):
        inv = tuple(int(i) for i in np.argsort(perm))
        back = t.transpose(perm).transpose(inv)
        for k, b in t.blocks.items():
            np.testing.assert_array_equal(back.blocks[k], b)
    p1, p2 = (2, 0, 3, 1), (1, 3, 0, 2)
    comp = tuple(p1[i] for i in p2)
    a, b = t.transpose(p1).transpose(p2), t.transpose(comp)
    for k in a.blocks:
        np.testing.assert_array_equal(a.blocks[k], b.blocks[k])
    signs = {k: np.sign(t.transpose((1, 0, 2, 3)).blocks[(k[1], k[0]) + k[2:]].ravel()[0] / v.ravel()[0])
             for k, v in t.blocks.items()}
    assert any(s < 0 for s in signs.values()), "no odd-odd exchange exercised"


def test_bosonic_transpose_is_plain():
    U1 = oc.U1()
    t = _rand(U1, [("a", 1), ("b", 1), ("c", -1)], np.array([-1, 0, 1]), 5)
    np.testing.assert_array_equal(t.transpose((2, 0, 1)).todense(), np.transpose(t.todense(), (2, 0, 1)))


def _contract(A, B, out=None):
    return ocn.contract(A, B, output_labels=out)


def _assert_same(x, y):
    assert set(x.blocks) == set(y.blocks)
    for k in x.blocks:
        np.testing.assert_allclose(x.blocks[k], y.blocks[k], rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_contraction_covariant_under_transposes(name):
    sym, ch = SYMS[name], CH[name]
    A = _rand(sym, [("p", 1), ("x", 1), ("y", -1), ("q", -1)], ch, 11)
    B = _rand(sym, [("y", 1), ("r", 1), ("x", -1), ("s", -1)], ch, 12)
    C = _contract(A, B)
    assert C.labels() == ("p", "q", "r", "s")
    assert C.blocks
    out = C.labels()
    for perm in [(1, 0, 2, 3), (3, 1, 2, 0), (2, 3, 0, 1), (0, 2, 1, 3)]:
        _assert_same(_contract(A.transpose(perm), B, out), C)
        _assert_same(_contract(A, B.transpose(perm), out), C)
    for out in [("q", "p", "r", "s"), ("s", "r", "q", "p"), ("r", "p", "s", "q")]:
        perm = tuple(C.labels().index(l) for l in out)
        _assert_same(_contract(A, B, out), C.transpose(perm))
    # conservation of the result
    for k in C.blocks:
        net = sum(l.flow * q for l, q in zip(C.legs, k))
        assert (net % 2 == 0) if name == "fp" else net == 0


def test_even_charges_give_bosonic_result():
    sym = FU1
    ch = np.array([-2, 0, 2])
    A = _rand(sym, [("p", 1), ("x", 1), ("q", -1)], ch, 1)
    B = _rand(sym, [("x", -1), ("r", 1)], ch, 2)
    Ab = oc.SymT(A.blocks, _legs(oc.U1(), [("p", 1), ("x", 1), ("q", -1)], ch), False)
    Bb = oc.SymT(B.blocks, _legs(oc.U1(), [("x", -1), ("r", 1)], ch), False)
    _assert_same(_contract(A.transpose((2, 1, 0)), B, ("q", "r", "p")),
                 _contract(Ab.transpose((2, 1, 0)), Bb, ("q", "r", "p")))


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_svd_reconstruction(name):
    sym, ch = SYMS[name], CH[name]
    t = _rand(sym, [("phys", 1), ("left", 1), ("right", -1)], ch, 7)
    U, s, Vh, _, _ = ol.truncated_svd_symmetric(t, ["phys", "left"], ["right"], "bond")
    Ud = U.todense() * np.asarray(s)[None, None, :]
    rec = np.einsum("plb,br->plr", Ud, Vh.todense())
    np.testing.assert_allclose(rec, t.todense(), atol=1e-12)


def test_dagger_twist():
    sym = FU1
    t = _rand(sym, [("a", 1), ("b", -1)], np.array([-1, 0, 1]), 9)
    td = t.dagger()
    assert [l.flow for l in td.legs] == [-1, 1]
    for k, b in t.blocks.items():
        tw = (-1) ** sum(abs(q) % 2 for q in k)
        np.testing.assert_array_equal(td.blocks[tuple(-q for q in k)], b * tw)
