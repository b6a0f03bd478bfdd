"""Fermionic sector signs on the device against the oracle (tensor.py:32-50,
783-816, 853-856; contractor.py:206-285, 462-470; linalg.py:171-186).

The transpose plan applies each block's Koszul sign; the pairwise planner
evaluates the reference's inversion-pair sign of every block pair and the
grouped GEMM applies it; dagger multiplies by the twist phases; the SVD
assembles sector matrices with the Koszul sign of the (left, right) order.
"""

import itertools

import numpy as np
import pytest

from oracle import contract as ocn
from oracle import core as oc
from oracle import linalg as ol
from tests.helpers import assert_same_structure, from_oracle, max_rel_err, to_oracle

pytestmark = pytest.mark.gpu

SYMS = {"fp": oc.FermionParity(), "fu1": oc.FermionicU1()}
CH = {"fp": np.array([0, 1, 1, 0, 1, 1]), "fu1": np.array([-1, 0, 1, 1, 2, -2, -1])}


def _rand(name, spec, seed):
    legs = tuple(oc.Leg(SYMS[name], CH[name], f, l) for l, f in spec)
    return oc.SymT.random_normal(legs, (seed, 0))


def _close(p, o, tol=1e-13):
    assert_same_structure(p, o)
    assert max_rel_err(p, o) < tol


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_transpose_signs(gpu, name):
    o = _rand(name, [("a", 1), ("b", 1), ("c", -1), ("d", -1)], 1)
    t = from_oracle(o)
    for perm in itertools.permutations(range(4)):
        p = t.transpose(perm)
        ref = o.transpose(perm)
        assert p.labels() == ref.labels()
        _close(p, ref, 0.0 + 1e-300)


@pytest.mark.parametrize("name", ["fp", "fu1"])
@pytest.mark.parametrize("case", [
    ("p x y q", "y r x s", None),
    ("p x y q", "y r x s", "s r q p"),
    ("x p y q", "r y s x", "q r p s"),
    ("x y", "y x", ""),
    ("p x", "x q", None),
    ("x p q", "x r", "r q p"),
])
def test_contract_signs(gpu, name, case):
    from tenax_amd import contract

    la, lb, out = case
    la, lb = la.split(), lb.split()
    cont = set(la) & set(lb)
    flows_a = {l: (1 if i % 2 == 0 else -1) for i, l in enumerate(la)}
    spec_a = [(l, flows_a[l]) for l in la]
    spec_b = [(l, -flows_a[l] if l in cont else (1 if i % 2 else -1)) for i, l in enumerate(lb)]
    oA, oB = _rand(name, spec_a, 3), _rand(name, spec_b, 4)
    outl = None if out is None else (out.split() if out else [])
    ref = ocn.contract(oA, oB, output_labels=outl)
    got = contract(from_oracle(oA), from_oracle(oB), output_labels=outl)
    if not ref.legs:
        want = sum(float(b.sum()) for b in ref.blocks.values())
        assert abs(float(got.data.sum()) - want) <= 1e-12 * max(1.0, abs(want))
        return
    assert tuple(got.labels()) == ref.labels()
    _close(got, ref)


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_contract_covariant_on_device(gpu, name):
    """contract(A.transpose(P), B) == contract(A, B) with the Koszul-signed transpose."""
    from tenax_amd import contract

    oA = _rand(name, [("p", 1), ("x", 1), ("y", -1), ("q", -1)], 11)
    oB = _rand(name, [("y", 1), ("r", 1), ("x", -1), ("s", -1)], 12)
    A, B = from_oracle(oA), from_oracle(oB)
    C = to_oracle(contract(A, B))
    for perm in [(3, 1, 2, 0), (2, 3, 0, 1), (1, 0, 3, 2)]:
        D = to_oracle(contract(A.transpose(perm), B, output_labels=["p", "q", "r", "s"]))
        for k in C.blocks:
            np.testing.assert_allclose(D.blocks[k], C.blocks[k], atol=1e-12)


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_three_operands_left_fold(gpu, name):
    from tenax_amd import contract

    oA = _rand(name, [("p", 1), ("x", 1)], 21)
    oB = _rand(name, [("x", -1), ("q", 1), ("y", 1)], 22)
    oC = _rand(name, [("y", -1), ("r", 1)], 23)
    ref = ocn.contract(ocn.contract(oA, oB), oC)
    got = contract(from_oracle(oA), from_oracle(oB), from_oracle(oC))
    assert tuple(got.labels()) == ref.labels()
    _close(got, ref)


def test_dagger_twist(gpu):
    o = _rand("fu1", [("a", 1), ("b", -1), ("c", 1)], 31)
    _close(from_oracle(o).dagger(), o.dagger(), 1e-300)


@pytest.mark.parametrize("name", ["fp", "fu1"])
def test_svd_koszul_assembly(gpu, name):
    """Left legs not leading: the sector matrices carry the Koszul sign of the
    reordering, so the singular values differ from the unsigned ones."""
    from tenax_amd.linalg import truncated_svd

    o = _rand(name, [("a", 1), ("b", 1), ("c", -1)], 41)
    t = from_oracle(o)
    U, s, Vh, _ = truncated_svd(t, ["c", "a"], ["b"], "bond")
    oU, os_, oVh, _, _ = ol.truncated_svd_symmetric(o, ["c", "a"], ["b"], "bond")
    np.testing.assert_allclose(np.sort(s.cpu().numpy()), np.sort(np.asarray(os_)), rtol=1e-12, atol=1e-13)
    # leading left legs reconstruct the tensor (test_fermionic.py:532-555)
    U, s, Vh, _ = truncated_svd(t, ["a", "b"], ["c"], "bond")
    Ud = U.todense() * s[None, None, :]
    rec = (Ud.reshape(-1, Ud.shape[-1]) @ Vh.todense()).reshape(o.todense().shape)
    np.testing.assert_allclose(rec.cpu().numpy(), o.todense(), atol=1e-12)


def test_fermionic_autograd_rejected(gpu):
    import torch

    from tenax_amd import contract

    oA = _rand("fp", [("p", 1), ("x", 1)], 51)
    oB = _rand("fp", [("x", -1), ("q", 1)], 52)
    A = from_oracle(oA)
    A = A._with_data(A.data.clone().requires_grad_(True))
    with pytest.raises(NotImplementedError, match="fermionic"):
        contract(A, from_oracle(oB))
    assert torch is not None
