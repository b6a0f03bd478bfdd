"""NetworkBlueprint on the device (netfile.py:325-527; test_netfile.py:273-390).

The two-site effective-Hamiltonian network ``abc,apqd,bpse,eqtf,dfg->cstg``
with charge-conserving block-sparse U(1) operands (legs mapped onto the
template with ``label_order``), launched with and without ORDER, against the
dense einsum of the densified operands; dense operands likewise.  For three
or more symmetric operands the reference's one-shot signature matching
returns an empty tensor (contractor.py:375-403); this build contracts
pairwise, so the dense einsum is the expected value.
"""

import numpy as np
import pytest

from oracle import core as oc
from tests.helpers import from_oracle, to_oracle

pytestmark = pytest.mark.gpu

NET = """
L: a, b, c
M1: a, p, q, d
A: b, p, s, e
M2: e, q, t, f
R: d, f, g
TOUT: c, s, t, g
"""


U1 = oc.U1()
CHARGES = {"bond": [-2, -1, -1, 0, 0, 0, 1, 1, 2], "mpo": [-1, 0, 0, 1, 0], "phys": [1, -1]}
KIND = dict(a="bond", c="bond", d="bond", g="bond", b="mpo", e="mpo", f="mpo",
            p="phys", q="phys", s="phys", t="phys")
# each contracted label leaves one tensor with flow +1 and enters the other
# with flow -1, so every pairwise intermediate conserves charge and the
# block-sparse result equals the dense einsum
FLOWS = {"L": "a+ b+ c-", "M1": "a- p+ q+ d+", "A": "b- p- s+ e+", "M2": "e- q- t+ f+", "R": "d- f- g-"}


def _named(o, labels):
    """The oracle tensor with its legs renamed positionally."""
    return oc.SymT(o.blocks, tuple(oc.Leg(l.sym, l.charges, l.flow, n) for l, n in zip(o.legs, labels)),
                   validate=False)


@pytest.fixture(scope="module")
def dmrg_net():
    """Charge-conserving U(1) operands L, M1, A, M2, R of the network (random
    blocks from the oracle generator), legs named ``L0, L1, ...``."""
    ops = []
    for n, (name, spec) in enumerate(FLOWS.items()):
        legs = tuple(oc.Leg(U1, CHARGES[KIND[tok[0]]], 1 if tok[1] == "+" else -1, f"{name}{k}")
                     for k, tok in enumerate(spec.split()))
        ops.append(oc.SymT.random_normal(legs, (17, n)))
    return tuple(ops)


def _expected(ops):
    dense = [o.todense() for o in ops]
    return np.einsum("abc,apqd,bpse,eqtf,dfg->cstg", *dense)


def _put(bp, ops):
    for name, o, lab in zip(["L", "M1", "A", "M2", "R"], ops,
                            [["a", "b", "c"], ["a", "p", "q", "d"], ["b", "p", "s", "e"], ["e", "q", "t", "f"],
                             ["d", "f", "g"]]):
        t = from_oracle(o)
        bp.put_tensor(name, t, label_order=list(t.labels()))
        assert list(bp._tensors[name].labels()) == lab


@pytest.mark.parametrize("order", [None, "((((L,M1),A),M2),R)", "(((L,M1),(A,M2)),R)"])
def test_dmrg_network_symmetric(gpu, dmrg_net, order):
    from tenax_amd import NetworkBlueprint

    bp = NetworkBlueprint(NET + (f"ORDER: {order}\n" if order else ""))
    _put(bp, dmrg_net)
    res = bp.launch()
    assert [str(l) for l in res.labels()] == ["c", "s", "t", "g"]
    got = to_oracle(res).todense()
    ref = _expected(dmrg_net)
    assert np.abs(ref).max() > 0
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_dmrg_network_dense(gpu):
    from tenax_amd import DenseTensor, FlowDirection, NetworkBlueprint, TensorIndex, U1Symmetry

    rng = np.random.default_rng(40)
    sym = U1Symmetry()

    def dense(shape, labels):
        idx = tuple(TensorIndex(sym, np.zeros(n, np.int32), FlowDirection.IN, l) for n, l in zip(shape, labels))
        return DenseTensor(rng.standard_normal(shape), idx)

    ops = [dense((2, 3, 2), "abc"), dense((2, 2, 2, 3), "apqd"), dense((3, 2, 2, 2), "bpse"),
           dense((2, 2, 2, 3), "eqtf"), dense((3, 3, 2), "dfg")]
    bp = NetworkBlueprint(NET)
    bp.put_tensors(dict(zip(["L", "M1", "A", "M2", "R"], ops)))
    res = bp.launch().todense().cpu().numpy()
    ref = np.einsum("abc,apqd,bpse,eqtf,dfg->cstg", *[o.todense().cpu().numpy() for o in ops])
    assert np.abs(res - ref).max() < 1e-12


def test_put_tensor_transposes_to_blueprint_order(gpu, dmrg_net):
    """A tensor carrying the blueprint's label set in another order is transposed."""
    from tenax_amd import NetworkBlueprint

    oL = _named(dmrg_net[0], ["a", "b", "c"])
    t = from_oracle(oL)
    bp = NetworkBlueprint("L: c, a, b\nX: a, b, c\nTOUT:")
    bp.put_tensor("L", t)
    assert [str(l) for l in bp._tensors["L"].labels()] == ["c", "a", "b"]
    assert np.allclose(to_oracle(bp._tensors["L"]).todense(), np.transpose(oL.todense(), (2, 0, 1)))
    with pytest.raises(ValueError, match="don't match blueprint"):
        bp.put_tensor("X", t.relabel("c", "z"))
    with pytest.raises(ValueError, match="rank"):
        bp.put_tensor("X", from_oracle(dmrg_net[1]))


def test_greedy_order_matches_left_fold(gpu, dmrg_net):
    from tenax_amd import contract_with_subscripts
    from tenax_amd.contraction.contractor import _labels_to_subscripts

    ts = [from_oracle(_named(o, lab)) for o, lab in zip(dmrg_net, ["abc", "apqd", "bpse", "eqtf", "dfg"])]
    subs, out_idx = _labels_to_subscripts(ts, list("cstg"))
    g = contract_with_subscripts(ts, subs, out_idx, "auto")
    l = contract_with_subscripts(ts, subs, out_idx, "left")
    a, b = to_oracle(g).todense(), to_oracle(l).todense()
    assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
