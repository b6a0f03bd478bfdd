"""GPU parity of the device planner and the grouped block GEMM against the oracle."""

import numpy as np
import pytest
import torch

from oracle import contract as ocn
from oracle import core as oc
from tests.helpers import assert_same_structure, from_oracle, max_rel_err, oracle_legs, product_indices, to_oracle

pytestmark = pytest.mark.gpu


def _legs(sym, rng, rank, maxq=2, dim=6):
    out = []
    for i in range(rank):
        if sym.n is None:
            ch = rng.integers(-maxq, maxq + 1, size=dim)
        else:
            ch = rng.integers(0, sym.n, size=dim)
        out.append(oc.Leg(sym, ch, oc.IN if rng.random() < 0.5 else oc.OUT, f"l{i}"))
    return tuple(out)


@pytest.mark.parametrize("sym", [oc.U1(), oc.Zn(2), oc.Zn(3)])
@pytest.mark.parametrize("rank", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("target", [None, 1])
def test_valid_blocks_device_matches_oracle(gpu, sym, rank, target):
    from tenax_amd.core.tensor import valid_blocks

    rng = np.random.default_rng(100 * rank + (sym.n or 0))
    legs = _legs(sym, rng, rank)
    keys, shapes = valid_blocks(product_indices(legs), target)
    ref = sorted(oc.compute_valid_blocks(legs, target))
    assert [tuple(int(q) for q in k) for k in keys] == ref
    for k, s in zip(keys, shapes):
        _, shp = oc.block_slices(legs, tuple(k))
        assert tuple(int(x) for x in s) == shp


def c2_legs(chi):
    from tenax_amd.workloads import PHYS_CHARGES, gaussian_bond_charges

    bond = gaussian_bond_charges(chi)
    U1 = oc.U1()
    A = (oc.Leg(U1, bond, oc.IN, "l"), oc.Leg(U1, PHYS_CHARGES, oc.IN, "p"), oc.Leg(U1, PHYS_CHARGES, oc.IN, "pp"),
         oc.Leg(U1, bond, oc.OUT, "r"))
    B = (oc.Leg(U1, bond, oc.IN, "r"), oc.Leg(U1, PHYS_CHARGES, oc.IN, "q"), oc.Leg(U1, PHYS_CHARGES, oc.IN, "qq"),
         oc.Leg(U1, bond, oc.OUT, "s"))
    return A, B


@pytest.fixture(scope="module")
def c2_pair():
    A, B = c2_legs(96)
    oA = oc.SymT.random_normal(A, (1, 0))
    oB = oc.SymT.random_normal(B, (1, 1))
    oC, _ = ocn.contract_symmetric([oA, oB], "lpmr,rqns->lpmqns")
    return oA, oB, oC


def test_contract_c2_f64_matches_oracle(gpu, c2_pair):
    from tenax_amd import contract

    oA, oB, oC = c2_pair
    C = contract(from_oracle(oA), from_oracle(oB))
    assert C.labels() == ("l", "p", "pp", "q", "qq", "s")
    assert_same_structure(C, oC)
    assert max_rel_err(C, oC) < 1e-13


@pytest.mark.parametrize("mode,tol", [("f32", 2e-5), ("split_f32", 2e-6)])
def test_contract_c2_reduced_precision(gpu, c2_pair, mode, tol):
    from tenax_amd import compute_mode, contract

    oA, oB, oC = c2_pair
    with compute_mode(mode):
        C = contract(from_oracle(oA), from_oracle(oB))
    assert_same_structure(C, oC)
    err = max_rel_err(C, oC)
    assert err < tol, err


def test_perblock_path_matches_grouped(gpu, c2_pair):
    from tenax_amd import contract, per_block_execution

    oA, oB, oC = c2_pair
    A, B = from_oracle(oA), from_oracle(oB)
    C1 = contract(A, B)
    with per_block_execution():
        C2 = contract(A, B)
    assert np.array_equal(C1.keys, C2.keys)
    d = (C1.data - C2.data).abs().max().item()
    assert d <= 1e-12 * C1.data.abs().max().item()


TILINGS = [(sh, kc, xc) for sh in (0, 1, 2, 3) for kc in (0, 1, 3, -1, -2) for xc in (0, 1)]


def _last_pair_plan():
    from tenax_amd.contraction import plans

    return [v for k, v in plans._cache.items() if k[0] == "pair"][-1][0]


@pytest.mark.parametrize("mode,tol", [("f64", 1e-13), ("f32", 2e-5), ("split_f32", 2e-6)])
def test_contract_tilings_agree(gpu, c2_pair, mode, tol):
    """Every tiling of the grouped GEMM (tile shape, split K -- including one K
    tile per task, which splits across pair boundaries -- and XCD order) gives
    the oracle's result; FP64 tilings agree with each other to rounding."""
    from tenax_amd import compute_mode, contract

    oA, oB, oC = c2_pair
    A, B = from_oracle(oA), from_oracle(oB)
    with compute_mode(mode):
        C0 = contract(A, B)
        plan = _last_pair_plan()
        for cfg in TILINGS:
            plan.set_tiling(0, *cfg)
            C = contract(A, B)
            assert plan.tiling(0)["shape"] == cfg[0]
            assert_same_structure(C, oC)
            assert max_rel_err(C, oC) < tol, cfg
            if mode == "f64":
                d = (C.data - C0.data).abs().max().item()
                assert d <= 1e-13 * C0.data.abs().max().item(), cfg
        plan.set_tiling(0, 0, 0, 0)
    with pytest.raises(Exception, match="tile shape"):
        plan.set_tiling(0, 4, 0, 0)


def test_balanced_split_schedule_matches_default(gpu):
    """The balanced two-wave split (kchunk = -2) on blocks whose K is much
    longer than the mean work per workgroup slot: a head of each long tile
    in the first wave, tail pieces after it, partials summed in K order;
    every tile shape and both task orders against the whole-K result."""
    from tenax_amd import contract

    U1 = oc.U1()
    rng = np.random.default_rng(8)
    k = np.repeat(np.array([0, 1], np.int32), [1500, 700])  # long K blocks
    A = (oc.Leg(U1, np.array([0] * 70 + [1] * 30, np.int32), oc.IN, "a"), oc.Leg(U1, k, oc.OUT, "k"))
    B = (oc.Leg(U1, k, oc.IN, "k"), oc.Leg(U1, np.array([0] * 90 + [1] * 20, np.int32), oc.OUT, "b"))
    oA, oB = oc.SymT.random_normal(A, (8, 1)), oc.SymT.random_normal(B, (8, 2))
    Ad, Bd = from_oracle(oA), from_oracle(oB)
    C0 = contract(Ad, Bd)
    plan = _last_pair_plan()
    ref = np.einsum("ak,kb->ab", oA.todense(), oB.todense())
    for sh in (0, 1, 2, 3):
        for xc in (0, 1):
            plan.set_tiling(0, sh, -2, xc)
            C = contract(Ad, Bd)
            np.testing.assert_allclose(C.todense().cpu().numpy(), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
            d = (C.data - C0.data).abs().max().item()
            assert d <= 1e-13 * C0.data.abs().max().item(), (sh, xc)
    plan.set_tiling(0, 0, 0, 0)


def test_contract_needs_transposes_and_swap(gpu):
    """Operands whose legs interleave: device transposes + swapped roles."""
    from tenax_amd import contract

    U1 = oc.U1()
    rng = np.random.default_rng(5)
    b1 = rng.integers(-2, 3, 7)
    b2 = rng.integers(-1, 2, 5)
    x = rng.integers(-1, 2, 4)
    A = (oc.Leg(U1, b1, oc.IN, "i"), oc.Leg(U1, x, oc.OUT, "x"), oc.Leg(U1, b2, oc.IN, "j"),
         oc.Leg(U1, rng.integers(-2, 3, 6), oc.OUT, "k"))
    B = (oc.Leg(U1, b2, oc.OUT, "j"), oc.Leg(U1, rng.integers(-1, 2, 3), oc.IN, "y"), oc.Leg(U1, b1, oc.OUT, "i"),
         oc.Leg(U1, rng.integers(-1, 2, 5), oc.IN, "z"))
    oA = oc.SymT.random_normal(A, (9,))
    oB = oc.SymT.random_normal(B, (10,))
    for out in [None, ("z", "x", "y", "k"), ("x", "z", "k", "y")]:
        C = contract(from_oracle(oA), from_oracle(oB), output_labels=out)
        lab = C.labels()
        sub = "ixjk,jyiz->" + "".join({"x": "x", "k": "k", "y": "y", "z": "z"}[l] for l in lab)
        ref = np.einsum(sub, oA.todense(), oB.todense())
        np.testing.assert_allclose(C.todense().cpu().numpy(), ref, atol=1e-12)


def test_contract_many_equals_one_by_one(gpu):
    """contract_many (every instance's operand permutations, block GEMMs and
    output permutations in shared launches) returns, instance by instance,
    exactly what contract() returns: mixed structures (U(1) and Z2, operands
    needing transposes on both sides or none, K-first A layouts, requested
    output orders), a dense pair and a fermionic pair taking the one-by-one
    path, and an instance with no matching blocks."""
    from tenax_amd import contract
    from tenax_amd.contraction import contract_many
    from tenax_amd.core.tensor import DenseTensor

    rng = np.random.default_rng(11)
    U1, Z2 = oc.U1(), oc.Zn(2)
    insts, outs = [], []
    for k in range(6):
        sym = U1 if k % 2 == 0 else Z2
        ch = (lambda n: rng.integers(-2, 3, n)) if sym is U1 else (lambda n: rng.integers(0, 2, n))
        bi, bj = ch(7 + k), ch(5 + k)
        A = (oc.Leg(sym, bi, oc.IN, "i"), oc.Leg(sym, ch(4), oc.OUT, "x"), oc.Leg(sym, bj, oc.IN, "j"),
             oc.Leg(sym, ch(6), oc.OUT, "k"))
        B = (oc.Leg(sym, bj, oc.OUT, "j"), oc.Leg(sym, ch(3), oc.IN, "y"), oc.Leg(sym, bi, oc.OUT, "i"),
             oc.Leg(sym, ch(5), oc.IN, "z"))
        if k == 3:  # A already [contracted..., free...]: K-first layout, no transpose
            A = (A[0], A[2], A[1], A[3])
        insts.append((from_oracle(oc.SymT.random_normal(A, (20 + k,))),
                      from_oracle(oc.SymT.random_normal(B, (30 + k,)))))
    # no matching charges at all: an empty result
    e1 = (oc.Leg(U1, np.array([5, 6]), oc.IN, "i"), oc.Leg(U1, np.array([0]), oc.OUT, "x"))
    e2 = (oc.Leg(U1, np.array([-1, 1]), oc.OUT, "i"), oc.Leg(U1, np.array([0]), oc.IN, "y"))
    insts.append((from_oracle(oc.SymT.random_normal(e1, (1,), target=None)),
                  from_oracle(oc.SymT.random_normal(e2, (2,), target=None))))
    for order in [None, ("z", "x", "y", "k")]:
        got = contract_many(insts[:-1], output_labels=order) + contract_many(insts[-1:])
        for (A, B), C in zip(insts[:-1], got[:-1]):
            ref = contract(A, B, output_labels=order)
            assert C.labels() == ref.labels()
            assert np.array_equal(C.keys, ref.keys) and np.array_equal(C.shapes, ref.shapes)
            assert torch.equal(C.data, ref.data)
        ref = contract(*insts[-1])
        assert got[-1].n_blocks == ref.n_blocks == 0
    # a dense pair and a fermionic pair go through contract() one by one
    d1 = torch.tensor(rng.standard_normal((3, 4)), device="cuda")
    d2 = torch.tensor(rng.standard_normal((4, 5)), device="cuda")
    from tenax_amd.core.index import FlowDirection, TensorIndex
    from tenax_amd.core.symmetry import U1Symmetry

    ix = lambda n, l: TensorIndex(U1Symmetry(), np.zeros(n, np.int32), FlowDirection.IN, l)
    D1, D2 = DenseTensor(d1, (ix(3, "a"), ix(4, "b"))), DenseTensor(d2, (ix(4, "b"), ix(5, "c")))
    FP = oc.FermionParity()
    bf = rng.integers(0, 2, 6)
    F1 = (oc.Leg(FP, rng.integers(0, 2, 4), oc.IN, "p"), oc.Leg(FP, bf, oc.OUT, "q"))
    F2 = (oc.Leg(FP, bf, oc.IN, "q"), oc.Leg(FP, rng.integers(0, 2, 3), oc.OUT, "r"))
    fA = from_oracle(oc.SymT.random_normal(F1, (40,), target=None))
    fB = from_oracle(oc.SymT.random_normal(F2, (41,), target=None))
    got = contract_many([(D1, D2), insts[0], (fA, fB)])
    assert torch.allclose(got[0].todense(), d1 @ d2, rtol=0, atol=1e-13)
    assert torch.equal(got[1].data, contract(*insts[0]).data)
    assert torch.equal(got[2].data, contract(fA, fB).data)


def test_full_contraction_matches_dense(gpu):
    """Every leg contracted (a scalar): the same block structure on both sides
    (one packed dot), different block sets and leg orders (common blocks
    gathered), a fermionic pair against the oracle's signed contraction, and
    the gradient of the scalar."""
    from tenax_amd import contract

    U1 = oc.U1()
    rng = np.random.default_rng(21)
    a, b, c = rng.integers(-2, 3, 6), rng.integers(-1, 2, 5), rng.integers(-2, 3, 7)
    L1 = (oc.Leg(U1, a, oc.IN, "i"), oc.Leg(U1, b, oc.OUT, "j"), oc.Leg(U1, c, oc.OUT, "k"))
    L2 = (oc.Leg(U1, a, oc.OUT, "i"), oc.Leg(U1, b, oc.IN, "j"), oc.Leg(U1, c, oc.IN, "k"))
    oA = oc.SymT.random_normal(L1, (3,))
    oB = oc.SymT.random_normal(L2, (4,))
    ref = float(np.einsum("ijk,ijk->", oA.todense(), oB.todense()))
    A, B = from_oracle(oA), from_oracle(oB)
    assert abs(float(contract(A, B).data[0]) - ref) <= 1e-12 * max(1.0, abs(ref))
    # B with its legs permuted and two of its blocks absent (common blocks gathered)
    oC = oc.SymT.random_normal((L2[2], L2[0], L2[1]), (5,))
    for key in sorted(oC.blocks)[1:3]:
        del oC.blocks[key]
    refc = float(np.einsum("ijk,kij->", oA.todense(), oC.todense()))
    assert abs(float(contract(A, from_oracle(oC)).data[0]) - refc) <= 1e-12 * max(1.0, abs(refc))
    # gradient of <A, B> with respect to A is B (same structure)
    Ag = A._with_data(A.data.clone().requires_grad_(True))
    contract(Ag, B).data.sum().backward()
    assert torch.equal(Ag.data.grad, B.data)
    # fermionic: the oracle's signed full contraction
    FP = oc.FermionParity()
    p, q = rng.integers(0, 2, 4), rng.integers(0, 2, 5)
    F1 = (oc.Leg(FP, p, oc.IN, "p"), oc.Leg(FP, q, oc.OUT, "q"))
    F2 = (oc.Leg(FP, q, oc.IN, "q"), oc.Leg(FP, p, oc.OUT, "p"))
    oF1, oF2 = oc.SymT.random_normal(F1, (6,)), oc.SymT.random_normal(F2, (7,))
    oS, _ = ocn.contract_symmetric([oF1, oF2], "pq,qp->")
    want = sum(float(np.sum(v)) for v in oS.blocks.values())
    got = float(contract(from_oracle(oF1), from_oracle(oF2)).data[0])
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("n", [2, 3])
def test_contract_zn(gpu, n):
    from tenax_amd import contract

    Z = oc.Zn(n)
    rng = np.random.default_rng(n)
    a = rng.integers(0, n, 6)
    b = rng.integers(0, n, 5)
    A = (oc.Leg(Z, a, oc.IN, "a"), oc.Leg(Z, rng.integers(0, n, 4), oc.IN, "p"), oc.Leg(Z, b, oc.OUT, "b"))
    B = (oc.Leg(Z, b, oc.IN, "b"), oc.Leg(Z, rng.integers(0, n, 3), oc.OUT, "c"))
    oA = oc.SymT.random_normal(A, (3,))
    oB = oc.SymT.random_normal(B, (4,))
    oC, _ = ocn.contract_symmetric([oA, oB], "apb,bc->apc")
    C = contract(from_oracle(oA), from_oracle(oB))
    assert_same_structure(C, oC)
    assert max_rel_err(C, oC) < 1e-13


def test_three_operand_contract_matches_dense(gpu):
    from tenax_amd import contract

    U1 = oc.U1()
    rng = np.random.default_rng(11)
    c1, c2, c3 = (rng.integers(-2, 3, 5) for _ in range(3))
    A = (oc.Leg(U1, c1, oc.IN, "a"), oc.Leg(U1, c2, oc.OUT, "b"))
    B = (oc.Leg(U1, c2, oc.IN, "b"), oc.Leg(U1, c3, oc.OUT, "c"))
    Cc = (oc.Leg(U1, c3, oc.IN, "c"), oc.Leg(U1, c1, oc.OUT, "d"))
    ts = [oc.SymT.random_normal(l, (i,)) for i, l in enumerate((A, B, Cc))]
    R = contract(*[from_oracle(t) for t in ts])
    ref = np.einsum("ab,bc,cd->ad", *[t.todense() for t in ts])
    np.testing.assert_allclose(R.todense().cpu().numpy(), ref, atol=1e-12)


def test_transpose_matches_oracle(gpu, c2_pair):
    oA, _, _ = c2_pair
    A = from_oracle(oA)
    for axes in [(3, 0, 1, 2), (1, 3, 0, 2), (0, 1, 2, 3)]:
        T = A.transpose(axes)
        oT = oA.transpose(axes)
        assert_same_structure(T, oT)
        assert max_rel_err(T, oT) == 0.0


@pytest.mark.parametrize("rank", [3, 4, 5])
def test_transpose_all_permutations_short_sectors(gpu, rank):
    """Every leg permutation of a Z2 tensor whose sectors are 3-25 wide (the
    shapes of chi <= 48 HOTRG tensors: tiles combine several short
    dimensions, edge tiles are partial, some blocks take the gather path)
    against the oracle's block transpose, bit for bit."""
    import itertools

    Z2 = oc.Zn(2)
    rng = np.random.default_rng(rank)
    dims = [(25, 23), (24, 24), (3, 8), (13, 12), (7, 2)][:rank]
    legs = tuple(oc.Leg(Z2, np.repeat(np.array([0, 1], np.int32), d), oc.IN if i % 2 else oc.OUT, f"l{i}")
                 for i, d in enumerate(dims))
    oA = oc.SymT.random_normal(legs, (rank, 7), target=0)
    A = from_oracle(oA)
    for axes in itertools.permutations(range(rank)):
        T = A.transpose(axes)
        oT = oA.transpose(axes)
        assert_same_structure(T, oT)
        assert max_rel_err(T, oT) == 0.0, axes


def test_from_dense_roundtrip_and_error(gpu):
    from tenax_amd import SymmetricTensor

    U1 = oc.U1()
    legs = (oc.Leg(U1, [-1, 0, 1, 0], oc.IN, "a"), oc.Leg(U1, [1, -1], oc.IN, "p"),
            oc.Leg(U1, [0, 1, -1, 2, 0], oc.OUT, "b"))
    o = oc.SymT.random_normal(legs, (2,))
    d = o.todense()
    T = SymmetricTensor.from_dense(d, product_indices(legs))
    assert_same_structure(T, o)
    np.testing.assert_array_equal(T.todense().cpu().numpy(), d)
    bad = d.copy()
    bad[0, 1, 0] = 1.0  # (q=-1, p=-1, q=0): -1 - 1 - 0 != 0 violates conservation
    with pytest.raises(ValueError, match="outside symmetry-allowed"):
        SymmetricTensor.from_dense(bad, product_indices(legs))


def test_zeros_and_random_normal_structure(gpu):
    from tenax_amd import PRNGKey, SymmetricTensor

    legs = c2_legs(48)[0]
    idx = product_indices(legs)
    Z = SymmetricTensor.zeros(idx)
    R = SymmetricTensor.random_normal(idx, PRNGKey(1))
    o = oc.SymT.random_normal(legs, (1,))
    assert_same_structure(Z, o)
    assert_same_structure(R, o)
    # identical Threefry words (tests/test_gpu_random.py); the normals differ
    # only through the device erfinv vs scipy's, within a few ulp
    assert max_rel_err(R, o) < 4 * np.finfo(float).eps
    assert float(Z.norm()) == 0.0


@pytest.mark.parametrize("k_first", [False, True])
@pytest.mark.parametrize("sym", [oc.U1(), oc.Zn(3)])
def test_contract_vjp_matches_dense_chain_rule(gpu, k_first, sym):
    """tnx_contract_vjp (through autograd of contract()): dL/dA and dL/dB of
    <W, A . B> against the dense chain rule, for A with its contracted legs
    last (M x K blocks) or first (K x M blocks), and only on stored blocks."""
    from tenax_amd import contract

    rng = np.random.default_rng(11 if k_first else 12)

    def leg(lbl, flow, n):
        ch = rng.integers(-2, 3, size=n) if sym.n is None else rng.integers(0, 3, size=n)
        return oc.Leg(sym, ch, flow, lbl)

    k1, k2 = leg("k", oc.OUT, 9), leg("j", oc.IN, 4)
    a, p = leg("a", oc.IN, 7), leg("p", oc.OUT, 3)
    A_legs = (k1, k2, a, p) if k_first else (a, p, k1, k2)
    B_legs = (k1.flip_flow(), leg("q", oc.IN, 5), k2.flip_flow(), leg("b", oc.OUT, 6))
    Ao, Bo = oc.SymT.random_normal(A_legs, (4, 1)), oc.SymT.random_normal(B_legs, (4, 2))
    A, B = from_oracle(Ao), from_oracle(Bo)
    a_leaf = A.data.detach().clone().requires_grad_(True)
    b_leaf = B.data.detach().clone().requires_grad_(True)
    Cp = contract(A._with_data(a_leaf), B._with_data(b_leaf))
    W = torch.as_tensor(rng.standard_normal(Cp.data.numel())).cuda()
    ga, gb = torch.autograd.grad(torch.dot(Cp.data, W), [a_leaf, b_leaf])
    Wd = to_oracle(Cp._with_data(W)).todense()
    Ad, Bd = Ao.todense(), Bo.todense()
    la, lb, lc = "".join(l.label for l in A_legs), "".join(l.label for l in B_legs), "".join(Cp.labels())
    ref_a = np.einsum(f"{lc},{lb}->{la}", Wd, Bd)
    ref_b = np.einsum(f"{lc},{la}->{lb}", Wd, Ad)
    got_a = to_oracle(A._with_data(ga)).todense()
    got_b = to_oracle(B._with_data(gb)).todense()
    # compare on stored blocks (entries outside every block have no gradient)
    ma = np.zeros(Ad.shape, bool)
    for k in Ao.blocks:
        ma[np.ix_(*[np.nonzero(l.charges == q)[0] for l, q in zip(A_legs, k)])] = True
    mb = np.zeros(Bd.shape, bool)
    for k in Bo.blocks:
        mb[np.ix_(*[np.nonzero(l.charges == q)[0] for l, q in zip(B_legs, k)])] = True
    scale = max(1.0, np.abs(ref_a).max(), np.abs(ref_b).max())
    assert np.abs((got_a - ref_a)[ma]).max() < 1e-12 * scale
    assert np.abs((got_b - ref_b)[mb]).max() < 1e-12 * scale
    assert np.all(got_a[~ma] == 0) and np.all(got_b[~mb] == 0)


def test_split_over_workspace_cap_is_refused_and_skipped(gpu, monkeypatch):
    """A split-K tiling whose partial tiles exceed the workspace cap (the
    balanced split of a 2048^3 one-block contraction in 128x128 tiles) raises
    TNX_ERR_CAPACITY, leaves the plan on its default tiling (which still
    contracts correctly), and is skipped as an autotuning candidate."""
    from tenax_amd import SymmetricTensor, _lib, contract
    from tenax_amd.contraction import autotune, contractor, plans
    from tenax_amd.core.index import FlowDirection, TensorIndex
    from tenax_amd.core.symmetry import U1Symmetry

    U1 = U1Symmetry()
    n = 2048
    z = np.zeros(n, np.int32)
    A = SymmetricTensor.random_normal((TensorIndex(U1, z, FlowDirection.IN, "l"),
                                       TensorIndex(U1, z, FlowDirection.OUT, "r")), (1, 0))
    B = SymmetricTensor.random_normal((TensorIndex(U1, z, FlowDirection.IN, "r"),
                                       TensorIndex(U1, z, FlowDirection.OUT, "s")), (1, 1))
    C = contract(A, B)
    ref = (A.data.view(n, n) @ B.data.view(n, n)).reshape(-1)
    plan = plans.pairwise_plan(A, B, 0, 1, 1, C.indices, 0, None)[0]
    with pytest.raises(_lib.TnxError, match="workspace exceeds its cap") as err:
        plan.set_tiling(0, 3, -2, 0)
    assert err.value.code == _lib.TNX_ERR_CAPACITY
    assert plan.tiling(0)["shape"] == 0 and plan.tiling(0)["kchunk"] == 0
    y = torch.empty_like(C.data)
    contractor._execute_pair(plan, A, B, y)
    assert float((y - ref).abs().max() / ref.abs().max()) <= 1e-13
    autotune.clear()
    monkeypatch.setattr(autotune, "CANDIDATES", [(3, -2, 0), (0, 0, 1)])
    cfg = autotune.tune(plan, 0, lambda: contractor._execute_pair(plan, A, B, y), plan.stats()["flops"], "test")
    assert cfg == (0, 0, 1)
    contractor._execute_pair(plan, A, B, y)
    assert float((y - ref).abs().max() / ref.abs().max()) <= 1e-13
    autotune.clear()


FUSED_ORDERS = [("s", "qq", "q", "pp", "p", "l"), ("q", "l", "qq", "p", "s", "pp"), ("p", "l", "s", "pp", "q", "qq"),
                ("l", "q", "qq", "s", "p", "pp")]


def _natural(fn):
    from tenax_amd.contraction import contractor

    old = contractor.FUSED_OUTPUT_ORDER
    contractor.FUSED_OUTPUT_ORDER = False
    try:
        return fn()
    finally:
        contractor.FUSED_OUTPUT_ORDER = old


@pytest.mark.parametrize("order", FUSED_ORDERS)
def test_fused_output_order_matches_oracle_and_transpose(gpu, c2_pair, order, monkeypatch):
    """contract(..., output_labels=order) with the GEMM writing the requested
    leg order itself (permuted-store plan): the oracle's blocks for that order,
    and bit for bit the natural result transposed, under every tiling (each
    work list checked against the operand and output block layouts)."""
    from tenax_amd import contract
    from tenax_amd.contraction import contractor, plans

    monkeypatch.setenv("TNX_VALIDATE", "1")
    plans.clear_plan_cache()
    oA, oB, _ = c2_pair
    A, B = from_oracle(oA), from_oracle(oB)
    sub = {"l": "l", "p": "p", "pp": "m", "q": "q", "qq": "n", "s": "s"}
    oC, _ = ocn.contract_symmetric([oA, oB], "lpmr,rqns->" + "".join(sub[x] for x in order))
    assert contractor.FUSED_OUTPUT_ORDER
    C = contract(A, B, output_labels=order)
    plan = _last_pair_plan()
    assert C.labels() == order
    assert_same_structure(C, oC)
    assert max_rel_err(C, oC) < 1e-13
    R = _natural(lambda: contract(A, B, output_labels=order))
    assert np.array_equal(C.keys, R.keys) and np.array_equal(C.shapes, R.shapes)
    assert np.array_equal(C.offsets, R.offsets)
    assert torch.equal(C.data, R.data)
    for cfg in [(sh, kc, xc) for sh in (0, 1, 2, 3) for kc in (0, 1, -1, -2) for xc in (0, 1)]:
        try:
            plan.set_tiling(0, *cfg)
        except Exception as e:  # a split over the workspace cap is refused, not run
            assert "workspace" in str(e), cfg
            continue
        C2 = contract(A, B, output_labels=order)
        d = (C2.data - R.data).abs().max().item()
        assert d <= 1e-13 * R.data.abs().max().item(), cfg
    plan.set_tiling(0, 0, 0, 0)


def test_fused_output_order_in_many_instance_launches(gpu, c2_pair):
    """contract_many with an output order: the instances' permuted-store plans
    merged into one launch, equal bit for bit to transposing each natural
    result; mixing a natural-order instance into the batch keeps it plain."""
    from tenax_amd.contraction import contract_many

    oA, oB, _ = c2_pair
    A, B = from_oracle(oA), from_oracle(oB)
    A2 = A._with_data(A.data * 0.5 + 1.0)
    order = FUSED_ORDERS[1]
    got = contract_many([(A, B), (A2, B)], output_labels=order)
    ref = _natural(lambda: contract_many([(A, B), (A2, B)], output_labels=order))
    for g, r in zip(got, ref):
        assert g.labels() == order and np.array_equal(g.keys, r.keys)
        assert torch.equal(g.data, r.data)
    nat = contract_many([(A, B), (A2, B)])
    from tenax_amd import contract

    for g, (a, b) in zip(nat, [(A, B), (A2, B)]):
        assert torch.equal(g.data, contract(a, b).data)


def test_permuted_store_limits(gpu):
    """A side needing more than three strides is refused by the planner and
    contract() transposes the natural result instead; a permuted-store plan
    has no adjoint and no per-block path."""
    from tenax_amd import _lib, contract, per_block_execution
    from tenax_amd.contraction import contractor, plans

    rng = np.random.default_rng(5)
    U1 = oc.U1()
    ch = lambda: rng.integers(-1, 2, size=4)
    la = tuple(oc.Leg(U1, ch(), oc.IN if i % 2 else oc.OUT, n) for i, n in enumerate("abcde"))
    lb = (oc.Leg(U1, la[4].charges, oc.IN if la[4].flow == oc.OUT else oc.OUT, "e"),
          oc.Leg(U1, ch(), oc.IN, "f"))
    oA = oc.SymT.random_normal(la, (1, 0))
    oB = oc.SymT.random_normal(lb, (1, 1))
    A, B = from_oracle(oA), from_oracle(oB)
    order = ("d", "c", "f", "b", "a")  # A's free legs a, b, c, d reversed: four runs
    oC, _ = ocn.contract_symmetric([oA, oB], "abcde,ef->dcfba")
    C = contract(A, B, output_labels=order)
    assert_same_structure(C, oC)
    assert max_rel_err(C, oC) < 1e-13
    assert not contractor._fused_order_fits((3, 2, 4, 1, 0), 4)
    nat = tuple(A.indices[:4]) + (B.indices[1],)
    with pytest.raises(_lib.TnxError, match="more than 3"):
        plans.pairwise_plan(A, B, 0, 1, 1, nat, _lib.MODE_PAIRWISE_FILTERED, None, out_perm=(3, 2, 4, 1, 0))
    # a fitting permutation: no adjoint, no per-block execution
    plan = plans.pairwise_plan(A, B, 0, 1, 1, nat, _lib.MODE_PAIRWISE_FILTERED, None, out_perm=(4, 0, 1, 2, 3))[0]
    lib = _lib.load()
    g = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.TnxError, match="no adjoint"):
        _lib.check(lib.tnx_contract_vjp(plan.ptr, _lib.dptr(g), _lib.dptr(A.data), _lib.dptr(B.data),
                                        _lib.dptr(A.data.clone()), None, 0, _lib.stream_handle()))
    with pytest.raises(_lib.TnxError, match="per-block"):
        _lib.check(lib.tnx_contract_execute_perblock(plan.ptr, _lib.dptr(A.data), _lib.dptr(B.data),
                                                     _lib.dptr(g), 0, _lib.stream_handle()))
    with per_block_execution():
        Cp = contract(A, B, output_labels=("f", "a", "b", "c", "d"))
    Cg = contract(A, B, output_labels=("f", "a", "b", "c", "d"))
    assert np.array_equal(Cp.keys, Cg.keys)
    assert (Cp.data - Cg.data).abs().max().item() <= 1e-12 * Cg.data.abs().max().item()


@pytest.mark.parametrize("tiling", [(0, 0, 1), (0, -2, 1), (0, -1, 0), (1, 0, 1), (3, -1, 1)])
def test_batched_launch_tilings_agree(gpu, c2_pair, tiling, monkeypatch):
    """A tiling set on the merged launch of contract_many (tnx_plan_set_tiling
    on the batch plan) gives the default launch's results, with natural and
    with fused output orders."""
    from tenax_amd.contraction import contract_many, contractor

    oA, oB, _ = c2_pair
    A, B = from_oracle(oA), from_oracle(oB)
    A2 = A._with_data(A.data * 0.5 + 1.0)
    for order in (None, FUSED_ORDERS[1]):
        monkeypatch.setattr(contractor, "BATCH_TILING", None)
        ref = contract_many([(A, B), (A2, B)], output_labels=order)
        monkeypatch.setattr(contractor, "BATCH_TILING", tiling)
        got = contract_many([(A, B), (A2, B)], output_labels=order)
        for g, r in zip(got, ref):
            assert np.array_equal(g.keys, r.keys)
            d = (g.data - r.data).abs().max().item()
            assert d <= 1e-13 * r.data.abs().max().item(), (tiling, order)


@pytest.mark.parametrize("sym_n,seed", [(3, 11), (None, 12), (2, 13)])
def test_fused_output_order_random_structures(gpu, sym_n, seed):
    """Random Z_n / U(1) operands with a charge target on A, contracted over
    two legs, every output order of the five free legs that fits the store
    descriptors: the oracle's blocks for that order and the natural result
    transposed, bit for bit."""
    import itertools

    from tenax_amd import contract
    from tenax_amd.contraction import contractor

    rng = np.random.default_rng(seed)
    sym = oc.U1() if sym_n is None else oc.Zn(sym_n)

    def ch(d):
        return rng.integers(-2, 3, size=d) if sym_n is None else rng.integers(0, sym_n, size=d)

    la = tuple(oc.Leg(sym, ch(5), oc.IN if i % 2 else oc.OUT, n) for i, n in enumerate("abcd"))
    lb = (oc.Leg(sym, la[2].charges, oc.OUT if la[2].flow == oc.IN else oc.IN, "c"),
          oc.Leg(sym, ch(4), oc.IN, "e"),
          oc.Leg(sym, la[3].charges, oc.OUT if la[3].flow == oc.IN else oc.IN, "d"),
          oc.Leg(sym, ch(3), oc.OUT, "f"),
          oc.Leg(sym, ch(4), oc.IN, "g"))
    oA = oc.SymT.random_normal(la, (seed, 0), target=1)
    oB = oc.SymT.random_normal(lb, (seed, 1))
    A, B = from_oracle(oA), from_oracle(oB)
    free = ("a", "b", "e", "f", "g")
    nat_pos = {l: i for i, l in enumerate(free)}
    tried = 0
    for order in itertools.permutations(free):
        if order == free or not contractor._fused_order_fits(tuple(nat_pos[l] for l in order), 2):
            continue
        tried += 1
        if tried % 7:  # every 7th fitting order keeps the test short
            continue
        oC, _ = ocn.contract_symmetric([oA, oB], "abcd,cedfg->" + "".join(order))
        C = contract(A, B, output_labels=order)
        assert_same_structure(C, oC)
        assert max_rel_err(C, oC) < 1e-13
        R = _natural(lambda: contract(A, B, output_labels=order))
        assert np.array_equal(C.keys, R.keys) and torch.equal(C.data, R.data), order
    assert tried > 20


def test_empty_operands_and_results(gpu):
    """Tensors without a single valid block, and contractions whose operands
    share no sector, through contract, transpose, add, inner, norm and
    max_abs: structure as the oracle's (zero blocks, the same legs), zero
    values, no launch on an empty work list (tensor.py:923-967, 296-319;
    contractor.py:444-476 iterate over no block combination)."""
    from tenax_amd import contract, inner

    U1 = oc.U1()
    rng = np.random.default_rng(5)
    # no valid block at all: every row charge 5 against column charges in [-1, 1]
    Ae = (oc.Leg(U1, np.array([5, 5, 5]), oc.IN, "i"), oc.Leg(U1, rng.integers(-1, 2, 4), oc.OUT, "j"))
    oAe = oc.SymT.random_normal(Ae, (1,), target=None)
    assert len(oAe.blocks) == 0
    B = (oc.Leg(U1, Ae[1].charges, oc.IN, "j"), oc.Leg(U1, rng.integers(-1, 2, 5), oc.OUT, "k"))
    oB = oc.SymT.random_normal(B, (2,), target=None)
    assert len(oB.blocks) > 0
    A_e, B_d = from_oracle(oAe), from_oracle(oB)
    assert A_e.n_blocks == 0 and A_e.data.numel() == 0

    oC, _ = ocn.contract_symmetric([oAe, oB], "ij,jk->ik")
    C = contract(A_e, B_d)
    assert_same_structure(C, oC)
    assert C.n_blocks == len(oC.blocks) == 0 and C.labels() == ("i", "k")
    assert torch.count_nonzero(C.todense()).item() == 0 and tuple(C.todense().shape) == (3, 5)

    # operands with blocks whose contracted sectors never meet: X holds only
    # the b = 1 sector, Y only b = 2
    X = (oc.Leg(U1, np.array([1, 1]), oc.IN, "a"), oc.Leg(U1, np.array([1, 2, 2]), oc.OUT, "b"))
    Y = (oc.Leg(U1, np.array([1, 2, 2]), oc.IN, "b"), oc.Leg(U1, np.array([2, 2, 2]), oc.OUT, "c"))
    oX, oY = oc.SymT.random_normal(X, (3,), target=None), oc.SymT.random_normal(Y, (4,), target=None)
    assert len(oX.blocks) > 0 and len(oY.blocks) > 0
    oZ, _ = ocn.contract_symmetric([oX, oY], "ab,bc->ac")
    Z = contract(from_oracle(oX), from_oracle(oY))
    assert_same_structure(Z, oZ)
    assert Z.n_blocks == len(oZ.blocks) == 0 and Z.labels() == ("a", "c")

    # transpose and vector algebra on the empty tensor
    T = A_e.transpose((1, 0))
    assert T.n_blocks == 0 and T.labels() == ("j", "i")
    assert float(A_e.norm()) == 0.0 and float(A_e.max_abs()) == 0.0
    assert float(inner(A_e, A_e)) == 0.0
    D = A_e + A_e
    assert D.n_blocks == 0 and D.data.numel() == 0


def test_large_and_sparse_u1_charges(gpu):
    """U(1) charges far from zero and far apart (sectors -100000, 0, 100000;
    sums reaching 2e5 on a fused output and a target of 2e5): the device planner's valid-block
    enumeration and key matching use the charge values themselves, so the
    result must carry the oracle's keys, shapes and values
    (contractor.py:291-484; core/index.py charge arithmetic on int32)."""
    from tenax_amd import contract

    U1 = oc.U1()
    rng = np.random.default_rng(21)
    qs = np.array([-100000, 0, 100000], np.int32)
    a = oc.Leg(U1, rng.choice(qs, 12), oc.IN, "a")
    b = oc.Leg(U1, rng.choice(qs, 10), oc.OUT, "b")
    k = oc.Leg(U1, rng.choice(qs, 11), oc.IN, "k")
    c = oc.Leg(U1, rng.choice(qs, 9), oc.OUT, "c")
    A = (a, b, oc.Leg(U1, k.charges, oc.OUT, "k"))
    B = (k, c)
    oA = oc.SymT.random_normal(A, (50,), target=None)
    oB = oc.SymT.random_normal(B, (51,), target=None)
    assert len(oA.blocks) > 0 and len(oB.blocks) > 0
    oC, _ = ocn.contract_symmetric([oA, oB], "abk,kc->abc")
    C = contract(from_oracle(oA), from_oracle(oB))
    assert_same_structure(C, oC)
    assert len(oC.blocks) > 0 and max_rel_err(C, oC) < 1e-13
    # a nonzero target charge far from zero
    oT = oc.SymT.random_normal((a, b), (52,), target=200000)
    assert len(oT.blocks) > 0
    T = from_oracle(oT)
    assert_same_structure(T, oT)
    assert max_rel_err(T.transpose((1, 0)).transpose((1, 0)), oT) == 0.0
    # the device enumeration of valid blocks at these charges and targets
    from tenax_amd.core.tensor import valid_blocks

    for legs, tgt in [((a, b, A[2]), None), ((a, b), 200000), ((a, b, A[2]), -100000)]:
        keys, shapes = valid_blocks(product_indices(legs), tgt)
        assert [tuple(int(q) for q in kk) for kk in keys] == sorted(oc.compute_valid_blocks(legs, tgt)), tgt
