"""The benchmark workloads at their own sizes on the GPU, against the FP64 oracle.

* L=20 Heisenberg chain, chi=64, 10 sweeps, seed 0 (reference dmrg.py:125-305):
  every sweep energy against the committed oracle trajectory
  (tests/golden/c1_oracle.json), bond charge multiplicities bit-exact, final
  energy against exact diagonalisation.
* Single chi=1024 U(1) contraction A[l,p,p',r] B[r,q,q',s] (contractor.py:291-484):
  sector structure bit-exact, FP64 values within 1e-13 relative; FP32 and
  split-FP32 errors stated.
* The vector algebra of the Krylov loop (tensor.py:296-319, 923-967) on
  chi=1024 operands with different key sets.
"""

import json
import os

import numpy as np
import pytest
import torch

from oracle import contract as ocn
from oracle import core as oc
from oracle import dmrg as od
from tests.helpers import assert_same_structure, from_oracle, max_rel_err, to_oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_oracle.json")


def _bond_multiplicities(res):
    out = []
    for i in range(19):
        t = res.mps.get_tensor(i)
        q, m = np.unique(np.asarray(t.indices[-1].charges), return_counts=True)
        out.append({str(int(a)): int(b) for a, b in zip(q, m)})
    return out


@pytest.fixture(scope="module")
def c1_gold():
    return json.load(open(GOLDEN))


def _run_c1(reorth):
    from tenax_amd import DMRGConfig, build_mpo_heisenberg, build_random_symmetric_mps, dmrg

    return dmrg(build_mpo_heisenberg(20), build_random_symmetric_mps(20, 4, seed=0),
                DMRGConfig(max_bond_dim=64, num_sweeps=10, lanczos_max_iter=50, lanczos_reorth=reorth))


def test_c1_reorth_trajectory_matches_oracle(gpu, c1_gold):
    """Re-orthogonalised Lanczos (FP64): every sweep energy to 1e-10, the same
    number of sweeps, the same bond multiplicities, truncation errors of the
    last sweep to 1e-9, and the ED ground energy to 1e-10."""
    ref = c1_gold["reorth"]
    res = _run_c1(True)
    print("C1 reorth energies", res.energies_per_sweep, flush=True)
    assert len(res.energies_per_sweep) == len(ref["energies_per_sweep"])
    d = np.abs(np.array(res.energies_per_sweep) - np.array(ref["energies_per_sweep"]))
    assert np.max(d) < 1e-10, d
    assert res.converged
    assert _bond_multiplicities(res) == ref["bond_multiplicities"]
    te = np.array(res.truncation_errors[-38:])
    np.testing.assert_allclose(te, ref["truncation_errors_last_sweep"], rtol=0, atol=1e-9)
    assert abs(res.energy - c1_gold["ed_anchor_2sz0"]) < 1e-10


def test_c1_plain_lanczos_reaches_oracle_energy(gpu, c1_gold):
    """The reference's Lanczos (no re-orthogonalisation, the sweep default):
    the first sweep follows a random start vector through 50 unorthogonalised
    steps, which amplifies rounding differences, so it is held to 1e-5; the
    later sweeps to 1e-8 and the converged energy to 1e-10 of the oracle's
    plain trajectory and of exact diagonalisation."""
    ref = c1_gold["no_reorth"]
    res = _run_c1(False)
    print("C1 plain energies", res.energies_per_sweep, flush=True)
    es, rs = np.array(res.energies_per_sweep), np.array(ref["energies_per_sweep"])
    n = min(len(es), len(rs))
    assert abs(es[0] - rs[0]) < 1e-5
    assert np.max(np.abs(es[1:n] - rs[1:n])) < 1e-8
    assert res.converged
    assert abs(res.energy - ref["energy"]) < 1e-10
    assert abs(res.energy - c1_gold["ed_anchor_2sz0"]) < 1e-10
    assert _bond_multiplicities(res) == ref["bond_multiplicities"]


def c2_operands(chi=1024):
    from tests.test_gpu_contract import c2_legs

    A, B = c2_legs(chi)
    return oc.SymT.random_normal(A, (1, 0)), oc.SymT.random_normal(B, (1, 1))


@pytest.fixture(scope="module")
def c2_full():
    oA, oB = c2_operands()
    oC, _ = ocn.contract_symmetric([oA, oB], "lpmr,rqns->lpmqns")
    return oA, oB, oC


def test_c2_chi1024_structure_and_f64_values(gpu, c2_full):
    from tenax_amd import contract

    oA, oB, oC = c2_full
    A, B = from_oracle(oA), from_oracle(oB)
    assert A.n_blocks == 80 and int(A.data.numel()) == 374202
    C = contract(A, B)
    assert C.n_blocks == 308 and int(C.data.numel()) == 1425846
    assert_same_structure(C, oC)
    # packed offsets: blocks back to back in sorted-key order
    sizes = np.prod(C.shapes.astype(np.int64), axis=1)
    np.testing.assert_array_equal(C.offsets, np.concatenate([[0], np.cumsum(sizes)[:-1]]))
    err = max_rel_err(C, oC)
    print(f"C2 chi=1024 f64 max rel err {err:.2e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("mode,tol", [("f32", 2e-5), ("split_f32", 2e-6)])
def test_c2_chi1024_reduced_precision_error(gpu, c2_full, mode, tol):
    from tenax_amd import compute_mode, contract

    oA, oB, oC = c2_full
    with compute_mode(mode):
        C = contract(from_oracle(oA), from_oracle(oB))
    assert_same_structure(C, oC)
    err = max_rel_err(C, oC)
    print(f"C2 chi=1024 {mode} max rel err {err:.2e}", flush=True)
    assert err < tol


def _drop_blocks(t, every):
    """The same tensor with every `every`-th block removed (a different key set)."""
    keep = {k: v for i, (k, v) in enumerate(sorted(t.blocks.items())) if i % every}
    return oc.SymT(keep, t.legs, validate=False)


def test_vector_algebra_union_keys_matches_oracle(gpu):
    """__add__ / __sub__ over the union of keys, inner over common keys, norm
    and max_abs (tensor.py:296-319, 863-868, 923-967) against the oracle, on
    chi=1024 operands whose key sets differ."""
    from tenax_amd import inner

    oA, _ = c2_operands()
    oB = _drop_blocks(oc.SymT.random_normal(oA.legs, (1, 7)), 3)
    oA2 = _drop_blocks(oA, 5)
    A2, B = from_oracle(oA2), from_oracle(oB)
    assert set(oA2.blocks) != set(oB.blocks)
    for got, ref in ((A2 + B, oA2 + oB), (A2 - B * 0.5, oA2 - oB * 0.5)):
        assert_same_structure(got, ref)
        po = to_oracle(got)
        for k, b in ref.blocks.items():
            np.testing.assert_array_equal(po.blocks[k], b)  # elementwise adds: exact
    ip = float(inner(A2, B))
    ir = oc.inner(oA2, oB)
    assert abs(ip - ir) <= 1e-13 * max(1.0, oA2.norm() * oB.norm())
    assert abs(float(A2.norm()) - oA2.norm()) <= 1e-14 * oA2.norm()
    assert float(B.max_abs()) == oB.max_abs()
    # inner of a tensor with itself is its squared norm
    assert abs(float(inner(B, B)) - oB.norm() ** 2) <= 1e-13 * oB.norm() ** 2
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def c3_centre():
    """The chi=2048 centre-bond problem of the L=100 chain (converged-state charge profile, seed 2)."""
    from tenax_amd.workloads import centre_bond_problem

    prob, (L, Wl, Wr, R, theta), _, _ = centre_bond_problem(100, 2048, 2)
    ops = tuple(to_oracle(t) for t in (L, Wl, Wr, R))
    oth = to_oracle(prob.extract(prob.present_after(1)))
    return prob, ops, oth


def test_c3_centre_matvec_matches_oracle(gpu, c3_centre):
    """One fused FP64 matvec at chi=2048 against the oracle's factored per-block
    products (same sum, same reached block set)."""
    prob, (oL, oWl, oWr, oR), oth = c3_centre
    assert int(prob.total) == 2986058 and len(prob.keys) == 46
    ref = od.FactoredMatvec(oL, oWl, oWr, oR)(oth)
    y = prob.matvec(prob.x, 0)
    from tenax_amd import SymmetricTensor

    got = to_oracle(SymmetricTensor._wrap(prob.indices, prob.keys, prob.shapes, prob.offsets, y))
    scale = max(float(np.abs(b).max()) for b in ref.blocks.values())
    err = max(float(np.abs(got.blocks[k] - b).max()) for k, b in ref.blocks.items()) / scale
    print(f"C3 centre matvec max rel err {err:.2e} ({len(ref.blocks)} blocks)", flush=True)
    assert err < 1e-13
    for k, b in got.blocks.items():
        if k not in ref.blocks:
            assert not np.any(b)


@pytest.mark.parametrize("reorth", [False, True])
def test_c3_centre_lanczos_matches_oracle(gpu, c3_centre, reorth):
    """The 50-step device Lanczos at the chi=2048 centre bond (dmrg.py:978-1045,
    FP64 MFMA matvecs) against the oracle's FP64 Lanczos on the same L, W, theta
    and R: |dE| < 1e-10 and the same Ritz vector."""
    prob, (oL, oWl, oWr, oR), oth = c3_centre
    x0 = prob.x.clone()
    try:
        e, n = prob.lanczos(50, 1e-12, reorth, 0)
        v = to_oracle(prob.extract(prob.present_after(n)))
    finally:
        prob.x.copy_(x0)
    e_ref, v_ref = od.lanczos(od.FactoredMatvec(oL, oWl, oWr, oR), oth, 50, 1e-12, reorth)
    print(f"C3 centre Lanczos reorth={reorth}: E={e:.13f} oracle {e_ref:.13f} dE={e - e_ref:.2e} steps={n}",
          flush=True)
    assert n == 50
    assert abs(e - e_ref) < 1e-10
    dot = sum(float(np.sum(v.blocks[k] * b)) for k, b in v_ref.blocks.items() if k in v.blocks)
    assert abs(abs(dot) - 1.0) < 1e-8
