"""GPU parity of the fused DMRG kernels and the device Lanczos against the FP64 oracle."""

import importlib

import numpy as np
import pytest
import torch

from oracle import contract as ocn
from oracle import dmrg as od
from tests.helpers import assert_same_structure, from_oracle, max_rel_err, to_oracle

pytestmark = pytest.mark.gpu

ED = {(4, 0): -1.616025403784438, (6, 0): -2.493577133887925, (4, 2): -0.957106781186548,
      (3, 1): -1.0, (5, 1): -1.9278862533179948}  # odd chains: the 2Sz = 1 sector


def _oracle_state(L, chi, seed, sweeps=1):
    """An oracle MPS after a few oracle sweeps (nontrivial block structure)."""
    mpo = od.build_mpo_heisenberg(L)
    mps = od.build_random_symmetric_mps(L, bond_dim=4, seed=seed)
    if sweeps:
        cfg = od.DMRGConfig(max_bond_dim=chi, num_sweeps=sweeps, lanczos_max_iter=8)
        _, _, mps, _, _ = od.dmrg(mpo, mps, cfg)
    return mpo, mps


@pytest.fixture(scope="module")
def state8():
    return _oracle_state(8, 12, 3)


def test_env_updates_match_oracle(gpu, state8):
    pd = importlib.import_module("tenax_amd.algorithms.dmrg")

    mpo, mps = state8
    L = len(mpo)
    oL = od.trivial_left_env()
    pL = pd._build_trivial_left_env_symmetric()
    for i in range(L - 1):
        oL = od.update_left_env(oL, mps[i], mpo[i])
        pL = pd._update_left_env_symmetric(pL, from_oracle(mps[i]), from_oracle(mpo[i]))
        assert_same_structure(pL, oL)
        assert max_rel_err(pL, oL) < 1e-12
    oR = od.trivial_right_env()
    pR = pd._build_trivial_right_env_symmetric()
    for i in range(L - 1, 0, -1):
        oR = od.update_right_env(oR, mps[i], mpo[i])
        pR = pd._update_right_env_symmetric(pR, from_oracle(mps[i]), from_oracle(mpo[i]))
        assert_same_structure(pR, oR)
        assert max_rel_err(pR, oR) < 1e-12


def _problem(mpo, mps, i):
    pd = importlib.import_module("tenax_amd.algorithms.dmrg")

    L = len(mpo)
    oL = od.trivial_left_env()
    for k in range(i):
        oL = od.update_left_env(oL, mps[k], mpo[k])
    oR = od.trivial_right_env()
    for k in range(L - 1, i + 1, -1):
        oR = od.update_right_env(oR, mps[k], mpo[k])
    oth = ocn.contract(mps[i], mps[i + 1])
    pad = None
    if oth.ndim == 3:
        pad = not oth.labels()[0].startswith("v")
        oth = od.pad(oth, pad)
    P = from_oracle
    theta, nd, pl = pd._theta_padded(P(mps[i]), P(mps[i + 1]))
    prob = pd.TwoSiteProblem(theta, P(oL), P(mpo[i]), P(mpo[i + 1]), P(oR))
    return prob, (oL, oth, mpo[i], mpo[i + 1], oR)


@pytest.mark.parametrize("site", [0, 3, 6])
def test_fused_matvec_matches_oracle(gpu, state8, site):
    pd = importlib.import_module("tenax_amd.algorithms.dmrg")

    mpo, mps = state8
    prob, (oL, oth, oWl, oWr, oR) = _problem(mpo, mps, site)
    ref = od.matvec(oL, oth, oWl, oWr, oR)
    y = prob.matvec(prob.x, 0)
    prob2 = prob
    out = pd.SymmetricTensor._wrap(prob.indices, prob.keys, prob.shapes, prob.offsets, y)
    po = to_oracle(out)
    for k, b in ref.blocks.items():
        np.testing.assert_allclose(po.blocks[k], b, atol=1e-12 * max(1.0, np.abs(b).max()))
    for k, b in po.blocks.items():
        if k not in ref.blocks:
            assert np.all(b == 0.0)
    # reach edges reproduce the reference's output key set
    pres = prob2.present_after(2)
    got = {tuple(int(q) for q in prob.keys[j]) for j in np.nonzero(pres)[0]}
    assert got == set(ref.blocks) | set(oth.blocks)
    # per-block path gives the same numbers
    y2 = prob.matvec(prob.x, 0, per_block=True)
    assert (y - y2).abs().max().item() <= 1e-12 * y.abs().max().item()


def test_matvec_tilings_agree(gpu, state8):
    """Both GEMM stages of the fused matvec under every tiling (shape, split K
    down to one K tile per task, XCD order) match the default tiling."""
    mpo, mps = state8
    prob, _ = _problem(mpo, mps, 3)
    y0 = prob.matvec(prob.x, 0).clone()
    for st in (1, 3):
        for sh in (0, 1, 2, 3):
            for kc in (0, 1, -1, -2):
                for xc in (0, 1):
                    prob.plan.set_tiling(st, sh, kc, xc)
                    y = prob.matvec(prob.x, 0)
                    err = (y - y0).abs().max().item()
                    assert err <= 1e-13 * y0.abs().max().item(), (st, sh, kc, xc)
        prob.plan.set_tiling(st, 0, 0, 0)
    with pytest.raises(Exception, match="stages 1 and 3"):
        prob.plan.set_tiling(2, 0, 0, 0)


@pytest.mark.parametrize("kc", [0, 1, -2])
def test_matvec_stage_by_stage_equals_matvec(gpu, state8, kc):
    """The matvec launched kernel by kernel through tnx_matvec_execute_stage
    (each GEMM with TNX_STAGE_GEMM_ONLY, then its reduction with
    TNX_STAGE_REDUCE_ONLY) gives the fused matvec's output bit for bit, with
    and without split K."""
    from tenax_amd import _lib

    mpo, mps = state8
    prob, _ = _problem(mpo, mps, 3)
    lib = _lib.load()
    for st in (1, 3):
        prob.plan.set_tiling(st, 0, kc, 1)
    try:
        y0 = prob.matvec(prob.x, 0).clone()
        L, Wl, Wr, R = prob.envs
        y = torch.full_like(prob.x, float("nan"))
        for st in (0, 1 | 16, 1 | 32, 2, 3 | 16, 3 | 32):
            _lib.check(lib.tnx_matvec_execute_stage(prob.plan.ptr, st, _lib.dptr(L.data), _lib.dptr(Wl.data),
                                                    _lib.dptr(Wr.data), _lib.dptr(R.data), _lib.dptr(prob.x),
                                                    _lib.dptr(y), 0, _lib.stream_handle()))
        torch.cuda.synchronize()
        assert torch.equal(y, y0)
        with pytest.raises(Exception, match="exclude each other"):
            _lib.check(lib.tnx_matvec_execute_stage(prob.plan.ptr, 3 | 48, _lib.dptr(L.data), _lib.dptr(Wl.data),
                                                    _lib.dptr(Wr.data), _lib.dptr(R.data), _lib.dptr(prob.x),
                                                    _lib.dptr(y), 0, _lib.stream_handle()))
        with pytest.raises(Exception, match="stage 1 or 3"):
            _lib.check(lib.tnx_matvec_execute_stage(prob.plan.ptr, 2 | 32, _lib.dptr(L.data), _lib.dptr(Wl.data),
                                                    _lib.dptr(Wr.data), _lib.dptr(R.data), _lib.dptr(prob.x),
                                                    _lib.dptr(y), 0, _lib.stream_handle()))
    finally:
        for st in (1, 3):
            prob.plan.set_tiling(st, 0, 0, 0)


def test_autotune_picks_a_candidate(gpu, monkeypatch):
    """The autotuner times the candidates on the plan's buffers and applies the fastest."""
    from tenax_amd.contraction import autotune

    autotune.clear()
    monkeypatch.setattr(autotune, "MIN_FLOPS", 0.0)
    mpo, mps = _oracle_state(8, 12, 3)
    prob, _ = _problem(mpo, mps, 2)
    prob._tune()  # the plan may come from the cache untuned: tune it (or reuse the choice)
    prob._tune()
    for st in (1, 3):
        t = prob.plan.tiling(st)
        assert (t["shape"], t["kchunk"], t["xcd"]) in autotune.CANDIDATES
    assert autotune.STATS["timed"] == 2  # one timing per stage signature
    assert autotune.STATS["reused"] >= 2  # later plans of the same work bucket reuse it
    ref = prob.matvec(prob.x, 0)
    prob.plan.set_tiling(1, 0, 0, 0)
    prob.plan.set_tiling(3, 0, 0, 0)
    y = prob.matvec(prob.x, 0)
    assert (y - ref).abs().max().item() <= 1e-13 * ref.abs().max().item()


@pytest.mark.parametrize("mode,tol", [("f32", 1e-5), ("split_f32", 1e-6)])
def test_fused_matvec_reduced_precision(gpu, state8, mode, tol):
    from tenax_amd import _lib

    mpo, mps = state8
    prob, _ = _problem(mpo, mps, 3)
    y64 = prob.matvec(prob.x, 0)
    y = prob.matvec(prob.x, _lib.COMPUTE[mode])
    err = (y - y64).abs().max().item() / y64.abs().max().item()
    assert err < tol, err


@pytest.mark.parametrize("reorth", [False, True])
def test_device_lanczos_matches_oracle(gpu, state8, reorth):
    mpo, mps = state8
    prob, (oL, oth, oWl, oWr, oR) = _problem(mpo, mps, 3)
    e_ref, v_ref = od.lanczos(lambda v: od.matvec(oL, v, oWl, oWr, oR), oth, 30, 1e-12, reorth)
    e, n = prob.lanczos(30, 1e-12, reorth, 0)
    assert n == 30
    assert abs(e - e_ref) < 1e-10
    got = prob.extract(prob.present_after(n))
    assert sorted(tuple(int(q) for q in k) for k in got.keys) == sorted(v_ref.blocks)
    # Ritz vectors agree up to sign
    po = to_oracle(got)
    dot = sum(float(np.sum(po.blocks[k] * v_ref.blocks[k])) for k in v_ref.blocks)
    assert abs(abs(dot) - 1.0) < 1e-8


@pytest.mark.parametrize("site", [0, 6])
def test_device_lanczos_stops_on_breakdown(gpu, state8, site):
    """Edge bonds span a Krylov space smaller than the step budget: the
    recurrence must stop at the step where beta falls below tol (dmrg.py:1012-1014),
    as the oracle does.  With re-orthogonalisation the breakdown beta is
    rounding residue far below tol, so the step is fixed by the algorithm;
    without it, beta after the space is exhausted is rounding noise (at site 6
    the device stops at 20 steps and the oracle runs all 30), which no
    implementation can be expected to reproduce."""
    mpo, mps = state8
    prob, (oL, oth, oWl, oWr, oR) = _problem(mpo, mps, site)
    calls = []

    def mv(v):
        calls.append(1)
        return od.matvec(oL, v, oWl, oWr, oR)

    e_ref, _ = od.lanczos(mv, oth, 30, 1e-8, True)
    e, n = prob.lanczos(30, 1e-8, True, 0)
    assert n == len(calls) < 30
    assert abs(e - e_ref) < 1e-10


@pytest.mark.parametrize("tol", [1e-8, 1e-12])
def test_device_plain_lanczos_after_breakdown_matches_oracle(gpu, state8, tol):
    """The reference's Lanczos (no re-orthogonalisation, dmrg.py:1004-1019) at
    site 6, whose Krylov space is exhausted before 30 steps: the step at which
    beta drops below tol is rounding noise (the two implementations may stop at
    different steps), but the lowest Ritz pair is not -- energy to 1e-10 and
    Ritz vector to 1e-8 of the oracle's plain solve."""
    mpo, mps = state8
    prob, (oL, oth, oWl, oWr, oR) = _problem(mpo, mps, 6)
    e_ref, v_ref = od.lanczos(lambda v: od.matvec(oL, v, oWl, oWr, oR), oth, 30, tol, False)
    e, n = prob.lanczos(30, tol, False, 0)
    print(f"site 6 plain Lanczos tol={tol}: device {n} steps, dE={e - e_ref:.2e}", flush=True)
    assert 1 <= n <= 30
    assert abs(e - e_ref) < 1e-10
    po = to_oracle(prob.extract(prob.present_after(n)))
    dot = sum(float(np.sum(po.blocks[k] * v_ref.blocks[k])) for k in v_ref.blocks if k in po.blocks)
    assert abs(abs(dot) - 1.0) < 1e-8


@pytest.mark.parametrize("reorth", [False, True])
def test_lanczos_graph_replays_match_eager(gpu, state8, reorth):
    """The captured solve (tnx_lanczos_graph_*) reproduces the eager device
    solve bit for bit on every replay, with eager work and synchronisations
    between replays; it follows a new start-vector buffer after an update, and
    refuses to replay after the plan's tiling changed until it is updated."""
    mpo, mps = state8
    prob, _ = _problem(mpo, mps, 3)
    x0 = prob.x.clone()
    e_ref, n_ref = prob.lanczos(30, 1e-12, reorth, 0)
    v_ref = prob.x.clone()
    prob.x.copy_(x0)
    g = prob.lanczos_graph(30, 1e-12, reorth, 0)
    try:
        for _ in range(3):
            prob.x.copy_(x0)
            prob.matvec(x0, 0)  # eager kernels on the same plan between replays
            torch.cuda.synchronize()
            g.launch()
            e, n = g.result()
            assert (e, n) == (e_ref, n_ref)
            assert torch.equal(prob.x, v_ref)
        x2 = x0.clone()
        g.update(x=x2)
        g.launch()
        e, n = g.result()
        assert (e, n) == (e_ref, n_ref)
        assert torch.equal(x2, v_ref)
        prob.plan.set_tiling(3, 0, 1, 1)
        with pytest.raises(Exception, match="tiling changed"):
            g.launch()
        prob.x.copy_(x0)
        e_t, n_t = prob.lanczos(30, 1e-12, reorth, 0)
        v_t = prob.x.clone()
        x2.copy_(x0)
        g.update()
        g.launch()
        e, n = g.result()
        assert (e, n) == (e_t, n_t)
        assert torch.equal(x2, v_t)
    finally:
        g.close()
        prob.plan.set_tiling(3, 0, 0, 0)


@pytest.mark.parametrize("L,tgt", [(4, 0), (6, 0), (4, 2), (3, 1), (5, 1)])
def test_dmrg_matches_ed(gpu, L, tgt):
    from tenax_amd import DMRGConfig, build_mpo_heisenberg, build_random_symmetric_mps, dmrg

    mpo = build_mpo_heisenberg(L)
    mps = build_random_symmetric_mps(L, bond_dim=4, seed=7, target_charge=tgt)
    res = dmrg(mpo, mps, DMRGConfig(max_bond_dim=16, num_sweeps=6, lanczos_max_iter=30,
                                    target_charge=tgt))
    assert abs(res.energy - ED[(L, tgt)]) < 1e-10


@pytest.mark.parametrize("reorth,tol_first", [(True, 1e-10), (False, 1e-5)])
def test_dmrg_matches_oracle_trajectory(gpu, reorth, tol_first):
    """Same initial tensors, same schedule: energies per sweep vs the FP64 oracle.

    Without re-orthogonalisation the 20-step Lanczos of the first sweep (random
    initial MPS) amplifies rounding differences of its start vector (the Gram
    factorisation resolves individual small-s vectors only to eps/(s/s_max)^2)
    to up to ~1e-6, so that sweep is held to 1e-5; later sweeps, and every
    sweep with FP64 re-orthogonalisation, to 1e-10."""
    from tenax_amd import DMRGConfig, build_mpo_heisenberg, build_random_symmetric_mps, dmrg

    L, chi = 10, 16
    cfg = dict(max_bond_dim=chi, num_sweeps=3, lanczos_max_iter=20)
    e_o, es_o, mps_o, errs_o, _ = od.dmrg(od.build_mpo_heisenberg(L), od.build_random_symmetric_mps(L, 4, seed=0),
                                          od.DMRGConfig(**cfg, lanczos_reorth=reorth))
    res = dmrg(build_mpo_heisenberg(L), build_random_symmetric_mps(L, 4, seed=0),
               DMRGConfig(**cfg, lanczos_reorth=reorth))
    d = np.abs(np.array(res.energies_per_sweep) - np.array(es_o))
    assert d[0] < tol_first and np.max(d[1:]) < 1e-10
    for i in range(L):
        t = res.mps.get_tensor(i)
        o = mps_o[i]
        assert sorted(tuple(int(q) for q in k) for k in t.keys) == sorted(o.blocks)


@pytest.mark.parametrize("reorth", [False, True])
def test_dmrg_with_captured_lanczos_equals_eager(gpu, reorth):
    """DMRGConfig(lanczos_graph=True): every two-site update's Lanczos solve is
    a replay of one captured graph rebound to the update's problem (bond
    dimensions growing along the chain, so the workspace grows and the launch
    topology changes); energies and the final MPS are bitwise those of the
    eager solves."""
    import torch

    import importlib

    from tenax_amd import DMRGConfig, build_mpo_heisenberg, build_random_symmetric_mps, dmrg

    pd = importlib.import_module("tenax_amd.algorithms.dmrg")
    L = 10
    mpo = build_mpo_heisenberg(L)
    runs = []
    for graph in (False, True):
        mps = build_random_symmetric_mps(L, bond_dim=4, seed=5)
        cfg = DMRGConfig(max_bond_dim=24, num_sweeps=3, lanczos_max_iter=30, lanczos_reorth=reorth,
                         lanczos_graph=graph, convergence_tol=0.0)
        runs.append(dmrg(mpo, mps, cfg))
    assert not pd._GRAPHS  # released at the end of the run
    eager, graph = runs
    assert eager.energies_per_sweep == graph.energies_per_sweep
    for i in range(L):
        a, b = eager.mps.get_tensor(i), graph.mps.get_tensor(i)
        assert np.array_equal(a.keys, b.keys) and torch.equal(a.data, b.data)


def test_dmrg_rejects_mixed_and_dense(gpu):
    from tenax_amd import DMRGConfig, TensorNetwork, build_mpo_heisenberg, build_random_symmetric_mps, dmrg
    from tenax_amd.core.tensor import DenseTensor

    mpo = build_mpo_heisenberg(4)
    mps = build_random_symmetric_mps(4, bond_dim=4, seed=0)
    dense = TensorNetwork()
    for i in range(4):
        t = mpo.get_tensor(i)
        dense.add_node(i, DenseTensor(t.todense(), t.indices))
    with pytest.raises(TypeError, match="uniform tensor types"):
        dmrg(dense, mps, DMRGConfig(max_bond_dim=8, num_sweeps=1))


def test_plan_with_live_graph_refuses_to_be_freed(gpu, state8):
    """A captured graph launches its plan's work lists: tnx_plan_free refuses
    the plan while the graph lives, and accepts it once the graph is freed."""
    from tenax_amd import _lib

    mpo, mps = state8
    prob, _ = _problem(mpo, mps, 2)
    lib = _lib.load()
    g = prob.lanczos_graph(10, 1e-12, False, 0)
    try:
        assert lib.tnx_plan_free(prob.plan.ptr) == _lib.TNX_ERR_ARG
        assert "captured graphs" in _lib.last_error()
    finally:
        g.close()
    # the plan stays usable (it was not freed); its owner releases it later
    e, n = prob.lanczos(10, 1e-12, False, 0)
    assert n == 10 and np.isfinite(e)
