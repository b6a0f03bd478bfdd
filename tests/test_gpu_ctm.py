"""GPU CTMRG, iPEPS energy and fixed-point gradient against the FP64 oracle.

What can be compared, and why:
* The reference's moves are not covariant under the projectors' eigenvector
  signs, so its CTM depends on them; both implementations fix them
  canonically.  Sweeps then agree step by step (checked from a common
  environment, and over fixed iteration counts from the initial one).
* The reference CTM does not converge for generic tensors: its corner legs
  swap horizontal / vertical roles in the top and right moves
  (_ctm_tensor.py:669-670, 712-713), so the map is chaotic (entry-wise
  differences of 1e-16 grow ~30x per sweep) and its own tests only check
  finiteness.  Trajectories are therefore compared over a few sweeps, and
  whole-run gradients are not compared with finite differences of the
  forward.  Instead: (i) the VJP of one CTM step, which the implicit
  backward is built from, against central differences of that step with its
  projectors frozen (they are stop-gradient in the reference); (ii) a D=1
  product state, where CTM is exact and the energy's gradient is known in
  closed form, through the full fixed-point backward (GMRES); (iii) the
  dense (torch) and block-sparse paths, whose backward runs through the
  engine's transpose / grouped-GEMM adjoints, give the same gradient.
"""

import numpy as np
import pytest
import torch

from oracle import ctm as occ
from oracle import contract as ocn
from tests.helpers import from_oracle, to_oracle

pytestmark = pytest.mark.gpu


def _A(seed, D=2, d=2, c4v=False):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((D, D, D, D, d))
    if c4v:
        ps = [(0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2), (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0),
              (3, 2, 1, 0)]
        A = sum(np.transpose(A, p + (4,)) for p in ps) / 8
    return A / np.linalg.norm(A)


def _ctm():
    import importlib

    return importlib.import_module("tenax_amd.algorithms.ctm")


@pytest.mark.parametrize("symmetric", [False, True])
def test_double_layer_matches_oracle(gpu, symmetric):
    C = _ctm()
    A = _A(1)
    a = C._build_double_layer_tensor(C.ipeps_site_tensor(A, symmetric)).todense().cpu().numpy()
    assert np.abs(a - occ.double_layer(A)).max() < 1e-14


@pytest.mark.parametrize("method", ["eigh", "qr"])
@pytest.mark.parametrize("symmetric", [False, True])
def test_sweep_step_parity(gpu, method, symmetric):
    """One sweep from the same (non-degenerate) environment: the identity-like
    initial environment has degenerate projector spectra, where rocSOLVER and
    LAPACK may keep different subspaces, so both start from the product's
    environment after three sweeps."""
    C = _ctm()
    A = _A(2)
    At = C.ipeps_site_tensor(A, symmetric)
    a = C._build_double_layer_tensor(At)
    env = C.initialize_ctm_tensor_env(At, 6)
    with torch.no_grad():
        for _ in range(3):
            env = C._ctm_tensor_sweep(env, a, 6, True, method)
        start = [t.todense().cpu().numpy() for t in env]
        env = C._ctm_tensor_sweep(env, a, 6, True, method)
    oenv = occ.sweep(start, occ.double_layer(A), 6, True, method)
    for t, o in zip(env[:4], oenv[:4]):
        s = np.sort(np.linalg.svd(t.todense().cpu().numpy(), compute_uv=False))
        so = np.sort(np.linalg.svd(o, compute_uv=False))
        assert np.abs(s - so).max() < 1e-10
    H = occ.heisenberg_gate()
    e = float(C.compute_energy_ctm_tensor(At, env, torch.as_tensor(H).cuda()))
    assert abs(e - occ.energy(A, oenv, H)) < 1e-10


@pytest.mark.parametrize("method", ["eigh", "qr"])
@pytest.mark.parametrize("symmetric", [False, True])
def test_fixed_iteration_energy_matches_oracle(gpu, method, symmetric):
    C = _ctm()
    A = _A(3)
    At = C.ipeps_site_tensor(A, symmetric)
    H = occ.heisenberg_gate()
    env = C.ctm_tensor(At, 8, max_iter=4, conv_tol=0.0, projector_method=method, qr_warmup_steps=2)
    e = float(C.compute_energy_ctm_tensor(At, env, torch.as_tensor(H).cuda()))
    eo = occ.energy(A, occ.ctm(A, 8, max_iter=4, conv_tol=0.0, method=method, qr_warmup_steps=2), H)
    assert abs(e - eo) < 1e-9


@pytest.mark.parametrize("symmetric", [False, True])
def test_ad_forward_matches_oracle(gpu, symmetric):
    C = _ctm()
    A = _A(4)
    H = occ.heisenberg_gate()
    cfg = C.CTMConfig(chi=4, max_iter=3, conv_tol=0.0, projector_method="qr")
    At = C.ipeps_site_tensor(A, symmetric)
    env = C.ctm_tensor_converge(At, cfg)
    e = float(C.compute_energy_ctm_tensor(At, env, torch.as_tensor(H).cuda()))
    oenv, _ = occ.ctm_fixed_point(A, 4, 3, 0.0, True, "qr")
    assert abs(e - occ.energy(A, oenv, H)) < 1e-9
    # the gauge-fixed environment is determined element-wise
    for t, o in zip(env, oenv):
        assert np.abs(t.todense().cpu().numpy() - o).max() < 1e-8


@pytest.mark.parametrize("symmetric", [False, True])
def test_step_vjp_matches_frozen_projector_differences(gpu, symmetric):
    C = _ctm()
    A = _A(5)
    cfg = C.CTMConfig(chi=6, max_iter=4, conv_tol=0.0, projector_method="qr")
    At = C.ipeps_site_tensor(A, symmetric)
    env, _ = C._ctm_tensor_fixed_point_impl(At, cfg)
    a0 = C._leaves([At])[0].detach()
    e0 = [x.detach() for x in C._leaves(env)]
    rng = np.random.default_rng(1)
    ua = torch.as_tensor(rng.standard_normal(a0.shape)).cuda()
    ue = [torch.as_tensor(rng.standard_normal(x.shape)).cuda() for x in e0]
    w = None
    fz = C.FROZEN_PROJECTORS

    def step(av, ev):
        out = C._ctm_tensor_step(C._with_leaves([At], [av])[0], C.CTMTensorEnv(*C._with_leaves(env, ev)), cfg)
        return C._leaves(out)

    try:
        fz.update(mode="record", list=[], i=0)
        ai = a0.clone().requires_grad_(True)
        ei = [x.clone().requires_grad_(True) for x in e0]
        out = step(ai, ei)
        w = [torch.as_tensor(rng.standard_normal(o.shape)).cuda() for o in out]
        ga, *ge = torch.autograd.grad(out, [ai] + ei, w, allow_unused=True)
        ad = float(torch.sum(ga * ua)) + sum(float(torch.sum(g * u)) for g, u in zip(ge, ue) if g is not None)
        h = 1e-6
        vals = []
        for sgn in (1, -1):
            fz.update(mode="replay", i=0)
            with torch.no_grad():
                o = step(a0 + sgn * h * ua, [x + sgn * h * u for x, u in zip(e0, ue)])
            vals.append(sum(float(torch.sum(oi * wi)) for oi, wi in zip(o, w)))
    finally:
        fz.update(mode=None, list=[], i=0)
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(ad - fd) < 1e-6 * max(1.0, abs(fd)), (ad, fd)


def test_product_state_gradient_closed_form(gpu):
    """D=1: CTM is exact, E(v) = 2 <vv|H|vv> / <v|v>^2 with v = A/|A|."""
    C = _ctm()
    H = occ.heisenberg_gate()
    Hm = torch.as_tensor(H.reshape(4, 4)).cuda()
    v0 = torch.tensor([0.8, -0.35], dtype=torch.float64).cuda()
    t = C.ipeps_site_tensor(v0.reshape(1, 1, 1, 1, 2))
    leaf = C._leaves([t])[0].detach().clone().requires_grad_(True)
    cfg = C.CTMConfig(chi=1, max_iter=20, conv_tol=1e-12, projector_method="qr")
    e = C.energy_loss(C._with_leaves([t], [leaf])[0], torch.as_tensor(H).cuda(), cfg)
    (g,) = torch.autograd.grad(e, [leaf])
    x = v0.clone().requires_grad_(True)
    vv = torch.kron(x, x)
    ref = 2 * (vv @ Hm @ vv) / (x @ x) ** 2
    (gr,) = torch.autograd.grad(ref, [x])
    assert abs(float(e) - float(ref)) < 1e-12
    assert torch.allclose(g.reshape(-1), gr, atol=1e-10, rtol=0)


def test_dense_and_block_gradients_agree(gpu):
    C = _ctm()
    A = _A(6)
    H = torch.as_tensor(occ.heisenberg_gate()).cuda()
    cfg = C.CTMConfig(chi=4, max_iter=3, conv_tol=0.0, projector_method="qr")
    gs = []
    for symmetric in (False, True):
        t = C.ipeps_site_tensor(A, symmetric)
        leaf = C._leaves([t])[0].detach().clone().requires_grad_(True)
        e = C.energy_loss(C._with_leaves([t], [leaf])[0], H, cfg)
        (g,) = torch.autograd.grad(e, [leaf])
        assert torch.isfinite(g).all() and float(g.abs().max()) > 1e-15
        gs.append(g.reshape(A.shape).cpu().numpy())
    assert np.abs(gs[0] - gs[1]).max() < 1e-9 * max(1.0, np.abs(gs[0]).max())


def test_block_contraction_adjoint(gpu):
    """d/dA and d/dB of <W, contract(A, B)> for U(1) block tensors vs the dense
    chain rule on the oracle's densified operands."""
    from oracle import core as oco
    from tenax_amd import contract

    rng = np.random.default_rng(3)
    sym = oco.U1()
    l1 = oco.Leg(sym, [-1, 0, 0, 1], oco.IN, "a")
    l2 = oco.Leg(sym, [-1, 1], oco.IN, "p")
    l3 = oco.Leg(sym, [-2, -1, 0, 0, 1, 2], oco.OUT, "k")
    l4 = oco.Leg(sym, [-1, 0, 1], oco.IN, "q")
    l5 = oco.Leg(sym, [-2, 0, 0, 2, 1], oco.OUT, "b")
    Ao = oco.SymT.random_normal((l1, l2, l3), (1, 2))
    Bo = oco.SymT.random_normal((l3.flip_flow(), l4, l5), (3, 4))
    A, B = from_oracle(Ao), from_oracle(Bo)
    a_leaf = A.data.detach().clone().requires_grad_(True)
    b_leaf = B.data.detach().clone().requires_grad_(True)
    Cp = contract(A._with_data(a_leaf), B._with_data(b_leaf))
    W = torch.as_tensor(rng.standard_normal(Cp.data.numel())).cuda()
    (ga, gb) = torch.autograd.grad(torch.dot(Cp.data, W), [a_leaf, b_leaf])
    # dense chain rule: dL/dA = sum_{q,b} Wd[a,p,q,b] B[k,q,b]
    Wd = to_oracle(Cp._with_data(W)).todense()
    Ad, Bd = Ao.todense(), Bo.todense()
    ref_a = np.einsum("apqb,kqb->apk", Wd, Bd)
    ref_b = np.einsum("apqb,apk->kqb", Wd, Ad)
    got_a = to_oracle(A._with_data(ga)).todense()
    got_b = to_oracle(B._with_data(gb)).todense()
    mask_a = np.abs(Ad) > 0
    mask_b = np.abs(Bd) > 0
    assert np.abs((got_a - ref_a)[mask_a]).max() < 1e-12
    assert np.abs((got_b - ref_b)[mask_b]).max() < 1e-12
    assert Cp.data.numel() and ocn is not None


def test_optimize_gs_ad_runs(gpu):
    """ipeps.py:2370-2434 loop: finite energies, deterministic (test_ad_utils.py:286-307)."""
    C = _ctm()
    H = occ.heisenberg_gate()
    cfg = C.iPEPSConfig(max_bond_dim=2, ctm=C.CTMConfig(chi=6, max_iter=8, conv_tol=1e-9, projector_method="qr"),
                        gs_learning_rate=0.05, gs_num_steps=3)
    r1 = C.optimize_gs_ad(H, C.ipeps_site_tensor(_A(8)), cfg)
    r2 = C.optimize_gs_ad(H, C.ipeps_site_tensor(_A(8)), cfg)
    assert np.isfinite(r1[2]) and abs(r1[2] - r2[2]) < 1e-12
