"""CTMRG on the Tensor protocol, its fixed-point gradient, iPEPS energy and AD
optimisation (reference src/tenax/algorithms/_ctm_tensor.py, ad_utils.py:139-411
and 597-764, ipeps.py:37-120, 1368-1549, 2265-2434).

Every contraction goes through ``contract()``: DenseTensor operands use torch
(rocBLAS), SymmetricTensor operands the block engine (grouped MFMA GEMM).
Gradients flow through torch autograd; for SymmetricTensor the engine's
transpose and grouped-GEMM nodes carry hand-written adjoints (a block
contraction's adjoint is two block contractions on the same kernels), so an
iPEPS gradient never leaves the device.

The fixed point (``ctm_tensor_converge``) is an autograd Function: forward runs
CTM with the QR gauge fix to convergence without recording; backward solves
(I - J^T) lambda = g by restarted GMRES, where J^T v is the VJP of ONE CTM
step at the fixed point, and returns dA = (d step / dA)^T lambda
(ad_utils.py:722-761).  The step's graph is built once and reused by every
GMRES product unless ``recompute=True``, which rebuilds it per product (the
reference's behaviour; memory of one step instead of its graph).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np
import torch

from tenax_amd.algorithms._tensor_utils import fuse_indices
from tenax_amd.contraction import contract
from tenax_amd.core import EPS
from tenax_amd.core.index import FlowDirection, TensorIndex
from tenax_amd.core.symmetry import U1Symmetry
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor, Tensor

IN = FlowDirection.IN
OUT = FlowDirection.OUT


@dataclass
class CTMConfig:
    """ipeps.py:37-57."""

    chi: int = 20
    max_iter: int = 100
    conv_tol: float = 1e-8
    renormalize: bool = True
    projector_method: str = "eigh"
    qr_warmup_steps: int = 3
    chi_I: int | None = None


@dataclass
class iPEPSConfig:
    """ipeps.py:60-89 (the fields the AD optimisation reads)."""

    max_bond_dim: int = 2
    num_imaginary_steps: int = 100
    dt: float = 0.01
    ctm: CTMConfig = field(default_factory=CTMConfig)
    svd_trunc_err: float | None = None
    gate_order: str = "sequential"
    unit_cell: str = "1x1"
    gs_optimizer: str = "adam"
    gs_learning_rate: float = 1e-3
    gs_num_steps: int = 200
    gs_conv_tol: float = 1e-8
    su_init: bool = False


class CTMTensorEnv(NamedTuple):
    """_ctm_tensor.py:38-54: corners (chi, chi), edges (chi, D^2, chi)."""

    C1: Tensor  # (c1_d, c1_r)
    C2: Tensor  # (c2_l, c2_d)
    C3: Tensor  # (c3_u, c3_l)
    C4: Tensor  # (c4_r, c4_u)
    T1: Tensor  # (t1_l, u2, t1_r)
    T2: Tensor  # (t2_u, r2, t2_d)
    T3: Tensor  # (t3_r, d2, t3_l)
    T4: Tensor  # (t4_d, l2, t4_u)


# (label_a, label_b, flow_a, flow_b, reference axis of A for the charges)
_CORNER_SPECS = {
    "C1": ("c1_d", "c1_r", IN, OUT, 1),
    "C2": ("c2_l", "c2_d", IN, OUT, 0),
    "C3": ("c3_u", "c3_l", OUT, IN, 1),
    "C4": ("c4_r", "c4_u", OUT, IN, 0),
}
# (chi label 1, D^2 label, chi label 2, flows, A axis for chi charges, A axis fused into D^2)
_EDGE_SPECS = {
    "T1": ("t1_l", "u2", "t1_r", IN, IN, OUT, 3, 0),
    "T2": ("t2_u", "r2", "t2_d", OUT, OUT, IN, 0, 3),
    "T3": ("t3_r", "d2", "t3_l", OUT, OUT, IN, 3, 1),
    "T4": ("t4_d", "l2", "t4_u", IN, IN, OUT, 1, 2),
}


def _fuse(T: Tensor, la, lb, label, flow) -> Tensor:
    labels = T.labels()
    return fuse_indices(T, labels.index(la), labels.index(lb), label, flow)


# ---------------------------------------------------------------- double layer
def _build_double_layer_tensor(A: Tensor) -> Tensor:
    """_ctm_tensor.py:346-365: contract(A, bar(A)) over phys, fuse (x, X) -> x2."""
    Ab = A.bar().relabels({"u": "U", "d": "D", "l": "L", "r": "R"})
    a8 = contract(A, Ab)
    x = _fuse(a8, "u", "U", "u2", IN)
    x = _fuse(x, "d", "D", "d2", OUT)
    x = _fuse(x, "l", "L", "l2", IN)
    return _fuse(x, "r", "R", "r2", OUT)


def _build_double_layer_open_tensor(A: Tensor) -> Tensor:
    """_ctm_tensor.py:368-384: physical legs (phys, phys_bra) left open."""
    Ab = A.bar().relabels({"u": "U", "d": "D", "l": "L", "r": "R", "phys": "phys_bra"})
    x = contract(A, Ab)
    x = _fuse(x, "u", "U", "u2", IN)
    x = _fuse(x, "d", "D", "d2", OUT)
    x = _fuse(x, "l", "L", "l2", IN)
    return _fuse(x, "r", "R", "r2", OUT)


# ---------------------------------------------------------------- initial environment
def _tile_charges(base, n):
    base = np.asarray(base, np.int32)
    if n <= len(base):
        return base[:n].copy()
    return np.tile(base, n // len(base) + 1)[:n].astype(np.int32)


def initialize_ctm_tensor_env(A: Tensor, chi: int) -> CTMTensorEnv:
    """_ctm_tensor.py:521-578: identity-like corners and edges T[i, :, i] = 1."""
    from tenax_amd.algorithms._tensor_utils import _compute_fused_charges

    D = A.indices[0].dim
    D2 = D * D
    m = min(chi, D2)
    dev = A.todense().device if isinstance(A, DenseTensor) else A.data.device
    eye = torch.zeros((chi, chi), dtype=torch.float64, device=dev)
    eye[:m, :m] = torch.eye(m, dtype=torch.float64, device=dev)
    edge = torch.zeros((chi, D2, chi), dtype=torch.float64, device=dev)
    for i in range(m):
        edge[i, :, i] = 1.0
    corners, edges = {}, {}
    if isinstance(A, SymmetricTensor):
        sym = A.indices[0].symmetry
        for name, (la, lb, fa, fb, ref) in _CORNER_SPECS.items():
            ix = A.indices[ref]
            q = _tile_charges(_compute_fused_charges(ix, ix.flip_flow(), fa, sym), chi)
            corners[name] = SymmetricTensor.from_dense(torch.eye(chi, dtype=torch.float64, device=dev),
                                                       (TensorIndex(sym, q, fa, la), TensorIndex(sym, q, fb, lb)))
        for name, (l1, l2, l3, f1, f2, f3, rc, rd) in _EDGE_SPECS.items():
            qc = _tile_charges(A.indices[rc].charges, chi)
            ix = A.indices[rd]
            qd = _compute_fused_charges(ix, ix.flip_flow(), f2, sym)
            edges[name] = SymmetricTensor.from_dense(
                edge, (TensorIndex(sym, qc, f1, l1), TensorIndex(sym, qd, f2, l2), TensorIndex(sym, qc, f3, l3)),
                tol=float("inf"))
    else:
        sym = U1Symmetry()
        z = lambda n: np.zeros(n, np.int32)  # noqa: E731
        for name, (la, lb, fa, fb, _) in _CORNER_SPECS.items():
            corners[name] = DenseTensor(eye.clone(), (TensorIndex(sym, z(chi), fa, la),
                                                      TensorIndex(sym, z(chi), fb, lb)))
        for name, (l1, l2, l3, f1, f2, f3, _, _) in _EDGE_SPECS.items():
            edges[name] = DenseTensor(edge.clone(), (TensorIndex(sym, z(chi), f1, l1), TensorIndex(sym, z(D2), f2, l2),
                                                     TensorIndex(sym, z(chi), f3, l3)))
    return CTMTensorEnv(corners["C1"], corners["C2"], corners["C3"], corners["C4"],
                        edges["T1"], edges["T2"], edges["T3"], edges["T4"])


# ---------------------------------------------------------------- projector
def _dense_of(T: Tensor) -> torch.Tensor:
    return T.todense()


# Projector recording / replay: the step's Jacobian treats P as a constant
# (stop-gradient, as the reference), so finite differences that check it must
# reuse the base point's projectors.  mode None | "record" | "replay".
FROZEN_PROJECTORS: dict = {"mode": None, "list": [], "i": 0}


def _compute_projector_tensor(C1g: Tensor, C4g: Tensor, chi: int, projector_method: str = "eigh") -> Tensor:
    """_ctm_tensor.py:195-296: top-chi eigenvectors of C1g C1g^T + C4g C4g^T
    (directly, or through QR of [C1g, C4g] and the small R R^T), no gradient,
    in the canonical basis of _canonical_top."""
    if projector_method not in ("eigh", "qr"):
        raise ValueError(f"Unknown projector_method={projector_method!r}; expected 'eigh' or 'qr'.")
    fz = FROZEN_PROJECTORS
    if fz["mode"] == "replay":
        P = fz["list"][fz["i"]]
        fz["i"] += 1
        return P
    P = _projector_raw(C1g, C4g, chi, projector_method)
    if fz["mode"] == "record":
        fz["list"].append(P)
    return P


# eigenvalues closer than CLUSTER_TOL x the largest form one cluster; inside
# it, projections of unit vectors shorter than PIVOT_TOL are skipped
CLUSTER_TOL = 1e-10
PIVOT_TOL = 1e-6


def _canonical_signs(P):
    """First entry with |entry| >= (1 - 1e-8) max|entry| of each column made
    positive (magnitudes equal up to rounding count as ties)."""
    a = P.abs()
    hit = (a >= (1.0 - 1e-8) * a.max(dim=0, keepdim=True).values).to(torch.int8)
    idx = torch.argmax(hit, dim=0)
    sg = torch.sign(P[idx, torch.arange(P.shape[1], device=P.device)])
    return P * torch.where(sg == 0, torch.ones_like(sg), sg)[None, :]


def _canonical_top(w, V, k):
    """Top-k eigenvectors in a canonical basis.

    The reference takes whatever basis its eigensolver returns, but its moves
    are not covariant under that choice: the right move relabels C2's
    projected leg c2_l and its unprojected leg c2_d (_ctm_tensor.py:669) and
    the top move then contracts c2_d with T2's projected leg, so even
    eigenvector signs change converged energies by ~1e-3, and the
    identity-like initial environment has exactly degenerate spectra.  Here
    every cluster of numerically equal eigenvalues reaching into the kept set
    gets the basis from Gram-Schmidt of its projector's columns in index order,
    and every column has its largest-|entry| (first within 1e-8 relative)
    positive, which makes CTM deterministic and the same on rocSOLVER and
    LAPACK."""
    w = w.flip(0)
    V = V.flip(1)
    n = w.shape[0]
    out = V[:, :k].clone()
    wh = w.cpu().numpy()
    scale = max(abs(float(wh[0])), 1e-300)
    i = 0
    while i < k:
        j = i + 1
        while j < n and abs(wh[j] - wh[j - 1]) <= CLUSTER_TOL * scale:
            j += 1
        if j - i > 1:
            Vc = V[:, i:j]
            basis = []
            for r in range(Vc.shape[0]):
                v = Vc @ Vc[r, :]
                for b in basis:
                    v = v - torch.dot(b, v) * b
                nv = float(torch.linalg.vector_norm(v))
                if nv > PIVOT_TOL:
                    basis.append(v / nv)
                if len(basis) == j - i:
                    break
            m = min(j, k) - i
            out[:, i:i + m] = torch.stack(basis[:m], dim=1)
        i = j
    return _canonical_signs(out)


def _projector_raw(C1g, C4g, chi, projector_method):
    with torch.no_grad():
        c1 = _dense_of(C1g).detach()
        c4 = _dense_of(C4g).detach()
        if projector_method == "qr":
            Q, R = torch.linalg.qr(torch.cat([c1, c4], dim=1))
            rho = R @ R.T
            rho = 0.5 * (rho + rho.T)
            w, V = torch.linalg.eigh(rho)
            k = min(chi, V.shape[1])
            P = _canonical_signs(Q @ _canonical_top(w, V, k))
        else:
            rho = c1 @ c1.T + c4 @ c4.T
            rho = 0.5 * (rho + rho.T)
            w, V = torch.linalg.eigh(rho)
            k = min(chi, V.shape[1])
            P = _canonical_top(w, V, k).contiguous()
    fidx = C1g.indices[C1g.labels().index("fused")]
    new = TensorIndex(fidx.symmetry, np.zeros(k, np.int32), OUT, "chi_new")
    if isinstance(C1g, SymmetricTensor):
        return SymmetricTensor.from_dense(P, (fidx, new), tol=float("inf"))
    return DenseTensor(P, (fidx, new))


def _apply_projector_tensor(P, Ca, Cb, Tg, fused_l, fused_r):
    """_ctm_tensor.py:299-338: P^T Ca, P^T Cb, P^T Tg P."""
    Pb = P.bar()
    Ca_new = contract(Pb, Ca)
    Cb_new = contract(Pb, Cb)
    step = contract(Pb.relabel("fused", fused_l), Tg)
    T_new = contract(step, P.relabels({"fused": fused_r, "chi_new": "chi_new_r"}))
    return Ca_new, Cb_new, T_new


# ---------------------------------------------------------------- moves
def _ctm_tensor_move_left(env, env_nb, a, chi, projector_method="eigh"):
    """_ctm_tensor.py:586-629: updates C1, T4, C4."""
    C1g = _fuse(contract(env.C1.relabel("c1_r", "t1_l"), env_nb.T1), "c1_d", "u2", "fused", IN)
    C4g = _fuse(contract(env.C4.relabel("c4_u", "t3_r"), env_nb.T3), "c4_r", "d2", "fused", IN)
    T4g = _fuse(_fuse(contract(env.T4, a), "t4_d", "u2", "fl", IN), "t4_u", "d2", "fr", OUT)
    P = _compute_projector_tensor(C1g, C4g, chi, projector_method)
    C1n, C4n, T4n = _apply_projector_tensor(P, C1g, C4g, T4g, "fl", "fr")
    return env._replace(C1=C1n.relabels({"chi_new": "c1_d", "t1_r": "c1_r"}),
                        C4=C4n.relabels({"chi_new": "c4_r", "t3_l": "c4_u"}),
                        T4=T4n.relabels({"chi_new": "t4_d", "chi_new_r": "t4_u", "r2": "l2"}))


def _ctm_tensor_move_right(env, env_nb, a, chi, projector_method="eigh"):
    """_ctm_tensor.py:632-672: updates C2, T2, C3."""
    C2g = _fuse(contract(env.C2.relabel("c2_l", "t1_r"), env_nb.T1), "c2_d", "u2", "fused", IN)
    C3g = _fuse(contract(env.C3.relabel("c3_u", "t3_l"), env_nb.T3), "c3_l", "d2", "fused", IN)
    T2g = _fuse(_fuse(contract(env.T2, a), "t2_u", "u2", "fl", IN), "t2_d", "d2", "fr", OUT)
    P = _compute_projector_tensor(C2g, C3g, chi, projector_method)
    C2n, C3n, T2n = _apply_projector_tensor(P, C2g, C3g, T2g, "fl", "fr")
    return env._replace(C2=C2n.relabels({"chi_new": "c2_l", "t1_l": "c2_d"}),
                        C3=C3n.relabels({"chi_new": "c3_u", "t3_r": "c3_l"}),
                        T2=T2n.relabels({"chi_new": "t2_u", "chi_new_r": "t2_d", "l2": "r2"}))


def _ctm_tensor_move_top(env, env_nb, a, chi, projector_method="eigh"):
    """_ctm_tensor.py:675-715: updates C1, T1, C2."""
    C1g = _fuse(contract(env.C1.relabel("c1_d", "t4_d"), env_nb.T4), "c1_r", "l2", "fused", IN)
    C2g = _fuse(contract(env.C2.relabel("c2_d", "t2_u"), env_nb.T2), "c2_l", "r2", "fused", IN)
    T1g = _fuse(_fuse(contract(env.T1, a), "t1_l", "l2", "fl", IN), "t1_r", "r2", "fr", OUT)
    P = _compute_projector_tensor(C1g, C2g, chi, projector_method)
    C1n, C2n, T1n = _apply_projector_tensor(P, C1g, C2g, T1g, "fl", "fr")
    return env._replace(C1=C1n.relabels({"chi_new": "c1_d", "t4_u": "c1_r"}),
                        C2=C2n.relabels({"chi_new": "c2_l", "t2_d": "c2_d"}),
                        T1=T1n.relabels({"chi_new": "t1_l", "chi_new_r": "t1_r", "d2": "u2"}))


def _ctm_tensor_move_bottom(env, env_nb, a, chi, projector_method="eigh"):
    """_ctm_tensor.py:718-759: updates C4, T3, C3."""
    C4g = _fuse(contract(env.C4.relabel("c4_r", "t4_u"), env_nb.T4), "c4_u", "l2", "fused", IN)
    C3g = _fuse(contract(env.C3.relabel("c3_l", "t2_d"), env_nb.T2), "c3_u", "r2", "fused", IN)
    T3g = _fuse(_fuse(contract(env.T3, a), "t3_r", "l2", "fl", IN), "t3_l", "r2", "fr", OUT)
    P = _compute_projector_tensor(C4g, C3g, chi, projector_method)
    C4n, C3n, T3n = _apply_projector_tensor(P, C4g, C3g, T3g, "fl", "fr")
    return env._replace(C4=C4n.relabels({"chi_new": "c4_r", "t4_d": "c4_u"}),
                        C3=C3n.relabels({"chi_new": "c3_u", "t2_u": "c3_l"}),
                        T3=T3n.relabels({"chi_new": "t3_r", "chi_new_r": "t3_l", "u2": "d2"}))


# ---------------------------------------------------------------- sweep
def _normalize_tensor(T: Tensor) -> Tensor:
    """_ctm_tensor.py:767-780: data / (max|data| + EPS)."""
    if isinstance(T, SymmetricTensor):
        d = T.data
        m = d.abs().max() if d.numel() else torch.zeros((), dtype=d.dtype, device=d.device)
        return T._with_data(d / (m + EPS))
    d = T.todense()
    return DenseTensor(d / (d.abs().max() + EPS), T.indices)


def _renormalize_tensor_env(env: CTMTensorEnv) -> CTMTensorEnv:
    return CTMTensorEnv(*(_normalize_tensor(t) for t in env))


def _ctm_tensor_sweep(env, a, chi, renormalize, projector_method="eigh"):
    """_ctm_tensor.py:796-810: left, right, top, bottom (+ renormalise)."""
    env = _ctm_tensor_move_left(env, env, a, chi, projector_method)
    env = _ctm_tensor_move_right(env, env, a, chi, projector_method)
    env = _ctm_tensor_move_top(env, env, a, chi, projector_method)
    env = _ctm_tensor_move_bottom(env, env, a, chi, projector_method)
    return _renormalize_tensor_env(env) if renormalize else env


def _ctm_sv_diff(sv_new, sv_old):
    """_ctm_tensor.py:861-865."""
    return torch.max(torch.abs(sv_new / (torch.sum(sv_new) + 1e-15) - sv_old / (torch.sum(sv_old) + 1e-15)))


def _corner_sv(env):
    return torch.linalg.svdvals(env.C1.todense().detach())


def ctm_tensor(A: Tensor, chi: int, max_iter: int = 100, conv_tol: float = 1e-8, renormalize: bool = True,
               projector_method: str = "eigh", qr_warmup_steps: int = 3) -> CTMTensorEnv:
    """_ctm_tensor.py:868-917."""
    if projector_method not in ("eigh", "qr"):
        raise ValueError(f"Unknown projector_method={projector_method!r}; expected 'eigh' or 'qr'.")
    with torch.no_grad():
        a = _build_double_layer_tensor(A)
        env = initialize_ctm_tensor_env(A, chi)
        if projector_method == "qr" and qr_warmup_steps > 0:
            w = min(qr_warmup_steps, max_iter)
            for _ in range(w):
                env = _ctm_tensor_sweep(env, a, chi, renormalize, "eigh")
            max_iter -= w
        prev = None
        for _ in range(max_iter):
            env = _ctm_tensor_sweep(env, a, chi, renormalize, projector_method)
            sv = _corner_sv(env)
            if prev is not None and float(_ctm_sv_diff(sv, prev)) < conv_tol:
                break
            prev = sv
    return env


# ---------------------------------------------------------------- gauge fix + AD fixed point
def _qr_pos(C):
    """QR with R's diagonal non-negative.  The reference takes the raw library
    factors, whose signs differ between LAPACK and rocSOLVER; the CTM's sign
    sensitivity makes a fixed convention necessary for reproducible results."""
    Q, R = torch.linalg.qr(C)
    d = torch.sign(torch.diagonal(R)).detach()
    d = torch.where(d == 0, torch.ones_like(d), d)
    return Q * d[None, :], R * d[:, None]


def _gauge_fix_dense(C1, C2, C3, C4, T1, T2, T3, T4):
    """ad_utils.py:190-229 on dense arrays (differentiable torch ops)."""
    Q1, R1 = _qr_pos(C1)
    T1n = torch.einsum("ab,bdc->adc", Q1.T, T1)
    T4n = torch.einsum("ab,bdc->adc", Q1.T, T4)
    Q2, R2 = _qr_pos(C2)
    T1n = torch.einsum("adb,bc->adc", T1n, Q2)
    T2n = torch.einsum("ab,bdc->adc", Q2.T, T2)
    Q3, R3 = _qr_pos(C3)
    T2n = torch.einsum("adb,bc->adc", T2n, Q3)
    T3n = torch.einsum("adb,bc->adc", T3, Q3)
    Q4, R4 = _qr_pos(C4)
    T3n = torch.einsum("ab,bdc->adc", Q4.T, T3n)
    T4n = torch.einsum("adb,bc->adc", T4n, Q4)
    return R1, R2, R3, R4, T1n, T2n, T3n, T4n


def _gauge_fix_ctm_tensor(env: CTMTensorEnv) -> CTMTensorEnv:
    """ad_utils.py:597-631: densify, QR gauge fix, re-wrap with the same legs."""
    fixed = _gauge_fix_dense(*(t.todense() for t in env))

    def wrap(data, orig):
        if isinstance(orig, SymmetricTensor):
            return SymmetricTensor.from_dense(data, orig.indices, tol=float("inf"))
        return DenseTensor(data, orig.indices)

    return CTMTensorEnv(*(wrap(d, o) for d, o in zip(fixed, env)))


def _leaves(x) -> list[torch.Tensor]:
    ts = x if isinstance(x, (tuple, list)) else (x,)
    return [t.data if isinstance(t, SymmetricTensor) else t.todense() for t in ts]


def _with_leaves(template, leaves):
    out = []
    for t, d in zip(template, leaves):
        out.append(t._with_data(d) if isinstance(t, SymmetricTensor) else DenseTensor(d, t.indices))
    return out


def _ctm_tensor_step(A: Tensor, env: CTMTensorEnv, config: CTMConfig) -> CTMTensorEnv:
    """ad_utils.py:634-656: one sweep + gauge fix."""
    a = _build_double_layer_tensor(A)
    env = _ctm_tensor_sweep(env, a, config.chi, config.renormalize, config.projector_method)
    return _gauge_fix_ctm_tensor(env)


def _ctm_tensor_fixed_point_impl(A: Tensor, config: CTMConfig):
    """ad_utils.py:659-685 (no QR warm-up, gauge fix every iteration)."""
    with torch.no_grad():
        env = initialize_ctm_tensor_env(A, config.chi)
        prev = None
        n = 0
        for n in range(1, config.max_iter + 1):
            env = _ctm_tensor_step(A, env, config)
            sv = _corner_sv(env)
            if prev is not None and float(_ctm_sv_diff(sv, prev)) < config.conv_tol:
                break
            prev = sv
    return env, n


def _flat(ts):
    return torch.cat([t.reshape(-1) for t in ts])


def _unflat(v, like):
    out, o = [], 0
    for t in like:
        n = t.numel()
        out.append(v[o:o + n].reshape(t.shape))
        o += n
    return out


def _gmres(matvec, b, x0, tol, restart=20, maxiter=50, progress=False):
    """Restarted GMRES (Saad-Schultz, modified Gram-Schmidt Arnoldi, Givens
    rotations); stops when ||b - A x|| <= tol ||b||, at most ``maxiter``
    restart cycles of ``restart`` Krylov vectors (the interface of
    jax.scipy.sparse.linalg.gmres used at ad_utils.py:745-752)."""
    bnorm = float(torch.linalg.vector_norm(b))
    x = x0.clone()
    if bnorm == 0.0:
        return torch.zeros_like(b), 0
    n_mv = 0
    for _ in range(maxiter):
        r = b - matvec(x)
        n_mv += 1
        beta = float(torch.linalg.vector_norm(r))
        if beta <= tol * bnorm:
            break
        V = [r / beta]
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k_used = 0
        for k in range(restart):
            w = matvec(V[k])
            n_mv += 1
            for i in range(k + 1):
                H[i, k] = float(torch.dot(w, V[i]))
                w = w - H[i, k] * V[i]
            H[k + 1, k] = float(torch.linalg.vector_norm(w))
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (1.0, 0.0) if den == 0.0 else (H[k, k] / den, H[k + 1, k] / den)
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k_used = k + 1
            if abs(g[k + 1]) <= tol * bnorm:
                break
            nv = float(torch.linalg.vector_norm(w))
            if nv == 0.0:
                break
            V.append(w / nv)
        y = np.linalg.solve(np.triu(H[:k_used, :k_used]), g[:k_used]) if k_used else np.zeros(0)
        for i in range(k_used):
            x = x + float(y[i]) * V[i]
        if progress:
            import sys

            print(f"[gmres] cycle done: {n_mv} products, residual {abs(g[k_used]) / bnorm:.3e}", file=sys.stderr,
                  flush=True)
        if abs(g[k_used]) <= tol * bnorm:
            break
    return x, n_mv


# statistics of the most recent backward (GMRES products, residual)
LAST_BACKWARD: dict = {}
# GMRES settings of the backward: Krylov size per restart and a cap on the
# number of restart cycles (None = the reference's min(max_iter, 50)); with
# ``progress`` a line per cycle goes to stderr.
GMRES_OPTIONS: dict = {"restart": 20, "maxiter": None, "progress": False}


class _CTMFixedPoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a_leaf, A_struct, config, recompute, box):
        A = _with_leaves([A_struct], [a_leaf.detach()])[0]
        env, n = _ctm_tensor_fixed_point_impl(A, config)
        ctx.meta = (A, env, config, recompute)
        box["env"] = env
        LAST_BACKWARD["forward_iterations"] = n
        return tuple(_leaves(env))

    @staticmethod
    def backward(ctx, *g):
        A, env, config, recompute = ctx.meta
        env_leaves = [x.detach() for x in _leaves(env)]
        g = [torch.zeros_like(e) if gi is None else gi for gi, e in zip(g, env_leaves)]

        def build():
            e_in = [x.clone().requires_grad_(True) for x in env_leaves]
            a_in = _leaves([A])[0].detach().clone().requires_grad_(True)
            with torch.enable_grad():
                out = _ctm_tensor_step(_with_leaves([A], [a_in])[0], CTMTensorEnv(*_with_leaves(env, e_in)), config)
                out_l = _leaves(out)
            return a_in, e_in, out_l

        state = {"graph": None if recompute else build()}

        def jt(vflat):
            gr = state["graph"] if state["graph"] is not None else build()
            _, e_in, out_l = gr
            vs = _unflat(vflat, env_leaves)
            res = torch.autograd.grad(out_l, e_in, vs, retain_graph=True, allow_unused=True)
            return _flat([torch.zeros_like(e) if r is None else r for r, e in zip(res, env_leaves)])

        b = _flat(g)
        cap = GMRES_OPTIONS["maxiter"]
        lam, n_mv = _gmres(lambda v: v - jt(v), b, b, config.conv_tol, restart=GMRES_OPTIONS["restart"],
                           maxiter=min(config.max_iter, 50) if cap is None else cap,
                           progress=GMRES_OPTIONS["progress"])
        resid = float(torch.linalg.vector_norm(b - (lam - jt(lam)))) / max(float(torch.linalg.vector_norm(b)), 1e-300)
        gr = state["graph"] if state["graph"] is not None else build()
        a_in, _, out_l = gr
        (dA,) = torch.autograd.grad(out_l, [a_in], _unflat(lam, env_leaves), allow_unused=True)
        LAST_BACKWARD.update(gmres_products=n_mv, relative_residual=resid)
        return (torch.zeros_like(a_in) if dA is None else dA), None, None, None, None


def ctm_tensor_converge(A: Tensor, config: CTMConfig, recompute: bool = False) -> CTMTensorEnv:
    """ad_utils.py:688-764: converged, gauge-fixed environment; differentiable
    in A through implicit differentiation of the fixed point."""
    box: dict = {}
    leaves = _CTMFixedPoint.apply(_leaves([A])[0], A, config, recompute, box)
    return CTMTensorEnv(*_with_leaves(box["env"], leaves))


# ---------------------------------------------------------------- energy
def _build_double_layer_open_dense(A: torch.Tensor) -> torch.Tensor:
    D, d = A.shape[0], A.shape[4]
    ao = torch.einsum("udlrs,UDLRt->uUdDlLrRst", A, A.conj())
    return ao.reshape(D * D, D * D, D * D, D * D, d, d)


def _rdm2x1(A, env, d):
    """ipeps.py:1384-1447."""
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    ao = _build_double_layer_open_dense(A)
    UL = torch.einsum("ab,buc->auc", C1, T1)
    UR = torch.einsum("cuf,fg->cug", T1, C2)
    LL = torch.einsum("gi,idj->gdj", C4, T3)
    LR = torch.einsum("jdk,mk->jdm", T3, C3)
    Lenv = torch.einsum("auc,axg,gdj->ucxdj", UL, T4, LL)
    Renv = torch.einsum("cuf,frm,jdm->curjd", UR, T2, LR)
    L1 = torch.einsum("ucxdj,udxrst->crjst", Lenv, ao)
    R2 = torch.einsum("curjd,udlrtv->cjltv", Renv, ao)
    rdm = torch.einsum("crjst,cjruv->stuv", L1, R2).permute(0, 2, 1, 3)
    m = rdm.reshape(d * d, d * d)
    m = 0.5 * (m + m.T)
    m = m / (torch.trace(m) + 1e-15)
    return m.reshape(d, d, d, d)


def _rdm1x2(A, env, d):
    """ipeps.py:1450-1516."""
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    ao = _build_double_layer_open_dense(A)
    top = torch.einsum("ab,buc,ce->aue", C1, T1, C2)
    row1 = torch.einsum("aue,alf,erg->ulfrg", top, T4, T2)
    site1 = torch.einsum("ulfrg,udlrst->dfgst", row1, ao)
    T4ao = torch.einsum("fmh,pqmnwx->fhpqnwx", T4, ao)
    s12 = torch.einsum("abcst,bhaqnwx->chqnstwx", site1, T4ao)
    s12r = torch.einsum("chqnstwx,cni->hqistwx", s12, T2)
    bot = torch.einsum("hj,jqk,ik->hqi", C4, T3, C3)
    rdm = torch.einsum("hqistwx,hqi->stwx", s12r, bot).permute(0, 2, 1, 3)
    m = rdm.reshape(d * d, d * d)
    m = 0.5 * (m + m.T)
    m = m / (torch.trace(m) + 1e-15)
    return m.reshape(d, d, d, d)


def compute_energy_ctm(A, env, hamiltonian_gate, d):
    """ipeps.py:1519-1549 on dense arrays: Tr(rho_h H) + Tr(rho_v H)."""
    H = torch.as_tensor(hamiltonian_gate, dtype=torch.float64, device=A.device).reshape(d, d, d, d)
    return torch.einsum("ijkl,ijkl->", _rdm2x1(A, env, d), H) + torch.einsum("ijkl,ijkl->", _rdm1x2(A, env, d), H)


def compute_energy_ctm_tensor(A: Tensor, env: CTMTensorEnv, hamiltonian_gate, d: int | None = None):
    """_ctm_tensor.py:1037-1069: densify, then the dense RDM energy."""
    Ad = A.todense()
    if d is None:
        d = Ad.shape[-1]
    H = hamiltonian_gate.todense() if isinstance(hamiltonian_gate, Tensor) else hamiltonian_gate
    return compute_energy_ctm(Ad, [t.todense() for t in env], H, d)


def heisenberg_gate(device=None) -> torch.Tensor:
    """Sz Sz + (S+ S- + S- S+)/2, shape (2, 2, 2, 2)."""
    sz = torch.diag(torch.tensor([0.5, -0.5], dtype=torch.float64))
    sp = torch.tensor([[0.0, 1.0], [0.0, 0.0]], dtype=torch.float64)
    sm = sp.T.contiguous()
    H = torch.kron(sz, sz) + 0.5 * (torch.kron(sp, sm) + torch.kron(sm, sp))
    return H.reshape(2, 2, 2, 2).to(device or torch.device("cuda", torch.cuda.current_device()))


def ipeps_site_tensor(data, symmetric: bool = False) -> Tensor:
    """Wrap a (D, D, D, D, d) array with trivial U(1) legs (u OUT, d IN, l OUT,
    r IN, phys IN), the reference tests' convention, as Dense- or SymmetricTensor."""
    t = torch.as_tensor(data, dtype=torch.float64)
    D, d = t.shape[0], t.shape[4]
    sym = U1Symmetry()
    z = np.zeros(D, np.int32)
    idx = (TensorIndex(sym, z, OUT, "u"), TensorIndex(sym, z, IN, "d"), TensorIndex(sym, z, OUT, "l"),
           TensorIndex(sym, z, IN, "r"), TensorIndex(sym, np.zeros(d, np.int32), IN, "phys"))
    return SymmetricTensor.from_dense(t, idx) if symmetric else DenseTensor(t, idx)


# ---------------------------------------------------------------- AD optimisation
def energy_loss(A: Tensor, gate, config: CTMConfig, recompute: bool = False):
    """optimize_gs_ad's loss (ipeps.py:2405-2410): normalise, converge, energy."""
    An = A * (1.0 / (A.norm() + 1e-10))
    env = ctm_tensor_converge(An, config, recompute)
    return compute_energy_ctm_tensor(An, env, gate)


def optimize_gs_ad(hamiltonian_gate, A_init, config: iPEPSConfig):
    """ipeps.py:2265-2434 (1-site): Adam (optax defaults b1=0.9, b2=0.999,
    eps=1e-8) on A, renormalised after every update; returns (A, env, E)."""
    if config.unit_cell != "1x1":
        raise NotImplementedError("only the 1x1 unit cell is on the accelerated path")
    gate = torch.as_tensor(hamiltonian_gate, dtype=torch.float64)
    if A_init is None:
        from tenax_amd.core.tensor import PRNGKey, normal

        D, d = config.max_bond_dim, gate.shape[0]
        A_init = normal(PRNGKey(0), (D, D, D, D, d))
    A = A_init if isinstance(A_init, Tensor) else ipeps_site_tensor(A_init)
    gate = gate.to(_leaves([A])[0].device)
    A = A * (1.0 / (A.norm() + 1e-10))
    x = _leaves([A])[0].detach().clone()
    m = torch.zeros_like(x)
    v = torch.zeros_like(x)
    lr, b1, b2, eps = config.gs_learning_rate, 0.9, 0.999, 1e-8
    best_e, best_x, prev = float("inf"), x.clone(), float("inf")
    for step in range(config.gs_num_steps):
        xa = x.clone().requires_grad_(True)
        E = energy_loss(_with_leaves([A], [xa])[0], gate, config.ctm)
        (g,) = torch.autograd.grad(E, [xa])
        e = float(E)
        if e < best_e:
            best_e, best_x = e, x.clone()
        if abs(e - prev) < config.gs_conv_tol:
            break
        prev = e
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** (step + 1))
        vh = v / (1 - b2 ** (step + 1))
        x = x - lr * mh / (torch.sqrt(vh) + eps)
        x = x / (torch.linalg.vector_norm(x) + 1e-10)
    Af = _with_leaves([A], [best_x])[0]
    Af = Af * (1.0 / (Af.norm() + 1e-10))
    with torch.no_grad():
        env = ctm_tensor_converge(Af, config.ctm)
        E = float(compute_energy_ctm_tensor(Af, env, gate))
    return Af, env, E
