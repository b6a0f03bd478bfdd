"""Oracle: Z2 / dense HOTRG for the 2D Ising model, FP64 NumPy.

Restates reference src/tenax/algorithms/hotrg.py:54-215 (the coarse-graining
schedule and both steps), _tensor_utils.py:65-80 (max-abs normalisation with
LOG_EPS = 1e-250, core/__init__.py:16), trg.py:173-232 (Ising site tensor,
Z2 eigenbasis) and trg.py:245-291 (Onsager free energy).

Symmetric tensors run through the oracle's own block contraction
(oracle.contract, contractor.py:291-484 semantics) and block SVD
(oracle.linalg, linalg.py:75-333).  ``order="reference"`` contracts in the
reference's order (the chi^6 merged pair first, hotrg.py:151-166);
``order="chi5"`` applies the first isometry before merging (same sum, chi^5
memory), which is the only order that fits chi >= 32 on a host.

The reference compresses the up pair with U and the down pair with Vh^dagger
of the SVD of the traced environment M; the bond between two coarse tensors
then carries conj(U) Vh, which is the identity only in special cases (M
symmetric positive semi-definite, as for the first Z2 Ising steps; a
symmetric M with a negative eigenvalue gives a sign there).  So untruncated
reference steps are not exact in general.  ``paired=True`` uses U for
both pairs (textbook HOTRG, exact without truncation) and exists to pin the
contraction machinery.

Pinned by tests/test_oracle_hotrg.py: brute-force partition functions of
2x1, 2x2, 4x2 and 4x4 tori (untruncated paired steps + final trace, and the
reference form for the first two Z2 steps), Onsager's closed form at T_c and its
high-temperature series, the Z2 tensor's 8 blocks, and the reference's own
accuracy tolerances.
Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import numpy as np

from oracle.contract import contract
from oracle.core import IN, OUT, Leg, SymT, U1, Zn
from oracle.linalg import truncated_svd_symmetric

LOG_EPS = 1e-250


# ---------------------------------------------------------------- site tensor
def ising_tensor_dense(beta, J=1.0):
    """trg.py:200-214: T_udlr = sum_s W[u,s] W[d,s] W[l,s] W[r,s], W = Q^(1/2)."""
    spins = np.array([1.0, -1.0])
    Q = np.exp(beta * J * np.outer(spins, spins))
    ev, V = np.linalg.eigh(Q)
    W = V @ np.diag(np.sqrt(ev)) @ V.T
    return np.einsum("us,ds,ls,rs->udlr", W, W, W, W)


def ising_tensor_z2(beta, J=1.0):
    """trg.py:216-232: T' = (H x H x H x H) T in the Z2 eigenbasis, legs
    (up IN, down OUT, left IN, right OUT), charges [0, 1]."""
    T = ising_tensor_dense(beta, J)
    H = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
    Tz = np.einsum("ua,vb,wc,xd,uvwx->abcd", H, H, H, H, T)
    sym = Zn(2)
    ch = np.array([0, 1])
    legs = (Leg(sym, ch, IN, "up"), Leg(sym, ch, OUT, "down"), Leg(sym, ch, IN, "left"), Leg(sym, ch, OUT, "right"))
    return SymT.from_dense(Tz, legs)


def ising_tensor_dense_symt(beta, J=1.0):
    """The dense tensor as a trivially-charged SymT (U(1), all charges 0): one
    block per tensor.  Used to pin the block machinery on exact tori; the
    reference's DenseTensor path is hotrg_dense (its SVD transposes legs first,
    the block SVD does not)."""
    T = ising_tensor_dense(beta, J)
    z = np.zeros(2, np.int32)
    sym = U1()
    legs = (Leg(sym, z, IN, "up"), Leg(sym, z, OUT, "down"), Leg(sym, z, IN, "left"), Leg(sym, z, OUT, "right"))
    return SymT({(0, 0, 0, 0): T}, legs, validate=False)


def onsager_free_energy(beta, J=1.0, n_pts=10000):
    """trg.py:245-291: f = -ln2/beta - I/(2 beta), I = (1/pi) int_0^pi
    ln((A + sqrt(A^2 - s^2))/2) dt, A = cosh(2K)^2 - sinh(2K) cos t, on the
    half-offset grid with the trapezoid rule."""
    if beta == 0:
        return -float(np.log(2.0))
    K = beta * J
    c2, s2 = np.cosh(2 * K), np.sinh(2 * K)
    dt = np.pi / n_pts
    t = np.linspace(dt / 2, np.pi - dt / 2, n_pts)
    A = c2 ** 2 - s2 * np.cos(t)
    integrand = np.log((A + np.sqrt(np.maximum(A ** 2 - s2 ** 2, 0.0))) / 2)
    integral = float(np.trapezoid(integrand, t)) / np.pi
    return -float(np.log(2.0)) / beta - integral / (2 * beta)


# ---------------------------------------------------------------- steps
def max_abs_normalize(T):
    """_tensor_utils.py:65-80."""
    m = T.max_abs()
    return T * (1.0 / (m + LOG_EPS)), float(np.log(m + LOG_EPS))


def _paired_vh(U, relabel):
    """Vh := U^dagger laid out as (a, pair...) on the other pair of legs."""
    Ud = U.dagger().relabels(relabel)
    return Ud.transpose((2, 0, 1))


def step_horizontal(T, chi, trunc_err=None, order="reference", paired=False, driver="gesdd"):
    """hotrg.py:109-162."""
    Tc = T.relabels({"up": "U", "down": "D", "left": "right", "right": "left"})
    M = contract(T, Tc)
    U, _, Vh, _, _ = truncated_svd_symmetric(M, ["up", "U"], ["down", "D"], "a", chi, trunc_err, driver)
    if paired:
        Vh = _paired_vh(U, {"up": "down", "U": "D"})
    Ud = U.dagger()
    Vb = Vh.relabel("a", "b").dagger()
    Tl = T.relabel("right", "k")
    Tr = T.relabels({"up": "U", "down": "D", "left": "k"})
    if order == "reference":
        Tm = contract(Tl, Tr)
        Tt = contract(Ud, Tm)
    else:
        Tt = contract(contract(Ud, Tl), Tr)
    Tn = contract(Tt, Vb, output_labels=("a", "b", "left", "right"))
    return max_abs_normalize(Tn.relabels({"a": "up", "b": "down"}))


def step_vertical(T, chi, trunc_err=None, order="reference", paired=False, driver="gesdd"):
    """hotrg.py:165-215."""
    Tc = T.relabels({"left": "L", "right": "R", "up": "down", "down": "up"})
    M = contract(T, Tc)
    U, _, Vh, _, _ = truncated_svd_symmetric(M, ["left", "L"], ["right", "R"], "a", chi, trunc_err, driver)
    if paired:
        Vh = _paired_vh(U, {"left": "right", "L": "R"})
    Ud = U.dagger()
    Vb = Vh.relabel("a", "b").dagger()
    Tt_ = T.relabel("down", "k")
    Tb = T.relabels({"left": "L", "right": "R", "up": "k"})
    if order == "reference":
        Tm = contract(Tt_, Tb)
        Tt = contract(Ud, Tm)
    else:
        Tt = contract(contract(Ud, Tt_), Tb)
    Tn = contract(Tt, Vb, output_labels=("up", "down", "a", "b"))
    return max_abs_normalize(Tn.relabels({"a": "left", "b": "right"}))


def hotrg(T, chi, num_steps, direction_order="alternating", trunc_err=None, order="reference",
          return_tensor=False, paired=False, driver="gesdd"):
    """hotrg.py:54-106: ln Z / N ~= sum_k log_norm_k / 2^(k+1)."""
    if direction_order not in ("alternating", "horizontal", "vertical"):
        raise ValueError(f"Invalid direction_order {direction_order!r}")
    total = 0.0
    norms = []
    for step in range(num_steps):
        horiz = (step % 2 == 0) if direction_order == "alternating" else direction_order == "horizontal"
        fn = step_horizontal if horiz else step_vertical
        T, ln = fn(T, chi, trunc_err, order, paired, driver)
        norms.append(ln)
        total += ln / 2.0 ** (step + 1)
    return (total, T, norms) if return_tensor else total


# ---------------------------------------------------------------- dense path
def _dense_svd_iso(mat4, chi, trunc_err=None):
    """linalg.py:745-825 for a dense 4-leg M: transpose to (l0, l1 | r0, r1),
    SVD, keep min(chi, n) (at least 1, with the truncation-error rule)."""
    a, b, c, d = mat4.shape
    U, s, Vh = np.linalg.svd(mat4.reshape(a * b, c * d), full_matrices=False)
    k = len(s)
    if trunc_err is not None:
        tot = float(np.sum(s ** 2))
        acc = 0.0
        for i in range(len(s) - 1, 0, -1):
            acc += float(s[i]) ** 2
            if acc / tot > trunc_err ** 2:
                k = i + 1
                break
    k = max(1, min(k, chi))
    return U[:, :k].reshape(a, b, k), Vh[:k].reshape(k, c, d)


def step_dense(T, chi, horizontal, trunc_err=None):
    """One dense HOTRG step (hotrg.py:109-215 on a DenseTensor): returns
    (normalised T_new, log norm)."""
    if horizontal:
        # T_new[a,b,l,r] = U[u,U,a] T[u,d,l,k] T[U,D,k,r] Vh[b,d,D]
        M = np.einsum("udlr,UDrl->uUdD", T, T)
        U, Vh = _dense_svd_iso(M, chi, trunc_err)
        X = np.tensordot(U, T, axes=([0], [0]))             # U a d l k
        Y = np.tensordot(X, T, axes=([0, 4], [0, 2]))       # a d l D r
        Tn = np.tensordot(Y, Vh, axes=([1, 3], [1, 2])).transpose(0, 3, 1, 2)
    else:
        # T_new[u,d,a,b] = U[l,L,a] T[u,k,l,r] T[k,d,L,R] Vh[b,r,R]
        M = np.einsum("udlr,duLR->lLrR", T, T)
        U, Vh = _dense_svd_iso(M, chi, trunc_err)
        X = np.tensordot(U, T, axes=([0], [2]))             # L a u k r
        Y = np.tensordot(X, T, axes=([0, 3], [2, 0]))       # a u r d R
        Tn = np.tensordot(Y, Vh, axes=([2, 4], [1, 2])).transpose(1, 2, 0, 3)
    m = float(np.abs(Tn).max())
    return Tn / (m + LOG_EPS), float(np.log(m + LOG_EPS))


def hotrg_dense(T, chi, num_steps, direction_order="alternating", trunc_err=None):
    """hotrg.py:54-106 on a DenseTensor (dense contractions, dense SVD with the
    legs transposed into (left | right) order first, linalg.py:745-750)."""
    total = 0.0
    for step in range(num_steps):
        horiz = (step % 2 == 0) if direction_order == "alternating" else direction_order == "horizontal"
        T, ln = step_dense(T, chi, horiz, trunc_err)
        total += ln / 2.0 ** (step + 1)
    return total


# ---------------------------------------------------------------- pinning helpers
def trace_final(T):
    """Tr over up=down and left=right of a 4-leg SymT (periodic closure)."""
    d = T.todense()
    return float(np.einsum("uull->", d))


def torus_lnz_bruteforce(nx, ny, beta, J=1.0):
    """ln Z of the nx x ny periodic Ising lattice by enumeration (each
    nearest-neighbour pair bond counted once per lattice link; for nx = 2 the
    two horizontal links of a row join the same pair, likewise ny = 2)."""
    n = nx * ny
    s = 1 - 2 * ((np.arange(2 ** n)[:, None] >> np.arange(n)[None, :]) & 1)
    s = s.reshape(-1, ny, nx).astype(np.float64)
    e = (s * np.roll(s, -1, axis=2)).sum(axis=(1, 2)) + (s * np.roll(s, -1, axis=1)).sum(axis=(1, 2))
    x = beta * J * e
    m = x.max()
    return float(m + np.log(np.exp(x - m).sum()))
