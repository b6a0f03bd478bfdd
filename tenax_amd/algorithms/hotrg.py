"""Higher-order TRG for 2D classical lattice models (reference
src/tenax/algorithms/hotrg.py:54-215), for one instance or many in lockstep.

Per step (horizontal; the vertical step is the same with the leg roles
exchanged):

1. traced environment ``M[u,d,U,D] = sum_{l,r} T[u,d,l,r] T[U,D,r,l]``
   (hotrg.py:134-138);
2. its truncated SVD over (up, U) | (down, D): isometry U for the up pair and
   Vh for the down pair (hotrg.py:140-149), keeping the SVD's pairing of the two;
3. ``T_new[a,b,l,r] = sum U*[u,U,a] T[u,d,l,k] T[U,D,k,r] Vh*[b,d,D]``
   (hotrg.py:151-166), then max-abs normalisation (log norm / 2^(step+1)
   accumulates into ln Z/N).

Step 3 is evaluated as ``(U* . T_left) . T_right`` followed by Vh*: the same
sum as the reference's ``U* . (T_left . T_right)``, with a chi^5 instead of a
chi^6 intermediate (at chi = 48 with Z2 that is 1 GB instead of 49 GB).
All contractions are block-sparse contractions of the device engine.  For
SymmetricTensor inputs the SVD is the batched Jordan-Wielandt eigh
factorisation (``absorb="isometries"`` in tenax_amd.linalg: both isometries
with the SVD's pairing and LAPACK-level accuracy for small singular values),
and ``hotrg_batched`` runs every instance's step in lockstep: the sector
eigenproblems of all instances share one batched eigh per step, and each
block contraction of the step is one grouped-GEMM launch over the block pairs
of every instance (contract_many).
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import torch

from tenax_amd.algorithms._tensor_utils import max_abs_normalize
from tenax_amd.contraction import contract_many
from tenax_amd.core.tensor import SymmetricTensor, Tensor
from tenax_amd.linalg import truncated_svd, truncated_svd_symmetric_many


@dataclass
class HOTRGConfig:
    """hotrg.py:36-51."""

    max_bond_dim: int = 16
    num_steps: int = 10
    direction_order: str = "alternating"
    svd_trunc_err: float | None = None


_VALID = ("alternating", "horizontal", "vertical")

# Isometries from the leading max_bond_dim singular triplets of each sector
# only (block subspace iteration, tenax_amd.linalg._leading_isometries) when
# the cut is the bond dimension alone; False: the full Jordan-Wielandt eigh.
LEADING = True

# Optional wall-clock accounting of the step phases (device-synchronised at
# phase boundaries while enabled; used by the benchmark).
TIMERS: dict = {"enabled": False}


def _tick(name, t0):
    if not TIMERS["enabled"]:
        return 0.0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if t0:
        TIMERS[name] = TIMERS.get(name, 0.0) + t1 - t0
    return t1


# leg roles per direction: (pair-left legs, pair-right legs, environment
# relabels, merge: first copy relabel, second copy relabels, output labels)
_H = dict(env=({"up": "U", "down": "D", "left": "right", "right": "left"}),
          left=["up", "U"], right=["down", "D"],
          first=("right", "k"), second={"up": "U", "down": "D", "left": "k"},
          out=("a", "b", "left", "right"), back={"a": "up", "b": "down"})
_V = dict(env=({"left": "L", "right": "R", "up": "down", "down": "up"}),
          left=["left", "L"], right=["right", "R"],
          first=("down", "k"), second={"left": "L", "right": "R", "up": "k"},
          out=("up", "down", "a", "b"), back={"a": "left", "b": "right"})


def _isometries(Ms, roles, max_bond_dim, svd_trunc_err):
    if all(isinstance(M, SymmetricTensor) for M in Ms):
        res = truncated_svd_symmetric_many(Ms, roles["left"], roles["right"], "a", max_bond_dim, svd_trunc_err,
                                           absorb="isometries", leading=LEADING)
        return [(U, Vh) for U, _, Vh, _ in res]
    out = []
    for M in Ms:
        U, _, Vh, _ = truncated_svd(M, roles["left"], roles["right"], "a", max_singular_values=max_bond_dim,
                                    max_truncation_err=svd_trunc_err)
        out.append((U, Vh))
    return out


def _next_a_layout(P, Q, nxt, nxt_out=None):
    """Leg order for the result of contract(P, Q) that the contraction with
    ``nxt`` takes as its A operand without a transpose: the legs it keeps
    (in the order of that contraction's output ``nxt_out``, or their own),
    then the legs it contracts with ``nxt``."""
    lp, lq = P.labels(), Q.labels()
    labels = [l for l in lp if l not in lq] + [l for l in lq if l not in lp]
    other = set(nxt.labels())
    cont = [l for l in labels if l in other]
    keep = [l for l in labels if l not in other]
    if nxt_out is not None:
        keep = [l for l in nxt_out if l in keep]
    return tuple(keep + cont)


def _steps(Ts, roles, max_bond_dim, svd_trunc_err):
    """One coarse-graining step of every instance, in lockstep: each
    contraction runs for all instances at once (contract_many: one grouped
    block-GEMM launch over every instance's block pairs, one launch per leg
    permutation phase).  The chi^5 intermediates are written by the grouped
    GEMM directly in the leg order the next contraction reads (no transpose
    of them)."""
    t = _tick("", 0.0)
    Ms = contract_many([(T, T.relabels(roles["env"])) for T in Ts])
    t = _tick("environment", t)
    isos = _isometries(Ms, roles, max_bond_dim, svd_trunc_err)
    t = _tick("svd", t)
    del Ms
    Uds = [U.dagger() for U, _ in isos]
    Vbs = [Vh.relabel("a", "b").dagger() for _, Vh in isos]
    out = []
    firsts = [T.relabel(*roles["first"]) for T in Ts]
    seconds = [T.relabels(roles["second"]) for T in Ts]
    for lo, hi in _chunks(list(zip(Uds, firsts))):
        x_order = _next_a_layout(Uds[lo], firsts[lo], seconds[lo])
        X = contract_many(list(zip(Uds[lo:hi], firsts[lo:hi])), output_labels=x_order)
        y_order = _next_a_layout(X[0], seconds[lo], Vbs[lo], roles["out"])
        Y = contract_many(list(zip(X, seconds[lo:hi])), output_labels=y_order)
        del X
        Tn = contract_many(list(zip(Y, Vbs[lo:hi])), output_labels=roles["out"])
        del Y
        out.extend(max_abs_normalize(T.relabels(roles["back"])) for T in Tn)
    _tick("coarse_grain", t)
    return out


# Instances whose chi^5 intermediates are contracted together: at most this
# many elements of the dense bound (the stored Z2 blocks hold about half),
# so that the intermediates of one group and their permuted copies stay
# within a few tens of GB of HBM.
CHUNK_ELEMS = 1 << 32


def _chunks(pairs):
    """[lo, hi) ranges of instances whose first intermediates (the dense size
    of the open legs of each pair) fit CHUNK_ELEMS together; every range holds
    at least one instance."""
    out, lo, acc = [], 0, 0
    for i, (A, B) in enumerate(pairs):
        la = {ix.label: ix.dim for ix in A.indices}
        lb = {ix.label: ix.dim for ix in B.indices}
        n = 1
        for lbl, d in list(la.items()) + list(lb.items()):
            if (lbl in la) != (lbl in lb):
                n *= d
        if i > lo and acc + n > CHUNK_ELEMS:
            out.append((lo, i))
            lo, acc = i, 0
        acc += n
    out.append((lo, len(pairs)))
    return out


def _hotrg_step_horizontal(T: Tensor, max_bond_dim: int, svd_trunc_err: float | None = None):
    """hotrg.py:109-162: compresses (up, U) and (down, D)."""
    return _steps([T], _H, max_bond_dim, svd_trunc_err)[0]


def _hotrg_step_vertical(T: Tensor, max_bond_dim: int, svd_trunc_err: float | None = None):
    """hotrg.py:165-215: compresses (left, L) and (right, R)."""
    return _steps([T], _V, max_bond_dim, svd_trunc_err)[0]


def _check(tensor, config):
    if config.direction_order not in _VALID:
        raise ValueError(f"Invalid direction_order {config.direction_order!r}. Must be one of {_VALID}.")
    if not isinstance(tensor, Tensor):
        raise TypeError(f"hotrg() requires a Tensor, got {type(tensor).__name__}")


def _roles(config, step):
    if config.direction_order == "alternating":
        return _H if step % 2 == 0 else _V
    return _H if config.direction_order == "horizontal" else _V


def hotrg_batched(tensors, config: HOTRGConfig, log_norms: list | None = None) -> list[float]:
    """ln Z / N for every instance (e.g. one Ising tensor per temperature).

    The instances advance in lockstep; per step their environments are
    factorised together (one batched eigh for all Z2 sectors of all
    instances).  ``log_norms``, if given, receives the per-step log norms of
    each instance.
    """
    for t in tensors:
        _check(t, config)
    Ts = list(tensors)
    totals = [0.0] * len(Ts)
    hist = [[] for _ in Ts]
    for step in range(config.num_steps):
        res = _steps(Ts, _roles(config, step), config.max_bond_dim, config.svd_trunc_err)
        for i, (T, ln) in enumerate(res):
            Ts[i] = T
            totals[i] += ln / 2.0 ** (step + 1)
            hist[i].append(ln)
    if log_norms is not None:
        log_norms.extend(hist)
    return totals


def hotrg(tensor: Tensor, config: HOTRGConfig):
    """ln Z / N estimate (hotrg.py:54-106) as a 0-d float64 tensor."""
    _check(tensor, config)
    return torch.tensor(hotrg_batched([tensor], config)[0], dtype=torch.float64)
