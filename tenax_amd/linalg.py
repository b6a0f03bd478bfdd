"""Block-diagonal SVD with global truncation for SymmetricTensor.

Mirrors reference src/tenax/linalg.py:34-333 and 665-825:

* sector q of a block = fused flow*charge of its left legs (linalg.py:34-69);
* each sector matrix is decomposed independently in FP64 on the GPU;
* singular values of all sectors are merged with a STABLE sort by -s
  (sector visit order, then index inside the sector; linalg.py:211), the
  first ``min(n, max_singular_values)`` are kept (at least one), with the
  optional truncation-error rule of linalg.py:220-235;
* the new bond's charges follow the global descending order (linalg.py:242),
  so charge-q positions on the bond are the sector's kept singular vectors in
  order; U blocks are keyed ``lk + (q,)``, Vh blocks ``(q,) + rk``.

Sector visit order here: first appearance in the sorted block order (the
reference visits dict insertion order; only exact ties across sectors can
tell the two apart).  Row/column sub-keys inside a sector are likewise in
sorted order.

Decomposition.  Without ``absorb`` each sector matrix goes through
torch.linalg.svd (rocSOLVER gesvd) in float64.  With ``absorb`` (the DMRG
split) only one factor has to be an isometry: the other one carries the
singular values.  The isometric side then comes from a Gram eigenproblem,
``M M^T = U diag(s^2) U^T`` (absorb="right") or ``M^T M = V diag(s^2) V^T``
(absorb="left"), and the carried side is the exact product ``U^T M`` or
``M V``; s is the row (column) norm of that product.  Eigenvectors are
orthonormal to working precision whatever the spectrum, the carried factor
needs no division by s, and ``U U^T M`` is the optimal rank-k approximation up
to eigenvector error eps*s_max^2/gap, which only matters for the truncation
when two singular values straddling the cut are degenerate to that level.
The eigenproblems of all sectors are solved in one batched torch.linalg.eigh
call: Gram matrices are scaled by their trace and padded to a common size
with -1 on the padded diagonal, so padded eigenpairs sit at -1 and the
sector's own ones in [0, 1] come out exactly decoupled.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

from tenax_amd.core.index import FlowDirection, Label, TensorIndex
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor, Tensor


def _sector_layout(t: SymmetricTensor, nl: int):
    """Group blocks of a tensor whose first nl legs are the row legs."""
    sym = t.indices[0].symmetry
    flows = np.array([int(ix.flow) for ix in t.indices[:nl]], np.int64)
    qs = (t.keys[:, :nl].astype(np.int64) * flows).sum(axis=1)
    n = sym.n_values()
    if n is not None:
        qs = np.mod(qs, n)
    sectors: dict[int, dict] = {}
    for i in range(t.n_blocks):
        q = int(qs[i])
        sec = sectors.setdefault(q, {"rows": {}, "cols": {}, "blocks": []})
        lk = tuple(int(x) for x in t.keys[i, :nl])
        rk = tuple(int(x) for x in t.keys[i, nl:])
        nr = int(np.prod(t.shapes[i, :nl]))
        nc = int(np.prod(t.shapes[i, nl:]))
        if lk not in sec["rows"]:
            sec["rows"][lk] = (nr, tuple(int(x) for x in t.shapes[i, :nl]))
        if rk not in sec["cols"]:
            sec["cols"][rk] = (nc, tuple(int(x) for x in t.shapes[i, nl:]))
        sec["blocks"].append((i, lk, rk))
    for sec in sectors.values():
        r0 = 0
        sec["row_off"] = {}
        for lk, (nr, _) in sec["rows"].items():
            sec["row_off"][lk] = r0
            r0 += nr
        c0 = 0
        sec["col_off"] = {}
        for rk, (nc, _) in sec["cols"].items():
            sec["col_off"][rk] = c0
            c0 += nc
        sec["R"], sec["C"] = r0, c0
    return sectors


def _pack(indices, items, device):
    """items: list of (key tuple, torch block); returns packed SymmetricTensor."""
    items = sorted(items, key=lambda kv: kv[0])
    r = len(indices)
    keys = np.array([k for k, _ in items], np.int32).reshape(len(items), r)
    shapes = np.array([tuple(b.shape) for _, b in items], np.int32).reshape(len(items), r)
    offs, tot = SymmetricTensor._layout(shapes)
    data = torch.cat([b.reshape(-1) for _, b in items]) if items else torch.zeros(0, dtype=torch.float64,
                                                                                   device=device)
    return SymmetricTensor._wrap(indices, keys, shapes, offs, data.contiguous())


# rocSOLVER's batched eigh costs about the same per call for a batch of 1 or
# of 12 (its tridiagonal reduction is a sequence of launches, ~20 us per row),
# so all sectors go into ONE padded batch unless that batch would exceed this.
_MAX_BATCH_BYTES = 1 << 30


def _batches(sizes):
    """Groups of sector indices sharing one padded eigh call, with the pad size."""
    P = -(-max(sizes) // 32) * 32
    if len(sizes) * P * P * 8 <= _MAX_BATCH_BYTES:
        return [(P, list(range(len(sizes))))]
    return [(n, [j]) for j, n in enumerate(sizes)]


def _gram_factors(mats, absorb):
    """Per sector matrix M: (U, s, Vh) with s absorbed on the ``absorb`` side.

    absorb="right": U from eigh(M M^T) (isometry), Vh = U^T M (= diag(s) Vh).
    absorb="left":  Vh from eigh(M^T M) (isometry), U = M V (= U diag(s)).
    Columns come sorted by descending s, s being the norms of the carried
    factor's rows (columns).  min(R, C) values per sector, as the SVD gives.
    """
    row_side = absorb == "right"
    grams = [M @ M.T if row_side else M.T @ M for M in mats]
    dev = mats[0].device
    vecs = [None] * len(mats)
    for P, js in _batches([g.shape[0] for g in grams]):
        stack = torch.zeros((len(js), P, P), dtype=torch.float64, device=dev)
        for b, j in enumerate(js):
            g = grams[j]
            n = g.shape[0]
            scale = torch.diagonal(g).sum().clamp_min(1e-300)
            stack[b, :n, :n] = g / scale
            if n < P:
                pad = torch.arange(n, P, device=dev)
                stack[b, pad, pad] = -1.0
        _, V = torch.linalg.eigh(stack)
        for b, j in enumerate(js):
            n = grams[j].shape[0]
            vecs[j] = V[b, :n, P - n:].flip(-1)
    out = []
    for M, V in zip(mats, vecs):
        k = min(M.shape)
        V = V[:, :k]
        if row_side:
            B = V.T @ M
            s = torch.linalg.vector_norm(B, dim=1)
        else:
            B = M @ V
            s = torch.linalg.vector_norm(B, dim=0)
        s, perm = torch.sort(s, descending=True, stable=True)
        if row_side:
            out.append((V[:, perm], s, B[perm, :]))
        else:
            out.append((B[:, perm], s, V[:, perm].T))
    return out


def truncated_svd_symmetric(tensor: SymmetricTensor, left_labels, right_labels, new_bond_label,
                            max_singular_values=None, max_truncation_err=None, normalize=False, absorb=None):
    """Returns (U, s, Vh, s_full); ``absorb`` in {None, "left", "right"} scales
    U or Vh by s along the new bond (dmrg.py:874-880 done in place)."""
    labels = tensor.labels()
    lax = [labels.index(l) for l in left_labels]
    rax = [labels.index(l) for l in right_labels]
    perm = tuple(lax + rax)
    t = tensor if perm == tuple(range(tensor.ndim)) else tensor.transpose(perm)
    nl = len(lax)
    dev = t.data.device
    sectors = _sector_layout(t, nl)
    results = {}
    s_list, q_list, mats = [], [], []
    for q, sec in sectors.items():
        R, C = sec["R"], sec["C"]
        if R == 0 or C == 0:
            continue
        M = torch.zeros((R, C), dtype=torch.float64, device=dev)
        for i, lk, rk in sec["blocks"]:
            nr = sec["rows"][lk][0]
            nc = sec["cols"][rk][0]
            r0, c0 = sec["row_off"][lk], sec["col_off"][rk]
            M[r0:r0 + nr, c0:c0 + nc] = t.block(i).reshape(nr, nc)
        mats.append(M)
        q_list.append(q)
    if not mats:
        raise ValueError("SVD of a tensor with no blocks")
    if absorb is None:
        for q, M in zip(q_list, mats):
            U, s, Vh = torch.linalg.svd(M, full_matrices=False)
            results[q] = (U, s, Vh)
            s_list.append(s)
    else:
        for q, fac in zip(q_list, _gram_factors(mats, absorb)):
            results[q] = fac
            s_list.append(fac[1])
    s_cat = torch.cat(s_list)
    s_sorted, order = torch.sort(s_cat, descending=True, stable=True)
    s_host = s_sorted.cpu().numpy()
    order = order.cpu().numpy()
    n_total = len(s_host)
    n_keep = n_total
    if max_truncation_err is not None and n_total > 0:
        tot = float(np.sum(s_host ** 2))
        if tot > 0:
            acc = 0.0
            for i in range(n_total - 1, 0, -1):
                acc += float(s_host[i]) ** 2
                if acc / tot > max_truncation_err ** 2:
                    n_keep = i + 1
                    break
    if max_singular_values is not None:
        n_keep = min(n_keep, int(max_singular_values))
    n_keep = max(1, min(n_keep, n_total))
    # map global position -> (sector, index)
    sec_of = np.concatenate([np.full(len(s), j, np.int64) for j, s in enumerate(s_list)])
    kept = order[:n_keep]
    bond_charges = np.array([q_list[sec_of[g]] for g in kept], np.int32)
    n_per = {q: 0 for q in q_list}
    for g in kept:
        n_per[q_list[sec_of[g]]] += 1
    s_final = s_sorted[:n_keep]
    if normalize and float(s_final.sum()) > 0:
        s_final = s_final / s_final.sum()
    sym = t.indices[0].symmetry
    bond_out = TensorIndex(sym, bond_charges, FlowDirection.OUT, new_bond_label)
    bond_in = TensorIndex(sym, bond_charges, FlowDirection.IN, new_bond_label)
    U_items, V_items = [], []
    for q in q_list:
        nq = n_per[q]
        if nq == 0:
            continue
        U, _, Vh = results[q]
        Uq = U[:, :nq]
        Vq = Vh[:nq, :]
        # with absorb the Gram factors already carry s on the absorbing side
        if normalize and absorb is not None and float(s_sorted[:n_keep].sum()) > 0:
            if absorb == "left":
                Uq = Uq / s_sorted[:n_keep].sum()
            else:
                Vq = Vq / s_sorted[:n_keep].sum()
        sec = sectors[q]
        for lk, (nr, shp) in sec["rows"].items():
            r0 = sec["row_off"][lk]
            U_items.append((lk + (q,), Uq[r0:r0 + nr, :].reshape(shp + (nq,))))
        for rk, (nc, shp) in sec["cols"].items():
            c0 = sec["col_off"][rk]
            V_items.append(((q,) + rk, Vq[:, c0:c0 + nc].reshape((nq,) + shp)))
    U_idx = tuple(t.indices[:nl]) + (bond_out,)
    V_idx = (bond_in,) + tuple(t.indices[nl:])
    return _pack(U_idx, U_items, dev), s_final, _pack(V_idx, V_items, dev), s_sorted


def svd(tensor: Tensor, left_labels: Sequence[Label], right_labels: Sequence[Label], new_bond_label: Label = "bond",
        max_singular_values: int | None = None, max_truncation_err: float | None = None, normalize: bool = False):
    """(U, s, Vh, s_full) of a SymmetricTensor or DenseTensor (linalg.py:665-825)."""
    labels = tensor.labels()
    for l in list(left_labels) + list(right_labels):
        if l not in labels:
            raise ValueError(f"Label {l!r} not in tensor labels {labels}")
    if set(left_labels) & set(right_labels):
        raise ValueError("left_labels and right_labels overlap")
    if len(left_labels) + len(right_labels) != len(labels):
        raise ValueError("left_labels + right_labels must cover every leg")
    if isinstance(tensor, SymmetricTensor):
        U, s, Vh, s_full = truncated_svd_symmetric(tensor, left_labels, right_labels, new_bond_label,
                                                   max_singular_values, max_truncation_err, normalize)
        return U, s, Vh, s_full
    # dense tensors: one FP64 SVD on the device
    lax = [labels.index(l) for l in left_labels]
    rax = [labels.index(l) for l in right_labels]
    d = tensor.todense().permute(*(lax + rax))
    ls = d.shape[: len(lax)]
    rs = d.shape[len(lax):]
    M = d.reshape(int(np.prod(ls)), int(np.prod(rs)))
    U, s, Vh = torch.linalg.svd(M, full_matrices=False)
    n = len(s)
    k = n if max_singular_values is None else max(1, min(n, int(max_singular_values)))
    if max_truncation_err is not None:
        sq = (s ** 2).cpu().numpy()
        tot = float(sq.sum())
        if tot > 0:
            acc = 0.0
            for i in range(n - 1, 0, -1):
                acc += float(sq[i])
                if acc / tot > max_truncation_err ** 2:
                    k = min(k, i + 1)
                    break
    s_full = s
    s = s[:k]
    if normalize and float(s.sum()) > 0:
        s = s / s.sum()
    sym = tensor.indices[0].symmetry
    zero = np.zeros(k, np.int32)
    U_t = DenseTensor(U[:, :k].reshape(tuple(ls) + (k,)),
                      tuple(tensor.indices[i] for i in lax) + (TensorIndex(sym, zero, FlowDirection.OUT, new_bond_label),))
    V_t = DenseTensor(Vh[:k, :].reshape((k,) + tuple(rs)),
                      (TensorIndex(sym, zero, FlowDirection.IN, new_bond_label),) + tuple(tensor.indices[i] for i in rax))
    return U_t, s, V_t, s_full


truncated_svd = svd
