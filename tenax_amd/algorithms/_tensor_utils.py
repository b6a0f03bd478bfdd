"""Tensor helpers shared by the coarse-graining and CTM algorithms
(reference src/tenax/algorithms/_tensor_utils.py)."""

from __future__ import annotations

import math

import numpy as np
import torch

from tenax_amd.core import LOG_EPS
from tenax_amd.core.index import TensorIndex
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor, Tensor


def max_abs_normalize(T: Tensor):
    """(T / (max|T| + LOG_EPS), log(max|T| + LOG_EPS)); _tensor_utils.py:65-80.

    The log norm comes back as a Python float (one device sync per call)."""
    norm = float(T.max_abs())
    return T * (1.0 / (norm + LOG_EPS)), math.log(norm + LOG_EPS)


def _compute_fused_charges(idx_a, idx_b, fused_flow, sym) -> np.ndarray:
    """Fused charge of basis pair (i, j) at position i*db + j:
    (flow_a q_a[i] + flow_b q_b[j]) * fused_flow, mod n for Z_n
    (_tensor_utils.py:172-206)."""
    qa = np.asarray(idx_a.charges, np.int64) * int(idx_a.flow)
    qb = np.asarray(idx_b.charges, np.int64) * int(idx_b.flow)
    q = (qa[:, None] + qb[None, :]).reshape(-1) * int(fused_flow)
    n = sym.n_values()
    if n is not None:
        q = np.mod(q, n)
    return q.astype(np.int32)


def fuse_indices(tensor: Tensor, axis_a: int, axis_b: int, fused_label, fused_flow) -> Tensor:
    """Fuse two legs into one at the position of the first (_tensor_utils.py:106-139).

    Dense: transpose the pair together (first axis slow) and reshape.
    Symmetric: the reference's block layout (_tensor_utils.py:209-325): fused
    charges from _compute_fused_charges; inside a fused sector the (q_a, q_b)
    sub-blocks follow the unique-charge loop order, each sub-block's two legs
    flattened row-major.  Differentiable (torch views, permutes and cat).
    """
    if isinstance(tensor, SymmetricTensor):
        return _fuse_symmetric(tensor, axis_a, axis_b, fused_label, fused_flow)
    a, b = sorted((axis_a, axis_b))
    nd = tensor.ndim
    other = [i for i in range(nd) if i not in (a, b)]
    perm = other[:a] + [a, b] + other[a:]
    data = tensor.todense().permute(*perm)
    shape = list(data.shape)
    data = data.reshape(shape[:a] + [shape[a] * shape[a + 1]] + shape[a + 2:])
    ip = [tensor.indices[i] for i in perm]
    sym = ip[a].symmetry
    fidx = TensorIndex(sym, _compute_fused_charges(ip[a], ip[a + 1], fused_flow, sym), fused_flow, fused_label)
    return DenseTensor(data, tuple(ip[:a]) + (fidx,) + tuple(ip[a + 2:]))


def _fuse_symmetric(T: SymmetricTensor, axis_a, axis_b, fused_label, fused_flow) -> SymmetricTensor:
    a, b = sorted((axis_a, axis_b))
    nd = T.ndim
    ia, ib = T.indices[a], T.indices[b]
    sym = ia.symmetry
    fidx = TensorIndex(sym, _compute_fused_charges(ia, ib, fused_flow, sym), fused_flow, fused_label)
    ua, ub = np.unique(ia.charges), np.unique(ib.charges)
    da = {int(q): int(np.sum(ia.charges == q)) for q in ua}
    db = {int(q): int(np.sum(ib.charges == q)) for q in ub}
    n = sym.n_values()
    groups: dict[int, list] = {}
    for qa in ua:
        for qb in ub:
            qf = (int(ia.flow) * int(qa) + int(ib.flow) * int(qb)) * int(fused_flow)
            if n is not None:
                qf %= n
            groups.setdefault(qf, []).append((int(qa), int(qb)))
    where, fdim = {}, {}
    for qf, pairs in groups.items():
        off = 0
        for qa, qb in pairs:
            where[(qa, qb)] = (qf, off)
            off += da[qa] * db[qb]
        fdim[qf] = off
    other = [i for i in range(nd) if i not in (a, b)]
    perm = other[:a] + [a, b] + other[a:]
    new_indices = [T.indices[i] for i in other]
    new_indices.insert(a, fidx)
    pieces: dict[tuple, list] = {}
    for i in range(T.n_blocks):
        key = [int(q) for q in T.keys[i]]
        qf, off = where[(key[a], key[b])]
        blk = T.block(i).permute(*perm)
        sh = list(blk.shape)
        blk = blk.reshape(sh[:a] + [sh[a] * sh[a + 1]] + sh[a + 2:])
        oth = [key[j] for j in other]
        nk = tuple(oth[:a] + [qf] + oth[a:])
        pieces.setdefault(nk, []).append((off, blk))
    keys = sorted(pieces)
    blocks, shapes = [], []
    for nk in keys:
        parts = sorted(pieces[nk], key=lambda x: x[0])
        ref = parts[0][1]
        full = list(ref.shape)
        full[a] = fdim[nk[a]]
        seq, pos = [], 0
        for off, blk in parts:
            if off > pos:
                gap = list(full)
                gap[a] = off - pos
                seq.append(torch.zeros(gap, dtype=blk.dtype, device=blk.device))
            seq.append(blk)
            pos = off + blk.shape[a]
        if pos < full[a]:
            gap = list(full)
            gap[a] = full[a] - pos
            seq.append(torch.zeros(gap, dtype=ref.dtype, device=ref.device))
        blocks.append(torch.cat(seq, dim=a) if len(seq) > 1 else seq[0])
        shapes.append(full)
    r = len(new_indices)
    karr = np.asarray(keys, np.int32).reshape(len(keys), r)
    sarr = np.asarray(shapes, np.int32).reshape(len(keys), r)
    offs, tot = SymmetricTensor._layout(sarr)
    data = torch.cat([x.reshape(-1) for x in blocks]) if blocks else T.data[:0]
    return SymmetricTensor._wrap(tuple(new_indices), karr, sarr, offs, data)
