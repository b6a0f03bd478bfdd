r"""Label-based contraction: ``contract(*tensors, output_labels=None, optimize="auto")``.

Mirrors reference src/tenax/contraction/contractor.py:48-569:

* ``_labels_to_subscripts`` (contractor.py:48-144): labels sorted by ``str``
  mapped to a-zA-Z, free labels in first-seen order, the same ValueErrors;
* ``contract`` / ``contract_with_subscripts`` (490-569): dense-only, symmetric-
  only or TypeError for a mix;
* ``_contract_symmetric`` (291-484): output target inferred from the operands'
  uniform Σflow·q (344-355), output sectors filtered by conservation (358-441).

The block work is MI355X-native: each pairwise step is one device plan
(matching, output sectors, offsets, GEMM tiles -- all on the GPU) and one
grouped-GEMM launch over every block pair.  Operands whose leg order does not
matricize directly go through the device transpose first.

Deliberate deviations (each pinned by the dense-einsum oracle in tests):
three or more symmetric operands contract pairwise (in the greedy order of
``plan_order``, or left to right for ``optimize="left"``) instead of
returning an empty tensor (reference quirk, contractor.py:375-403), and a
failing block product raises instead of being skipped (contractor.py:458-459).
"""

from __future__ import annotations

import contextlib
import string
from collections import Counter
from collections.abc import Sequence

import numpy as np
import torch

from tenax_amd import _lib
from tenax_amd.contraction import plans
from tenax_amd.core.index import Label, TensorIndex
from tenax_amd.core.symmetry import U1Symmetry
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor, Tensor, _transpose_raw

_CHARS = string.ascii_lowercase + string.ascii_uppercase

_state = {"compute": "f64", "grouped": True}

# Algorithmic work of the pairwise block contractions executed while
# ``enabled`` (flops = sum of 2*M*N*K over block GEMMs, from the plans).
WORK = {"enabled": False, "flops": 0.0, "launches": 0}


def set_compute(mode: str):
    """Block-GEMM arithmetic: "f64" (FP64 MFMA), "f32" or "split_f32"."""
    if mode not in _lib.COMPUTE:
        raise ValueError(f"unknown compute mode {mode!r}; expected one of {sorted(_lib.COMPUTE)}")
    _state["compute"] = mode


def get_compute() -> str:
    return _state["compute"]


@contextlib.contextmanager
def compute_mode(mode: str):
    old = _state["compute"]
    set_compute(mode)
    try:
        yield
    finally:
        _state["compute"] = old


@contextlib.contextmanager
def per_block_execution(enabled: bool = True):
    """Run pairwise contractions with one GEMM launch per block pair (comparison path)."""
    old = _state["grouped"]
    _state["grouped"] = not enabled
    try:
        yield
    finally:
        _state["grouped"] = old


# ---------------------------------------------------------------- subscripts
def _labels_to_subscripts(tensors: Sequence[Tensor], output_labels: Sequence[Label] | None = None):
    counts: Counter = Counter()
    first: dict = {}
    for t in tensors:
        for ix in t.indices:
            counts[ix.label] += 1
            first.setdefault(ix.label, ix)
    for lbl, c in counts.items():
        if c > 2:
            raise ValueError(f"Label {lbl!r} appears {c} times across tensors. "
                             f"Labels must appear at most 2 times (one per tensor to contract).")
    free = [l for l, c in counts.items() if c == 1]
    ordered = sorted(counts.keys(), key=str)
    if len(ordered) > 52:
        raise ValueError(f"Too many unique labels ({len(ordered)}) for einsum encoding. "
                         f"Maximum supported is 52 (a-z + A-Z).")
    ch = {l: _CHARS[i] for i, l in enumerate(ordered)}
    subs = ["".join(ch[ix.label] for ix in t.indices) for t in tensors]
    if output_labels is None:
        seen, out = set(), []
        for t in tensors:
            for ix in t.indices:
                if ix.label in free and ix.label not in seen:
                    out.append(ix.label)
                    seen.add(ix.label)
        output_labels = out
    else:
        fs = set(free)
        for lbl in output_labels:
            if lbl not in fs:
                raise ValueError(f"output_labels contains {lbl!r} which is not a free label. "
                                 f"Free labels are: {free}")
    s = ",".join(subs) + "->" + "".join(ch[l] for l in output_labels)
    return s, tuple(first[l] for l in output_labels)


# ---------------------------------------------------------------- dense
def _trivial_legs(indices):
    """Legs of the same dimensions, flows and labels with every charge 0: a
    dense tensor is then one block of the block engine."""
    sym = U1Symmetry()
    return tuple(TensorIndex(sym, np.zeros(ix.dim, np.int32), ix.flow, ix.label) for ix in indices)


def _as_one_block(t: DenseTensor) -> SymmetricTensor:
    idx = _trivial_legs(t.indices)
    r = len(idx)
    shape = np.array([[ix.dim for ix in idx]], np.int32).reshape(1, r)
    return SymmetricTensor._wrap(idx, np.zeros((1, r), np.int32), shape, np.zeros(1, np.int64),
                                 t.todense().reshape(-1))


def _contract_dense(tensors, subscripts, output_indices, optimize="auto"):
    """contractor.py:150-200 on the block engine: every float64 operand is one
    block with trivial charges, and the network is contracted pairwise by the
    grouped GEMM (differentiable through its adjoint).  Operands of another
    dtype, or with a trace inside one operand, go through torch.einsum
    (rocBLAS)."""
    lhs, so = subscripts.split("->")
    subs = lhs.split(",")
    dense_ok = all(t.todense().dtype == torch.float64 and t.todense().numel() > 0 for t in tensors) and all(
        len(set(sb)) == len(sb) for sb in subs)
    if not dense_ok:
        arrays = [t.todense() for t in tensors]
        return DenseTensor(torch.einsum(subscripts, *arrays), output_indices)
    blocks = [_as_one_block(t) for t in tensors]
    out = _contract_symmetric(blocks, subscripts, _trivial_legs(output_indices), optimize)
    shape = tuple(ix.dim for ix in output_indices)
    if out.n_blocks == 0:
        return DenseTensor(torch.zeros(shape, dtype=torch.float64, device=out.data.device), output_indices)
    return DenseTensor(out.data.reshape(shape), output_indices)


# ---------------------------------------------------------------- symmetric
def _uniform_target(t: SymmetricTensor):
    if t.n_blocks == 0:
        return None
    flows = np.array([int(ix.flow) for ix in t.indices], np.int64)
    v = (t.keys.astype(np.int64) * flows).sum(axis=1)
    return int(v[0]) if np.all(v == v[0]) else None


def _execute_pair(plan, A, B, out):
    if WORK["enabled"]:
        WORK["flops"] += plan.stats()["flops"]
        WORK["launches"] += 1
    lib = _lib.load()
    fn = lib.tnx_contract_execute if _state["grouped"] else lib.tnx_contract_execute_perblock
    _lib.check(fn(plan.ptr, _lib.dptr(A.data.detach()), _lib.dptr(B.data.detach()), _lib.dptr(out),
                  _lib.COMPUTE[_state["compute"]], _lib.stream_handle()))


class _PairExecFn(torch.autograd.Function):
    """Grouped block GEMM C = A . B with its adjoint.

    Layouts: A = [free_a..., cont...] (or [cont..., free_a...]), B = [cont...,
    free_b...], C = [free_a..., free_b...].  The backward is one library call,
    tnx_contract_vjp on the forward plan: dA = G . B^T and dB = A^T . G, each
    laid out in its operand's own block table, from work lists the library
    derives once from the forward pair list; blocks no pair touches get zero.
    """

    @staticmethod
    def forward(ctx, a_data, b_data, plan, A, B, C, sa, sb, sc):
        if C.data.numel():
            _execute_pair(plan, A, B, C.data)
        ctx.meta = (plan, A, B)
        return C.data

    @staticmethod
    def backward(ctx, g):
        plan, A, B = ctx.meta
        ga = torch.empty_like(A.data) if ctx.needs_input_grad[0] else None
        gb = torch.empty_like(B.data) if ctx.needs_input_grad[1] else None
        if ga is None and gb is None:
            return None, None, None, None, None, None, None, None, None
        g = g.contiguous()
        lib = _lib.load()
        _lib.check(lib.tnx_contract_vjp(plan.ptr, _lib.dptr(g), _lib.dptr(A.data), _lib.dptr(B.data),
                                        _lib.dptr(ga) if ga is not None else None,
                                        _lib.dptr(gb) if gb is not None else None,
                                        _lib.COMPUTE[_state["compute"]], _lib.stream_handle()))
        if WORK["enabled"]:
            WORK["flops"] += 2.0 * plan.stats()["flops"]
            WORK["launches"] += 2
        return ga, gb, None, None, None, None, None, None, None


def _inversion_pairs(input_subs: Sequence[str], output: str) -> list[tuple[str, str]]:
    """Leg pairs whose exchange can sign a fermionic contraction
    (contractor.py:206-285).  The input legs, concatenated, are carried to
    the order [output legs..., each contracted leg's two occurrences] (first
    occurrence, then second, contracted legs in order of first appearance);
    every inversion of that reordering is returned as a pair of subscript
    characters.  A block combination takes a factor -1 for each pair whose two
    charges are both odd."""
    seq = "".join(input_subs)
    counts = Counter(seq)
    contracted = list(dict.fromkeys(c for c in seq if counts[c] >= 2))
    slot = {c: i for i, c in enumerate(output)}
    for i, c in enumerate(contracted):
        slot[c] = len(output) + i
    rank, seen = [], Counter()
    for c in seq:
        rank.append(2 * slot[c] + (seen[c] if counts[c] >= 2 else 0))
        seen[c] += 1
    return [(seq[i], seq[j]) for i in range(len(seq)) for j in range(i + 1, len(seq)) if rank[i] > rank[j]]


def _is_graded(T) -> bool:
    return bool(T.indices) and T.indices[0].symmetry.is_fermionic


# The grouped GEMM writes a requested output leg order directly (a
# permuted-store plan) when each side of the product -- the A free legs
# (rows) and the B free legs (columns) -- falls into at most this many runs of
# legs that stay adjacent in the requested order; otherwise the natural
# result is transposed.  FUSED_OUTPUT_ORDER = False always transposes.
FUSED_OUTPUT_ORDER = True
_STORE_LEGS = 3  # STORE_LEGS of the library

# Tiling (shape, kchunk, xcd) of the merged grouped-GEMM launch of
# contract_many / contract_pairs_batched (tnx_plan_set_tiling on the batch
# plan); None keeps the default 64x64 whole-K list.
BATCH_TILING = None


def _fused_order_fits(out_axes, nfa: int) -> bool:
    pos = {leg: j for j, leg in enumerate(out_axes)}
    n = len(out_axes)
    for lo, hi in ((0, nfa), (nfa, n)):
        runs = sum(1 for i in range(lo, hi) if i == lo or pos[i] != pos[i - 1] + 1)
        if runs > _STORE_LEGS:
            return False
    return True


class _PairJob:
    """The host-side decisions of one pairwise contraction: which operand
    plays A, the leg permutations that bring A to [free, contracted] (or
    [contracted, free]) and B to [contracted, free], the planner's arguments
    and the permutation of the GEMM's natural output order to the requested
    one."""

    __slots__ = ("A", "B", "a_axes", "b_axes", "a_kfirst", "nc", "nat", "nat_indices", "out_axes", "mode",
                 "target", "sign_pairs", "graded", "nfa")

    def __init__(self, A, B, sa, sb, so, out_indices, mode, target):
        if len(set(sa)) != len(sa) or len(set(sb)) != len(sb):
            raise NotImplementedError("traces within one operand are not supported")
        cont = [c for c in sa if c in sb]
        free_a = [c for c in sa if c not in cont]
        free_b = [c for c in sb if c not in cont]
        if sorted(so) != sorted(free_a + free_b):
            raise ValueError(f"output subscripts {so!r} are not the free legs of {sa!r},{sb!r}")
        self.graded = _is_graded(A)
        sign_chars = []
        if self.graded:
            if mode != _lib.MODE_PAIRWISE_FILTERED:
                raise NotImplementedError("fermionic signs are defined for the contract() semantics only")
            if torch.is_grad_enabled() and (A.data.requires_grad or B.data.requires_grad):
                raise NotImplementedError("autograd through fermionic block contractions is not supported")
            sign_chars = _inversion_pairs([sa, sb], so)
        ix_of = dict(zip(so, out_indices))
        # Let the operand whose free leg comes first in the output play "A".
        if free_a and free_b and so[0] in free_b:
            A, B, sa, sb = B, A, sb, sa
            cont = [c for c in sa if c in sb]
            free_a, free_b = free_b, free_a
        fa = [c for c in so if c in free_a]
        fb = [c for c in so if c in free_b]
        self.nat = "".join(fa + fb)
        # A layout
        self.a_axes = None
        if list(sa) == fa + cont:
            self.a_kfirst = 0
        elif list(sa) == cont + fa:
            self.a_kfirst = 1
        else:
            self.a_axes = tuple(sa.index(c) for c in fa + cont)
            sa, self.a_kfirst = "".join(fa + cont), 0
        # B layout: always [contracted (in A's order)..., free...]
        self.b_axes = None
        if list(sb) != cont + fb:
            self.b_axes = tuple(sb.index(c) for c in cont + fb)
            sb = "".join(cont + fb)
        self.A, self.B = A, B
        self.nc = len(cont)
        self.nat_indices = tuple(ix_of[c] for c in self.nat)
        pos = {c: i for i, c in enumerate(sa)}
        for i, c in enumerate(sb):
            pos.setdefault(c, len(sa) + i)
        self.sign_pairs = [(pos[x], pos[y]) for x, y in sign_chars]
        self.out_axes = tuple(self.nat.index(c) for c in so) if self.nat != so else None
        self.nfa = len(fa)
        self.mode, self.target = mode, target

    def fused_order(self) -> bool:
        """Whether the GEMM should write the requested output order itself."""
        return (self.out_axes is not None and FUSED_OUTPUT_ORDER and not self.graded and _state["grouped"]
                and _fused_order_fits(self.out_axes, self.nfa))


def contract_pair(A: SymmetricTensor, B: SymmetricTensor, sa: str, sb: str, so: str,
                  out_indices: tuple[TensorIndex, ...], mode: int, target) -> SymmetricTensor:
    """One pairwise block-sparse contraction with einsum-style subscripts.

    ``out_indices`` are the legs of ``so`` in order.  mode = filtered
    (conservation, contractor.py semantics) or reachable (dmrg.py:1048-1151).
    For graded (fermionic) groups each block pair takes the exchange sign of
    the reference (contractor.py:462-470), evaluated on the device; the
    operand and output reshuffles below then move data without signs.
    """
    if not so:
        if len(set(sa)) != len(sa) or len(set(sb)) != len(sb):
            raise NotImplementedError("traces within one operand are not supported")
        if sorted(sa) != sorted(sb):
            raise ValueError(f"output subscripts {so!r} are not the free legs of {sa!r},{sb!r}")
        sign_chars = _inversion_pairs([sa, sb], so) if _is_graded(A) else []
        return _full_contraction(A, B, sa, sb, sign_chars)
    job = _PairJob(A, B, sa, sb, so, out_indices, mode, target)
    graded = job.graded
    move = (lambda T, axes: _transpose_raw(T, axes, graded=False)) if graded else (lambda T, axes: T.transpose(axes))
    A = job.A if job.a_axes is None else move(job.A, job.a_axes)
    B = job.B if job.b_axes is None else move(job.B, job.b_axes)
    differentiated = torch.is_grad_enabled() and (A.data.requires_grad or B.data.requires_grad)
    if job.fused_order() and not differentiated:
        # the GEMM stores the requested leg order directly
        plan, keys, shapes, offs, tot = plans.pairwise_plan(A, B, job.a_kfirst, 1, job.nc, job.nat_indices, mode,
                                                            target, job.sign_pairs, out_perm=job.out_axes)
        out = torch.empty(tot, dtype=torch.float64, device=A.data.device)
        if tot:
            _execute_pair(plan, A, B, out)
        return SymmetricTensor._wrap(tuple(job.nat_indices[p] for p in job.out_axes), keys, shapes, offs, out)
    plan, keys, shapes, offs, tot = plans.pairwise_plan(A, B, job.a_kfirst, 1, job.nc, job.nat_indices, mode,
                                                        target, job.sign_pairs)
    out = torch.empty(tot, dtype=torch.float64, device=A.data.device)
    C = SymmetricTensor._wrap(job.nat_indices, keys, shapes, offs, out)
    if differentiated:
        Ad = A._with_data(A.data.detach())
        Bd = B._with_data(B.data.detach())
        C = C._with_data(_PairExecFn.apply(A.data, B.data, plan, Ad, Bd, C, None, None, job.nat))
    elif tot:
        _execute_pair(plan, A, B, out)
    if job.out_axes is not None:
        C = move(C, job.out_axes)
    return C


def _batched_transposes(tensors, axes_list):
    """Leg permutations of several tensors (no signs) in one launch
    (tnx_batch_transpose); returns the permuted tensors, their data
    consecutive in one buffer."""
    lib = _lib.load()
    items = [plans.transpose_plan(T, axes, False) for T, axes in zip(tensors, axes_list)]
    starts = np.concatenate([[0], np.cumsum([it[4] for it in items])]).astype(np.int64)
    dev = tensors[0].data.device
    buf = torch.empty(int(starts[-1]), dtype=torch.float64, device=dev)
    outs = [buf[starts[k]:starts[k + 1]] for k in range(len(items))]
    n = len(items)
    P = (_lib.C.c_void_p * n)(*[it[0].ptr for it in items])
    I = (_lib.C.c_void_p * n)(*[_lib.dptr(T.data.detach()) for T in tensors])
    O = (_lib.C.c_void_p * n)(*[_lib.dptr(o) for o in outs])
    batch = _lib.C.c_void_p()
    _lib.check(lib.tnx_batch_transpose(P, n, I, O, _lib.stream_handle(), _lib.C.byref(batch)))
    bplan = _lib.Plan(batch)
    _lib.check(lib.tnx_batch_execute(bplan.ptr, 0, _lib.stream_handle()))
    return [SymmetricTensor._wrap(tuple(T.indices[i] for i in axes), it[1], it[2], it[3], o)
            for T, axes, it, o in zip(tensors, axes_list, items, outs)]


def contract_pairs_batched(jobs) -> list[SymmetricTensor]:
    """Several independent pairwise contractions (_PairJob) with one launch per
    phase over all of them: the operand permutations, the grouped block GEMM
    (tnx_batch_pairwise; one launch per A layout) and the output
    permutations.  Each result equals contract_pair's on the same job."""
    lib = _lib.load()
    for sel, attr in ((lambda j: j.a_axes is not None, "A"), (lambda j: j.b_axes is not None, "B")):
        todo = [j for j in jobs if sel(j)]
        if todo:
            axes = [j.a_axes if attr == "A" else j.b_axes for j in todo]
            moved = _batched_transposes([getattr(j, attr) for j in todo], axes)
            for j, T in zip(todo, moved):
                setattr(j, attr, T)
                if attr == "A":
                    j.a_axes = None
                else:
                    j.b_axes = None
    fused = [j.fused_order() for j in jobs]
    planned = [plans.pairwise_plan(j.A, j.B, j.a_kfirst, 1, j.nc, j.nat_indices, j.mode, j.target, j.sign_pairs,
                                   out_perm=j.out_axes if f else None)
               for j, f in zip(jobs, fused)]
    starts = np.concatenate([[0], np.cumsum([p[4] for p in planned])]).astype(np.int64)
    dev = jobs[0].A.data.device
    buf = torch.empty(int(starts[-1]), dtype=torch.float64, device=dev)
    res = [SymmetricTensor._wrap(tuple(j.nat_indices[q] for q in j.out_axes) if f else j.nat_indices,
                                 p[1], p[2], p[3], buf[starts[k]:starts[k + 1]])
           for k, (j, p, f) in enumerate(zip(jobs, planned, fused))]
    for kfirst in (0, 1):
        ks = [k for k, j in enumerate(jobs) if j.a_kfirst == kfirst and planned[k][4] > 0]
        if not ks:
            continue
        n = len(ks)
        P = (_lib.C.c_void_p * n)(*[planned[k][0].ptr for k in ks])
        Aa = (_lib.C.c_void_p * n)(*[_lib.dptr(jobs[k].A.data.detach()) for k in ks])
        Bb = (_lib.C.c_void_p * n)(*[_lib.dptr(jobs[k].B.data.detach()) for k in ks])
        Cc = (_lib.C.c_void_p * n)(*[_lib.dptr(res[k].data) for k in ks])
        batch = _lib.C.c_void_p()
        _lib.check(lib.tnx_batch_pairwise(P, n, Aa, Bb, Cc, _lib.stream_handle(), _lib.C.byref(batch)))
        bplan = _lib.Plan(batch)
        if BATCH_TILING is not None:
            try:
                bplan.set_tiling(0, *BATCH_TILING)
            except _lib.TnxError as e:  # a split over the workspace cap: keep the default list
                if e.code != _lib.TNX_ERR_CAPACITY:
                    raise
        if WORK["enabled"]:
            WORK["flops"] += bplan.stats()["flops"]
            WORK["launches"] += 1
        _lib.check(lib.tnx_batch_execute(bplan.ptr, _lib.COMPUTE[_state["compute"]], _lib.stream_handle()))
    fin = [k for k, j in enumerate(jobs) if j.out_axes is not None and not fused[k]]
    if fin:
        moved = _batched_transposes([res[k] for k in fin], [jobs[k].out_axes for k in fin])
        for k, T in zip(fin, moved):
            res[k] = T
    return res


def contract_many(operands, output_labels: Sequence[Label] | None = None) -> list[Tensor]:
    """``[contract(A, B, output_labels=output_labels) for A, B in operands]``
    with every instance in the same launches (one grouped block-GEMM launch
    for all instances' block pairs, one launch per operand / output
    permutation phase): the many-instance HOTRG step (hotrg.py:109-215 for
    every temperature at once).  Instances that are not plain
    SymmetricTensor pairs (dense, fermionic, differentiated, full
    contractions) are contracted one by one by contract()."""
    out = [None] * len(operands)
    jobs, where = [], []
    for k, (A, B) in enumerate(operands):
        plain = (isinstance(A, SymmetricTensor) and isinstance(B, SymmetricTensor) and not _is_graded(A)
                 and not (torch.is_grad_enabled() and (A.data.requires_grad or B.data.requires_grad)))
        if plain:
            subscripts, out_idx = _labels_to_subscripts((A, B), output_labels)
            lhs, so = subscripts.split("->")
            sa, sb = lhs.split(",")
            plain = bool(so) and len(set(sa)) == len(sa) and len(set(sb)) == len(sb)
        if not plain:
            out[k] = contract(A, B, output_labels=output_labels)
            continue
        total = 0
        for t in (A, B):
            u = _uniform_target(t)
            if u is not None:
                total += u
        jobs.append(_PairJob(A, B, sa, sb, so, tuple(out_idx), _lib.MODE_PAIRWISE_FILTERED,
                             total if total != 0 else None))
        where.append(k)
    if jobs:
        for k, C in zip(where, contract_pairs_batched(jobs)):
            out[k] = C
    return out


def _full_contraction(A, B, sa, sb, sign_chars=()):
    """All legs contracted: the sum over common keys of <A block, B block>
    (with the fermionic exchange sign per block).  One dot over the packed
    buffers when both operands hold the same blocks and no sign is negative;
    otherwise the common blocks (A's signed) are gathered by one batched copy
    each and dotted once."""
    Bp = B if sb == sa else _transpose_raw(B, tuple(sb.index(c) for c in sa), graded=False)
    ka, kb = A.key_index(), Bp.key_index()
    common = [(i, kb[k], k) for k, i in ka.items() if k in kb]
    signs = [(-1.0 if sum((dict(zip(sa, k))[x] & 1) * (dict(zip(sa, k))[y] & 1) for x, y in sign_chars) % 2
              else 1.0) for _, _, k in common]
    dev = A.data.device
    same = (len(common) == A.n_blocks == Bp.n_blocks and np.array_equal(A.shapes, Bp.shapes)
            and np.array_equal(A.offsets, Bp.offsets) and all(i == j for i, j, _ in common))
    if not common:
        tot = torch.zeros((), dtype=torch.float64, device=dev)
    elif same and all(v > 0 for v in signs):
        tot = torch.dot(A.data, Bp.data)
    elif torch.is_grad_enabled() and (A.data.requires_grad or Bp.data.requires_grad):
        # differentiable form (the batched copy below is not): block views
        tot = torch.zeros((), dtype=torch.float64, device=dev)
        for (i, j, _), sg in zip(common, signs):
            d = torch.dot(A.block(i).reshape(-1), Bp.block(j).reshape(-1))
            tot = tot - d if sg < 0 else tot + d
    else:
        sizes = [int(np.prod(A.shapes[i])) for i, _, _ in common]
        starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        ga = torch.empty(int(starts[-1]), dtype=torch.float64, device=dev)
        gb = torch.empty_like(ga)
        ad, bd = A.data.detach(), Bp.data.detach()
        _lib.copy2d([(ad, int(A.offsets[i]), n, ga, int(o), n, 1, n, sg)
                     for (i, _, _), n, o, sg in zip(common, sizes, starts, signs)], dev)
        _lib.copy2d([(bd, int(Bp.offsets[j]), n, gb, int(o), n, 1, n, 1.0)
                     for (_, j, _), n, o in zip(common, sizes, starts)], dev)
        tot = torch.dot(ga, gb)
    return SymmetricTensor._wrap((), np.zeros((1, 0), np.int32), np.zeros((1, 0), np.int32),
                                 np.zeros(1, np.int64), tot.reshape(1))


def _contract_symmetric(tensors, subscripts, output_indices, optimize="auto"):
    lhs, so = subscripts.split("->")
    subs = lhs.split(",")
    if len(tensors) == 1:
        (T,), (s,) = tensors, subs
        if len(set(s)) != len(s):
            raise NotImplementedError("traces within one operand are not supported")
        return T.transpose(tuple(s.index(c) for c in so))
    c2ix = {}
    for t, s in zip(tensors, subs):
        for c, ix in zip(s, t.indices):
            c2ix.setdefault(c, ix)
    total = 0
    for t in tensors:
        u = _uniform_target(t)
        if u is not None:
            total += u
    target = total if total != 0 else None
    dims = {c: ix.dim for c, ix in c2ix.items()}
    # graded groups fold left to right: each step's exchange sign is then
    # relative to the operands' own order
    left = optimize == "left" or _is_graded(tensors[0])
    pairs = [(0, 1)] * (len(tensors) - 1) if left else plan_order(subs, dims)
    items = list(zip(tensors, subs))
    for i, j in pairs:
        (a, sa), (b, sb) = items[i], items[j]
        rest = [x[1] for k, x in enumerate(items) if k not in (i, j)]
        later = set("".join(rest)) | set(so)
        keep = list(dict.fromkeys(c for c in sa + sb if (c not in sa or c not in sb) and c in later))
        step_out = "".join(c for c in so if c in keep) + "".join(c for c in keep if c not in so)
        last = len(items) == 2
        tgt = target if last else _partial_target(a, b)
        # an open leg keeps the flow of the operand it sits on (a label's
        # first occurrence in the network may be on a tensor not yet folded in)
        legs = dict(zip(sb, b.indices))
        legs.update(zip(sa, a.indices))
        if last:
            legs.update(zip(so, output_indices))
        cur = contract_pair(a, b, sa, sb, step_out, tuple(legs[c] for c in step_out),
                            _lib.MODE_PAIRWISE_FILTERED, tgt)
        items = [x for k, x in enumerate(items) if k not in (i, j)]
        items.insert(min(i, j), (cur, step_out))
    cur, cs = items[0]
    if cs != so:
        cur = cur.transpose(tuple(cs.index(c) for c in so))
    return cur


def plan_order(subs: Sequence[str], dims: dict) -> list[tuple[int, int]]:
    """Greedy pairwise contraction order (the role of opt_einsum's path search,
    contractor.py:150-166): repeatedly contract the connected pair whose dense
    cost (product of the dimensions of all legs involved) is smallest, ties
    to the earliest pair; disconnected networks fall back to the first two
    operands.  Returns positions into the shrinking operand list (the result
    of a step takes the smaller of the two positions)."""
    cur = [set(s) for s in subs]
    out = []
    while len(cur) > 1:
        best = None
        for i in range(len(cur)):
            for j in range(i + 1, len(cur)):
                if not (cur[i] & cur[j]):
                    continue
                cost = 1
                for c in cur[i] | cur[j]:
                    cost *= dims.get(c, 1)
                if best is None or cost < best[0]:
                    best = (cost, i, j)
        i, j = (0, 1) if best is None else best[1:]
        others = set().union(*(cur[k] for k in range(len(cur)) if k not in (i, j))) if len(cur) > 2 else set()
        merged = (cur[i] ^ cur[j]) | ((cur[i] & cur[j]) & others)
        cur = [x for k, x in enumerate(cur) if k not in (i, j)]
        cur.insert(i, merged)
        out.append((i, j))
    return out


def _partial_target(a, b):
    t = 0
    for x in (a, b):
        u = _uniform_target(x)
        if u is not None:
            t += u
    return t if t != 0 else None


def contract(*tensors: Tensor, output_labels: Sequence[Label] | None = None, optimize: str = "auto") -> Tensor:
    """Contract tensors over shared labels (contractor.py:490-531)."""
    if not tensors:
        raise ValueError("contract() requires at least one tensor")
    subscripts, out_idx = _labels_to_subscripts(tensors, output_labels)
    if len(tensors) == 1:
        lhs, rhs = subscripts.split("->")
        if lhs == rhs:
            return tensors[0]
    return contract_with_subscripts(tensors, subscripts, out_idx, optimize)


def contract_with_subscripts(tensors: Sequence[Tensor], subscripts: str, output_indices, optimize: str = "auto"):
    """Contract with explicit subscripts (contractor.py:534-569)."""
    if all(isinstance(t, DenseTensor) for t in tensors):
        return _contract_dense(list(tensors), subscripts, tuple(output_indices), optimize)
    if all(isinstance(t, SymmetricTensor) for t in tensors):
        return _contract_symmetric(list(tensors), subscripts, tuple(output_indices), optimize)
    types = [type(t).__name__ for t in tensors]
    raise TypeError(f"Cannot mix DenseTensor and SymmetricTensor in a single contraction. "
                    f"Got types: {types}. Convert all tensors to the same type first.")


def blockwise_contract(tensors, subscripts: str, output_indices) -> SymmetricTensor:
    """Reachability semantics of dmrg._blockwise_contract (dmrg.py:1048-1151).

    Equality matching on shared subscript characters, flows ignored, no
    conservation filter; folded pairwise left to right on the device.
    """
    lhs, so = subscripts.split("->")
    subs = lhs.split(",")
    c2ix = {}
    for t, s in zip(tensors, subs):
        for c, ix in zip(s, t.indices):
            c2ix.setdefault(c, ix)
    for c, ix in zip(so, output_indices):
        c2ix[c] = ix
    cur, cs = tensors[0], subs[0]
    for k in range(1, len(tensors)):
        nxt, ns = tensors[k], subs[k]
        later = set("".join(subs[k + 1:])) | set(so)
        keep = list(dict.fromkeys(c for c in cs + ns if (c not in cs or c not in ns) and c in later))
        step_out = "".join(c for c in so if c in keep) + "".join(c for c in keep if c not in so)
        cur = contract_pair(cur, nxt, cs, ns, step_out, tuple(c2ix[c] for c in step_out),
                            _lib.MODE_MULTI_REACHABLE, None)
        cs = step_out
    if cs != so:
        cur = cur.transpose(tuple(cs.index(c) for c in so))
    return SymmetricTensor._wrap(tuple(output_indices), cur.keys, cur.shapes, cur.offsets, cur.data)
