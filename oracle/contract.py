"""Oracle: label contraction of dict-of-blocks tensors (FP64 NumPy).

Restates reference src/tenax/contraction/contractor.py:48-144 (labels to
subscripts), 291-484 (pairwise conservation-filtered engine) and
src/tenax/algorithms/dmrg.py:1048-1151 (multi-operand reachability engine).
Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import itertools
import string
from collections import Counter

import numpy as np

from oracle.core import SymT, compute_valid_blocks, is_fermionic, parity

_CHARS = string.ascii_lowercase + string.ascii_uppercase


def labels_to_subscripts(tensors, output_labels=None):
    """contractor.py:48-144."""
    counts = Counter()
    first = {}
    for t in tensors:
        for l in t.legs:
            counts[l.label] += 1
            first.setdefault(l.label, l)
    for lbl, c in counts.items():
        if c > 2:
            raise ValueError(f"Label {lbl!r} appears {c} times across tensors.")
    free = [l for l, c in counts.items() if c == 1]
    all_labels = sorted(counts.keys(), key=str)
    if len(all_labels) > 52:
        raise ValueError("Too many unique labels")
    ch = {l: _CHARS[i] for i, l in enumerate(all_labels)}
    subs = ["".join(ch[l.label] for l in t.legs) for t in tensors]
    if output_labels is None:
        seen, order = set(), []
        for t in tensors:
            for l in t.legs:
                if l.label in free and l.label not in seen:
                    order.append(l.label)
                    seen.add(l.label)
        output_labels = order
    else:
        for lbl in output_labels:
            if lbl not in set(free):
                raise ValueError(f"output_labels contains {lbl!r} which is not a free label.")
    s = ",".join(subs) + "->" + "".join(ch[l] for l in output_labels)
    return s, tuple(first[l] for l in output_labels)


def inversion_pairs(input_subs, output_part):
    """contractor.py:206-285: the concatenated input legs are sent to
    [output legs..., contracted legs (first occurrence, second occurrence) in
    order of first appearance]; every inversion is a character pair."""
    chars = [c for s in input_subs for c in s]
    cnt = Counter(chars)
    contracted = {c for c, n in cnt.items() if n >= 2}
    target = list(output_part)
    for c in chars:
        if c in contracted and c not in target:
            target.append(c)
    pos = {c: i for i, c in enumerate(target)}
    used = Counter()
    ranks = []
    for c in chars:
        ranks.append(2 * pos[c] + (used[c] if c in contracted else 0))
        used[c] += 1
    return [(chars[i], chars[j]) for i in range(len(chars)) for j in range(i + 1, len(chars))
            if ranks[i] > ranks[j]]


def contract_symmetric(tensors, subscripts):
    """contractor.py:291-484: signature buckets, product, filter, and the
    fermionic sign of each combination (462-470).

    Returns (SymT, combos) where combos lists (output_key, input keys) in the
    order the reference loop visits them (common signatures in sorted order
    here; the reference iterates a set, so only per-key sums are comparable).
    """
    inp, outp = subscripts.split("->")
    isubs = inp.split(",")
    c2leg = {}
    for t, s in zip(tensors, isubs):
        for c, l in zip(s, t.legs):
            c2leg[c] = l
    out_legs = tuple(c2leg[c] for c in outp)
    cnt = Counter(inp.replace(",", ""))
    contracted = sorted(c for c, n in cnt.items() if n >= 2)
    # target inference (344-355)
    total = 0
    for t in tensors:
        if t.blocks:
            ts = {sum(l.flow * int(q) for l, q in zip(t.legs, k)) for k in t.blocks}
            if len(ts) == 1:
                total += ts.pop()
    target = total if total != 0 else None
    valid = set(compute_valid_blocks(out_legs, target))
    sym0 = tensors[0].legs[0].sym
    ferm = is_fermionic(sym0)
    inv = inversion_pairs(isubs, outp) if ferm else []
    by_sig = []
    for t, s in zip(tensors, isubs):
        pos = sorted(((contracted.index(c), p) for p, c in enumerate(s) if c in contracted))
        idx = {}
        for k, a in t.blocks.items():
            idx.setdefault(tuple(int(k[p]) for _, p in pos), []).append((k, a))
        by_sig.append(idx)
    common = set(by_sig[0])
    for m in by_sig[1:]:
        common &= set(m)
    out = {}
    combos = []
    for sig in sorted(common):
        for combo in itertools.product(*[m[sig] for m in by_sig]):
            cc = {}
            ok = True
            for (k, _), s in zip(combo, isubs):
                for c, q in zip(s, k):
                    if c in cc and cc[c] != int(q):
                        ok = False
                        break
                    cc[c] = int(q)
                if not ok:
                    break
            if not ok:
                continue
            okey = tuple(cc.get(c, 0) for c in outp)
            if okey not in valid:
                continue
            r = _einsum(subscripts, [a for _, a in combo])
            if ferm:
                sg = 1
                for ci, cj in inv:
                    if parity(sym0, cc[ci]) and parity(sym0, cc[cj]):
                        sg = -sg
                r = r * sg
            combos.append((okey, tuple(k for k, _ in combo)))
            out[okey] = out[okey] + r if okey in out else r
    return SymT(out, out_legs, validate=target is None), combos


_PATHS: dict = {}


def _einsum(subscripts, arrs):
    """np.einsum with the contraction path cached by block shapes (the role of
    the reference's shape-keyed opt_einsum expression cache, dmrg.py:1099-1107)."""
    key = (subscripts, tuple(a.shape for a in arrs))
    path = _PATHS.get(key)
    if path is None:
        path = np.einsum_path(subscripts, *arrs, optimize="optimal")[0]
        _PATHS[key] = path
    return np.einsum(subscripts, *arrs, optimize=path)


def blockwise_contract(tensors, subscripts, out_legs, cache=None):
    """dmrg.py:1048-1151: depth-first reachability, no conservation filter."""
    inp, outp = subscripts.split("->")
    isubs = inp.split(",")
    acc = {}
    cc = {}
    arrs = []
    items = [list(t.blocks.items()) for t in tensors]

    def rec(i):
        if i == len(tensors):
            okey = tuple(cc.get(c, 0) for c in outp)
            acc.setdefault(okey, []).append(_einsum(subscripts, arrs))
            return
        s = isubs[i]
        for k, a in items[i]:
            added = []
            ok = True
            for c, q in zip(s, k):
                q = int(q)
                if c in cc:
                    if cc[c] != q:
                        ok = False
                        break
                else:
                    cc[c] = q
                    added.append(c)
            if ok:
                arrs.append(a)
                rec(i + 1)
                arrs.pop()
            for c in added:
                del cc[c]

    rec(0)
    out = {}
    for k, lst in acc.items():
        tot = lst[0]
        for a in lst[1:]:
            tot = tot + a
        out[k] = tot
    return SymT(out, out_legs, validate=False)


def blockwise_keys(key_lists, subscripts):
    """Output key set of dmrg.py:1094-1136's recursion, on keys only."""
    inp, outp = subscripts.split("->")
    isubs = inp.split(",")
    out = set()
    cc = {}

    def rec(i):
        if i == len(key_lists):
            out.add(tuple(cc.get(c, 0) for c in outp))
            return
        for k in key_lists[i]:
            added = []
            ok = True
            for c, q in zip(isubs[i], k):
                q = int(q)
                if c in cc:
                    if cc[c] != q:
                        ok = False
                        break
                else:
                    cc[c] = q
                    added.append(c)
            if ok:
                rec(i + 1)
            for c in added:
                del cc[c]

    rec(0)
    return out


def extract_blocks(dense, legs, keys):
    """Blocks of a dense array at the given keys (tensor.py:700-713 gather)."""
    from oracle.core import block_slices

    out = {}
    for k in keys:
        masks, shape = block_slices(legs, k)
        if all(s > 0 for s in shape):
            out[k] = dense[np.ix_(*[np.where(m)[0] for m in masks])].copy()
    return out


def blockwise_contract_dense(tensors, subscripts, out_legs):
    """Same result as blockwise_contract: output keys from the key-level
    recursion, values from one dense einsum of the densified operands (the
    sum over block combinations restricted to those keys)."""
    keys = blockwise_keys([list(t.blocks) for t in tensors], subscripts)
    dense = _einsum(subscripts, [t.todense() for t in tensors])
    return SymT(extract_blocks(dense, out_legs, keys), out_legs, validate=False)


def contract(*tensors, output_labels=None):
    """contractor.py:490-531 for symmetric operands."""
    if not tensors:
        raise ValueError("contract() requires at least one tensor")
    subs, _ = labels_to_subscripts(tensors, output_labels)
    if len(tensors) == 1:
        lhs, rhs = subs.split("->")
        if lhs == rhs:
            return tensors[0]
    return contract_symmetric(list(tensors), subs)[0]
