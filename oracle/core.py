"""Oracle: charge algebra, leg metadata and dict-of-blocks symmetric tensors.

Restates reference src/tenax/core/symmetry.py:182-255, 311-440,
core/index.py:27-195 and core/tensor.py:32-50, 53-188, 296-319, 505-974 in
NumPy FP64.
Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import numpy as np

IN = 1
OUT = -1


# ---------------------------------------------------------------- symmetry
class U1:
    """core/symmetry.py:182-213: fuse = add, dual = negate, identity 0."""

    n = None

    def fuse(self, a, b):
        return a + b

    def dual(self, q):
        return -q

    def __eq__(self, o):
        return isinstance(o, U1)

    def __hash__(self):
        return hash("U1")


class Zn:
    """core/symmetry.py:216-255: fuse = add mod n, dual = -q mod n."""

    def __init__(self, n):
        if n < 2:
            raise ValueError(f"n must be >= 2, got {n}")
        self.n = n

    def fuse(self, a, b):
        return (a + b) % self.n

    def dual(self, q):
        return (-q) % self.n

    def __eq__(self, o):
        return isinstance(o, Zn) and o.n == self.n

    def __hash__(self):
        return hash(("Zn", self.n))


class FermionParity(Zn):
    """core/symmetry.py:311-371: Z2 with fermionic exchange; parity = q."""

    fermionic = True

    def __init__(self):
        super().__init__(2)

    def __eq__(self, o):
        return isinstance(o, FermionParity)

    def __hash__(self):
        return hash("FermionParity")


class FermionicU1(U1):
    """core/symmetry.py:374-440: U(1) with parity |q| mod 2 (default grading)."""

    fermionic = True

    def __eq__(self, o):
        return isinstance(o, FermionicU1)

    def __hash__(self):
        return hash("FermionicU1")


def is_fermionic(sym) -> bool:
    return bool(getattr(sym, "fermionic", False))


def parity(sym, q) -> int:
    """symmetry.py:99-109, 345-346, 426-427: 0 for bosonic groups, |q| mod 2 otherwise."""
    return abs(int(q)) % 2 if is_fermionic(sym) else 0


def koszul_sign(parities, perm) -> int:
    """tensor.py:32-50: -1 per inversion of ``perm`` between two odd elements
    (parities indexed by the ORIGINAL positions)."""
    sign = 1
    for i in range(len(perm)):
        for j in range(i + 1, len(perm)):
            if perm[i] > perm[j] and parities[perm[i]] and parities[perm[j]]:
                sign = -sign
    return sign


def fuse_many(sym, qs):
    """core/symmetry.py:73-87: left fold of fuse."""
    r = qs[0]
    for q in qs[1:]:
        r = sym.fuse(r, q)
    return r


class Leg:
    """core/index.py:43-195: int32 charges per basis state, flow, label."""

    __slots__ = ("sym", "charges", "flow", "label")

    def __init__(self, sym, charges, flow, label=""):
        self.sym = sym
        self.charges = np.asarray(charges).astype(np.int32)
        self.flow = int(flow)
        self.label = label

    @property
    def dim(self):
        return len(self.charges)

    def dual(self):
        # index.py:88-103: dual charges and flipped flow
        return Leg(self.sym, self.sym.dual(self.charges), -self.flow, self.label)

    def flip_flow(self):
        # index.py:105-122: flip flow, keep charges
        return Leg(self.sym, self.charges, -self.flow, self.label)

    def relabel(self, lbl):
        return Leg(self.sym, self.charges, self.flow, lbl)


# ------------------------------------------------------- sector enumeration
def compute_valid_blocks(legs, target=None):
    """core/tensor.py:53-165: incremental partial-sum enumeration.

    Returns keys in the reference's propagation order (a list).
    """
    if not legs:
        return [()]
    sym = legs[0].sym
    tgt = 0 if target is None else target
    uniq = [sorted(set(l.charges.tolist())) for l in legs]
    n = len(legs)
    if n == 1:
        return [(q,) for q in uniq[0] if legs[0].flow * q == tgt]
    partial = {}
    for q in uniq[0]:
        partial.setdefault(legs[0].flow * q, []).append((q,))
    for i in range(1, n - 1):
        f = legs[i].flow
        nxt = {}
        for q in uniq[i]:
            for prev, combos in partial.items():
                nf = int(sym.fuse(np.int64(prev), np.int64(f * q)))
                ext = [c + (q,) for c in combos]
                if nf in nxt:
                    nxt[nf].extend(ext)
                else:
                    nxt[nf] = ext
        partial = nxt
    fl = legs[-1].flow
    out = []
    if sym.n is None:
        last = set(uniq[-1])
        for prev, combos in partial.items():
            ql = (tgt - prev) * fl
            if ql in last:
                for c in combos:
                    out.append(c + (ql,))
    else:
        for q in uniq[-1]:
            for prev, combos in partial.items():
                if int(sym.fuse(np.int64(prev), np.int64(fl * q))) == tgt:
                    for c in combos:
                        out.append(c + (q,))
    return out


def block_slices(legs, key):
    """core/tensor.py:168-188: positions where charges == q, per leg."""
    masks = tuple(l.charges == q for l, q in zip(legs, key))
    return masks, tuple(int(m.sum()) for m in masks)


# ------------------------------------------------------------ generator
def normal(seed_words, shape):
    """N(0,1) block for the key PRNGKey(seed_words[0]) folded with the later
    words (jax.random.PRNGKey/fold_in/normal as used by core/tensor.py:655-660
    and dmrg.py:1484): the Threefry-2x32 stream of oracle/random.py."""
    from oracle import random as orand

    return orand.normal(orand.key_from_words(tuple(seed_words)), tuple(shape))


# ------------------------------------------------------ symmetric tensor
class SymT:
    """core/tensor.py:505-974: dict[key -> float64 block] + legs."""

    def __init__(self, blocks, legs, validate=True):
        self.legs = tuple(legs)
        self.blocks = dict(blocks)
        if validate:
            self._validate()

    def _validate(self):
        # tensor.py:551-568
        if not self.legs:
            return
        sym = self.legs[0].sym
        for key in self.blocks:
            v = fuse_many(sym, [np.int64(l.flow * q) for l, q in zip(self.legs, key)])
            if int(v) != 0:
                raise ValueError(f"Block {key} violates charge conservation")

    @property
    def ndim(self):
        return len(self.legs)

    def labels(self):
        return tuple(l.label for l in self.legs)

    @classmethod
    def zeros(cls, legs, target=None):
        # tensor.py:596-627
        blocks = {}
        for key in compute_valid_blocks(legs, target):
            _, shape = block_slices(legs, key)
            if all(s > 0 for s in shape):
                blocks[key] = np.zeros(shape)
        return cls(blocks, legs, validate=not (target not in (None, 0)))

    @classmethod
    def random_normal(cls, legs, seed_words, target=None, stddev=1.0):
        # tensor.py:629-668: block i of the sorted valid keys gets fold_in(key, i)
        blocks = {}
        for i, key in enumerate(sorted(compute_valid_blocks(legs, target))):
            _, shape = block_slices(legs, key)
            if all(s > 0 for s in shape):
                blocks[key] = normal(tuple(seed_words) + (i,), shape) * stddev
        return cls(blocks, legs, validate=not (target not in (None, 0)))

    @classmethod
    def from_dense(cls, data, legs, tol=1e-12):
        # tensor.py:670-732
        data = np.asarray(data, dtype=np.float64)
        full = np.zeros(data.shape, dtype=bool)
        blocks = {}
        for key in sorted(compute_valid_blocks(legs)):
            masks, shape = block_slices(legs, key)
            if not all(s > 0 for s in shape):
                continue
            grid = np.ix_(*[np.where(m)[0] for m in masks])
            blocks[key] = data[grid].copy()
            full[grid] = True
        if tol != float("inf"):
            outside = data[~full]
            if np.any(np.abs(outside) > tol):
                raise ValueError("data has non-zero elements outside symmetry-allowed sectors")
        return cls(blocks, legs)

    def todense(self):
        # tensor.py:760-781
        out = np.zeros(tuple(l.dim for l in self.legs))
        for key, b in self.blocks.items():
            masks, _ = block_slices(self.legs, key)
            out[np.ix_(*[np.where(m)[0] for m in masks])] = b
        return out

    def conj(self):
        return SymT({k: np.conj(v) for k, v in self.blocks.items()}, self.legs, False)

    def dagger(self):
        # tensor.py:783-816: dual keys and legs, conj data, and for fermionic
        # groups the product of twist phases (-1)^parity(q) over the key
        sym = self.legs[0].sym
        nb = {}
        for k, v in self.blocks.items():
            tw = (-1) ** sum(parity(sym, q) for q in k)
            nb[tuple(int(sym.dual(np.int64(q))) for q in k)] = np.conj(v) * tw
        return SymT(nb, tuple(l.dual() for l in self.legs), False)

    def bar(self):
        # tensor.py:823-830
        return SymT({k: np.conj(v) for k, v in self.blocks.items()},
                    tuple(l.flip_flow() for l in self.legs), False)

    def transpose(self, axes):
        # tensor.py:832-861: fermionic blocks take their Koszul sign (853-856)
        sym = self.legs[0].sym
        nb = {}
        for k, v in self.blocks.items():
            sg = koszul_sign([parity(sym, q) for q in k], axes) if is_fermionic(sym) else 1
            nb[tuple(k[i] for i in axes)] = np.transpose(v, axes) * sg
        return SymT(nb, tuple(self.legs[i] for i in axes), False)

    def norm(self):
        # tensor.py:863-868
        if not self.blocks:
            return 0.0
        return float(np.sqrt(sum(float(np.sum(np.abs(v) ** 2)) for v in self.blocks.values())))

    def relabels(self, mapping):
        return SymT(self.blocks, tuple(l.relabel(mapping[l.label]) if l.label in mapping else l
                                       for l in self.legs), False)

    def relabel(self, old, new):
        if old not in self.labels():
            raise KeyError(old)
        return self.relabels({old: new})

    def __mul__(self, s):
        return SymT({k: v * s for k, v in self.blocks.items()}, self.legs, False)

    __rmul__ = __mul__

    def __add__(self, o):
        # tensor.py:923-952: union of keys
        nb = {}
        for k in set(self.blocks) | set(o.blocks):
            if k in self.blocks and k in o.blocks:
                nb[k] = self.blocks[k] + o.blocks[k]
            elif k in self.blocks:
                nb[k] = self.blocks[k]
            else:
                nb[k] = o.blocks[k]
        return SymT(nb, self.legs, False)

    def __sub__(self, o):
        return self + o * -1.0

    def max_abs(self):
        if not self.blocks:
            return 0.0
        return float(max(np.max(np.abs(v)) for v in self.blocks.values()))


def inner(a, b):
    """core/tensor.py:296-319: sum over common keys of conj(a)*b."""
    t = 0.0
    for k in a.blocks:
        if k in b.blocks:
            t += float(np.sum(np.conj(a.blocks[k]) * b.blocks[k]))
    return t
