"""Finite-automaton MPO builder with U(1) bond charges (host side, once per run).

Mirrors reference src/tenax/algorithms/auto_mpo.py:38-48 (spin operators),
93-205 (bond states: done = 0, one state per in-flight term, vacuum = last;
W[state_l, :, :, state_r] entries), 245-318 (operator and bond charges) and
321-386 (wrap as SymmetricTensor with legs (w_l IN, top IN, bot OUT, w_r OUT)).
The W matrices are small host arrays; the SymmetricTensor blocks are gathered
on the device by ``SymmetricTensor.from_dense``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from tenax_amd.core.index import FlowDirection, TensorIndex
from tenax_amd.core.symmetry import U1Symmetry
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor
from tenax_amd.network.network import TensorNetwork


def spin_half_ops() -> dict[str, np.ndarray]:
    return {
        "Sz": np.array([[0.5, 0.0], [0.0, -0.5]]),
        "Sp": np.array([[0.0, 1.0], [0.0, 0.0]]),
        "Sm": np.array([[0.0, 0.0], [1.0, 0.0]]),
        "Id": np.eye(2),
    }


def spin_one_ops() -> dict[str, np.ndarray]:
    r2 = np.sqrt(2.0)
    return {
        "Sz": np.diag([1.0, 0.0, -1.0]),
        "Sp": np.array([[0.0, r2, 0.0], [0.0, 0.0, r2], [0.0, 0.0, 0.0]]),
        "Sm": np.array([[0.0, 0.0, 0.0], [r2, 0.0, 0.0], [0.0, r2, 0.0]]),
        "Id": np.eye(3),
    }


@dataclass(frozen=True)
class HamiltonianTerm:
    coefficient: complex
    ops: tuple


def _operator_charge(op: np.ndarray, phys: np.ndarray) -> int:
    """Charge transfer phys[i] - phys[j] shared by all nonzero op[i, j]."""
    nz = np.argwhere(np.abs(op) > 1e-14)
    if len(nz) == 0:
        return 0
    qs = {int(phys[i]) - int(phys[j]) for i, j in nz}
    if len(qs) != 1:
        a, b = sorted(qs)[:2]
        raise ValueError(f"Operator mixes charge sectors: found transfers {a} and {b}.")
    return qs.pop()


class AutoMPO:
    """Symbolic Hamiltonian -> MPO (auto_mpo.py:453-635)."""

    def __init__(self, L: int, d: int = 2, site_ops: dict | None = None):
        if L < 1:
            raise ValueError(f"L must be >= 1, got {L}")
        self.L, self.d = L, d
        if site_ops is not None:
            self._ops = site_ops
        elif d == 2:
            self._ops = spin_half_ops()
        elif d == 3:
            self._ops = spin_one_ops()
        else:
            raise ValueError(f"No default site_ops for d={d}; provide site_ops explicitly.")
        self._terms: list[HamiltonianTerm] = []

    def add_term(self, coeff, *args):
        if len(args) == 0 or len(args) % 2:
            raise ValueError(f"args must be alternating (op_name, site) pairs; got {len(args)} argument(s).")
        pairs = []
        for k in range(0, len(args), 2):
            name, site = args[k], args[k + 1]
            if not isinstance(name, str):
                raise TypeError(f"Expected operator name (str) at position {k}, got {type(name).__name__}.")
            if name not in self._ops:
                raise KeyError(f"Operator '{name}' not in site_ops. Available: {sorted(self._ops)}")
            if not isinstance(site, int) or not 0 <= site < self.L:
                raise ValueError(f"Site {site!r} out of range [0, {self.L}).")
            pairs.append((site, self._ops[name]))
        pairs.sort(key=lambda p: p[0])
        sites = [s for s, _ in pairs]
        if len(set(sites)) != len(sites):
            raise ValueError(f"Duplicate sites in term: {sites}")
        self._terms.append(HamiltonianTerm(complex(coeff), tuple(pairs)))

    def __iadd__(self, args):
        self.add_term(args[0], *args[1:])
        return self

    def n_terms(self) -> int:
        return len(self._terms)

    def _bond_states(self):
        out = []
        for j in range(self.L - 1):
            st = {}
            for tid, t in enumerate(self._terms):
                sites = [s for s, _ in t.ops]
                if min(sites) <= j < max(sites):
                    st[tid] = len(st) + 1
            out.append(st)
        return out

    def bond_dims(self):
        return [len(b) + 2 for b in self._bond_states()]

    def _w_matrices(self, bstates):
        L, d = self.L, self.d
        ident = np.eye(d)
        cplx = any(t.coefficient.imag != 0.0 for t in self._terms)
        dt = np.complex128 if cplx else np.float64
        ws = []
        for i in range(L):
            Dl = 1 if i == 0 else len(bstates[i - 1]) + 2
            Dr = 1 if i == L - 1 else len(bstates[i]) + 2
            W = np.zeros((Dl, d, d, Dr), dt)
            vac_l = 0 if i == 0 else Dl - 1
            vac_r = Dr - 1
            if L > 1:
                if i == 0:
                    W[0, :, :, vac_r] = ident
                elif i == L - 1:
                    W[0, :, :, 0] = ident
                else:
                    W[0, :, :, 0] = ident
                    W[vac_l, :, :, vac_r] = ident
            for tid, t in enumerate(self._terms):
                od = dict(t.ops)
                lo, hi = min(od), max(od)
                if i < lo or i > hi:
                    continue
                op = od.get(i, ident)
                c = t.coefficient if cplx else t.coefficient.real
                if lo == hi:
                    W[vac_l, :, :, 0] += c * op
                elif i == lo:
                    W[vac_l, :, :, bstates[i][tid]] += c * op
                elif i == hi:
                    W[bstates[i - 1][tid], :, :, 0] += op
                else:
                    W[bstates[i - 1][tid], :, :, bstates[i][tid]] = op
            ws.append(W)
        return ws

    def to_mpo(self, compress: bool = False, compress_tol: float = 1e-12, dtype=None, symmetric: bool = False,
               phys_charges=None) -> TensorNetwork:
        if not self._terms:
            raise ValueError("No terms; call add_term() before to_mpo().")
        if compress:
            raise NotImplementedError("MPO compression is not on the accelerated path")
        bst = self._bond_states()
        ws = self._w_matrices(bst)
        L = self.L
        sym = U1Symmetry()
        mpo = TensorNetwork(name="AutoMPO_sym" if symmetric else "AutoMPO")
        if symmetric:
            if phys_charges is None:
                phys = {2: np.array([1, -1], np.int32), 3: np.array([2, 0, -2], np.int32)}.get(self.d)
                if phys is None:
                    raise ValueError(f"No default phys_charges for d={self.d}; provide phys_charges explicitly.")
            else:
                phys = np.asarray(phys_charges, np.int32)
            opq = [[_operator_charge(op, phys) for _, op in t.ops] for t in self._terms]
            bcs = []
            for j in range(L - 1):
                ch = np.zeros(len(bst[j]) + 2, np.int32)
                for tid, si in bst[j].items():
                    ch[si] = sum(q for (s, _), q in zip(self._terms[tid].ops, opq[tid]) if s <= j)
                bcs.append(ch)
        for i, W in enumerate(ws):
            ll = "w_left_0" if i == 0 else f"w{i - 1}_{i}"
            rl = "w_right" if i == L - 1 else f"w{i}_{i + 1}"
            if symmetric:
                lc = np.zeros(W.shape[0], np.int32) if i == 0 else bcs[i - 1]
                rc = np.zeros(W.shape[3], np.int32) if i == L - 1 else bcs[i]
                idx = (TensorIndex(sym, lc, FlowDirection.IN, ll), TensorIndex(sym, phys, FlowDirection.IN, f"mpo_top_{i}"),
                       TensorIndex(sym, phys, FlowDirection.OUT, f"mpo_bot_{i}"),
                       TensorIndex(sym, rc, FlowDirection.OUT, rl))
                mpo.add_node(i, SymmetricTensor.from_dense(W.real.astype(np.float64), idx))
            else:
                z = np.zeros(self.d, np.int32)
                idx = (TensorIndex(sym, np.zeros(W.shape[0], np.int32), FlowDirection.IN, ll),
                       TensorIndex(sym, z, FlowDirection.IN, f"mpo_top_{i}"),
                       TensorIndex(sym, z, FlowDirection.OUT, f"mpo_bot_{i}"),
                       TensorIndex(sym, np.zeros(W.shape[3], np.int32), FlowDirection.OUT, rl))
                mpo.add_node(i, DenseTensor(W, idx))
        for i in range(L - 1):
            shared = set(mpo.get_tensor(i).labels()) & set(mpo.get_tensor(i + 1).labels())
            for lbl in sorted(shared, key=str):
                mpo.connect(i, lbl, i + 1, lbl)
        return mpo


def build_auto_mpo(terms_spec, L: int, d: int = 2, site_ops=None, compress: bool = False,
                   compress_tol: float = 1e-12, dtype=None, symmetric: bool = False, phys_charges=None):
    """auto_mpo.py:643-689."""
    a = AutoMPO(L=L, d=d, site_ops=site_ops)
    for t in terms_spec:
        a.add_term(t[0], *t[1:])
    return a.to_mpo(compress=compress, compress_tol=compress_tol, symmetric=symmetric, phys_charges=phys_charges)
