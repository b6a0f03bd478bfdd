"""Conversions between the product's packed tensors and the oracle's dict tensors."""

import numpy as np

from oracle import core as oc


def _oracle_sym(sym):
    if sym.is_fermionic:
        return oc.FermionicU1() if sym.n_values() is None else oc.FermionParity()
    return oc.U1() if sym.n_values() is None else oc.Zn(sym.n_values())


def oracle_legs(indices):
    return tuple(oc.Leg(_oracle_sym(ix.symmetry), ix.charges, int(ix.flow), ix.label) for ix in indices)


def to_oracle(t):
    """Product SymmetricTensor -> oracle SymT (host float64 copies)."""
    host = t.data.detach().cpu().numpy()
    blocks = {}
    for i, k in enumerate(t.keys):
        shp = tuple(int(s) for s in t.shapes[i])
        o = int(t.offsets[i])
        blocks[tuple(int(q) for q in k)] = host[o:o + int(np.prod(shp))].reshape(shp).copy()
    return oc.SymT(blocks, oracle_legs(t.indices), validate=False)


def product_indices(legs):
    from tenax_amd import FermionicU1, FermionParity, FlowDirection, TensorIndex, U1Symmetry, ZnSymmetry

    out = []
    for l in legs:
        if oc.is_fermionic(l.sym):
            sym = FermionicU1() if l.sym.n is None else FermionParity()
        else:
            sym = U1Symmetry() if l.sym.n is None else ZnSymmetry(l.sym.n)
        out.append(TensorIndex(sym, l.charges, FlowDirection(l.flow), l.label))
    return tuple(out)


def from_oracle(s):
    """Oracle SymT -> product SymmetricTensor (sorted packing)."""
    import torch

    from tenax_amd import SymmetricTensor

    idx = product_indices(s.legs)
    keys = sorted(s.blocks)
    r = len(idx)
    shapes = np.array([s.blocks[k].shape for k in keys], np.int32).reshape(len(keys), r)
    offs, tot = SymmetricTensor._layout(shapes)
    host = np.concatenate([s.blocks[k].reshape(-1) for k in keys]) if keys else np.zeros(0)
    data = torch.from_numpy(host.astype(np.float64)).cuda()
    return SymmetricTensor._wrap(idx, np.array(keys, np.int32).reshape(len(keys), r), shapes, offs, data)


def assert_same_structure(p, o):
    """Bit-exact sector structure: same sorted keys and block shapes."""
    ok = sorted(o.blocks)
    pk = [tuple(int(q) for q in k) for k in p.keys]
    assert pk == ok, f"keys differ: product {len(pk)} vs oracle {len(ok)}"
    for i, k in enumerate(pk):
        assert tuple(int(s) for s in p.shapes[i]) == o.blocks[k].shape, f"shape of {k}"


def max_rel_err(p, o):
    po = to_oracle(p)
    num = 0.0
    den = 0.0
    for k, b in o.blocks.items():
        num = max(num, float(np.max(np.abs(po.blocks[k] - b))) if b.size else 0.0)
        den = max(den, float(np.max(np.abs(b))) if b.size else 0.0)
    return num / max(den, 1e-300)
