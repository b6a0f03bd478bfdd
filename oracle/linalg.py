"""Oracle: block-diagonal SVD with global truncation (FP64 LAPACK).

Restates reference src/tenax/linalg.py:34-69 (sector grouping) and 75-333
(per-sector SVD, global stable sort by -s, bond charges in global order),
and src/tenax/algorithms/_tensor_utils.py:44-62 (block-wise bond scaling).
Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import numpy as np

from oracle.core import IN, OUT, Leg, SymT, fuse_many, is_fermionic, koszul_sign, parity


def group_blocks_by_bond_charge(t, left_axes, right_axes):
    """linalg.py:34-69: sector q = fused flow*charge of the left legs."""
    sym = t.legs[0].sym
    g = {}
    for key, blk in t.blocks.items():
        q = int(fuse_many(sym, [np.int64(t.legs[i].flow * int(key[i])) for i in left_axes]))
        g.setdefault(q, []).append((tuple(key[i] for i in left_axes),
                                    tuple(key[i] for i in right_axes), blk))
    return g


def truncated_svd_symmetric(t, left_labels, right_labels, new_bond_label,
                            max_singular_values=None, max_truncation_err=None, driver="gesdd"):
    """linalg.py:75-333. Returns (U, s, Vh, s_full, sectors).

    ``sectors`` lists (q, n_rows, n_cols) per SVD'd sector in visit order.
    ``driver``: LAPACK routine (numpy's default gesdd, or gesvd), used to
    measure how much a result depends on the SVD implementation.
    """
    lab = t.labels()
    la = [lab.index(x) for x in left_labels]
    ra = [lab.index(x) for x in right_labels]
    grouped = group_blocks_by_bond_charge(t, la, ra)
    res = {}
    sectors = []
    for q, entries in grouped.items():
        lseen, rseen = {}, {}
        for lk, rk, _ in entries:
            lseen.setdefault(lk, len(lseen))
            rseen.setdefault(rk, len(rseen))
        lsz = [int(np.prod([np.sum(t.legs[ax].charges == c) for ax, c in zip(la, lk)])) for lk in lseen]
        rsz = [int(np.prod([np.sum(t.legs[ax].charges == c) for ax, c in zip(ra, rk)])) for rk in rseen]
        R, C = sum(lsz), sum(rsz)
        if R == 0 or C == 0:
            continue
        M = np.zeros((R, C))
        sym = t.legs[0].sym
        for lk, rk, blk in entries:
            li, ri = lseen[lk], rseen[rk]
            r0, c0 = sum(lsz[:li]), sum(rsz[:ri])
            b = blk.reshape(lsz[li], rsz[ri])
            if is_fermionic(sym):
                # linalg.py:171-186: Koszul sign of the reordering to (left, right)
                full = [0] * len(t.legs)
                for ax, c in zip(la, lk):
                    full[ax] = c
                for ax, c in zip(ra, rk):
                    full[ax] = c
                b = b * koszul_sign([parity(sym, c) for c in full], la + ra)
            M[r0:r0 + lsz[li], c0:c0 + rsz[ri]] = b
        if driver == "gesdd":
            U, s, Vh = np.linalg.svd(M, full_matrices=False)
        else:
            import scipy.linalg

            U, s, Vh = scipy.linalg.svd(M, full_matrices=False, lapack_driver=driver)
        res[q] = (U, s, Vh, list(lseen), list(rseen), lsz, rsz)
        sectors.append((q, R, C))
    pairs = []
    for q, (_, s, *_r) in res.items():
        for i, v in enumerate(s):
            pairs.append((float(v), q, i))
    pairs.sort(key=lambda x: -x[0])  # stable
    s_full = np.array([v for v, _, _ in pairs])
    n_total = len(pairs)
    n_keep = n_total
    if max_truncation_err is not None and n_total > 0:
        tot = sum(x[0] ** 2 for x in pairs)
        if tot > 0:
            tr = 0.0
            for i in range(n_total - 1, 0, -1):
                tr += pairs[i][0] ** 2
                if tr / tot > max_truncation_err ** 2:
                    n_keep = i + 1
                    break
    if max_singular_values is not None:
        n_keep = min(n_keep, max_singular_values)
    n_keep = max(1, min(n_keep, n_total))
    kept = pairs[:n_keep]
    bond = np.array([q for _, q, _ in kept], dtype=np.int32)
    s_fin = np.array([v for v, _, _ in kept])
    cols = {}
    for g, (_, q, i) in enumerate(kept):
        cols.setdefault(q, []).append((g, i))
    sym = t.legs[0].sym
    Ul = tuple(t.legs[i] for i in la) + (Leg(sym, bond, OUT, new_bond_label),)
    Vl = (Leg(sym, bond, IN, new_bond_label),) + tuple(t.legs[i] for i in ra)
    Ub, Vb = {}, {}
    for q, cl in cols.items():
        U, _, Vh, lks, rks, lsz, rsz = res[q]
        idx = [i for _, i in cl]
        Ut, Vt = U[:, idx], Vh[idx, :]
        ro = 0
        for li, lk in enumerate(lks):
            shp = tuple(int(np.sum(t.legs[ax].charges == c)) for ax, c in zip(la, lk))
            Ub[lk + (q,)] = Ut[ro:ro + lsz[li], :].reshape(shp + (len(idx),))
            ro += lsz[li]
        co = 0
        for ri, rk in enumerate(rks):
            shp = tuple(int(np.sum(t.legs[ax].charges == c)) for ax, c in zip(ra, rk))
            Vb[(q,) + rk] = Vt[:, co:co + rsz[ri]].reshape((len(idx),) + shp)
            co += rsz[ri]
    tgt = 0
    if t.blocks:
        k0 = next(iter(t.blocks))
        tgt = sum(l.flow * int(q) for l, q in zip(t.legs, k0))
    return (SymT(Ub, Ul, validate=tgt == 0), s_fin, SymT(Vb, Vl, validate=tgt == 0),
            s_full, sectors)


def scale_bond_axis(t, label, scale):
    """_tensor_utils.py:17-62: block-wise scaling by s at charge positions."""
    ax = t.labels().index(label)
    leg = t.legs[ax]
    nb = {}
    for key, blk in t.blocks.items():
        pos = np.where(leg.charges == key[ax])[0]
        n = blk.shape[ax]
        shp = [1] * t.ndim
        shp[ax] = n
        nb[key] = blk * np.asarray(scale)[pos[:n]].reshape(shp)
    return SymT(nb, t.legs, validate=False)
