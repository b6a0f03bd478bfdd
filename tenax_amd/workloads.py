"""Synthetic leg structures of the benchmark workloads.

Bond-charge multiplicities m_q ∝ exp(-q²/18) for q in [-12, 12], normalised to
a total dimension chi by largest-remainder rounding (ties broken by ascending
q).  At chi = 1024 this gives 21 non-empty sectors (±11, ±12 round to zero)
with m_0 = 136; at chi = 2048, m_0 = 272.
"""

from __future__ import annotations

import numpy as np


def gaussian_multiplicities(chi: int, qmax: int = 12, width2: float = 18.0) -> dict[int, int]:
    qs = np.arange(-qmax, qmax + 1)
    w = np.exp(-(qs.astype(np.float64) ** 2) / width2)
    raw = chi * w / w.sum()
    base = np.floor(raw).astype(np.int64)
    rem = int(chi - base.sum())
    frac = raw - base
    order = sorted(range(len(qs)), key=lambda i: (-frac[i], qs[i]))
    for i in order[:rem]:
        base[i] += 1
    return {int(q): int(m) for q, m in zip(qs, base) if m > 0}


def gaussian_bond_charges(chi: int, qmax: int = 12, width2: float = 18.0) -> np.ndarray:
    """Leg charge array: each charge repeated m_q times, ascending."""
    m = gaussian_multiplicities(chi, qmax, width2)
    return np.concatenate([np.full(v, q, np.int32) for q, v in sorted(m.items())])


PHYS_CHARGES = np.array([-1, 1], np.int32)


def dmrg_bond_multiplicities(n_left: int, L: int, chi: int, width2: float = 18.0) -> dict[int, int]:
    """Charge multiplicities of MPS bond after ``n_left`` sites of a spin-1/2 chain.

    Charges are 2Sz of the left block in the total-Sz = 0 sector: parity of
    n_left, |q| <= min(n_left, L - n_left).  Each charge is capped by the exact
    Schmidt-rank bound min(C(n_left, (n_left+q)/2), C(n_right, (n_right-q)/2)); the
    remaining budget up to chi follows exp(-q^2/width2) with largest-remainder
    rounding (ties by ascending q).  Models a converged DMRG bond of dimension chi.
    """
    from math import comb

    n_right = L - n_left
    nmax = min(n_left, n_right)
    qs = [q for q in range(-nmax, nmax + 1) if (q - n_left) % 2 == 0]
    cap = {q: min(comb(n_left, (n_left + q) // 2), comb(n_right, (n_right - q) // 2)) for q in qs}
    if sum(cap.values()) <= chi:
        return {q: c for q, c in cap.items() if c > 0}
    m = {q: 0 for q in qs}
    free = set(qs)
    budget = chi
    while budget > 0 and free:
        w = {q: np.exp(-(q * q) / width2) for q in free}
        tot = sum(w.values())
        raw = {q: budget * w[q] / tot for q in free}
        add = {q: int(np.floor(raw[q])) for q in free}
        rem = budget - sum(add.values())
        for q in sorted(free, key=lambda q: (-(raw[q] - add[q]), q))[:rem]:
            add[q] += 1
        over = False
        for q in list(free):
            room = cap[q] - m[q]
            if add[q] >= room:
                m[q] += room
                budget -= room
                free.discard(q)
                over = True
        if not over:
            for q in free:
                m[q] += add[q]
            budget = 0
    return {q: v for q, v in m.items() if v > 0}


def dmrg_bond_charges(n_left: int, L: int, chi: int, width2: float = 18.0) -> np.ndarray:
    m = dmrg_bond_multiplicities(n_left, L, chi, width2)
    return np.concatenate([np.full(v, q, np.int32) for q, v in sorted(m.items())])


def synthetic_dmrg_mps(L: int, chi: int, seed: int = 2, width2: float = 18.0):
    """Random U(1) MPS whose bonds carry the converged-state charge profile.

    Site i has legs (v_{i-1,i} IN, p_i IN, v_{i,i+1} OUT) (boundary sites drop
    the missing bond), physical charges [1, -1], total 2Sz = 0, blocks N(0,1)
    from PRNGKey(seed + i): the layout of build_random_symmetric_mps with bond
    dimension up to chi.  Each site is scaled to squared norm sqrt(D_l * D_r)
    (D = 1 for a missing bond), which makes it an isometry on average in both
    directions (E[sum_p A_p^T A_p] = I for D_l = D_r), so environments built
    from these sites keep O(1) entries over the whole chain instead of
    shrinking by ~1/chi per site.
    """
    from tenax_amd.core.index import FlowDirection, TensorIndex
    from tenax_amd.core.symmetry import U1Symmetry
    from tenax_amd.core.tensor import PRNGKey, SymmetricTensor
    from tenax_amd.network.network import TensorNetwork

    sym = U1Symmetry()
    phys = np.array([1, -1], np.int32)
    bonds = [dmrg_bond_charges(i + 1, L, chi, width2) for i in range(L - 1)]
    mps = TensorNetwork(name=f"synthetic_MPS_L{L}_chi{chi}")
    for i in range(L):
        if i == 0:
            idx = (TensorIndex(sym, phys, FlowDirection.IN, f"p{i}"),
                   TensorIndex(sym, bonds[0], FlowDirection.OUT, f"v{i}_{i + 1}"))
        elif i == L - 1:
            idx = (TensorIndex(sym, bonds[i - 1], FlowDirection.IN, f"v{i - 1}_{i}"),
                   TensorIndex(sym, phys, FlowDirection.IN, f"p{i}"))
        else:
            idx = (TensorIndex(sym, bonds[i - 1], FlowDirection.IN, f"v{i - 1}_{i}"),
                   TensorIndex(sym, phys, FlowDirection.IN, f"p{i}"),
                   TensorIndex(sym, bonds[i], FlowDirection.OUT, f"v{i}_{i + 1}"))
        t = SymmetricTensor.random_normal(idx, PRNGKey(seed + i), target=0 if i == L - 1 else None)
        dl = len(bonds[i - 1]) if i > 0 else 1
        dr = len(bonds[i]) if i < L - 1 else 1
        mps.add_node(i, t * (float(dl * dr) ** 0.25 / float(t.norm())))
    for i in range(L - 1):
        mps.connect(i, f"v{i}_{i + 1}", i + 1, f"v{i}_{i + 1}")
    return mps


def centre_bond_problem(L: int, chi: int, seed: int = 2):
    """The two-site problem at the centre bond (L/2 - 1, L/2) of the Heisenberg
    chain with the converged-state charge profile (synthetic_dmrg_mps): left and
    right environments built on the device from the synthetic sites, each
    rescaled to max|.| = 1 per site (the synthetic sites are not canonical; the
    scale keeps FP32 variants in range and changes no structure), theta
    normalised.  Returns (problem, (left, Wl, Wr, right, theta), mps, mpo)."""
    import importlib

    from tenax_amd.algorithms.dmrg import build_mpo_heisenberg

    pd = importlib.import_module("tenax_amd.algorithms.dmrg")
    mpo_tn = build_mpo_heisenberg(L)
    mps_tn = synthetic_dmrg_mps(L, chi, seed=seed)
    mpo = [mpo_tn.get_tensor(i) for i in range(L)]
    mps = [mps_tn.get_tensor(i) for i in range(L)]
    c = L // 2 - 1
    left = pd._build_trivial_left_env_symmetric()
    for i in range(c):
        left = pd._update_left_env_symmetric(left, mps[i], mpo[i])
        left = left * (1.0 / float(left.max_abs()))
    right = pd._build_trivial_right_env_symmetric()
    for i in range(L - 1, c + 1, -1):
        right = pd._update_right_env_symmetric(right, mps[i], mpo[i])
        right = right * (1.0 / float(right.max_abs()))
    theta, _, _ = pd._theta_padded(mps[c], mps[c + 1])
    theta = theta * (1.0 / float(theta.norm()))
    prob = pd.TwoSiteProblem(theta, left, mpo[c], mpo[c + 1], right)
    return prob, (left, mpo[c], mpo[c + 1], right, theta), mps, mpo
