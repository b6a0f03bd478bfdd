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
