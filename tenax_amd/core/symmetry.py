"""Charge algebra of abelian symmetries (host metadata, NumPy integers).

Mirrors the public interface of reference src/tenax/core/symmetry.py:29-255:
``BaseSymmetry.fuse/dual/identity/n_values/fuse_many/is_conserved``, ``U1Symmetry``
(fuse = +, dual = -q, unbounded) and ``ZnSymmetry`` (fuse = + mod n).  The
device planner receives the group as ``sym_n`` (0 for U(1), n for Z_n).
"""

from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np


class BaseSymmetry(ABC):
    """Abelian charge group acting on integer charge arrays (symmetry.py:29-179)."""

    @abstractmethod
    def fuse(self, charges_a: np.ndarray, charges_b: np.ndarray) -> np.ndarray: ...

    @abstractmethod
    def dual(self, charges: np.ndarray) -> np.ndarray: ...

    @abstractmethod
    def identity(self) -> int: ...

    @abstractmethod
    def n_values(self) -> int | None: ...

    @property
    def sym_n(self) -> int:
        """Group order handed to libtnx: 0 for U(1), n for Z_n."""
        n = self.n_values()
        return 0 if n is None else int(n)

    def fuse_many(self, charge_list):
        if not charge_list:
            raise ValueError("charge_list must be non-empty")
        acc = charge_list[0]
        for c in charge_list[1:]:
            acc = self.fuse(acc, c)
        return acc

    @property
    def is_fermionic(self) -> bool:
        return False

    def parity(self, charges):
        return np.zeros_like(charges)

    def is_conserved(self, charges_per_leg, flows, target=None) -> bool:
        want = self.identity() if target is None else target
        net = 0
        for f, q in zip(flows, charges_per_leg):
            net += int(f) * int(q)
        n = self.n_values()
        if n is not None:
            net %= n
        return net == want


class U1Symmetry(BaseSymmetry):
    """U(1): integer charges, fusion by addition (symmetry.py:182-213)."""

    def fuse(self, charges_a, charges_b):
        return charges_a + charges_b

    def dual(self, charges):
        return -charges

    def identity(self) -> int:
        return 0

    def n_values(self):
        return None

    def __eq__(self, other):
        return isinstance(other, U1Symmetry)

    def __hash__(self):
        return hash("U1Symmetry")

    def __repr__(self):
        return "U1Symmetry()"


class ZnSymmetry(BaseSymmetry):
    """Z_n: charges mod n, fusion by addition mod n (symmetry.py:216-255)."""

    def __init__(self, n: int):
        if n < 2:
            raise ValueError(f"n must be >= 2, got {n}")
        self.n = int(n)

    def fuse(self, charges_a, charges_b):
        return (charges_a + charges_b) % self.n

    def dual(self, charges):
        return (-charges) % self.n

    def identity(self) -> int:
        return 0

    def n_values(self) -> int:
        return self.n

    def __eq__(self, other):
        return isinstance(other, ZnSymmetry) and other.n == self.n

    def __hash__(self):
        return hash(("ZnSymmetry", self.n))

    def __repr__(self):
        return f"ZnSymmetry({self.n})"
