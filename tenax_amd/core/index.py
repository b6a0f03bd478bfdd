"""Leg metadata: charges per basis state, flow and label.

Mirrors reference src/tenax/core/index.py:27-195 (``FlowDirection``,
``TensorIndex`` with ``dim/dual/flip_flow/relabel/is_dual_of/compatible_with``).
Adds ``sectors()``: the sorted distinct charges and their multiplicities, the
per-leg descriptor the device planner consumes (tnx_leg in include/tnx.h).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from tenax_amd.core.symmetry import BaseSymmetry

Label = str | int


class FlowDirection(IntEnum):
    IN = 1
    OUT = -1


@dataclass(frozen=True, slots=True)
class TensorIndex:
    symmetry: BaseSymmetry
    charges: np.ndarray
    flow: FlowDirection
    label: Label = ""
    _sectors: tuple = field(default=None, init=False, repr=False, compare=False, hash=False)

    def __post_init__(self):
        ch = np.asarray(self.charges)
        if ch.ndim != 1:
            raise ValueError(f"charges must be 1-D, got shape {ch.shape}")
        if ch.dtype != np.int32:
            ch = ch.astype(np.int32)
        object.__setattr__(self, "charges", ch)
        if not isinstance(self.flow, FlowDirection):
            object.__setattr__(self, "flow", FlowDirection(int(self.flow)))

    @property
    def dim(self) -> int:
        return int(self.charges.shape[0])

    def sectors(self):
        """(sorted distinct charges int32, multiplicities int32)."""
        s = self._sectors
        if s is None:
            u, c = np.unique(self.charges, return_counts=True)
            s = (u.astype(np.int32), c.astype(np.int32))
            object.__setattr__(self, "_sectors", s)
        return s

    def positions(self, q: int) -> np.ndarray:
        """Basis positions carrying charge q, in index order (tensor.py:168-188)."""
        return np.nonzero(self.charges == q)[0]

    def dual(self) -> "TensorIndex":
        return TensorIndex(self.symmetry, self.symmetry.dual(self.charges),
                           FlowDirection(-int(self.flow)), self.label)

    def flip_flow(self) -> "TensorIndex":
        return TensorIndex(self.symmetry, self.charges, FlowDirection(-int(self.flow)), self.label)

    def relabel(self, new_label: Label) -> "TensorIndex":
        return TensorIndex(self.symmetry, self.charges, self.flow, new_label)

    def is_dual_of(self, other: "TensorIndex") -> bool:
        if type(self.symmetry) is not type(other.symmetry):
            return False
        if self.flow == other.flow or self.dim != other.dim:
            return False
        return bool(np.array_equal(self.charges, self.symmetry.dual(other.charges)))

    def compatible_with(self, other: "TensorIndex") -> bool:
        return (type(self.symmetry) is type(other.symmetry) and self.dim == other.dim
                and self.flow != other.flow)

    def __hash__(self):
        return hash((self.symmetry, self.charges.tobytes(), int(self.flow), self.label))

    def __eq__(self, other):
        if not isinstance(other, TensorIndex):
            return NotImplemented
        return (type(self.symmetry) is type(other.symmetry) and self.symmetry == other.symmetry
                and np.array_equal(self.charges, other.charges) and self.flow == other.flow
                and self.label == other.label)

    def __repr__(self):
        return (f"TensorIndex(sym={self.symmetry!r}, dim={self.dim}, "
                f"flow={self.flow.name}, label={self.label!r})")
