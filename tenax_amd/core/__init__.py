"""Core tensor and symmetry classes (mirrors reference src/tenax/core/__init__.py)."""

from tenax_amd.core.index import FlowDirection, Label, TensorIndex
from tenax_amd.core.symmetry import BaseSymmetry, FermionicU1, FermionParity, U1Symmetry, ZnSymmetry
from tenax_amd.core.tensor import BlockKey, DenseTensor, PRNGKey, SymmetricTensor, Tensor, fold_in, inner

# zero-guards used by the algorithms (reference core/__init__.py:15-16)
EPS = 1e-15
LOG_EPS = 1e-250

__all__ = [
    "BaseSymmetry", "U1Symmetry", "ZnSymmetry", "FermionParity", "FermionicU1", "FlowDirection", "Label", "TensorIndex", "Tensor",
    "DenseTensor", "SymmetricTensor", "BlockKey", "PRNGKey", "fold_in", "inner", "EPS", "LOG_EPS",
]
