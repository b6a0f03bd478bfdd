"""2D Ising model inputs for the tensor renormalisation group algorithms
(reference src/tenax/algorithms/trg.py:173-291)."""

from __future__ import annotations

import numpy as np

from tenax_amd.core.index import FlowDirection, TensorIndex
from tenax_amd.core.symmetry import U1Symmetry, ZnSymmetry
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor, Tensor


def _ising_dense(beta: float, J: float) -> np.ndarray:
    # T_udlr = sum_s W[u,s] W[d,s] W[l,s] W[r,s], W the matrix square root of
    # Q = exp(beta J s s') (trg.py:200-214); host FP64, 16 numbers
    spins = np.array([1.0, -1.0])
    Q = np.exp(beta * J * np.outer(spins, spins))
    ev, vec = np.linalg.eigh(Q)
    W = (vec * np.sqrt(ev)) @ vec.T
    return np.einsum("us,ds,ls,rs->udlr", W, W, W, W)


def compute_ising_tensor(beta: float, J: float = 1.0, symmetric: bool = False) -> Tensor:
    """Site tensor of the 2D Ising partition function, legs (up, down, left, right).

    symmetric=True: the tensor in the Z2 eigenbasis (Hadamard on every leg),
    a Z2 SymmetricTensor with 8 blocks, legs (IN, OUT, IN, OUT), charges [0, 1]
    (trg.py:216-232).  Otherwise a DenseTensor with trivial U(1) legs (233-243).
    """
    T = _ising_dense(beta, J)
    if symmetric:
        H = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
        Tz = np.einsum("ua,vb,wc,xd,uvwx->abcd", H, H, H, H, T)
        sym = ZnSymmetry(2)
        ch = np.array([0, 1], dtype=np.int32)
        idx = (TensorIndex(sym, ch, FlowDirection.IN, label="up"),
               TensorIndex(sym, ch, FlowDirection.OUT, label="down"),
               TensorIndex(sym, ch, FlowDirection.IN, label="left"),
               TensorIndex(sym, ch, FlowDirection.OUT, label="right"))
        return SymmetricTensor.from_dense(Tz, idx)
    sym = U1Symmetry()
    z = np.zeros(2, dtype=np.int32)
    idx = (TensorIndex(sym, z, FlowDirection.IN, label="up"),
           TensorIndex(sym, z, FlowDirection.OUT, label="down"),
           TensorIndex(sym, z, FlowDirection.IN, label="left"),
           TensorIndex(sym, z, FlowDirection.OUT, label="right"))
    return DenseTensor(T, idx)


def ising_free_energy_exact(beta: float, J: float = 1.0) -> float:
    """Onsager's free energy per site f = -ln Z / (N beta) (trg.py:245-291).

    ln Z/N = ln 2 + (1/(2 pi)) int_0^pi ln((A + sqrt(A^2 - s^2)) / 2) dt with
    A = cosh(2K)^2 - sinh(2K) cos t, s = sinh(2K), K = beta J (the inner
    integral of Onsager's double integral done in closed form), evaluated by
    the trapezoid rule on 10^4 points offset by half a step from both ends.
    """
    if beta == 0:
        return -float(np.log(2.0))
    K = beta * J
    c2, s2 = np.cosh(2 * K), np.sinh(2 * K)
    n = 10000
    h = np.pi / n
    t = np.linspace(h / 2, np.pi - h / 2, n)
    A = c2 * c2 - s2 * np.cos(t)
    g = np.log((A + np.sqrt(np.maximum(A * A - s2 * s2, 0.0))) / 2)
    integral = float(np.trapezoid(g, t)) / np.pi
    return -float(np.log(2.0)) / beta - integral / (2 * beta)
