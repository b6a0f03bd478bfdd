"""GPU HOTRG (Z2 block-sparse and dense) against the FP64 oracle and Onsager.

Gauge: the product's isometries may differ from LAPACK's by a sign per
singular vector (and a rotation inside exactly degenerate subspaces).  Signs
leave |T| and hence every max-abs log norm unchanged, so ln Z/N is compared
directly; tensors are compared through |entries|.

Conditioning: every step, taken from the same input tensor, agrees with the
oracle to ~1e-15 in its log norm (checked along the product's own trajectory
for all 20 steps).  Whole runs are not determined to that level by the
reference algorithm itself: from step ~11 on (beta = 0.3, chi = 16) the kept
spectrum holds pairs degenerate to 1e-16 s_max and the cut falls between
values 8e-9 s_max apart, so which vectors survive depends on the SVD
implementation's rounding.  Measured: the oracle moves by 7e-7 in ln Z/N when
only its LAPACK driver changes (gesdd -> gesvd), by < 1e-13 under 1e-15 input
noise, and the product's run differs from the oracle's by up to 1.6e-5.
Whole runs are compared at RUN_TOL = 1e-4, a hundredth of the reference's own
Onsager tolerance.
"""


import numpy as np
import pytest

from oracle import hotrg as oh
from tests.helpers import assert_same_structure, to_oracle

pytestmark = pytest.mark.gpu

RUN_TOL = 1e-4


def _trajectory_parity(T, chi, steps, direction_order="alternating"):
    """Step the product; at each step run the oracle step on the same input."""
    from tenax_amd.algorithms.hotrg import _hotrg_step_horizontal, _hotrg_step_vertical

    worst = 0.0
    for step in range(steps):
        horiz = (step % 2 == 0) if direction_order == "alternating" else direction_order == "horizontal"
        f, fo = (_hotrg_step_horizontal, oh.step_horizontal) if horiz else (_hotrg_step_vertical, oh.step_vertical)
        To = to_oracle(T)
        T, ln = f(T, chi)
        _, lno = fo(To, chi, order="chi5")
        worst = max(worst, abs(ln - lno))
    return worst


def _abs_blocks_close(p, o, tol):
    po = to_oracle(p)
    assert_same_structure(p, o)
    for k, b in o.blocks.items():
        assert np.abs(np.abs(po.blocks[k]) - np.abs(b)).max() < tol


@pytest.mark.parametrize("beta", [0.3, 0.44])
def test_single_steps_match_oracle(gpu, beta):
    from tenax_amd.algorithms.hotrg import _hotrg_step_horizontal, _hotrg_step_vertical
    from tenax_amd.algorithms.trg import compute_ising_tensor

    T = compute_ising_tensor(beta, symmetric=True)
    To = oh.ising_tensor_z2(beta)
    for step in range(4):
        f, fo = (_hotrg_step_horizontal, oh.step_horizontal) if step % 2 == 0 else (_hotrg_step_vertical,
                                                                                  oh.step_vertical)
        T, ln = f(T, 8)
        To, lno = fo(To, 8, order="chi5")
        assert abs(ln - lno) < 1e-12
        _abs_blocks_close(T, To, 1e-11)
        To = to_oracle(T)  # continue from the product's gauge


@pytest.mark.parametrize("beta,tol", [(0.2, 0.005), (0.3, 0.01), (0.44, 0.02), (0.6, 0.005)])
def test_z2_hotrg_matches_oracle_and_onsager(gpu, beta, tol):
    """test_hotrg.py:138-199, 304-317 (chi=16, 20 steps) plus oracle parity."""
    from tenax_amd.algorithms.hotrg import HOTRGConfig, hotrg
    from tenax_amd.algorithms.trg import compute_ising_tensor, ising_free_energy_exact

    assert _trajectory_parity(compute_ising_tensor(beta, symmetric=True), 16, 20) < 1e-13
    lnz = float(hotrg(compute_ising_tensor(beta, symmetric=True), HOTRGConfig(max_bond_dim=16, num_steps=20)))
    assert abs(lnz - oh.hotrg(oh.ising_tensor_z2(beta), 16, 20, order="chi5")) < RUN_TOL
    ex = ising_free_energy_exact(beta)
    assert abs(ex - oh.onsager_free_energy(beta)) < 1e-14
    assert abs(-lnz / beta - ex) / abs(ex) < tol


def test_dense_hotrg_matches_oracle(gpu):
    """DenseTensor path: every step matches the oracle's dense step from the
    same input; the whole run meets the reference's accuracy test (the dense
    chi=8 run cuts through degenerate triplets, where the oracle alone moves
    by 3e-4 between LAPACK drivers)."""
    from tenax_amd.algorithms.hotrg import HOTRGConfig, _hotrg_step_horizontal, _hotrg_step_vertical, hotrg
    from tenax_amd.algorithms.trg import compute_ising_tensor, ising_free_energy_exact
    from tenax_amd.core.tensor import DenseTensor

    T = compute_ising_tensor(0.3)
    assert isinstance(T, DenseTensor)
    for step in range(10):
        f = _hotrg_step_horizontal if step % 2 == 0 else _hotrg_step_vertical
        To = T.todense().cpu().numpy()
        T, ln = f(T, 8)
        _, lno = oh.step_dense(To, 8, step % 2 == 0)
        assert abs(ln - lno) < 1e-13
        assert isinstance(T, DenseTensor)
    lnz = float(hotrg(compute_ising_tensor(0.3), HOTRGConfig(max_bond_dim=16, num_steps=20)))
    ex = ising_free_energy_exact(0.3)
    assert abs(-lnz / 0.3 - ex) / abs(ex) < 0.01


@pytest.mark.parametrize("order", ["horizontal", "vertical"])
def test_single_direction(gpu, order):
    from tenax_amd.algorithms.hotrg import HOTRGConfig, hotrg
    from tenax_amd.algorithms.trg import compute_ising_tensor

    lnz = float(hotrg(compute_ising_tensor(0.3, symmetric=True),
                      HOTRGConfig(max_bond_dim=4, num_steps=4, direction_order=order)))
    assert abs(lnz - oh.hotrg(oh.ising_tensor_z2(0.3), 4, 4, direction_order=order)) < RUN_TOL
    assert _trajectory_parity(compute_ising_tensor(0.3, symmetric=True), 4, 4, order) < 1e-13


def test_blocks_preserved(gpu):
    """test_hotrg.py:288-302: Z2 sectors never collapse."""
    from tenax_amd.algorithms.hotrg import _hotrg_step_horizontal
    from tenax_amd.algorithms.trg import compute_ising_tensor
    from tenax_amd.core.tensor import SymmetricTensor

    T = compute_ising_tensor(0.3, symmetric=True)
    assert T.n_blocks == 8
    for _ in range(3):
        T, _ = _hotrg_step_horizontal(T, 8)
        assert isinstance(T, SymmetricTensor) and T.n_blocks >= 8


def test_batched_equals_single(gpu):
    """Lockstep instances (one shared eigh per step) step exactly like single ones."""
    from tenax_amd.algorithms.hotrg import HOTRGConfig, _steps, _H, _V, hotrg_batched
    from tenax_amd.algorithms.trg import compute_ising_tensor

    betas = [1 / 2.0, 1 / 2.27, 1 / 2.6]
    Ts = [compute_ising_tensor(b, symmetric=True) for b in betas]
    for step in range(6):
        roles = _H if step % 2 == 0 else _V
        many = _steps(Ts, roles, 12, None)
        for T, (Tm, lnm) in zip(Ts, many):
            _, ln1 = _steps([T], roles, 12, None)[0]
            assert abs(lnm - ln1) < 1e-13
        Ts = [t for t, _ in many]
    cfg = HOTRGConfig(max_bond_dim=12, num_steps=12)
    vals = hotrg_batched([compute_ising_tensor(b, symmetric=True) for b in betas], cfg)
    for b, v in zip(betas, vals):
        assert abs(v - oh.hotrg(oh.ising_tensor_z2(b), 12, 12, order="chi5")) < RUN_TOL
