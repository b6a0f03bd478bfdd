"""HOTRG host-side behaviour that needs no GPU (hotrg.py:54-106 checks)."""

import numpy as np
import pytest

from tenax_amd.algorithms.hotrg import HOTRGConfig, hotrg
from tenax_amd.algorithms.trg import ising_free_energy_exact
from oracle import hotrg as oh


def test_config_defaults():
    c = HOTRGConfig()
    assert (c.max_bond_dim, c.num_steps, c.direction_order, c.svd_trunc_err) == (16, 10, "alternating", None)


def test_rejects_non_tensor():
    with pytest.raises(TypeError, match="requires a Tensor"):
        hotrg(np.random.default_rng(0).random((2, 2, 2, 2)), HOTRGConfig(max_bond_dim=4, num_steps=3))


def test_rejects_bad_direction():
    with pytest.raises(ValueError, match="Invalid direction_order"):
        hotrg(np.zeros((2, 2, 2, 2)), HOTRGConfig(direction_order="diagonal"))


@pytest.mark.parametrize("beta", [0.0, 0.2, 0.4406868, 0.6])
def test_onsager_matches_oracle(beta):
    assert abs(ising_free_energy_exact(beta) - oh.onsager_free_energy(beta)) < 1e-14
