"""Algorithms on the accelerated path (mirrors reference src/tenax/algorithms)."""

from tenax_amd.algorithms._tensor_utils import max_abs_normalize
from tenax_amd.algorithms.auto_mpo import AutoMPO, build_auto_mpo, spin_half_ops, spin_one_ops
from tenax_amd.algorithms.dmrg import (
    DMRGConfig,
    DMRGResult,
    SweepOps,
    build_mpo_heisenberg,
    build_random_symmetric_mps,
    compute_mps_sector,
    dmrg,
    validate_mps_sector,
)
from tenax_amd.algorithms.ctm import (
    CTMConfig,
    CTMTensorEnv,
    compute_energy_ctm_tensor,
    ctm_tensor,
    ctm_tensor_converge,
    initialize_ctm_tensor_env,
    iPEPSConfig,
    optimize_gs_ad,
)
from tenax_amd.algorithms.hotrg import HOTRGConfig, hotrg, hotrg_batched
from tenax_amd.algorithms.trg import compute_ising_tensor, ising_free_energy_exact

__all__ = ["AutoMPO", "build_auto_mpo", "spin_half_ops", "spin_one_ops", "DMRGConfig", "DMRGResult", "SweepOps",
           "build_mpo_heisenberg", "build_random_symmetric_mps", "compute_mps_sector", "dmrg",
           "validate_mps_sector", "HOTRGConfig", "hotrg", "hotrg_batched", "compute_ising_tensor",
           "ising_free_energy_exact", "max_abs_normalize", "CTMConfig", "CTMTensorEnv", "compute_energy_ctm_tensor",
           "ctm_tensor", "ctm_tensor_converge", "initialize_ctm_tensor_env", "iPEPSConfig", "optimize_gs_ad"]
