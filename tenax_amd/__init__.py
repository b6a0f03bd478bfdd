"""tenax_amd: MI355X-native block-sparse SymmetricTensor contraction and the
DMRG contractions built on it, behind the reference tenax API (contract(),
NetworkBlueprint.launch(), the Tensor protocol, dmrg()).

Compute runs in libtnx.so (hand-written HIP for gfx950, C ABI in include/).
There is no CPU fallback: without the library or a GPU the compute entry
points raise ``tenax_amd._lib.TnxUnavailable``.
"""

from tenax_amd.contraction import (
    blockwise_contract,
    compute_mode,
    contract,
    contract_with_subscripts,
    get_compute,
    per_block_execution,
    set_compute,
)
from tenax_amd.core import (
    EPS,
    LOG_EPS,
    BaseSymmetry,
    BlockKey,
    DenseTensor,
    FermionicU1,
    FermionParity,
    FlowDirection,
    Label,
    PRNGKey,
    SymmetricTensor,
    Tensor,
    TensorIndex,
    U1Symmetry,
    ZnSymmetry,
    fold_in,
    inner,
)

__version__ = "0.1.0"
from tenax_amd.algorithms import (  # noqa: E402
    AutoMPO,
    DMRGConfig,
    DMRGResult,
    HOTRGConfig,
    SweepOps,
    build_auto_mpo,
    build_mpo_heisenberg,
    build_random_symmetric_mps,
    compute_ising_tensor,
    compute_mps_sector,
    dmrg,
    hotrg,
    hotrg_batched,
    ising_free_energy_exact,
    validate_mps_sector,
)
from tenax_amd.linalg import svd, truncated_svd  # noqa: E402
from tenax_amd.network import NetworkBlueprint, TensorNetwork, parse_netfile  # noqa: E402
