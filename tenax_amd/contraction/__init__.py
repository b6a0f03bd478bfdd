"""Label-based contraction on the device (mirrors reference src/tenax/contraction)."""

from tenax_amd.contraction.contractor import (
    blockwise_contract,
    compute_mode,
    contract,
    contract_many,
    contract_with_subscripts,
    get_compute,
    per_block_execution,
    plan_order,
    set_compute,
)

__all__ = ["contract", "contract_many", "contract_with_subscripts", "blockwise_contract", "set_compute", "get_compute",
           "compute_mode", "per_block_execution", "plan_order"]
