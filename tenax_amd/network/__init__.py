"""Network containers (mirrors reference src/tenax/network)."""

from tenax_amd.network.network import TensorNetwork

__all__ = ["TensorNetwork"]
