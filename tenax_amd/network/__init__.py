"""Network containers and .net templates (mirrors reference src/tenax/network)."""

from tenax_amd.network.netfile import NetworkBlueprint, parse_netfile
from tenax_amd.network.network import TensorNetwork

__all__ = ["TensorNetwork", "NetworkBlueprint", "parse_netfile"]
