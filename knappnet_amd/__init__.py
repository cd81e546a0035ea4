"""knappnet_amd: exact Brandes betweenness on an AMD Instinct MI355X.

A drop-in for knappnet's ``net_betweenness`` C ABI (crates/capi), computed by
hand-written HIP kernels for gfx950 in ``libnetbc.so``. ``engine.Graph``
mirrors the reference's ctypes harness; ``generate`` builds the seeded
workloads.
"""
from .engine import EngineError, Graph, device_info  # noqa: F401

__all__ = ["Graph", "EngineError", "device_info"]
