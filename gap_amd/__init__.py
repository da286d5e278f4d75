"""MI355X (gfx950) Newton-Raphson power-flow backend.

Native pieces live in gap_amd/lib (built by ``python -m gap_amd.build``):
libgap_hip.so (HIP kernels + C ABI, include/gap_hip.h) and
libgap_hip_plugins.so (C++ plugin classes, include/gap_hip_plugins.h); the
pybind11 module ``gap_amd.gap_solver`` mirrors the reference Python surface.
"""
from .network import Network, PQ, PV, SLACK, SOURCE, LOADGEN, SHUNT  # noqa: F401

__all__ = ["Network", "PQ", "PV", "SLACK", "SOURCE", "LOADGEN", "SHUNT"]
__version__ = "0.1.0"
