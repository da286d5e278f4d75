"""MI355X (gfx950) Newton-Raphson power-flow backend.

Native pieces live in gap_amd/lib (built by ``python -m gap_amd.build``):
libgap_hip.so (HIP kernels + C ABI, include/gap_hip.h) and
libgap_hip_plugins.so (C++ plugin classes, include/gap_hip_plugins.h); the
pybind11 module ``gap_amd.gap_solver`` mirrors the reference Python surface.
"""
# One HIP runtime per process: torch's libraries load their bundled
# libamdhip64 by a name that does not match the soname ours resolve to, so when
# both are used, torch has to be loaded first and ours then binds to the same
# runtime (two runtimes in one process do not share the device).
try:
    import torch  # noqa: F401
except ImportError:  # torch is optional for the C ABI and the pybind module
    pass

from .network import Network, PQ, PV, SLACK, SOURCE, LOADGEN, SHUNT  # noqa: F401,E402

__all__ = ["Network", "PQ", "PV", "SLACK", "SOURCE", "LOADGEN", "SHUNT"]
__version__ = "0.1.0"
