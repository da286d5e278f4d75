"""The reference's Python package surface (bindings/__init__.py), served by
the MI355X HIP backend: ``import gap_amd.gap as gap``.

Everything comes from the compiled module gap_amd.gap_solver.  The reference
package imports a CPU solver class its module never binds
(bindings/__init__.py:36); this build has no CPU backend, so only the device
solver is exported (HipNewtonRaphson, also under the reference's GPU class name
GPUNewtonRaphson), and create_solver(BackendType.CPU) raises.
"""
from __future__ import annotations

from .. import gap_solver as _m
from ..gap_solver import (  # noqa: F401
    ApplianceData,
    ApplianceType,
    BackendType,
    BranchData,
    BranchType,
    BusData,
    BusType,
    GPUNewtonRaphson,
    HipAdmittanceMatrix,
    HipNewtonRaphson,
    IAdmittanceMatrix,
    IIOModule,
    IPowerFlowSolver,
    IterationState,
    NetworkData,
    PgmJsonIO,
    PowerFlowConfig,
    PowerFlowResult,
    SparseMatrix,
    create_admittance_matrix,
    create_io_module,
    create_solver,
    get_available_backends,
    is_cuda_available,
    is_hip_available,
    load_network_from_json,
    save_results_to_json,
    solve_simple_power_flow,
)

__version__ = _m.__version__


def get_build_info() -> dict:
    """How this build was made (reference bindings/__init__.py:55-64)."""
    return {"version": __version__, "cuda_available": False, "hip_available": bool(is_hip_available()),
            "backends": get_available_backends(), "python_bindings": True}


def print_build_info() -> None:
    info = get_build_info()
    print("GAP Power Flow Calculator - Python Bindings (MI355X HIP backend)")
    print(f"Version: {info['version']}")
    print(f"HIP Support: {'Available' if info['hip_available'] else 'Not Available'}")
    print(f"Available Backends: {', '.join(str(b) for b in info['backends'])}")


__all__ = [name for name in dir() if not name.startswith("_")]
