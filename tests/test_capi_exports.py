"""The native libraries load on a host without a GPU and export every symbol
the headers in include/ declare; no compute is attempted here."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gaphip|create_gpu)_\w+)\s*\(", text)))


def test_core_exports_every_declared_function():
    from gap_amd import capi
    lib = ctypes.CDLL(str(capi.CORE_SO))
    names = declared("gap_hip.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(capi.SIGNATURES), "ctypes table out of sync with the header"


@pytest.mark.parametrize("so", ["libgap_hip_plugins.so", "libgap_admittance_gpu.so", "libgap_lu_gpu.so",
                                "libgap_powerflow_gpu.so"])
def test_plugin_factories_exported(so):
    from gap_amd import capi
    lib = ctypes.CDLL(str(capi.LIB_DIR / so))
    for n in declared("gap_hip_plugins.h"):
        assert hasattr(lib, n), (so, n)


def test_abi_version_and_no_device_paths():
    from gap_amd import capi
    L = capi.lib()
    assert L.gaphip_abi_version() == 4
    if capi.device_available():
        pytest.skip("a device is visible; the no-device error path is exercised elsewhere")
    net = __import__("gap_amd.gridgen", fromlist=["x"]).small_mesh(5, seed=0)
    with pytest.raises(capi.GapHipError) as e:
        capi.Plan([net])
    assert e.value.code == capi.E_DEVICE


def test_pybind_module_imports():
    import gap_amd.gap_solver as gs
    assert hasattr(gs, "solve_simple_power_flow")
    assert gs.is_cuda_available() is False
    if not gs.is_hip_available():
        with pytest.raises(RuntimeError, match="Simple power flow solution failed"):
            gs.solve_simple_power_flow([[0, 1e4, 2, 1.0, 0, 0, 0]], [], backend="hip")
