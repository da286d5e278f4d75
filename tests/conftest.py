import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    if not pyoracle.ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return pyoracle.Oracle()


def golden_cases(prefix=""):
    """Single-network golden solves (not the LU vectors or the oracle fixtures)."""
    return sorted(p.stem for p in GOLDEN.glob(f"{prefix}*.npz") if not p.stem.startswith(("lu_", "batch")) and not p.stem.endswith("_oracle"))


def finite_golden_cases():
    """Golden cases whose reference solution is finite (not the NaN-load ones)."""
    return [n for n in golden_cases() if not n.startswith("fail_nan")]


def max_dv(a, b):
    """max |a - b| over complex arrays; an entry NaN in both counts as equal,
    NaN in only one of them as infinitely far."""
    a, b = np.asarray(a), np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if np.any(na != nb):
        return np.inf
    d = np.abs(np.where(na, 0, a - b))
    return float(d.max()) if d.size else 0.0


def same_nan(a, b):
    """True when a and b are NaN at the same positions."""
    return bool(np.array_equal(np.isnan(np.asarray(a)), np.isnan(np.asarray(b))))


def load_golden(name):
    from gap_amd.network import Network
    d = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    net = Network(**{k[4:]: d[k] for k in d.files if k.startswith("net_")})
    return net, d


@pytest.fixture(scope="session")
def device():
    from gap_amd import capi
    if not capi.device_available():
        raise RuntimeError("GPU test run without a visible gfx950 device")
    return True
