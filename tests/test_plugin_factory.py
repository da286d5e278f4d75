"""The plugins loaded the way the reference's BackendFactory loads them
(src/core/backend_factory.cpp:32-44, 56-67, 80-91): gap_amd/lib/gap_hip_plugin_check,
built against include/gap and linked against nothing of this project, dlopens
libgap_{admittance,lu,powerflow}_gpu.so by bare name (found through
LD_LIBRARY_PATH), takes create_gpu_* as factory functions returning interface
pointers, hands the LU plugin to the solver with set_lu_solver and solves a
golden case through the interfaces; its result must be the reference's."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import load_golden

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "gap_amd" / "lib"
EXE = LIB / "gap_hip_plugin_check"


def write_network(path, net, base, tol, max_it):
    r = lambda x: repr(float(x))  # shortest round-trip decimal of the double
    with open(path, "w") as f:
        f.write(f"{net.n_bus} {net.n_branch} {r(base)} {r(tol)} {max_it}\n")
        for i in range(net.n_bus):
            f.write(f"{int(net.bus_id[i])} {r(net.u_rated[i])} {int(net.bus_type[i])} {r(net.u_pu[i])} "
                    f"{r(net.p_w[i])} {r(net.q_var[i])}\n")
        for k in range(net.n_branch):
            f.write(f"{int(net.br_id[k])} {int(net.br_from[k])} {int(net.br_to[k])} {int(net.br_status[k])} "
                    f"{r(net.br_r[k])} {r(net.br_x[k])} {r(net.br_g[k])} {r(net.br_b[k])}\n")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["mesh200_seed0", "pgm_radial_3feeder_8nodepf", "one_based_ids"])
def test_reference_factory_load_path(device, tmp_path, name):
    assert EXE.exists(), "gap_hip_plugin_check is not built (python -m gap_amd.build)"
    net, g = load_golden(name)
    assert net.n_app == 0
    f = tmp_path / "net.txt"
    write_network(f, net, float(g["base_power"]), float(g["tolerance"]), int(g["max_iterations"]))
    env = dict(os.environ, LD_LIBRARY_PATH=f"{LIB}:{os.environ.get('LD_LIBRARY_PATH', '')}")
    out = subprocess.run([str(EXE), str(f)], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == "backends 2 2 2"  # BackendType::GPU_HIP from every plugin
    conv, its, _ = lines[1].split()
    v = np.array([complex(float(a), float(b)) for a, b in (ln.split() for ln in lines[2:2 + net.n_bus])])
    assert int(conv) == int(bool(g["converged"])) and int(its) == int(g["iterations"])
    assert np.abs(v - g["v"]).max() <= 1e-8
    tag, ok, err = lines[2 + net.n_bus].split()
    assert tag == "lu" and ok == "1" and float(err) < 1e-9
