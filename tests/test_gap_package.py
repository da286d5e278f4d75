"""The reference's Python package surface (bindings/__init__.py,
core_bindings.cpp, io_bindings.cpp, gap_bindings.cpp:55-76) over the HIP
backend: ``gap_amd.gap``."""
import json
from pathlib import Path

import numpy as np
import pytest

import gap_amd.gap as gap
from gap_amd import capi
from gap_amd.network import Network
from gap_amd.pgm import read_pgm_input

ROOT = Path(__file__).resolve().parents[1]
INPUTS = sorted((ROOT / "tests" / "golden" / "pgm_inputs").glob("*.json"))
PGM_LINES = json.loads((ROOT / "tests" / "golden" / "pgm_lines.json").read_text())

REFERENCE_NAMES = ["ApplianceType", "BackendType", "BranchType", "BusType", "NetworkData", "PowerFlowConfig",
                   "PowerFlowResult", "create_solver", "get_available_backends", "is_cuda_available",
                   "BusData", "BranchData", "ApplianceData", "IIOModule", "create_io_module",
                   "load_network_from_json", "save_results_to_json", "solve_simple_power_flow", "GPUNewtonRaphson"]


def test_reference_names_present():
    for name in REFERENCE_NAMES:
        assert hasattr(gap, name), name
    info = gap.get_build_info()
    assert info["python_bindings"] and info["cuda_available"] is False


def test_cpu_backend_is_reported_unavailable():
    with pytest.raises(RuntimeError, match="not available"):
        gap.create_solver(gap.BackendType.CPU)
    assert gap.BackendType.CPU not in gap.get_available_backends()


@pytest.mark.parametrize("path", INPUTS[:3], ids=[p.stem for p in INPUTS[:3]])
def test_io_module_reads_the_harness_network(path):
    io = gap.create_io_module()
    assert io.validate_input_format(str(path))
    net = io.read_network_data(str(path))
    bus, br, _ = read_pgm_input(path, 1e6)
    ref = Network.from_rows(bus, br)
    assert net.num_buses == ref.n_bus and net.num_branches == ref.n_branch
    assert [b.id for b in net.buses] == list(ref.bus_id)
    assert [b.active_power for b in net.buses] == list(ref.p_w)
    assert [b.r1 for b in net.branches] == list(ref.br_r)
    assert [b.b1 for b in net.branches] == list(ref.br_b)
    assert all(b.i_n == 1000.0 for b in net.branches)  # the fixtures' rated current


def test_io_module_rejects_non_pgm(tmp_path):
    f = tmp_path / "x.json"
    f.write_text('{"foo": 1}')
    assert not gap.create_io_module().validate_input_format(str(f))


def test_data_classes_are_writable():
    b = gap.BranchData()
    b.id, b.from_bus, b.to_bus, b.r1, b.x1, b.branch_type = 7, 1, 2, 0.1, 0.2, gap.BranchType.TRAFO
    net = gap.NetworkData()
    net.branches = [b]
    assert net.branches[0].branch_type == gap.BranchType.TRAFO and net.branches[0].r1 == 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("path", INPUTS, ids=[p.stem for p in INPUTS])
def test_solver_objects_end_to_end(device, tmp_path, path):
    case = "pgm_" + path.stem
    net = gap.load_network_from_json(str(path))
    y = gap.create_admittance_matrix(gap.BackendType.GPU_HIP).build_admittance_matrix(net)
    solver = gap.create_solver(gap.BackendType.GPU_HIP)
    assert solver.get_backend_type() == gap.BackendType.GPU_HIP
    cfg = gap.PowerFlowConfig()
    cfg.tolerance, cfg.max_iterations, cfg.base_power = 1e-8, 50, 1e6
    solver.enable_state_capture(True)
    res = solver.solve_power_flow(net, y, cfg)
    assert res.converged
    assert len(solver.get_iteration_states()) == res.iterations
    bus, br, _ = read_pgm_input(path, 1e6)
    v = capi.Plan([Network.from_rows(bus, br)]).solve(1e-8, 50, 1e6)["v"][0]
    assert np.abs(np.array(res.bus_voltages) - v).max() < 1e-12
    flows = solver.branch_flows(net)
    ref = PGM_LINES[case]
    by_id = {b.id: b for b in flows}
    for j, lid in enumerate(ref["id"]):
        # power-grid-model's loading: max(i_from, i_to) / i_n
        assert by_id[lid].loading == pytest.approx(ref["loading"][j], rel=1e-6)
        assert by_id[lid].p_from == pytest.approx(ref["p_from"][j], rel=1e-6, abs=1e-6 * max(map(abs, ref["p_from"])))
    out = tmp_path / "r.json"
    gap.save_results_to_json(str(out), net, res.bus_voltages, res.converged, res.iterations)
    doc = json.loads(out.read_text())
    assert [n["id"] for n in doc["data"]["node"]] == [b.id for b in net.buses]


@pytest.mark.gpu
def test_branch_flows_need_a_solve_of_that_network(device):
    net = gap.load_network_from_json(str(INPUTS[0]))
    solver = gap.HipNewtonRaphson()
    with pytest.raises(RuntimeError, match="no solve_power_flow"):
        solver.branch_flows(net)
