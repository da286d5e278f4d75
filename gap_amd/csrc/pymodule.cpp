// pybind11 module ``gap_solver``: the Python surface of the reference bindings
// (bindings/solver_bindings.cpp:21-217, bindings/gap_bindings.cpp:30-76),
// served by the MI355X plugin classes.  The reference CPU backend is not part
// of this build, so backend="cpu" reports itself unavailable the way the
// reference does for a missing backend (solver_bindings.cpp:127-131).
#include <pybind11/complex.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "backends.hpp"
#include "gap_hip.h"

namespace py = pybind11;
using namespace gap;

namespace {

bool is_hip_backend(std::string const& b) {
    return b == "hip" || b == "HIP" || b == "gpu" || b == "GPU" || b == "cuda" || b == "CUDA";
}

int parse_precision(std::string const& p) {
    if (p == "fp64") return GAPHIP_FP64;
    if (p == "mixed") return GAPHIP_MIXED;
    throw std::runtime_error("precision must be 'fp64' or 'mixed'");
}

std::vector<std::vector<double>> solve_simple(std::vector<std::vector<double>> const& bus_data,
                                              std::vector<std::vector<double>> const& branch_data, double tolerance,
                                              int max_iterations, bool verbose, double base_power,
                                              std::string const& backend, std::string const& precision) {
    try {
        NetworkData net;
        net.buses.reserve(bus_data.size());
        for (auto const& row : bus_data) {
            if (row.size() < 7)
                throw std::runtime_error(
                    "Bus data must have at least 7 elements: [id, u_rated, bus_type, u_pu, u_angle, p, q]");
            BusData b;
            b.id = static_cast<int>(row[0]);
            b.u_rated = row[1];
            b.bus_type = static_cast<BusType>(static_cast<int>(row[2]));
            b.energized = true;
            b.u_pu = row[3];
            b.u_angle = row[4];
            b.active_power = row[5];
            b.reactive_power = row[6];
            net.buses.push_back(b);
        }
        net.branches.reserve(branch_data.size());
        for (auto const& row : branch_data) {
            if (row.size() < 5)
                throw std::runtime_error("Branch data must have at least 5 elements: [id, from_bus, to_bus, r, x]");
            BranchData br;
            br.id = static_cast<int>(row[0]);
            br.from_bus = static_cast<int>(row[1]);
            br.to_bus = static_cast<int>(row[2]);
            br.status = 1;
            br.r1 = row[3];
            br.x1 = row[4];
            br.b1 = row.size() > 5 ? row[5] : 0.0;
            net.branches.push_back(br);
        }
        net.num_buses = static_cast<int>(net.buses.size());
        net.num_branches = static_cast<int>(net.branches.size());
        net.num_appliances = 0;

        solver::PowerFlowConfig cfg;
        cfg.tolerance = tolerance;
        cfg.max_iterations = max_iterations;
        cfg.verbose = verbose;
        cfg.base_power = base_power;

        if (!is_hip_backend(backend) || !gaphip_device_available())
            throw std::runtime_error("Requested backend '" + backend + "' is not available");
        hip::HipAdmittanceMatrix ybuilder;
        auto y = ybuilder.build_admittance_matrix(net);
        hip::HipNewtonRaphson nr(parse_precision(precision));
        auto res = nr.solve_power_flow(net, *y, cfg);
        std::vector<std::vector<double>> out;
        out.reserve(res.bus_voltages.size() + 1);
        for (auto const& v : res.bus_voltages) out.push_back({v.real(), v.imag(), std::abs(v), std::arg(v)});
        out.push_back({res.converged ? 1.0 : 0.0, static_cast<double>(res.iterations), res.final_mismatch, 0.0});
        return out;
    } catch (std::exception const& e) {
        throw std::runtime_error(std::string("Simple power flow solution failed: ") + e.what());
    }
}

}  // namespace

PYBIND11_MODULE(gap_solver, m) {
    m.doc() = "GAP power flow, MI355X (gfx950) HIP backend";
    m.attr("__version__") = "0.1.0";
    m.attr("__cuda_available__") = py::bool_(false);

    py::enum_<BusType>(m, "BusType")
        .value("PQ", BusType::PQ)
        .value("PV", BusType::PV)
        .value("SLACK", BusType::SLACK);
    py::enum_<BackendType>(m, "BackendType")
        .value("CPU", BackendType::CPU)
        .value("GPU_CUDA", BackendType::GPU_CUDA)
        .value("GPU_HIP", BackendType::GPU_HIP);

    py::class_<solver::PowerFlowConfig>(m, "PowerFlowConfig")
        .def(py::init<>())
        .def_readwrite("tolerance", &solver::PowerFlowConfig::tolerance)
        .def_readwrite("max_iterations", &solver::PowerFlowConfig::max_iterations)
        .def_readwrite("use_flat_start", &solver::PowerFlowConfig::use_flat_start)
        .def_readwrite("acceleration_factor", &solver::PowerFlowConfig::acceleration_factor)
        .def_readwrite("verbose", &solver::PowerFlowConfig::verbose)
        .def_readwrite("base_power", &solver::PowerFlowConfig::base_power);
    py::class_<solver::PowerFlowResult>(m, "PowerFlowResult")
        .def(py::init<>())
        .def_readwrite("bus_voltages", &solver::PowerFlowResult::bus_voltages)
        .def_readwrite("converged", &solver::PowerFlowResult::converged)
        .def_readwrite("iterations", &solver::PowerFlowResult::iterations)
        .def_readwrite("final_mismatch", &solver::PowerFlowResult::final_mismatch)
        .def_readwrite("bus_injections", &solver::PowerFlowResult::bus_injections);

    m.def("solve_simple_power_flow", &solve_simple, py::arg("bus_data"), py::arg("branch_data"),
          py::arg("tolerance") = 1e-6, py::arg("max_iterations") = 50, py::arg("verbose") = false,
          py::arg("base_power") = 100e6, py::arg("backend") = "hip", py::arg("precision") = "fp64",
          "Solve power flow from basic Python rows.\n\n"
          "bus row: [id, u_rated (V), type (0 PQ, 1 PV, 2 SLACK), u_pu, u_angle, p (W), q (VAr)]\n"
          "branch row: [id, from_id, to_id, r (ohm), x (ohm), (b (S))]\n"
          "Returns one row [re, im, |V|, angle] per bus and a last row [converged, iterations, "
          "final_mismatch, 0].  backend: 'hip' (aliases 'gpu', 'cuda'); precision: 'fp64' or 'mixed'.");
    m.def("get_available_backends", []() {
        std::vector<BackendType> b;
        if (gaphip_device_available()) b.push_back(BackendType::GPU_HIP);
        return b;
    });
    m.def("is_cuda_available", []() { return false; });
    m.def("is_hip_available", []() { return gaphip_device_available() != 0; });
}
