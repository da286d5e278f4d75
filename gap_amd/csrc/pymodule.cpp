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
#include "pgm_io.hpp"

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
    // data model (core_bindings.cpp:15-116)
    py::enum_<ApplianceType>(m, "ApplianceType")
        .value("SOURCE", ApplianceType::SOURCE)
        .value("LOADGEN", ApplianceType::LOADGEN)
        .value("SHUNT", ApplianceType::SHUNT);
    py::enum_<BranchType>(m, "BranchType")
        .value("LINE", BranchType::LINE)
        .value("TRAFO", BranchType::TRAFO)
        .value("GENERIC", BranchType::GENERIC);
    py::class_<BusData>(m, "BusData")
        .def(py::init<>())
        .def_readwrite("id", &BusData::id)
        .def_readwrite("u_rated", &BusData::u_rated)
        .def_readwrite("bus_type", &BusData::bus_type)
        .def_readwrite("energized", &BusData::energized)
        .def_readwrite("u", &BusData::u)
        .def_readwrite("u_pu", &BusData::u_pu)
        .def_readwrite("u_angle", &BusData::u_angle)
        .def_readwrite("active_power", &BusData::active_power)
        .def_readwrite("reactive_power", &BusData::reactive_power)
        .def("__repr__", [](BusData const& b) {
            return "BusData(id=" + std::to_string(b.id) + ", type=" + bus_type_to_string(b.bus_type) + ")";
        });
    py::class_<BranchData>(m, "BranchData")
        .def(py::init<>())
        .def_readwrite("id", &BranchData::id)
        .def_readwrite("from_bus", &BranchData::from_bus)
        .def_readwrite("from_status", &BranchData::from_status)
        .def_readwrite("to_bus", &BranchData::to_bus)
        .def_readwrite("to_status", &BranchData::to_status)
        .def_readwrite("status", &BranchData::status)
        .def_readwrite("branch_type", &BranchData::branch_type)
        .def_readwrite("p_from", &BranchData::p_from)
        .def_readwrite("p_to", &BranchData::p_to)
        .def_readwrite("q_from", &BranchData::q_from)
        .def_readwrite("q_to", &BranchData::q_to)
        .def_readwrite("s_from", &BranchData::s_from)
        .def_readwrite("s_to", &BranchData::s_to)
        .def_readwrite("i_from", &BranchData::i_from)
        .def_readwrite("i_from_angle", &BranchData::i_from_angle)
        .def_readwrite("i_to", &BranchData::i_to)
        .def_readwrite("i_to_angle", &BranchData::i_to_angle)
        .def_readwrite("loading", &BranchData::loading)
        .def_readwrite("r1", &BranchData::r1)
        .def_readwrite("x1", &BranchData::x1)
        .def_readwrite("g1", &BranchData::g1)
        .def_readwrite("b1", &BranchData::b1)
        .def_readwrite("k", &BranchData::k)
        .def_readwrite("theta", &BranchData::theta)
        .def_readwrite("sn", &BranchData::sn)
        .def_readwrite("i_n", &BranchData::i_n)
        .def("__repr__", [](BranchData const& b) {
            return "BranchData(id=" + std::to_string(b.id) + ", " + std::to_string(b.from_bus) + " -> " +
                   std::to_string(b.to_bus) + ")";
        });
    py::class_<ApplianceData>(m, "ApplianceData")
        .def(py::init<>())
        .def_readwrite("id", &ApplianceData::id)
        .def_readwrite("node", &ApplianceData::node)
        .def_readwrite("status", &ApplianceData::status)
        .def_readwrite("type", &ApplianceData::type)
        .def_readwrite("p", &ApplianceData::p)
        .def_readwrite("q", &ApplianceData::q)
        .def_readwrite("i", &ApplianceData::i)
        .def_readwrite("s", &ApplianceData::s)
        .def_readwrite("u_ref", &ApplianceData::u_ref)
        .def_readwrite("u_ref_angle", &ApplianceData::u_ref_angle)
        .def_readwrite("sk", &ApplianceData::sk)
        .def_readwrite("rx_ratio", &ApplianceData::rx_ratio)
        .def_readwrite("p_specified", &ApplianceData::p_specified)
        .def_readwrite("q_specified", &ApplianceData::q_specified)
        .def_readwrite("g1", &ApplianceData::g1)
        .def_readwrite("b1", &ApplianceData::b1);
    py::class_<NetworkData>(m, "NetworkData")
        .def(py::init<>())
        .def_readwrite("buses", &NetworkData::buses)
        .def_readwrite("branches", &NetworkData::branches)
        .def_readwrite("appliances", &NetworkData::appliances)
        .def_readwrite("num_buses", &NetworkData::num_buses)
        .def_readwrite("num_branches", &NetworkData::num_branches)
        .def_readwrite("num_appliances", &NetworkData::num_appliances)
        .def_readwrite("version", &NetworkData::version)
        .def_readwrite("type", &NetworkData::type)
        .def_readwrite("is_batch", &NetworkData::is_batch)
        .def("validate_pgm_compliance", &NetworkData::validate_pgm_compliance,
             "True when every bus has u_rated > 0 and every branch r1 >= 0, x1 != 0, i_n > 0");
    m.def("appliance_type_to_string", &appliance_type_to_string, py::arg("type"));
    m.def("bus_type_to_string", &bus_type_to_string, py::arg("type"));
    py::class_<SparseMatrix>(m, "SparseMatrix")
        .def(py::init<>())
        .def_readwrite("row_ptr", &SparseMatrix::row_ptr)
        .def_readwrite("col_idx", &SparseMatrix::col_idx)
        .def_readwrite("values", &SparseMatrix::values)
        .def_readwrite("num_rows", &SparseMatrix::num_rows)
        .def_readwrite("num_cols", &SparseMatrix::num_cols)
        .def_readwrite("nnz", &SparseMatrix::nnz);
    py::class_<solver::IterationState>(m, "IterationState")
        .def_readonly("iteration", &solver::IterationState::iteration)
        .def_readonly("max_mismatch", &solver::IterationState::max_mismatch);

    // the plugin interfaces, served by the HIP classes (the reference binds no
    // solver interface, so create_solver's result is unusable there)
    py::class_<admittance::IAdmittanceMatrix, std::shared_ptr<admittance::IAdmittanceMatrix>>(m, "IAdmittanceMatrix")
        .def("build_admittance_matrix", &admittance::IAdmittanceMatrix::build_admittance_matrix,
             py::arg("network_data"))
        .def("update_admittance_matrix", &admittance::IAdmittanceMatrix::update_admittance_matrix,
             py::arg("matrix"), py::arg("branch_changes"))
        .def("get_backend_type", &admittance::IAdmittanceMatrix::get_backend_type);
    py::class_<hip::HipAdmittanceMatrix, admittance::IAdmittanceMatrix, std::shared_ptr<hip::HipAdmittanceMatrix>>(
        m, "HipAdmittanceMatrix")
        .def(py::init<>());
    py::class_<solver::ILUSolver, std::shared_ptr<solver::ILUSolver>>(m, "ILUSolver")
        .def("factorize", &solver::ILUSolver::factorize, py::arg("matrix"))
        .def("solve", &solver::ILUSolver::solve, py::arg("rhs"))
        .def("update_factorization", &solver::ILUSolver::update_factorization, py::arg("matrix"))
        .def("get_backend_type", &solver::ILUSolver::get_backend_type)
        .def("is_factorized", &solver::ILUSolver::is_factorized);
    py::class_<hip::HipLUSolver, solver::ILUSolver, std::shared_ptr<hip::HipLUSolver>>(m, "HipLUSolver")
        .def(py::init<>());
    m.def(
        "create_lu_solver",
        [](BackendType backend) -> std::shared_ptr<solver::ILUSolver> {
            if (backend == BackendType::CPU)
                throw std::runtime_error("Requested backend 'CPU' is not available (this build has GPU_HIP)");
            return std::make_shared<hip::HipLUSolver>();
        },
        py::arg("backend"), "LU solver of the backend (the reference factory's create_lu_solver).");
    py::class_<solver::IPowerFlowSolver, std::shared_ptr<solver::IPowerFlowSolver>>(m, "IPowerFlowSolver")
        .def("solve_power_flow", &solver::IPowerFlowSolver::solve_power_flow, py::arg("network_data"),
             py::arg("admittance_matrix"), py::arg("config"))
        .def("calculate_mismatches", &solver::IPowerFlowSolver::calculate_mismatches, py::arg("network_data"),
             py::arg("bus_voltages"), py::arg("admittance_matrix"))
        .def("get_backend_type", &solver::IPowerFlowSolver::get_backend_type)
        .def("enable_state_capture", &solver::IPowerFlowSolver::enable_state_capture, py::arg("enable"))
        .def("get_iteration_states", &solver::IPowerFlowSolver::get_iteration_states)
        .def("clear_iteration_states", &solver::IPowerFlowSolver::clear_iteration_states);
    py::class_<hip::HipNewtonRaphson, solver::IPowerFlowSolver, std::shared_ptr<hip::HipNewtonRaphson>>(
        m, "HipNewtonRaphson")
        .def(py::init([](std::string const& precision) {
                 return std::make_shared<hip::HipNewtonRaphson>(parse_precision(precision));
             }),
             py::arg("precision") = "fp64")
        .def("branch_flows", &hip::HipNewtonRaphson::branch_flows, py::arg("network_data"),
             "Branches of network_data with p/q/s/i at both ends and loading filled at the last solution.");
    m.attr("GPUNewtonRaphson") = m.attr("HipNewtonRaphson");  // the name of the reference's GPU solver class
    m.def(
        "create_solver",
        [](BackendType backend) -> std::shared_ptr<solver::IPowerFlowSolver> {
            if (backend == BackendType::CPU)
                throw std::runtime_error("Requested backend 'CPU' is not available (this build has GPU_HIP)");
            if (!gaphip_device_available()) throw std::runtime_error("No gfx950 GPU visible");
            return std::make_shared<hip::HipNewtonRaphson>(GAPHIP_FP64);
        },
        py::arg("backend"), "Power flow solver of the backend (GPU_HIP; GPU_CUDA names the same device backend).");
    m.def(
        "create_admittance_matrix",
        [](BackendType backend) -> std::shared_ptr<admittance::IAdmittanceMatrix> {
            if (backend == BackendType::CPU)
                throw std::runtime_error("Requested backend 'CPU' is not available (this build has GPU_HIP)");
            return std::make_shared<hip::HipAdmittanceMatrix>();
        },
        py::arg("backend"));

    // I/O (io_bindings.cpp:15-60): PGM JSON
    py::class_<io::IIOModule, std::shared_ptr<io::IIOModule>>(m, "IIOModule")
        .def("read_network_data", &io::IIOModule::read_network_data, py::arg("filename"))
        .def("write_results", &io::IIOModule::write_results, py::arg("filename"), py::arg("bus_voltages"),
             py::arg("converged"), py::arg("iterations"))
        .def("validate_input_format", &io::IIOModule::validate_input_format, py::arg("filename"));
    py::class_<pgm::PgmJsonIO, io::IIOModule, std::shared_ptr<pgm::PgmJsonIO>>(m, "PgmJsonIO")
        .def(py::init<double>(), py::arg("base_power") = 1e6)
        .def("write_output", &pgm::PgmJsonIO::write_output, py::arg("filename"), py::arg("network_data"),
             py::arg("bus_voltages"), py::arg("branches"), py::arg("converged"), py::arg("iterations"),
             py::arg("final_mismatch"));
    m.def("create_io_module", []() -> std::shared_ptr<io::IIOModule> { return std::make_shared<pgm::PgmJsonIO>(); });
    m.def(
        "load_network_from_json", [](std::string const& f) { return pgm::PgmJsonIO().read_network_data(f); },
        py::arg("filename"));
    m.def(
        "save_results_to_json",
        [](std::string const& f, NetworkData const& net, ComplexVector const& v, bool converged, int iterations) {
            // write_results needs the network's ids and ratings: pass the network read
            pgm::PgmJsonIO io;
            io.write_output(f, net, v, net.branches, converged, iterations, 0.0);
        },
        py::arg("filename"), py::arg("network_data"), py::arg("bus_voltages"), py::arg("converged"),
        py::arg("iterations"));

    m.def("get_available_backends", []() {
        std::vector<BackendType> b;
        if (gaphip_device_available()) b.push_back(BackendType::GPU_HIP);
        return b;
    });
    m.def("is_cuda_available", []() { return false; });
    m.def("is_hip_available", []() { return gaphip_device_available() != 0; });
}
