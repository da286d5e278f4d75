// gap_hip: command-line power flow on the MI355X backend, the counterpart of
// the reference CLI (src/main/main.cpp:48-105 options, :145-178 flow) for PGM
// JSON input.  The network is read by PgmJsonIO, the Y-bus built by
// HipAdmittanceMatrix, solved by HipNewtonRaphson (the objects the plugin
// factories return for "-b gpu"), and the result written as a PGM-style output
// dataset with node voltages and injections and line flows.
//
//   gap_hip -i input.json [-o output.json] [-b hip|gpu|cuda] [-t tol] [-m max_iter]
//           [--base-power VA] [--precision fp64|mixed] [--dump-rows] [-v]
//
// --dump-rows prints the bus and branch rows the input converts to (JSON) and
// exits without touching the GPU.  Exit status: 0 converged, 1 not converged,
// 2 usage or input error, 3 no device.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "backends.hpp"
#include "gap_hip.h"
#include "pgm_io.hpp"

namespace {

void usage(char const* prog) {
    std::printf(
        "usage: %s -i input.json [-o output.json] [-b hip|gpu|cuda] [-t tolerance] [-m max_iterations]\n"
        "          [--base-power VA] [--precision fp64|mixed] [--dump-rows] [-v]\n"
        "  -i, --input       PGM JSON input dataset (node, line, sym_load, source)\n"
        "  -o, --output      write the PGM-style output dataset here\n"
        "  -b, --backend     hip (gpu and cuda are accepted as names of the same backend)\n"
        "  -t, --tolerance   max |mismatch| in pu (default 1e-8)\n"
        "  -m, --max-iter    Newton iteration limit (default 50)\n"
        "  --base-power      S_base in VA (default 1e6, the PGM harness's)\n"
        "  --precision       fp64 (default) or mixed (FP32 factors, FP64 residual)\n"
        "  --dump-rows       print the converted bus / branch rows and exit\n"
        "  -v, --verbose     print the iteration count and final mismatch\n",
        prog);
}

std::string rows_json(gap::pgm::Rows const& r) {
    auto block = [](std::vector<std::vector<double>> const& rows) {
        std::string s = "[";
        for (size_t k = 0; k < rows.size(); ++k) {
            s += k ? ",[" : "[";
            for (size_t c = 0; c < rows[k].size(); ++c) s += (c ? "," : "") + gap::pgm::fmt(rows[k][c]);
            s += "]";
        }
        return s + "]";
    };
    return "{\"bus\":" + block(r.bus) + ",\"branch\":" + block(r.branch) + "}";
}

}  // namespace

int main(int argc, char** argv) {
    std::string input, output, backend = "hip", precision = "fp64";
    double tolerance = 1e-8, base_power = 1e6;
    int max_it = 50;
    bool dump = false, verbose = false;
    for (int k = 1; k < argc; ++k) {
        std::string const a = argv[k];
        auto next = [&]() -> char const* {
            if (k + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", a.c_str());
                std::exit(2);
            }
            return argv[++k];
        };
        try {
            if (a == "-h" || a == "--help") {
                usage(argv[0]);
                return 0;
            } else if (a == "-i" || a == "--input") {
                input = next();
            } else if (a == "-o" || a == "--output") {
                output = next();
            } else if (a == "-b" || a == "--backend") {
                backend = next();
            } else if (a == "-t" || a == "--tolerance") {
                tolerance = std::stod(next());
            } else if (a == "-m" || a == "--max-iter") {
                max_it = std::stoi(next());
            } else if (a == "--base-power") {
                base_power = std::stod(next());
            } else if (a == "--precision") {
                precision = next();
            } else if (a == "--dump-rows") {
                dump = true;
            } else if (a == "-v" || a == "--verbose") {
                verbose = true;
            } else {
                std::fprintf(stderr, "unknown option %s\n", a.c_str());
                usage(argv[0]);
                return 2;
            }
        } catch (std::exception const&) {
            std::fprintf(stderr, "bad value for %s\n", a.c_str());
            return 2;
        }
    }
    if (input.empty()) {
        usage(argv[0]);
        return 2;
    }
    if (backend != "hip" && backend != "HIP" && backend != "gpu" && backend != "GPU" && backend != "cuda" &&
        backend != "CUDA") {
        std::fprintf(stderr, "Requested backend '%s' is not available (this build has hip)\n", backend.c_str());
        return 2;
    }
    if (precision != "fp64" && precision != "mixed") {
        std::fprintf(stderr, "precision must be fp64 or mixed\n");
        return 2;
    }
    gap::pgm::PgmJsonIO io(base_power);
    gap::NetworkData net;
    try {
        net = io.read_network_data(input);
    } catch (std::exception const& e) {
        std::fprintf(stderr, "input error: %s\n", e.what());
        return 2;
    }
    if (dump) {
        std::cout << rows_json(io.last_rows()) << std::endl;
        return 0;
    }
    if (!gaphip_device_available()) {
        std::fprintf(stderr, "no gfx950 GPU visible\n");
        return 3;
    }
    try {
        gap::hip::HipAdmittanceMatrix ybuilder;
        auto y = ybuilder.build_admittance_matrix(net);
        gap::hip::HipNewtonRaphson solver(precision == "mixed" ? GAPHIP_MIXED : GAPHIP_FP64);
        gap::solver::PowerFlowConfig cfg;
        cfg.tolerance = tolerance;
        cfg.max_iterations = max_it;
        cfg.base_power = base_power;
        cfg.verbose = verbose;
        auto res = solver.solve_power_flow(net, *y, cfg);
        auto flows = solver.branch_flows(net);
        if (!output.empty())
            io.write_output(output, net, res.bus_voltages, flows, res.converged, res.iterations, res.final_mismatch);
        std::printf("{\"converged\":%s,\"iterations\":%d,\"final_mismatch\":%s,\"buses\":%zu,\"branches\":%zu}\n",
                    res.converged ? "true" : "false", res.iterations, gap::pgm::fmt(res.final_mismatch).c_str(),
                    net.buses.size(), net.branches.size());
        return res.converged ? 0 : 1;
    } catch (std::exception const& e) {
        std::fprintf(stderr, "power flow failed: %s\n", e.what());
        return 2;
    }
}
