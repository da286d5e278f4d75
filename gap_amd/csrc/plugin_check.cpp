// gap_hip_plugin_check: loads the three backend libraries the way the
// reference's BackendFactory does (src/core/backend_factory.cpp:32-44, 56-67,
// 80-91) -- dlopen by bare name with RTLD_LAZY, so they resolve through
// LD_LIBRARY_PATH or the rpath, dlsym of create_gpu_admittance_matrix,
// create_gpu_lu_solver and create_gpu_powerflow_solver cast to factory
// functions returning interface pointers, each result owned by a unique_ptr --
// and runs a power flow through the interfaces only: Y-bus from the admittance
// plugin, the LU plugin handed to the solver with set_lu_solver, then
// solve_power_flow.  Links against nothing of this project: a check that a
// caller compiled against the reference's headers can use the plugins.
//
// Usage: gap_hip_plugin_check NETWORK_FILE
// NETWORK_FILE (text):  n_bus n_branch base_power tolerance max_iterations
//                       n_bus lines    id u_rated type u_pu p_w q_var
//                       n_branch lines id from to status r1 x1 g1 b1
// Prints the backend types, "converged iterations final_mismatch", one line
// "re im" per bus, and the residual of an LU solve of the Y-bus through the LU
// plugin's vtable.
#include <dlfcn.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>

#include "gap/admittance/admittance_interface.h"
#include "gap/core/types.h"
#include "gap/solver/lu_solver_interface.h"
#include "gap/solver/powerflow_interface.h"

using namespace gap;

template <typename T>
std::unique_ptr<T> load(char const* lib, char const* sym) {
    void* handle = dlopen(lib, RTLD_LAZY);
    if (!handle) throw std::runtime_error(std::string("dlopen ") + lib + ": " + dlerror());
    typedef T* (*CreateFunc)();
    CreateFunc create = (CreateFunc)dlsym(handle, sym);
    if (!create) throw std::runtime_error(std::string("dlsym ") + sym + " not found in " + lib);
    return std::unique_ptr<T>(create());
}

int main(int argc, char** argv) {
    try {
        if (argc != 2) throw std::runtime_error("usage: gap_hip_plugin_check NETWORK_FILE");
        std::ifstream in(argv[1]);
        int nb = 0, nr = 0, max_it = 0;
        double base = 0.0, tol = 0.0;
        if (!(in >> nb >> nr >> base >> tol >> max_it)) throw std::runtime_error("bad header");
        NetworkData net;
        for (int i = 0; i < nb; ++i) {
            BusData b;
            int type = 0;
            in >> b.id >> b.u_rated >> type >> b.u_pu >> b.active_power >> b.reactive_power;
            b.bus_type = static_cast<BusType>(type);
            net.buses.push_back(b);
        }
        for (int k = 0; k < nr; ++k) {
            BranchData br;
            int status = 0;
            in >> br.id >> br.from_bus >> br.to_bus >> status >> br.r1 >> br.x1 >> br.g1 >> br.b1;
            br.status = br.from_status = br.to_status = static_cast<int8_t>(status);
            net.branches.push_back(br);
        }
        if (!in) throw std::runtime_error("truncated network file");
        net.num_buses = nb;
        net.num_branches = nr;
        net.num_appliances = 0;

        auto adm = load<admittance::IAdmittanceMatrix>("libgap_admittance_gpu.so", "create_gpu_admittance_matrix");
        auto lu = load<solver::ILUSolver>("libgap_lu_gpu.so", "create_gpu_lu_solver");
        auto pf = load<solver::IPowerFlowSolver>("libgap_powerflow_gpu.so", "create_gpu_powerflow_solver");
        std::printf("backends %d %d %d\n", static_cast<int>(adm->get_backend_type()),
                    static_cast<int>(lu->get_backend_type()), static_cast<int>(pf->get_backend_type()));
        std::shared_ptr<solver::ILUSolver> shared_lu(lu.release());
        pf->set_lu_solver(shared_lu);

        std::unique_ptr<SparseMatrix> y = adm->build_admittance_matrix(net);
        solver::PowerFlowConfig cfg;
        cfg.tolerance = tol;
        cfg.max_iterations = max_it;
        cfg.base_power = base;
        solver::PowerFlowResult res = pf->solve_power_flow(net, *y, cfg);
        std::printf("%d %d %.17g\n", res.converged ? 1 : 0, res.iterations, res.final_mismatch);
        for (Complex const& v : res.bus_voltages) std::printf("%.17g %.17g\n", v.real(), v.imag());

        // the LU plugin through its interface: Y x = Y 1 must give x = 1
        if (!shared_lu->factorize(*y)) throw std::runtime_error("LU plugin could not factorize the Y-bus");
        ComplexVector rhs(nb, Complex(0.0, 0.0));
        for (int i = 0; i < nb; ++i)
            for (int q = y->row_ptr[i]; q < y->row_ptr[i + 1]; ++q) rhs[i] += y->values[q];
        ComplexVector x = shared_lu->solve(rhs);
        double err = 0.0;
        for (Complex const& xi : x) err = std::max(err, std::abs(xi - Complex(1.0, 0.0)));
        std::printf("lu %d %.3e\n", shared_lu->is_factorized() ? 1 : 0, err);
        return 0;
    } catch (std::exception const& e) {
        std::fprintf(stderr, "gap_hip_plugin_check: %s\n", e.what());
        return 1;
    }
}
