// ref_harness.cpp — drives the reference's OWN CPU power-flow sources, compiled
// where they lie under the reference tree (see oracle/Makefile), so that tests
// can pin the C restatement (pforacle.c) and the HIP path against them.
//
// TEST INFRASTRUCTURE ONLY: the output library lands in oracle/_ref/ and is
// loaded by tests/ and bench.py's CPU-baseline leg, never by the product.
//
// The reference translation units are included as-is; this file only builds a
// NetworkData the way bindings/solver_bindings.cpp:64-111 does, switches the
// logger off (include/gap/logging/log_config.h:48-51), and exposes flat extern "C"
// entry points.  The JSON reader and backend factory are not used: json_io.cpp
// does not compile against libstdc++ 11.

#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "admittance/cpu/cpu_admittance_matrix.cpp"
#include "gap/logging/log_config.h"
#include "solver/lu/cpu/cpu_lu_solver.cpp"
#include "solver/powerflow/cpu/cpu_newton_raphson.cpp"

#include "pf_net.h"

using namespace gap;

namespace {

void quiet() { gap::logging::LogConfig::forBenchmarking(); }

NetworkData to_network(pf_net const* in) {
    NetworkData net;
    net.buses.reserve(in->n_bus);
    for (int i = 0; i < in->n_bus; ++i) {
        BusData b{};
        b.id = in->bus_id[i];
        b.u_rated = in->u_rated[i];
        b.bus_type = static_cast<BusType>(in->bus_type[i]);
        b.energized = true;
        b.u_pu = in->u_pu[i];
        b.u_angle = 0.0;
        b.active_power = in->p_w[i];
        b.reactive_power = in->q_var[i];
        net.buses.push_back(b);
    }
    net.branches.reserve(in->n_branch);
    for (int i = 0; i < in->n_branch; ++i) {
        BranchData br{};
        br.id = in->br_id[i];
        br.from_bus = in->br_from[i];
        br.to_bus = in->br_to[i];
        br.from_status = 1;
        br.to_status = 1;
        br.status = static_cast<int8_t>(in->br_status[i] != 0);
        br.r1 = in->br_r[i];
        br.x1 = in->br_x[i];
        br.g1 = in->br_g ? in->br_g[i] : 0.0;
        br.b1 = in->br_b ? in->br_b[i] : 0.0;
        net.branches.push_back(br);
    }
    for (int i = 0; i < in->n_app; ++i) {
        ApplianceData a{};
        a.id = in->app_id[i];
        a.node = in->app_node[i];
        a.status = static_cast<int8_t>(in->app_status[i]);
        a.type = static_cast<ApplianceType>(in->app_type[i]);
        a.p_specified = in->app_p[i];
        a.q_specified = in->app_q[i];
        a.g1 = in->app_g[i];
        a.b1 = in->app_b[i];
        net.appliances.push_back(a);
    }
    net.num_buses = static_cast<int>(net.buses.size());
    net.num_branches = static_cast<int>(net.branches.size());
    net.num_appliances = static_cast<int>(net.appliances.size());
    return net;
}

// Wraps the reference CPU LU so the Jacobian handed to factorize() on a chosen
// Newton iteration can be copied out (pattern and values).
class RecordingLU : public solver::ILUSolver {
  public:
    explicit RecordingLU(int capture_call) : capture_call_(capture_call) {}
    bool factorize(SparseMatrix const& m) override {
        if (calls_ == capture_call_) captured_ = m;
        ++calls_;
        return inner_.factorize(m);
    }
    ComplexVector solve(ComplexVector const& rhs) override { return inner_.solve(rhs); }
    bool update_factorization(SparseMatrix const& m) override { return inner_.update_factorization(m); }
    BackendType get_backend_type() const noexcept override { return BackendType::CPU; }
    bool is_factorized() const noexcept override { return inner_.is_factorized(); }
    SparseMatrix const& captured() const { return captured_; }
    int calls() const { return calls_; }

  private:
    solver::CPULUSolver inner_;
    int capture_call_;
    int calls_ = 0;
    SparseMatrix captured_{};
};

}  // namespace

extern "C" {

// Y-bus in siemens (cpu_admittance_matrix.cpp:44-202).  Returns nnz, or -1 if
// the capacity is too small.  values: interleaved (re, im).
int gapref_build_ybus(pf_net const* in, int* row_ptr, int* col_idx, double* values, int cap) {
    quiet();
    NetworkData net = to_network(in);
    admittance::CPUAdmittanceMatrix builder;
    auto y = builder.build_admittance_matrix(net);
    if (y->nnz > cap) return -1;
    std::memcpy(row_ptr, y->row_ptr.data(), sizeof(int) * (y->num_rows + 1));
    std::memcpy(col_idx, y->col_idx.data(), sizeof(int) * y->nnz);
    std::memcpy(values, y->values.data(), sizeof(double) * 2 * y->nnz);
    return y->nnz;
}

// Full reference solve (cpu_newton_raphson.cpp:33-218) on the reference Y-bus.
//   v_out       : 2*n_bus doubles (re, im)
//   traj        : max|F| of each iteration (state capture, powerflow_interface.h:28-35), cap traj_cap
//   j_call      : which factorize() call's Jacobian to copy out (0 = first), or -1 for none
//   j_row_ptr   : n_var+1; j_col_idx / j_val: up to j_cap entries (real parts)
// Returns 0 on success, 1 if the Jacobian capacity was too small, -1 on exception.
int gapref_solve(pf_net const* in, double tol, int max_it, double base_power, double* v_out,
                 int* iterations, int* converged, double* final_mismatch, double* traj,
                 int traj_cap, int* n_traj, int j_call, int* j_n, int* j_row_ptr, int* j_col_idx,
                 double* j_val, int j_cap, int* j_nnz) {
    try {
        quiet();
        NetworkData net = to_network(in);
        admittance::CPUAdmittanceMatrix builder;
        auto y = builder.build_admittance_matrix(net);
        solver::CPUNewtonRaphson nr_impl;
        // state capture is public on the interface, private on the class
        solver::IPowerFlowSolver& nr = nr_impl;
        auto lu = std::make_shared<RecordingLU>(j_call);
        nr.set_lu_solver(lu);
        nr.enable_state_capture(traj != nullptr);
        solver::PowerFlowConfig cfg;
        cfg.tolerance = tol;
        cfg.max_iterations = max_it;
        cfg.base_power = base_power;
        auto res = nr.solve_power_flow(net, *y, cfg);
        for (int i = 0; i < in->n_bus; ++i) {
            v_out[2 * i] = res.bus_voltages[i].real();
            v_out[2 * i + 1] = res.bus_voltages[i].imag();
        }
        *iterations = res.iterations;
        *converged = res.converged ? 1 : 0;
        *final_mismatch = res.final_mismatch;
        if (traj) {
            auto const& st = nr.get_iteration_states();
            int k = 0;
            for (auto const& s : st) {
                if (k < traj_cap) traj[k] = s.max_mismatch;
                ++k;
            }
            *n_traj = k;
        }
        int rc = 0;
        if (j_call >= 0 && j_n) {
            auto const& J = lu->captured();
            *j_n = J.num_rows;
            *j_nnz = J.nnz;
            if (J.nnz > j_cap) {
                rc = 1;
            } else if (J.num_rows > 0) {
                std::memcpy(j_row_ptr, J.row_ptr.data(), sizeof(int) * (J.num_rows + 1));
                std::memcpy(j_col_idx, J.col_idx.data(), sizeof(int) * J.nnz);
                for (int k = 0; k < J.nnz; ++k) j_val[k] = J.values[k].real();
            }
        }
        return rc;
    } catch (...) {
        return -1;
    }
}

// Public mismatch entry (cpu_newton_raphson.cpp:225-240): watts, S_base fixed at 100 MVA.
int gapref_mismatches(pf_net const* in, double const* v_ri, double const* y_ri, int const* y_row_ptr,
                      int const* y_col_idx, int y_nnz, double* out, int cap) {
    quiet();
    NetworkData net = to_network(in);
    ComplexVector v(in->n_bus);
    for (int i = 0; i < in->n_bus; ++i) v[i] = Complex(v_ri[2 * i], v_ri[2 * i + 1]);
    SparseMatrix y;
    y.num_rows = y.num_cols = in->n_bus;
    y.nnz = y_nnz;
    if (y_nnz > 0) {
        y.row_ptr.assign(y_row_ptr, y_row_ptr + in->n_bus + 1);
        y.col_idx.assign(y_col_idx, y_col_idx + y_nnz);
        y.values.resize(y_nnz);
        for (int k = 0; k < y_nnz; ++k) y.values[k] = Complex(y_ri[2 * k], y_ri[2 * k + 1]);
    }
    solver::CPUNewtonRaphson nr;
    auto f = nr.calculate_mismatches(net, v, y);
    if (static_cast<int>(f.size()) > cap) return -1;
    for (size_t k = 0; k < f.size(); ++k) out[k] = f[k];
    return static_cast<int>(f.size());
}

// Reference CPU LU (cpu_lu_solver.cpp:91-185): factorize + solve one complex system.
// Returns 0 ok, 1 factorize failed, 2 solve threw.
int gapref_lu_solve(int n, int const* row_ptr, int const* col_idx, double const* val_ri, int nnz,
                    double const* rhs_ri, double* x_ri) {
    quiet();
    SparseMatrix a;
    a.num_rows = a.num_cols = n;
    a.nnz = nnz;
    a.row_ptr.assign(row_ptr, row_ptr + n + 1);
    a.col_idx.assign(col_idx, col_idx + nnz);
    a.values.resize(nnz);
    for (int k = 0; k < nnz; ++k) a.values[k] = Complex(val_ri[2 * k], val_ri[2 * k + 1]);
    solver::CPULUSolver lu;
    if (!lu.factorize(a)) return 1;
    ComplexVector b(n);
    for (int i = 0; i < n; ++i) b[i] = Complex(rhs_ri[2 * i], rhs_ri[2 * i + 1]);
    try {
        auto x = lu.solve(b);
        for (int i = 0; i < n; ++i) {
            x_ri[2 * i] = x[i].real();
            x_ri[2 * i + 1] = x[i].imag();
        }
    } catch (...) {
        return 2;
    }
    return 0;
}

}  // extern "C"
