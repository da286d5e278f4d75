// C++20 plugin classes of the MI355X backend, written against the C ABI only
// (include/gap_hip.h); built with the host compiler, no HIP headers.
//
//   HipAdmittanceMatrix : gap::admittance::IAdmittanceMatrix
//   HipLUSolver         : gap::solver::ILUSolver
//   HipNewtonRaphson    : gap::solver::IPowerFlowSolver
//
// and the extern "C" factories create_gpu_{admittance_matrix,lu_solver,
// powerflow_solver} that the reference backend factory resolves by name
// (src/core/backend_factory.cpp:34-41, 58-63, 82-87).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "backends.hpp"
#include "gap_hip.h"

namespace gap::hip {

namespace {

[[noreturn]] void raise(int rc, char const* where) {
    throw std::runtime_error(std::string(where) + ": " + gaphip_last_error() + " (code " + std::to_string(rc) + ")");
}

}  // namespace

NetArrays::NetArrays(NetworkData const& net) {
    for (auto const& b : net.buses) {
        bus_id.push_back(b.id);
        u_rated.push_back(b.u_rated);
        bus_type.push_back(static_cast<int>(b.bus_type));
        u_pu.push_back(b.u_pu);
        p_w.push_back(b.active_power);
        q_var.push_back(b.reactive_power);
    }
    for (auto const& br : net.branches) {
        br_id.push_back(br.id);
        br_from.push_back(br.from_bus);
        br_to.push_back(br.to_bus);
        br_status.push_back(br.status ? 1 : 0);
        br_r.push_back(br.r1);
        br_x.push_back(br.x1);
        br_g.push_back(br.g1);
        br_b.push_back(br.b1);
    }
    for (auto const& a : net.appliances) {
        app_id.push_back(a.id);
        app_node.push_back(a.node);
        app_status.push_back(a.status);
        app_type.push_back(static_cast<int>(a.type));
        app_p.push_back(a.p_specified);
        app_q.push_back(a.q_specified);
        app_g.push_back(a.g1);
        app_b.push_back(a.b1);
    }
}

gaphip_net NetArrays::view() const {
    gaphip_net v{};
    v.n_bus = static_cast<int>(bus_id.size());
    v.bus_id = bus_id.data();
    v.u_rated = u_rated.data();
    v.bus_type = bus_type.data();
    v.u_pu = u_pu.data();
    v.p_w = p_w.data();
    v.q_var = q_var.data();
    v.n_branch = static_cast<int>(br_id.size());
    v.br_id = br_id.data();
    v.br_from = br_from.data();
    v.br_to = br_to.data();
    v.br_status = br_status.data();
    v.br_r = br_r.data();
    v.br_x = br_x.data();
    v.br_g = br_g.data();
    v.br_b = br_b.data();
    v.n_app = static_cast<int>(app_id.size());
    v.app_id = app_id.data();
    v.app_node = app_node.data();
    v.app_status = app_status.data();
    v.app_type = app_type.data();
    v.app_p = app_p.data();
    v.app_q = app_q.data();
    v.app_g = app_g.data();
    v.app_b = app_b.data();
    return v;
}

bool NetArrays::same_topology(NetArrays const& o) const {
    return bus_id == o.bus_id && bus_type == o.bus_type && u_rated == o.u_rated && u_pu == o.u_pu &&
           br_id == o.br_id && br_from == o.br_from && br_to == o.br_to && br_status == o.br_status &&
           br_g == o.br_g && br_b == o.br_b && app_id == o.app_id && app_node == o.app_node &&
           app_status == o.app_status && app_type == o.app_type && app_p == o.app_p && app_q == o.app_q &&
           app_g == o.app_g && app_b == o.app_b;
}

// ------------------------------------------------------------ admittance
std::unique_ptr<SparseMatrix> HipAdmittanceMatrix::build_admittance_matrix(NetworkData const& network_data) {
    NetArrays a(network_data);
    gaphip_net v = a.view();
    int const cap = v.n_bus + 2 * v.n_branch;
    auto m = std::make_unique<SparseMatrix>();
    m->num_rows = m->num_cols = v.n_bus;
    m->row_ptr.resize(v.n_bus + 1);
    m->col_idx.resize(cap);
    std::vector<double> ri(2 * static_cast<size_t>(std::max(cap, 1)));
    int nnz = 0;
    int rc = gaphip_ybus_build(&v, m->row_ptr.data(), m->col_idx.data(), ri.data(), cap, &nnz);
    if (rc) raise(rc, "build_admittance_matrix");
    m->col_idx.resize(nnz);
    m->values.resize(nnz);
    for (int k = 0; k < nnz; ++k) m->values[k] = Complex(ri[2 * k], ri[2 * k + 1]);
    m->nnz = nnz;
    return m;
}

// Incremental update with the semantics of the reference CPU builder
// (cpu_admittance_matrix.cpp:204-301): for each changed branch the diagonal
// entries get +-(y + g1 + j b1) + +-(j b1/2) and the first matching
// off-diagonal entries get -+y; indices are 0-based unless they exceed the
// matrix, then 1-based.  Runs on the host: it touches O(changes) entries.
std::unique_ptr<SparseMatrix> HipAdmittanceMatrix::update_admittance_matrix(
    SparseMatrix const& matrix, std::vector<BranchData> const& branch_changes) {
    auto out = std::make_unique<SparseMatrix>(matrix);
    auto add_at = [&](int row, int col, Complex d, bool subtract) {
        for (int q = out->row_ptr[row]; q < out->row_ptr[row + 1]; ++q)
            if (out->col_idx[q] == col) {
                if (subtract)
                    out->values[q] -= d;
                else
                    out->values[q] += d;
                return;
            }
    };
    for (auto const& br : branch_changes) {
        int f = br.from_bus, t = br.to_bus;
        if (f > 0 && t > 0 && (f > out->num_rows || t > out->num_rows)) {
            f -= 1;
            t -= 1;
        }
        if (f < 0 || f >= out->num_rows || t < 0 || t >= out->num_rows) continue;
        double z2 = br.r1 * br.r1 + br.x1 * br.x1;
        Complex ys(0.0, 0.0);
        if (z2 > 1e-12) ys = Complex(br.r1 / z2, -br.x1 / z2);
        Complex total = ys + Complex(br.g1, br.b1);
        Complex sh(0.0, br.b1 / 2.0);
        if (!br.status) {
            total = -total;
            ys = -ys;
            sh = -sh;
        }
        add_at(f, f, total + sh, false);
        add_at(t, t, total + sh, false);
        add_at(f, t, ys, true);
        add_at(t, f, ys, true);
    }
    return out;
}

// ----------------------------------------------------------------- LU
HipLUSolver::HipLUSolver() {
    int rc = gaphip_lu_create(&lu_);
    if (rc) raise(rc, "HipLUSolver");
}

HipLUSolver::~HipLUSolver() { gaphip_lu_destroy(lu_); }

bool HipLUSolver::factorize(SparseMatrix const& matrix) {
    n_ = 0;
    if (matrix.num_rows != matrix.num_cols || matrix.nnz <= 0 || matrix.values.empty()) return false;
    std::vector<double> ri(2 * static_cast<size_t>(matrix.nnz));
    for (int k = 0; k < matrix.nnz; ++k) {
        ri[2 * k] = matrix.values[k].real();
        ri[2 * k + 1] = matrix.values[k].imag();
    }
    int rc = gaphip_lu_factorize(lu_, matrix.num_rows, matrix.num_cols, matrix.row_ptr.data(),
                                 matrix.col_idx.data(), ri.data(), matrix.nnz);
    if (rc) return false;
    n_ = matrix.num_rows;
    return true;
}

ComplexVector HipLUSolver::solve(ComplexVector const& rhs) {
    if (!gaphip_lu_is_factorized(lu_)) throw std::runtime_error("Matrix not factorized. Call factorize() first.");
    if (static_cast<int>(rhs.size()) != n_) throw std::runtime_error("RHS size does not match matrix dimension");
    std::vector<double> b(2 * rhs.size()), x(2 * rhs.size());
    for (size_t i = 0; i < rhs.size(); ++i) {
        b[2 * i] = rhs[i].real();
        b[2 * i + 1] = rhs[i].imag();
    }
    int rc = gaphip_lu_solve(lu_, b.data(), x.data(), n_);
    if (rc) raise(rc, "solve");
    ComplexVector out(rhs.size());
    for (size_t i = 0; i < rhs.size(); ++i) out[i] = Complex(x[2 * i], x[2 * i + 1]);
    return out;
}

bool HipLUSolver::update_factorization(SparseMatrix const& matrix) { return factorize(matrix); }

bool HipLUSolver::is_factorized() const noexcept { return gaphip_lu_is_factorized(lu_) != 0; }

// -------------------------------------------------------- Newton-Raphson
HipNewtonRaphson::~HipNewtonRaphson() {
    if (plan_) gaphip_plan_destroy(plan_);
}

gaphip_plan* HipNewtonRaphson::plan_for(NetworkData const& net) {
    NetArrays a(net);
    if (plan_ && cached_ && cached_->same_topology(a)) {
        if (a.p_w != cached_->p_w || a.q_var != cached_->q_var) {
            int rc = gaphip_plan_set_injections(plan_, 0, a.p_w.data(), a.q_var.data());
            if (rc) raise(rc, "set_injections");
        }
        if (a.br_r != cached_->br_r || a.br_x != cached_->br_x) {
            int rc = gaphip_plan_set_branch_rx(plan_, 0, a.br_r.data(), a.br_x.data());
            if (rc) raise(rc, "set_branch_rx");
        }
        *cached_ = std::move(a);
        return plan_;
    }
    if (plan_) {
        gaphip_plan_destroy(plan_);
        plan_ = nullptr;
    }
    cached_ = std::make_unique<NetArrays>(std::move(a));
    gaphip_net v = cached_->view();
    int rc = gaphip_plan_create(1, &v, precision_, &plan_);
    if (rc) {
        plan_ = nullptr;
        raise(rc, "plan_create");
    }
    return plan_;
}

void HipNewtonRaphson::give_y(SparseMatrix const& y, int per_unit) {
    std::vector<double> ri(2 * static_cast<size_t>(y.nnz));
    for (int k = 0; k < y.nnz; ++k) {
        ri[2 * k] = y.values[k].real();
        ri[2 * k + 1] = y.values[k].imag();
    }
    int rc = gaphip_plan_set_ybus(plan_, 0, y.row_ptr.data(), y.col_idx.data(), ri.data(), y.nnz, per_unit);
    if (rc) raise(rc, "admittance matrix");
}

solver::PowerFlowResult HipNewtonRaphson::solve_power_flow(NetworkData const& network_data,
                                                           SparseMatrix const& admittance_matrix,
                                                           solver::PowerFlowConfig const& config) {
    flows_valid_ = false;
    plan_for(network_data);
    give_y(admittance_matrix, 0);
    gaphip_solve_opts o{config.tolerance, config.max_iterations, config.base_power};
    int const n = static_cast<int>(network_data.buses.size());
    std::vector<double> v(2 * static_cast<size_t>(n));
    int its = 0, conv = 0;
    double fm = 0.0;
    int rc = gaphip_solve(plan_, &o, v.data(), &its, &conv, &fm);
    if (rc) raise(rc, "solve_power_flow");
    flows_valid_ = true;
    solver::PowerFlowResult r;
    r.converged = conv != 0;
    r.iterations = its;
    r.final_mismatch = fm;
    r.bus_voltages.resize(n);
    for (int i = 0; i < n; ++i) r.bus_voltages[i] = Complex(v[2 * i], v[2 * i + 1]);
    if (capture_) {
        states_.clear();
        std::vector<double> tr(static_cast<size_t>(std::max(config.max_iterations, 1)) + 1);
        int nt = 0;
        if (gaphip_trajectory(plan_, 0, tr.data(), static_cast<int>(tr.size()), &nt) == 0) {
            for (int k = 0; k < nt && k < static_cast<int>(tr.size()); ++k) {
                solver::IterationState s;
                s.iteration = k;
                s.max_mismatch = tr[k];
                states_.push_back(std::move(s));
            }
        }
    }
    return r;
}

std::vector<BranchData> HipNewtonRaphson::branch_flows(NetworkData const& network_data) {
    if (!flows_valid_ || !plan_ || !cached_ || !cached_->same_topology(NetArrays(network_data)))
        throw std::runtime_error("branch_flows: no solve_power_flow of this network precedes it");
    size_t const nb = network_data.branches.size();
    std::vector<double> f(GAPHIP_FLOW_FIELDS * std::max<size_t>(nb, 1));
    int rc = gaphip_branch_flows(plan_, f.data());
    if (rc) raise(rc, "branch_flows");
    std::vector<BranchData> out(network_data.branches);
    for (size_t k = 0; k < nb; ++k) {
        double const* r = f.data() + GAPHIP_FLOW_FIELDS * k;
        BranchData& b = out[k];
        b.p_from = r[0];
        b.p_to = r[1];
        b.q_from = r[2];
        b.q_to = r[3];
        b.s_from = r[4];
        b.s_to = r[5];
        b.i_from = r[6];
        b.i_from_angle = r[7];
        b.i_to = r[8];
        b.i_to_angle = r[9];
        // power-grid-model's definition: current over rated current for a
        // line, apparent power over rated power for a transformer
        if (b.branch_type == BranchType::TRAFO)
            b.loading = b.sn > 0.0 ? std::max(b.s_from, b.s_to) / b.sn : 0.0;
        else
            b.loading = b.i_n > 0.0 ? std::max(b.i_from, b.i_to) / b.i_n : 0.0;
    }
    return out;
}

std::vector<Float> HipNewtonRaphson::calculate_mismatches(NetworkData const& network_data,
                                                          ComplexVector const& bus_voltages,
                                                          SparseMatrix const& admittance_matrix) {
    // public entry of the reference: S_base fixed at 100 MVA, Y used as given
    // whatever its pattern, result in watts / VAr (cpu_newton_raphson.cpp:225-240)
    double const base = 100e6;
    int const n = static_cast<int>(network_data.buses.size());
    if (static_cast<int>(bus_voltages.size()) != n) throw std::runtime_error("voltage vector size mismatch");
    // the reference's emptiness test (:301-305): zero calculated power
    SparseMatrix const& y = admittance_matrix;
    bool const empty_y = y.nnz == 0 || y.row_ptr.empty() || y.col_idx.empty() || y.values.empty();
    int const rows = empty_y ? 0 : static_cast<int>(y.row_ptr.size()) - 1;
    int const nnz = empty_y ? 0 : static_cast<int>(std::min(y.col_idx.size(), y.values.size()));
    std::vector<double> yri(2 * static_cast<size_t>(nnz));
    for (int k = 0; k < nnz; ++k) {
        yri[2 * k] = y.values[k].real();
        yri[2 * k + 1] = y.values[k].imag();
    }
    std::vector<double> v(2 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        v[2 * i] = bus_voltages[i].real();
        v[2 * i + 1] = bus_voltages[i].imag();
    }
    NetArrays const arrays(network_data);
    gaphip_net const view = arrays.view();
    std::vector<double> f(2 * static_cast<size_t>(n) + 1);
    int m = 0;
    int rc = gaphip_power_mismatches(&view, v.data(), rows, empty_y ? nullptr : y.row_ptr.data(),
                                     empty_y ? nullptr : y.col_idx.data(), yri.data(), nnz, base, f.data(),
                                     static_cast<int>(f.size()), &m);
    if (rc) raise(rc, "calculate_mismatches");
    std::vector<Float> out(m);
    for (int k = 0; k < m; ++k) out[k] = f[k] * base;
    return out;
}

}  // namespace gap::hip

extern "C" {

void* create_gpu_admittance_matrix(void) {
    return static_cast<gap::admittance::IAdmittanceMatrix*>(new gap::hip::HipAdmittanceMatrix());
}
void* create_gpu_lu_solver(void) { return static_cast<gap::solver::ILUSolver*>(new gap::hip::HipLUSolver()); }
void* create_gpu_powerflow_solver(void) {
    return static_cast<gap::solver::IPowerFlowSolver*>(new gap::hip::HipNewtonRaphson());
}

}  // extern "C"
