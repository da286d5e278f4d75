// Plugin classes of the MI355X backend (see backends.cpp).
#pragma once

#include <memory>
#include <vector>

#include "gap/admittance/admittance_interface.h"
#include "gap/solver/lu_solver_interface.h"
#include "gap/solver/powerflow_interface.h"
#include "gap_hip.h"

namespace gap::hip {

// NetworkData flattened into the arrays gaphip_net points at.
struct NetArrays {
    std::vector<int> bus_id, bus_type, br_id, br_from, br_to, br_status, app_id, app_node, app_status, app_type;
    std::vector<double> u_rated, u_pu, p_w, q_var, br_r, br_x, br_g, br_b, app_p, app_q, app_g, app_b;
    explicit NetArrays(NetworkData const& net);
    gaphip_net view() const;
    // everything except bus injections and branch r1 / x1
    bool same_topology(NetArrays const& o) const;
};

class HipAdmittanceMatrix final : public admittance::IAdmittanceMatrix {
  public:
    std::unique_ptr<SparseMatrix> build_admittance_matrix(NetworkData const& network_data) override;
    std::unique_ptr<SparseMatrix> update_admittance_matrix(SparseMatrix const& matrix,
                                                           std::vector<BranchData> const& branch_changes) override;
    BackendType get_backend_type() const noexcept override { return BackendType::GPU_HIP; }
};

class HipLUSolver final : public solver::ILUSolver {
  public:
    HipLUSolver();
    ~HipLUSolver() override;
    HipLUSolver(HipLUSolver const&) = delete;
    HipLUSolver& operator=(HipLUSolver const&) = delete;
    bool factorize(SparseMatrix const& matrix) override;
    ComplexVector solve(ComplexVector const& rhs) override;
    bool update_factorization(SparseMatrix const& matrix) override;
    BackendType get_backend_type() const noexcept override { return BackendType::GPU_HIP; }
    bool is_factorized() const noexcept override;

  private:
    gaphip_lu* lu_ = nullptr;
    int n_ = 0;
};

// The Newton loop, Jacobian and LU all run on the device inside one plan, so
// an ILUSolver handed to set_lu_solver is kept but not called.
class HipNewtonRaphson final : public solver::IPowerFlowSolver {
  public:
    explicit HipNewtonRaphson(int precision = GAPHIP_FP64) : precision_(precision) {}
    ~HipNewtonRaphson() override;
    HipNewtonRaphson(HipNewtonRaphson const&) = delete;
    HipNewtonRaphson& operator=(HipNewtonRaphson const&) = delete;
    solver::PowerFlowResult solve_power_flow(NetworkData const& network_data, SparseMatrix const& admittance_matrix,
                                             solver::PowerFlowConfig const& config) override;
    void set_lu_solver(std::shared_ptr<solver::ILUSolver> lu_solver) override { lu_ = std::move(lu_solver); }
    std::vector<Float> calculate_mismatches(NetworkData const& network_data, ComplexVector const& bus_voltages,
                                            SparseMatrix const& admittance_matrix) override;
    BackendType get_backend_type() const noexcept override { return BackendType::GPU_HIP; }
    void enable_state_capture(bool enable) override { capture_ = enable; }
    std::vector<solver::IterationState> const& get_iteration_states() const override { return states_; }
    void clear_iteration_states() override { states_.clear(); }

    // Copies of network_data's branches with the BranchData result fields
    // (p/q/s/i at both ends, and loading: current over i_n for a line, apparent
    // power over sn for a transformer) filled at the solution of the last
    // solve_power_flow on that network (types.h:233-243; the reference never
    // fills them).  No counterpart in the reference.  Throws when no solve of
    // this network precedes it or another call has used the plan since.
    std::vector<BranchData> branch_flows(NetworkData const& network_data);

  private:
    gaphip_plan* plan_for(NetworkData const& net);
    void give_y(SparseMatrix const& y, int per_unit);
    int precision_;
    gaphip_plan* plan_ = nullptr;
    std::unique_ptr<NetArrays> cached_;
    std::shared_ptr<solver::ILUSolver> lu_;
    bool capture_ = false;
    bool flows_valid_ = false;  // the plan holds the last solve's solution
    std::vector<solver::IterationState> states_;
};

}  // namespace gap::hip
