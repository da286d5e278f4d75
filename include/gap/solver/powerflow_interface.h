// Power-flow plugin interface; same types and virtual signatures as the
// reference (include/gap/solver/powerflow_interface.h:13-108 of the reference).
#pragma once

#include <memory>
#include <vector>

#include "gap/core/types.h"
#include "gap/solver/lu_solver_interface.h"

namespace gap::solver {

struct PowerFlowConfig {
    Float tolerance = 1e-6;
    int max_iterations = 50;
    bool use_flat_start = true;
    Float acceleration_factor = 1.0;  // accepted, not used: the update is the 0.9-damped one
    bool verbose = false;
    Float base_power = 100e6;  // VA
};

struct IterationState {
    int iteration = 0;
    std::vector<Complex> voltages;
    std::vector<Complex> currents;
    std::vector<Complex> power_injections;
    std::vector<double> mismatches;
    double max_mismatch = 0.0;
};

struct PowerFlowResult {
    bool converged = false;
    int iterations = 0;
    Float final_mismatch = 0.0;
    ComplexVector bus_voltages;
    ComplexVector bus_injections;
};

class IPowerFlowSolver {
  public:
    virtual ~IPowerFlowSolver() = default;
    virtual PowerFlowResult solve_power_flow(NetworkData const& network_data,
                                             SparseMatrix const& admittance_matrix,
                                             PowerFlowConfig const& config = PowerFlowConfig{}) = 0;
    virtual void set_lu_solver(std::shared_ptr<ILUSolver> lu_solver) = 0;
    virtual std::vector<Float> calculate_mismatches(NetworkData const& network_data,
                                                    ComplexVector const& bus_voltages,
                                                    SparseMatrix const& admittance_matrix) = 0;
    virtual BackendType get_backend_type() const noexcept = 0;
    virtual void enable_state_capture(bool enable) { (void)enable; }
    virtual std::vector<IterationState> const& get_iteration_states() const {
        static std::vector<IterationState> const empty;
        return empty;
    }
    virtual void clear_iteration_states() {}
};

}  // namespace gap::solver
