// Layout of the plugin boundary's types: compiled by tests/test_abi_layout.py
// once against the reference's headers and once against include/gap, run, and
// the two outputs compared line by line.  A plugin built against include/gap
// is handed these objects by the reference's code (and hands them back), so
// sizes, alignments, field offsets, enumerator values and the vtable slot of
// every virtual method must agree.  Test code only; nothing here ships.
#include <cstddef>
#include <cstdio>

#include "gap/admittance/admittance_interface.h"
#include "gap/core/types.h"
#include "gap/io/io_interface.h"
#include "gap/solver/lu_solver_interface.h"
#include "gap/solver/powerflow_interface.h"

using namespace gap;

// vtable slot of a virtual member function: in the Itanium C++ ABI a pointer
// to a virtual member holds 1 + the slot's byte offset in the vtable
template <typename M>
long vslot(M m) {
    union {
        M pmf;
        struct {
            std::ptrdiff_t ptr, adj;
        } raw;
    } u;
    u.pmf = m;
    return static_cast<long>((u.raw.ptr - 1) / static_cast<std::ptrdiff_t>(sizeof(void*)));
}

#define SZ(T) std::printf("size %s %zu %zu\n", #T, sizeof(T), alignof(T))
#define OFF(T, f) std::printf("offset %s.%s %zu %zu\n", #T, #f, offsetof(T, f), sizeof(T::f))
#define EN(E, v) std::printf("enum %s::%s %d\n", #E, #v, static_cast<int>(E::v))
#define VS(C, f) std::printf("vslot %s::%s %ld\n", #C, #f, vslot(&C::f))

int main() {
    SZ(SparseMatrix);
    OFF(SparseMatrix, row_ptr); OFF(SparseMatrix, col_idx); OFF(SparseMatrix, values);
    OFF(SparseMatrix, num_rows); OFF(SparseMatrix, num_cols); OFF(SparseMatrix, nnz);
    SZ(BusData);
    OFF(BusData, id); OFF(BusData, u_rated); OFF(BusData, bus_type); OFF(BusData, energized); OFF(BusData, u);
    OFF(BusData, u_pu); OFF(BusData, u_angle); OFF(BusData, active_power); OFF(BusData, reactive_power);
    SZ(BranchData);
    OFF(BranchData, id); OFF(BranchData, from_bus); OFF(BranchData, from_status); OFF(BranchData, to_bus);
    OFF(BranchData, to_status); OFF(BranchData, status); OFF(BranchData, branch_type); OFF(BranchData, p_from);
    OFF(BranchData, p_to); OFF(BranchData, q_from); OFF(BranchData, q_to); OFF(BranchData, s_from);
    OFF(BranchData, s_to); OFF(BranchData, i_from); OFF(BranchData, i_from_angle); OFF(BranchData, i_to);
    OFF(BranchData, i_to_angle); OFF(BranchData, loading); OFF(BranchData, r1); OFF(BranchData, x1);
    OFF(BranchData, g1); OFF(BranchData, b1); OFF(BranchData, k); OFF(BranchData, theta); OFF(BranchData, sn);
    OFF(BranchData, i_n);
    SZ(ApplianceData);
    OFF(ApplianceData, id); OFF(ApplianceData, node); OFF(ApplianceData, status); OFF(ApplianceData, type);
    OFF(ApplianceData, p); OFF(ApplianceData, q); OFF(ApplianceData, i); OFF(ApplianceData, s);
    OFF(ApplianceData, u_ref); OFF(ApplianceData, u_ref_angle); OFF(ApplianceData, sk);
    OFF(ApplianceData, rx_ratio); OFF(ApplianceData, p_specified); OFF(ApplianceData, q_specified);
    OFF(ApplianceData, g1); OFF(ApplianceData, b1);
    SZ(NetworkData);
    OFF(NetworkData, buses); OFF(NetworkData, branches); OFF(NetworkData, appliances);
    OFF(NetworkData, num_buses); OFF(NetworkData, num_branches); OFF(NetworkData, num_appliances);
    OFF(NetworkData, version); OFF(NetworkData, type); OFF(NetworkData, is_batch);
    SZ(solver::PowerFlowConfig);
    OFF(solver::PowerFlowConfig, tolerance); OFF(solver::PowerFlowConfig, max_iterations);
    OFF(solver::PowerFlowConfig, use_flat_start); OFF(solver::PowerFlowConfig, acceleration_factor);
    OFF(solver::PowerFlowConfig, verbose); OFF(solver::PowerFlowConfig, base_power);
    SZ(solver::PowerFlowResult);
    OFF(solver::PowerFlowResult, converged); OFF(solver::PowerFlowResult, iterations);
    OFF(solver::PowerFlowResult, final_mismatch); OFF(solver::PowerFlowResult, bus_voltages);
    OFF(solver::PowerFlowResult, bus_injections);
    SZ(solver::IterationState);
    OFF(solver::IterationState, iteration); OFF(solver::IterationState, voltages);
    OFF(solver::IterationState, currents); OFF(solver::IterationState, power_injections);
    OFF(solver::IterationState, mismatches); OFF(solver::IterationState, max_mismatch);
    EN(BusType, PQ); EN(BusType, PV); EN(BusType, SLACK);
    EN(ApplianceType, SOURCE); EN(ApplianceType, LOADGEN); EN(ApplianceType, SHUNT);
    EN(BranchType, LINE); EN(BranchType, TRAFO); EN(BranchType, GENERIC);
    EN(BackendType, CPU); EN(BackendType, GPU_CUDA);
    SZ(admittance::IAdmittanceMatrix);
    VS(admittance::IAdmittanceMatrix, build_admittance_matrix);
    VS(admittance::IAdmittanceMatrix, update_admittance_matrix);
    VS(admittance::IAdmittanceMatrix, get_backend_type);
    SZ(solver::ILUSolver);
    VS(solver::ILUSolver, factorize); VS(solver::ILUSolver, solve); VS(solver::ILUSolver, update_factorization);
    VS(solver::ILUSolver, get_backend_type); VS(solver::ILUSolver, is_factorized);
    SZ(solver::IPowerFlowSolver);
    VS(solver::IPowerFlowSolver, solve_power_flow); VS(solver::IPowerFlowSolver, set_lu_solver);
    VS(solver::IPowerFlowSolver, calculate_mismatches); VS(solver::IPowerFlowSolver, get_backend_type);
    VS(solver::IPowerFlowSolver, enable_state_capture); VS(solver::IPowerFlowSolver, get_iteration_states);
    VS(solver::IPowerFlowSolver, clear_iteration_states);
    SZ(io::IIOModule);
    VS(io::IIOModule, read_network_data); VS(io::IIOModule, write_results); VS(io::IIOModule, validate_input_format);
    return 0;
}
