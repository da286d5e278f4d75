/*
 * gap_hip_plugins.h — plugin factory symbols of the MI355X backend.
 *
 * libgap_hip_plugins.so (also reachable under the reference's library names
 * libgap_admittance_gpu.so, libgap_lu_gpu.so and libgap_powerflow_gpu.so)
 * exports the same three factories the reference backend factory resolves
 * with dlsym (src/core/backend_factory.cpp:34-41, 58-63, 82-87).  Each returns
 * a new object of the matching C++ interface, owned by the caller and deleted
 * through its virtual destructor:
 *   create_gpu_admittance_matrix -> gap::admittance::IAdmittanceMatrix*
 *   create_gpu_lu_solver         -> gap::solver::ILUSolver*
 *   create_gpu_powerflow_solver  -> gap::solver::IPowerFlowSolver*
 * They are declared here with void* returns so C callers can bind them; C++
 * callers cast to the interface type, as the reference factory does.
 */
#ifndef GAP_HIP_PLUGINS_H
#define GAP_HIP_PLUGINS_H

#ifdef __cplusplus
extern "C" {
#endif

void* create_gpu_admittance_matrix(void);
void* create_gpu_lu_solver(void);
void* create_gpu_powerflow_solver(void);

#ifdef __cplusplus
}
#endif
#endif
