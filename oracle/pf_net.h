/*
 * pf_net.h — flat, pointer-and-size description of a power-flow network,
 * shared by the two CPU checkers under oracle/:
 *   - pforacle.c   : the FP64 C restatement of the reference CPU path;
 *   - ref_harness.cpp : a thin driver around the reference's own CPU sources
 *                       (compiled from /root/reference into oracle/_ref/).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gap_amd/, include/)
 * includes this header or links anything built from oracle/.
 *
 * Field meaning follows the reference data model, include/gap/core/types.h:
 *   BusData        types.h:204-219   (id, u_rated [V], bus_type, u_pu, active/reactive power [W, VAr])
 *   BranchData     types.h:224-274   (id, from/to bus *ids*, status, r1/x1 [ohm], g1/b1 [S])
 *   ApplianceData  types.h:284-310   (id, node id, status, type, p/q_specified [W, VAr], g1/b1 [S])
 * Bus/branch rows are exactly what bindings/solver_bindings.cpp:64-111 builds
 * from the Python row API; appliances are what the C++ model adds on top
 * (SOURCE/LOADGEN power, SHUNT admittance).
 */
#ifndef PF_NET_H
#define PF_NET_H

#ifdef __cplusplus
extern "C" {
#endif

enum { PF_PQ = 0, PF_PV = 1, PF_SLACK = 2 };             /* types.h:30-34 */
enum { PF_SOURCE = 0, PF_LOADGEN = 1, PF_SHUNT = 2 };     /* types.h:39-43 */

typedef struct pf_net {
    int n_bus;
    const int* bus_id;
    const double* u_rated;
    const int* bus_type;
    const double* u_pu;
    const double* p_w;    /* BusData::active_power  (generator sign, W)   */
    const double* q_var;  /* BusData::reactive_power (generator sign, VAr) */

    int n_branch;
    const int* br_id;
    const int* br_from;   /* bus id */
    const int* br_to;     /* bus id */
    const int* br_status;
    const double* br_r;
    const double* br_x;
    const double* br_g;
    const double* br_b;

    int n_app;
    const int* app_id;
    const int* app_node;  /* bus id */
    const int* app_status;
    const int* app_type;
    const double* app_p;
    const double* app_q;
    const double* app_g;
    const double* app_b;
} pf_net;

#ifdef __cplusplus
}
#endif
#endif
