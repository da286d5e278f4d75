/*
 * gap_hip.h — C ABI of the MI355X Newton-Raphson power-flow backend.
 *
 * Plain C: opaque handles, pointers and sizes only.  The C++ plugin classes
 * (IAdmittanceMatrix / ILUSolver / IPowerFlowSolver, exported through the
 * create_gpu_* factories of gap_hip_plugins.h) and the Python module are
 * written on top of these functions; the HIP kernels sit underneath.
 *
 * Every function returns 0 on success and a negative GAPHIP_E* code on
 * failure; gaphip_last_error() then holds a message (thread-local).
 *
 * Units and conventions follow the reference data model
 * (include/gap/core/types.h): volts, ohms, siemens, watts, VAr, powers in
 * generator sign; complex numbers are interleaved (re, im) doubles.
 */
#ifndef GAP_HIP_H
#define GAP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define GAPHIP_ABI_VERSION 4

enum {
    GAPHIP_OK = 0,
    GAPHIP_E_ARG = -1,      /* invalid argument / inconsistent sizes            */
    GAPHIP_E_DEVICE = -2,   /* no usable GPU or a HIP runtime error             */
    GAPHIP_E_CAPACITY = -3, /* caller-provided output buffer too small          */
    GAPHIP_E_SINGULAR = -4, /* numerically zero pivot during factorisation      */
    GAPHIP_E_STATE = -5     /* call out of order (e.g. solve before factorize)  */
};

enum { GAPHIP_PQ = 0, GAPHIP_PV = 1, GAPHIP_SLACK = 2 };           /* types.h:30-34 */
enum { GAPHIP_SOURCE = 0, GAPHIP_LOADGEN = 1, GAPHIP_SHUNT = 2 };  /* types.h:39-43 */
enum { GAPHIP_FP64 = 0, GAPHIP_MIXED = 1 }; /* factor/solve precision: FP64, or FP32 with FP64 residuals */

/* One network.  Bus, branch and appliance fields mirror BusData (types.h:204-219),
 * BranchData (types.h:224-274) and ApplianceData (types.h:284-310); bus and
 * appliance endpoints are ids, resolved as in cpu_admittance_matrix.cpp:56-73.
 * br_g / br_b may be NULL (zero). */
typedef struct gaphip_net {
    int n_bus;
    const int* bus_id;
    const double* u_rated;
    const int* bus_type;
    const double* u_pu;
    const double* p_w;
    const double* q_var;

    int n_branch;
    const int* br_id;
    const int* br_from;
    const int* br_to;
    const int* br_status;
    const double* br_r;
    const double* br_x;
    const double* br_g;
    const double* br_b;

    int n_app;
    const int* app_id;
    const int* app_node;
    const int* app_status;
    const int* app_type;
    const double* app_p;
    const double* app_q;
    const double* app_g;
    const double* app_b;
} gaphip_net;

int gaphip_abi_version(void);
const char* gaphip_last_error(void);

/* 1 when a gfx950 device is visible.  Replaces the CUDA availability probe
 * BackendFactory::check_cuda_availability (src/core/backend_factory.cpp:134-160). */
int gaphip_device_available(void);

/* ---------------------------------------------------------------- Y-bus
 * Admittance matrix in siemens, CSR with int32 indices, assembled on the GPU.
 * Replaces IAdmittanceMatrix::build_admittance_matrix
 * (include/gap/admittance/admittance_interface.h:21-22; semantics of
 * src/admittance/cpu/cpu_admittance_matrix.cpp:44-202, bit-identical values).
 * row_ptr: n_bus+1; col_idx / values_ri: capacity `cap` entries (values 2*cap doubles). */
int gaphip_ybus_build(const gaphip_net* net, int* row_ptr, int* col_idx, double* values_ri, int cap,
                      int* nnz_out);

/* ------------------------------------------------------ power-flow plan
 * A plan holds the symbolic work for a batch of independent networks
 * (Y-bus assembly map, Jacobian pattern, fill-reducing order, supernodal
 * fronts and level schedule) plus their device-resident data.  It is the
 * analysis step that IPowerFlowSolver::solve_power_flow
 * (include/gap/solver/powerflow_interface.h:59-61) and ILUSolver::factorize
 * (include/gap/solver/lu_solver_interface.h:21-26) redo on every call in the
 * reference; here it runs once per topology. */
typedef struct gaphip_plan gaphip_plan;

typedef struct gaphip_plan_stats {
    int n_inst;
    long long n_bus, n_branch, ybus_nnz, n_var, jac_nnz;
    long long n_fronts, n_levels, front_entries, factor_entries, max_front;
    double analysis_ms;
    /* fronts by size: at most 2, 4, 8, 16, 32, 64, 128 slots, and above */
    long long front_hist[8];
} gaphip_plan_stats;

int gaphip_plan_create(int n_inst, const gaphip_net* nets, int precision, gaphip_plan** out);

/* Plan-time choices of the symbolic analysis and the grid schedule's task
 * layout (a plan must be created again to change them); results agree to
 * rounding whatever the choice.  relax_abs / relax_frac: relaxed supernode
 * amalgamation -- a child supernode that precedes its parent in postorder is
 * merged while the explicit zeros added stay within
 * max(relax_abs, relax_frac * merged front entries), in node units (defaults
 * 4, 0.10).  grid_min_class: smallest lanes-per-front class of a grid task, 0
 * to 4 for 4, 8, 16, 32, 64 lanes (default 0).  grid_pack: 1 (default) lets
 * smaller fronts fill the free lane groups of a larger class's last task of a
 * level.  No counterpart in the reference. */
typedef struct gaphip_plan_options {
    int relax_abs;
    double relax_frac;
    int grid_min_class;
    int grid_pack;
} gaphip_plan_options;
/* gaphip_plan_create with the given options (NULL: the defaults). */
int gaphip_plan_create_opt(int n_inst, const gaphip_net* nets, int precision, const gaphip_plan_options* options,
                           gaphip_plan** out);
/* Host-only symbolic analysis of a batch (no device touched): fills the same
 * statistics as gaphip_plan_stats_get so fill, front memory and the level
 * count can be sized before anything is allocated. */
int gaphip_analyze(int n_inst, const gaphip_net* nets, gaphip_plan_stats* out);
int gaphip_plan_destroy(gaphip_plan* plan);
int gaphip_plan_stats_get(const gaphip_plan* plan, gaphip_plan_stats* out);

/* Replace the specified bus injections (BusData active/reactive power, W / VAr)
 * of instance `inst` before the next solve; pointers of length n_bus of that instance. */
int gaphip_plan_set_injections(gaphip_plan* plan, int inst, const double* p_w, const double* q_var);
/* Replace branch r1 / x1 (ohm) of instance `inst` (length n_branch). */
int gaphip_plan_set_branch_rx(gaphip_plan* plan, int inst, const double* r, const double* x);
/* Incremental admittance update on the device (IAdmittanceMatrix::
 * update_admittance_matrix, include/gap/admittance/admittance_interface.h:29-30,
 * for the plan-resident Y-bus): new r1 / x1 (ohm) for `n` branches of instance
 * `inst` (indices within the instance).  Once a solve has assembled the plan's
 * Y-bus, exactly the Y slots those branches contribute to are recomputed on the
 * device with the full assembly's summation order, so the result is
 * bit-identical to assembling the changed network afresh (not the reference's
 * delta update, which double-counts b1, cpu_admittance_matrix.cpp:247-266).
 * Solves re-assemble the whole Y-bus as before.  Instances whose Y-bus the
 * caller supplied (gaphip_plan_set_ybus) only get the new r1 / x1. */
int gaphip_plan_update_branches(gaphip_plan* plan, int inst, int n, const int* branch, const double* r1,
                                const double* x1);
/* The plan's Y-bus of instance `inst`: per unit (siemens * u_rated[0]^2 / S_base
 * of the last solve), instance-local CSR (row_ptr n_bus + 1, col_idx and
 * values_ri (re, im pairs) of *nnz entries, which is always set).  Before the
 * first solve GAPHIP_E_STATE; cap < nnz GAPHIP_E_CAPACITY.  Any output pointer
 * may be NULL. */
int gaphip_plan_get_ybus(const gaphip_plan* plan, int inst, int* row_ptr, int* col_idx, double* values_ri, int cap,
                         int* nnz);
/* Use a caller-supplied Y-bus for instance `inst` instead of assembling it
 * (IPowerFlowSolver::solve_power_flow receives Y from the caller,
 * powerflow_interface.h:59-61).  The pattern must equal the plan's Y-bus
 * pattern.  per_unit = 0: siemens, scaled by (u_rated[0]^2)/base_power at use
 * (cpu_newton_raphson.cpp:46-58); per_unit = 1: used as given.
 * Passing row_ptr = NULL returns the instance to on-device assembly. */
int gaphip_plan_set_ybus(gaphip_plan* plan, int inst, const int* row_ptr, const int* col_idx,
                         const double* values_ri, int nnz, int per_unit);

/* Device schedule of the Newton loop.  LEVELS: one launch per phase and per
 * (elimination-tree level, front size class) over all networks, host check of
 * convergence every iteration.  GRID: one workgroup per network runs every
 * iteration of that network inside a single launch (needs every front of the
 * plan to have at most 64 slots).  AUTO (the default) picks GRID when the plan
 * allows it.  `effective` (may be NULL) receives the schedule solves will use.
 * No counterpart in the reference: its CPU solver has no device schedule. */
enum { GAPHIP_SCHED_AUTO = 0, GAPHIP_SCHED_LEVELS = 1, GAPHIP_SCHED_GRID = 2 };
int gaphip_plan_set_schedule(gaphip_plan* plan, int schedule, int* effective);

/* Launch mode of the level schedule's Newton step (Jacobian, factorisation,
 * front solves, update: hundreds of launches per iteration on a large grid).
 * on = 1: the step is captured once per plan into a HIP graph and the graph
 * is replayed every iteration, one launch per step; on = 0: the kernels are
 * launched one by one.  Default: on, except for plans with a front of 4,096
 * slots or more (their look-ahead overlaps two streams, which a replayed
 * graph overlaps less).  Results are the same bit for bit; the
 * mismatch and convergence check stay direct launches (the host reads the
 * count of networks still iterating after them).  Profiling
 * (gaphip_profile_enable) uses direct launches.  No counterpart in the
 * reference. */
int gaphip_plan_set_graph(gaphip_plan* plan, int on);

/* Run-time schedule knobs of a plan; results do not depend on them.
 * Grid schedule: a network whose elimination tree holds at least
 * grid_big_work front slots plus factor entries gets grid_big_waves
 * wavefronts per workgroup, the others grid_waves (each 1, 2, 4 or 8); the two
 * groups run concurrently, the second launched a few microseconds after the
 * first.  grid_big_work < 0 puts every network in the first group.
 * Level schedule: level_graph 1 replays the Newton step from a captured HIP
 * graph, 0 launches its kernels one by one (see gaphip_plan_set_graph);
 * fronts of at least big_two_level_min slots are factored in outer panels of
 * big_outer pivots (a multiple of 64, at most 1024), each as 64-wide inner
 * panels with one trailing update of K = big_outer per outer panel, and the
 * next panel's update overlapped; smaller fronts above 128 slots use 64-wide
 * panels.  Defaults: 2, 4, 40000 (GAPHIP_MIXED) or 2, 8, -1 (GAPHIP_FP64); graph
 * on unless a front reaches big_two_level_min; 256; 4096.
 * gap_amd.tuning.autotune() times candidates on a plan and sets the fastest.
 * No counterpart in the reference. */
typedef struct gaphip_tuning {
    int grid_waves;
    int grid_big_waves;
    long long grid_big_work;
    int level_graph;
    int big_outer;
    int big_two_level_min;
} gaphip_tuning;
int gaphip_plan_get_tuning(const gaphip_plan* plan, gaphip_tuning* out);
int gaphip_plan_set_tuning(gaphip_plan* plan, const gaphip_tuning* tuning);

typedef struct gaphip_solve_opts {
    double tolerance;   /* PowerFlowConfig::tolerance (powerflow_interface.h:14)       */
    int max_iterations; /* PowerFlowConfig::max_iterations (powerflow_interface.h:15)  */
    double base_power;  /* PowerFlowConfig::base_power, VA (powerflow_interface.h:19)  */
} gaphip_solve_opts;

/* Newton-Raphson from flat start on every instance of the plan
 * (cpu_newton_raphson.cpp:33-218: 0.9-damped, clamped polar update;
 * iterations = iter+1 at convergence, max_iterations otherwise).
 * Any output pointer may be NULL (results then stay on the device):
 *   v_ri            2 * sum(n_bus) doubles, instances concatenated
 *   iterations      n_inst ints;  converged n_inst ints;  final_mismatch n_inst doubles */
int gaphip_solve(gaphip_plan* plan, const gaphip_solve_opts* opts, double* v_ri, int* iterations,
                 int* converged, double* final_mismatch);

/* Per-iteration max|F| of instance `inst` from the last solve (state capture,
 * powerflow_interface.h:28-35).  Returns the count through n_out. */
int gaphip_trajectory(const gaphip_plan* plan, int inst, double* max_mismatch, int cap, int* n_out);

/* Jacobian of instance `inst` at voltages v_ri (2*n_bus), in the reference
 * layout (rows: P then Q per non-slack bus; columns: angles then magnitudes;
 * |J| > 1e-12 kept, cpu_newton_raphson.cpp:509-765), evaluated by the GPU
 * kernel with the plan's per-unit Y-bus (base_power as given). */
int gaphip_jacobian(gaphip_plan* plan, int inst, const double* v_ri, double base_power, int* n_var,
                    int* row_ptr, int* col_idx, double* values, int cap, int* nnz_out);

/* Mismatch vector of instance `inst` at v_ri (per unit of base_power), same
 * layout as cpu_newton_raphson.cpp:243-292. */
int gaphip_mismatches(gaphip_plan* plan, int inst, const double* v_ri, double base_power,
                      double* f, int cap, int* m_out);

/* Mismatch vector of one network at v_ri (2*n_bus) for a caller-given
 * admittance matrix of any pattern, used as given (no per-unit scaling),
 * with S_spec over base_power: the evaluation behind the public
 * IPowerFlowSolver::calculate_mismatches (include/gap/solver/powerflow_interface.h:46-48;
 * cpu_newton_raphson.cpp:225-337, which multiplies the result by its fixed
 * 100 MVA).  y_rows = row_ptr entries - 1; y_nnz = the shorter of the column
 * and value arrays.  As the reference walks the matrix: rows past y_rows and
 * an empty matrix (y_nnz = 0) give zero current, a row stops at the end of
 * the arrays, columns >= n_bus are skipped (negative ones are E_ARG).
 * Rows: P of every non-slack bus, then Q if PQ, in bus order; *m_out is set. */
int gaphip_power_mismatches(const gaphip_net* net, const double* v_ri, int y_rows, const int* y_row_ptr,
                            const int* y_col_idx, const double* y_ri, int y_nnz, double base_power, double* f,
                            int cap, int* m_out);

/* The same evaluation with no device allocation: no counterpart in the
 * reference, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as gaphip_power_mismatches (its kernels and
 * arithmetic, bit for bit); every byte of device memory it touches lies in one
 * static array per GPU, sized for the limits below (about 8.4 MB, held until
 * the process ends).  Larger inputs return GAPHIP_E_CAPACITY.  It calls no
 * hipMalloc / hipFree and creates no stream: copies and kernels run on the
 * null stream and the call returns when F is on the host.  Calls on one GPU
 * share its array: this library serialises them with a lock, and each call
 * overwrites the previous call's contents. */
#define GAPHIP_NOALLOC_MAX_BUS 32768
#define GAPHIP_NOALLOC_MAX_NNZ 262144
#define GAPHIP_NOALLOC_MAX_APP 32768
int gaphip_power_mismatches_noalloc(const gaphip_net* net, const double* v_ri, int y_rows, const int* y_row_ptr,
                                    const int* y_col_idx, const double* y_ri, int y_nnz, double base_power,
                                    double* f, int cap, int* m_out);

/* Adjoint sensitivities at the converged state of the last solve (FP64):
 * solves J^T lambda = dL/dx with the factors of J at the solution and returns
 * dL/d(bus p_w, bus q_var) [per W / VAr] and dL/d(branch r1, x1) [per ohm] for
 *   L = w_loss * (total active branch loss, pu) + w_dev * sum_PQ(|V_i| - 1)^2
 *       + sum_i g_vm[i] |V_i| + sum_i g_va[i] arg V_i          (g_* may be NULL).
 * Arrays are concatenated over instances.  loss_out (n_inst) may be NULL.
 * GAPHIP_E_STATE when no solve has run, when a network of the last solve did
 * not converge, or when injections, branch data or the Y-bus were changed
 * (set_injections, set_branch_rx, update_branches, set_ybus) after it: the
 * gradients belong to a converged solve of the current inputs. */
int gaphip_adjoint(gaphip_plan* plan, double w_loss, double w_dev, const double* g_vm,
                   const double* g_va, double* grad_p, double* grad_q, double* grad_r,
                   double* grad_x, double* loss_out);

/* Branch flows at the last solution of every instance: per branch, in
 * instance-concatenated branch order, the ten BranchData result fields
 * (include/gap/core/types.h:233-243, which the reference declares but never
 * fills) in struct order: p_from, p_to, q_from, q_to (W, VAr), s_from, s_to
 * (VA), i_from, i_from_angle, i_to, i_to_angle (A, rad).  Computed from the
 * branch's own block of the Y-bus (series 1/(r1 + j x1), g1 + j b1/2 at each
 * end, the builder's model, cpu_admittance_matrix.cpp:121-135) in the
 * network's per-unit system; currents in A of its voltage base u_rated[0].
 * Out-of-service branches get zeros; `loading` needs a rating BranchData does
 * not carry and is not computed.  out: 10 * total branches doubles.
 * Before any solve, or after a change of branch data or the Y-bus since the
 * last solve: GAPHIP_E_STATE. */
#define GAPHIP_FLOW_FIELDS 10
int gaphip_branch_flows(gaphip_plan* plan, double* out);

/* Timing hook for measurement: after each solve, the summed device time and
 * launch count of each kernel family, and the algorithmic HBM bytes they move.
 * kernel ids: see gaphip_kernel_name(). Enabled by gaphip_profile_enable(plan, 1). */
int gaphip_profile_enable(gaphip_plan* plan, int on);
int gaphip_profile_get(const gaphip_plan* plan, int kernel, double* total_ms, long long* launches,
                       double* bytes);
const char* gaphip_kernel_name(int kernel);
int gaphip_kernel_count(void);

/* ------------------------------------------------------- sparse LU
 * General sparse LU on the GPU for the ILUSolver plugin
 * (include/gap/solver/lu_solver_interface.h:12-48): complex CSR in,
 * factorize / solve / refactor, error contract of cpu_lu_solver.cpp:99-108,
 * 157-164 (non-square or empty -> E_ARG, solve before factorize -> E_STATE).
 * The matrix is factored without row exchanges on a minimum-degree order
 * of pattern(A + A^T); a numerically zero pivot returns GAPHIP_E_SINGULAR. */
typedef struct gaphip_lu gaphip_lu;
int gaphip_lu_create(gaphip_lu** out);
int gaphip_lu_destroy(gaphip_lu* lu);
int gaphip_lu_factorize(gaphip_lu* lu, int n_rows, int n_cols, const int* row_ptr,
                        const int* col_idx, const double* values_ri, int nnz);
int gaphip_lu_solve(gaphip_lu* lu, const double* rhs_ri, double* x_ri, int n);
int gaphip_lu_is_factorized(const gaphip_lu* lu);

#ifdef __cplusplus
}
#endif
#endif
