/*
 * pforacle.h — FP64 CPU restatement of the reference Newton-Raphson power-flow
 * path, written in C99.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP path: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; nothing in
 * the product (gap_amd/, include/) links or calls it.
 *
 * Semantics followed (reference file:line):
 *   pfo_build_ybus   cpu_admittance_matrix.cpp:44-202
 *   pfo_solve        cpu_newton_raphson.cpp:33-218 (driver), 243-292 (mismatch),
 *                    296-337 (S = V conj(YV)), 346-377 (S_spec), 382-433 (correction),
 *                    438-504 (damped, clamped polar update), 509-727 (Jacobian),
 *                    732-765 (|J| > 1e-12 compaction)
 *   pfo_mismatches   cpu_newton_raphson.cpp:225-240 (watts, S_base fixed 100 MVA)
 *   pfo_lu_solve     any exact LU meets the contract of cpu_lu_solver.cpp:91-185;
 *                    this one is a left-looking sparse LU with threshold partial
 *                    pivoting and a minimum-degree column order.
 * The dense Jacobian of the reference is restated sparsely: an entry can only be
 * non-zero where Y has a structural entry, and duplicates of a Y entry keep the
 * last one (dense overwrite, cpu_newton_raphson.cpp:585-597).
 */
#ifndef PFORACLE_H
#define PFORACLE_H

#include "pf_net.h"

#ifdef __cplusplus
extern "C" {
#endif

int pfo_build_ybus(const pf_net* net, int* row_ptr, int* col_idx, double* val_ri, int cap);

int pfo_solve(const pf_net* net, double tol, int max_it, double base_power, double* v_out,
              int* iterations, int* converged, double* final_mismatch, double* traj, int traj_cap,
              int* n_traj, int j_call, int* j_n, int* j_row_ptr, int* j_col_idx, double* j_val,
              int j_cap, int* j_nnz);

int pfo_mismatches(const pf_net* net, const double* v_ri, const double* y_ri, const int* y_row_ptr,
                   const int* y_col_idx, int y_nnz, double* out, int cap);

int pfo_lu_solve(int n, const int* row_ptr, const int* col_idx, const double* val_ri, int nnz,
                 const double* rhs_ri, double* x_ri);

/* Jacobian (reference layout) of the per-unit Y at voltages v: returns nnz or -1. */
int pfo_jacobian(const pf_net* net, const double* v_ri, const double* ypu_ri, const int* y_row_ptr,
                 const int* y_col_idx, int* j_n, int* j_row_ptr, int* j_col_idx, double* j_val,
                 int j_cap);

#ifdef __cplusplus
}
#endif
#endif
