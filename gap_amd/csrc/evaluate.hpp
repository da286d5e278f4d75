// One-off power mismatch of a network for a caller-given admittance matrix:
// the evaluation behind IPowerFlowSolver::calculate_mismatches
// (cpu_newton_raphson.cpp:225-377).  The matrix may have any pattern and is
// walked with the reference's bounds rules on the device:
//   * rows at or past y_rows, and every row when y_nnz = 0, give zero current
//     (:301-314);
//   * a row stops at the end of the arrays (q >= y_nnz, :318-321) and skips
//     columns >= n_bus (:323-326);
//   * a negative row pointer or column, which the reference would read out of
//     bounds, sets the error flag instead.
// Rows of F: P of every non-slack bus, then Q if PQ, in bus order (:269-289).
// S_spec (:346-377): the bus's own injection when either part is non-zero,
// plus every in-service SOURCE / LOADGEN appliance whose node equals the bus
// id, in appliance order, each part divided by base_power.
//
// All device memory of an evaluation is one region laid out by EvalLayout:
// the staged inputs (row pointers, columns, values, V, bus and appliance
// fields), the row maps, the output and the error flag.  gaphip_power_mismatches
// allocates the region per call; gaphip_power_mismatches_noalloc uses a static
// device array of a fixed maximum size.
#pragma once

#include "kernels.hpp"

namespace gaphip {

struct EvalArgs {
    int n_bus, y_rows, y_nnz, n_app;
    double base;
    const int* rp;   // min(y_rows, n_bus) + 1 entries (unused when y_nnz = 0)
    const int* ci;   // y_nnz
    const double2* y;
    const double2* v;
    const int* bus_id;
    const int* bus_type;
    const double* p_w;
    const double* q_var;
    const int* app_node;
    const int* app_status;
    const int* app_type;
    const double* app_p;
    const double* app_q;
    int* peq;  // per bus: its P row or -1
    int* qeq;  // per bus: its Q row or -1
    double* f;
    int* err;
};

// Row maps by a scan of the bus types: one workgroup of kEvalScanThreads,
// each thread a contiguous chunk of buses.  Also clears the error flag (the
// region keeps the previous evaluation's contents).
constexpr int kEvalScanThreads = 1024;
__global__ void __launch_bounds__(kEvalScanThreads) k_eval_rows(EvalArgs a) {
    __shared__ int cnt[kEvalScanThreads];
    int const t = threadIdx.x;
    int const chunk = (a.n_bus + kEvalScanThreads - 1) / kEvalScanThreads;
    int const b0 = min(t * chunk, a.n_bus), b1 = min(b0 + chunk, a.n_bus);
    int c = 0;
    for (int i = b0; i < b1; ++i) {
        int const ty = a.bus_type[i];
        c += ty == 2 ? 0 : (ty == 0 ? 2 : 1);  // SLACK none, PQ P and Q, PV P (types.h:30-34)
    }
    cnt[t] = c;
    __syncthreads();
    // inclusive scan over the thread counts (Hillis-Steele in LDS)
    for (int off = 1; off < kEvalScanThreads; off <<= 1) {
        int const add = t >= off ? cnt[t - off] : 0;
        __syncthreads();
        cnt[t] += add;
        __syncthreads();
    }
    int row = cnt[t] - c;
    for (int i = b0; i < b1; ++i) {
        int const ty = a.bus_type[i];
        a.peq[i] = ty == 2 ? -1 : row;
        row += ty == 2 ? 0 : 1;
        a.qeq[i] = ty == 0 ? row : -1;
        row += ty == 0 ? 1 : 0;
    }
    if (t == 0) a.err[0] = 0;
}

__global__ void __launch_bounds__(256) k_eval_mismatch(EvalArgs a) {
#pragma clang fp contract(off)
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_bus) return;
    double cr = 0.0, cim = 0.0;
    if (i < a.y_rows && a.y_nnz > 0) {
        for (int q = a.rp[i]; q < a.rp[i + 1] && q < a.y_nnz; ++q) {
            if (q < 0) {
                a.err[0] = 1;
                break;
            }
            int const k = a.ci[q];
            if (k < 0) {
                a.err[0] = 1;
                break;
            }
            if (k >= a.n_bus) continue;
            double2 const yv = a.y[q];
            double2 const vv = a.v[k];
            double const pr = yv.x * vv.x - yv.y * vv.y;
            double const pi = yv.x * vv.y + yv.y * vv.x;
            cr += pr;
            cim += pi;
        }
    }
    double2 const vi = a.v[i];
    double const c = cr, d = -cim;
    double const sr = vi.x * c - vi.y * d;
    double const si = vi.x * d + vi.y * c;
    double pr = 0.0, pim = 0.0;
    if (a.p_w[i] != 0.0 || a.q_var[i] != 0.0) {
        pr = a.p_w[i];
        pim = a.q_var[i];
    }
    int const id = a.bus_id[i];
    for (int k = 0; k < a.n_app; ++k)
        if (a.app_node[k] == id && a.app_status[k] == 1 && (a.app_type[k] == 0 || a.app_type[k] == 1)) {
            pr += a.app_p[k];
            pim += a.app_q[k];
        }
    pr = pr / a.base;
    pim = pim / a.base;
    if (a.peq[i] >= 0) a.f[a.peq[i]] = sr - pr;
    if (a.qeq[i] >= 0) a.f[a.qeq[i]] = si - pim;
}

// Byte offsets of one evaluation's region (each array 256-byte aligned).
struct EvalLayout {
    size_t rp, ci, y, v, bus_id, bus_type, p_w, q_var, app_node, app_status, app_type, app_p, app_q, peq, qeq, f,
        err, total;
    static constexpr size_t up(size_t x) { return (x + 255) & ~size_t(255); }
    constexpr EvalLayout(long long n_bus, long long rows, long long nnz, long long n_app)
        : rp(0), ci(0), y(0), v(0), bus_id(0), bus_type(0), p_w(0), q_var(0), app_node(0), app_status(0),
          app_type(0), app_p(0), app_q(0), peq(0), qeq(0), f(0), err(0), total(0) {
        size_t o = 0;
        auto take = [&](size_t bytes) {
            size_t const at = o;
            o += up(bytes > 0 ? bytes : 1);
            return at;
        };
        size_t const nb = static_cast<size_t>(n_bus), na = static_cast<size_t>(n_app);
        rp = take(sizeof(int) * (static_cast<size_t>(rows) + 1));
        ci = take(sizeof(int) * static_cast<size_t>(nnz));
        y = take(sizeof(double2) * static_cast<size_t>(nnz));
        v = take(sizeof(double2) * nb);
        bus_id = take(sizeof(int) * nb);
        bus_type = take(sizeof(int) * nb);
        p_w = take(sizeof(double) * nb);
        q_var = take(sizeof(double) * nb);
        app_node = take(sizeof(int) * na);
        app_status = take(sizeof(int) * na);
        app_type = take(sizeof(int) * na);
        app_p = take(sizeof(double) * na);
        app_q = take(sizeof(double) * na);
        peq = take(sizeof(int) * nb);
        qeq = take(sizeof(int) * nb);
        f = take(sizeof(double) * 2 * nb);
        err = take(sizeof(int));
        total = o;
    }
};

}  // namespace gaphip
