// HIP kernels of the gfx950 Newton-Raphson power-flow path.
//
// All instances of a batch are packed end to end (buses, Y slots, Newton
// variables, Jacobian slots, fronts), so one launch serves every instance;
// per-instance state (converged flag, failed factorisation) gates the work.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gaphip {

constexpr int kWave = 64;

// ------------------------------------------------------------------ Y-bus
// One thread per CSR slot walks its contribution list in the reference order
// (shunts, then branches; cpu_admittance_matrix.cpp:86-154).  Contraction is
// off so r*r + x*x and the complex sums round exactly like the host code.
struct YbusArgs {
    int nnz;
    const int* seg;
    const int* code;
    const int* slot_inst;
    const double* br_r;
    const double* br_x;
    const double* br_g;
    const double* br_b;
    const double* app_g;
    const double* app_b;
    const double* zb;  // per instance, (u_rated[0]^2)/S_base (cpu_newton_raphson.cpp:46-58)
    double2* y_pu;
    double2* y_si;     // optional: siemens values
};

__global__ void __launch_bounds__(256) k_ybus(YbusArgs a) {
#pragma clang fp contract(off)
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nnz) return;
    double re = 0.0, im = 0.0;
    for (int s = a.seg[k]; s < a.seg[k + 1]; ++s) {
        int c = a.code[s];
        int kind = c & 3, idx = c >> 2;
        if (kind == 0) {
            re += a.app_g[idx];
            im += a.app_b[idx];
        } else {
            double r = a.br_r[idx], x = a.br_x[idx];
            double z2 = r * r + x * x;
            double yr = 0.0, yi = 0.0;
            if (z2 > 1e-12) {
                yr = r / z2;
                yi = -x / z2;
            }
            if (kind == 2) {
                re = -yr;
                im = -yi;
            } else {
                double bar = yr + a.br_g[idx], bai = yi + 0.0;
                double tr = bar + 0.0, ti = bai + a.br_b[idx] / 2.0;
                re += tr;
                im += ti;
            }
        }
    }
    if (a.y_si) a.y_si[k] = make_double2(re, im);
    double zb = a.zb[a.slot_inst[k]];
    a.y_pu[k] = make_double2(re * zb, im * zb);
}

// Caller-supplied Y-bus: y_pu = given * scale (scale = Z_base, or 1 when the
// values are already per unit), slots [begin, end).
__global__ void __launch_bounds__(256) k_yscale(int begin, int end, const double2* given, double scale,
                                                double2* y_pu) {
    int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    double2 g = given[k];
    y_pu[k] = make_double2(g.x * scale, g.y * scale);
}

// ------------------------------------------------------------ flat start
struct FlatArgs {
    int n_bus;
    const double* flat;  // u_pu for slack / PV, 0.98 for PQ (cpu_newton_raphson.cpp:77-88)
    double2* v;
};
__global__ void __launch_bounds__(256) k_flat(FlatArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_bus) a.v[i] = make_double2(a.flat[i], 0.0);
}

// --------------------------------------- SpMV + power + mismatch + max|F|
// I = Y V, S = V conj(I) (cpu_newton_raphson.cpp:296-337), F = S - S_spec on
// the P / Q equation rows (:269-289) and the per-instance max |F| (:115-120),
// reduced per wavefront and folded with a 64-bit atomic max on the IEEE bits
// (non-negative doubles order like their bit patterns).
struct MisArgs {
    int n_bus;
    const int* rp;
    const int* ci;
    const double2* y;
    const double2* v;
    const double2* sspec;
    const int* peq;
    const int* qeq;
    const int* bus_inst;
    const int* done;
    double2* s;
    double2* vpol;  // (|V|, arg V) for the Jacobian
    double* f;
    unsigned long long* maxf;  // this iteration's row of per-instance maxima
};

__device__ inline unsigned long long dbits(double x) { return __double_as_longlong(x); }

__global__ void __launch_bounds__(256) k_mismatch(MisArgs a) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = i < a.n_bus;
    int inst = live ? a.bus_inst[i] : -1;
    if (live && a.done[inst]) live = false;
    double m = 0.0;
    if (live) {
        double cr = 0.0, cim = 0.0;
        for (int q = a.rp[i]; q < a.rp[i + 1]; ++q) {
            double2 yv = a.y[q];
            double2 vv = a.v[a.ci[q]];
            double pr = yv.x * vv.x - yv.y * vv.y;
            double pi = yv.x * vv.y + yv.y * vv.x;
            cr += pr;
            cim += pi;
        }
        double2 vi = a.v[i];
        double c = cr, d = -cim;
        double sr = vi.x * c - vi.y * d;
        double si = vi.x * d + vi.y * c;
        a.s[i] = make_double2(sr, si);
        a.vpol[i] = make_double2(hypot(vi.x, vi.y), atan2(vi.y, vi.x));
        double2 sp = a.sspec[i];
        int pe = a.peq[i], qe = a.qeq[i];
        if (pe >= 0) {
            double fp = sr - sp.x;
            a.f[pe] = fp;
            m = fabs(fp);
        }
        if (qe >= 0) {
            double fq = si - sp.y;
            a.f[qe] = fq;
            m = fmax(m, fabs(fq));
        }
    }
    // wave-uniform instance: reduce in registers, one atomic per wave
    int lane = threadIdx.x & (kWave - 1);
    int inst0 = __shfl(inst, 0);
    bool uniform = __all(inst == inst0);
    if (uniform) {
        for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if (lane == 0 && inst0 >= 0 && m > 0.0) atomicMax(&a.maxf[inst0], dbits(m));
    } else if (live && m > 0.0) {
        atomicMax(&a.maxf[inst], dbits(m));
    }
}

// ------------------------------------------------- convergence bookkeeping
struct CheckArgs {
    int n_inst;
    int iter;
    double tol;
    const unsigned long long* maxf;  // this iteration
    int* done;
    int* iters;
    int* lu_fail;
    double* final_mismatch;
    int* active;  // this iteration's counter of instances still iterating
};
__global__ void __launch_bounds__(256) k_check(CheckArgs a) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n_inst) return;
    if (a.done[g]) return;
    double m = __longlong_as_double(static_cast<long long>(a.maxf[g]));
    a.final_mismatch[g] = m;
    a.lu_fail[g] = 0;
    if (m < a.tol) {
        a.done[g] = 1;
        a.iters[g] = a.iter + 1;
    } else {
        atomicAdd(a.active, 1);
    }
}

// --------------------------------------------------------- Jacobian fill
// One thread per unique Y pair (i, j) with i non-slack: the 2x2 block of
// dP/dtheta, dP/d|V|, dQ/dtheta, dQ/d|V| from one sincos, formulas of
// cpu_newton_raphson.cpp:546-557 / 640-725 (G, B from the last duplicate slot).
struct JacArgs {
    int n_items;
    const int* it_i;
    const int* it_j;
    const int* it_y;
    const int* it_pos;  // 4 per item
    const int* bus_inst;
    const int* done;
    const double2* y;
    const double2* s;
    const double2* vpol;
};

template <typename T>
__global__ void __launch_bounds__(256) k_jacobian(JacArgs a, T* jv) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_items) return;
    int i = a.it_i[t], j = a.it_j[t];
    if (a.done[a.bus_inst[i]]) return;
    double2 yv = a.y[a.it_y[t]];
    double G = yv.x, B = yv.y;
    double2 pi = a.vpol[i];
    double Vi = pi.x < 1e-6 ? 1.0 : pi.x;
    double pt, pv, qt, qv;
    if (i == j) {
        double2 si = a.s[i];
        double P = si.x, Q = si.y;
        pt = -Q - B * Vi * Vi;
        pv = (P / Vi) + G * Vi;
        qt = P - G * Vi * Vi;
        qv = (Q / Vi) - B * Vi;
    } else {
        double2 pj = a.vpol[j];
        double Vj = pj.x < 1e-6 ? 1.0 : pj.x;
        double th = pi.y - pj.y;
        double sn, cs;
        sincos(th, &sn, &cs);
        pt = Vi * Vj * (G * sn - B * cs);
        pv = Vi * (G * cs + B * sn);
        qt = -Vi * Vj * (G * cs + B * sn);
        qv = Vi * (G * sn - B * cs);
    }
    const int* p4 = a.it_pos + 4 * t;
    if (p4[0] >= 0) jv[p4[0]] = static_cast<T>(pt);
    if (p4[1] >= 0) jv[p4[1]] = static_cast<T>(pv);
    if (p4[2] >= 0) jv[p4[2]] = static_cast<T>(qt);
    if (p4[3] >= 0) jv[p4[3]] = static_cast<T>(qv);
}

// ------------------------------------------------------- multifrontal LU
// A front is a dense m x m row-major block: pivots first (p of them), then the
// structure slots shared with ancestors.  Factorisation leaves L (unit,
// below the diagonal) and U in the first p rows / columns and the Schur
// complement in the trailing block, which the parent reads during assembly.
// Arrowhead targets are packed as (row << 16) | col.
struct FrontArgs {
    const int* level_front;
    const int* f_m;
    const int* f_p;
    const long long* f_off;
    const long long* f_voff;
    const int* f_slot_ptr;
    const int* f_child_ptr;
    const int* child;
    const int* f_rel_off;
    const int* rel;
    const int* f_arrow_ptr;
    const int* arrow_src;
    const int* arrow_dst;
    const int* f_inst;
    const int* done;  // may be null
    int* fail;        // per instance; may be null
    const int* slot_row;  // per slot: index into the right-hand side
    const int* slot_col;  // per slot: index into the solution
};

template <typename T>
__device__ inline T tabs(T x) {
    return x < T(0) ? -x : x;
}

// Fronts of up to MM slots, one workgroup (NT threads) per front, the whole
// front staged in LDS (leading dimension MM + 1).  Thread t owns column t
// (t < MM), so every pivot step is one pass and one barrier.
template <typename T, int MM, int NT>
__global__ void __launch_bounds__(NT) k_front_factor_lds(FrontArgs a, int begin, const T* jv, T* front) {
    static_assert(NT >= MM, "one thread per column");
    constexpr int LD = MM + 1;
    __shared__ T F[MM * LD];
    int const s = a.level_front[begin + blockIdx.x];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    int const t = threadIdx.x;
    for (int e = t; e < m * LD; e += NT) F[e] = T(0);
    __syncthreads();
    for (int q = a.f_arrow_ptr[s] + t; q < a.f_arrow_ptr[s + 1]; q += NT) {
        int d = a.arrow_dst[q];
        F[(d >> 16) * LD + (d & 0xffff)] = jv[a.arrow_src[q]];
    }
    __syncthreads();
    for (int cq = a.f_child_ptr[s]; cq < a.f_child_ptr[s + 1]; ++cq) {
        int const c = a.child[cq];
        int const mc = a.f_m[c], pc = a.f_p[c], nu = mc - pc;
        const T* Fc = front + a.f_off[c] + static_cast<long long>(pc) * mc + pc;
        const int* rl = a.rel + a.f_rel_off[c];
        if (t < nu) {
            int const col = rl[t];
            for (int r = 0; r < nu; ++r) F[rl[r] * LD + col] += Fc[r * mc + t];
        }
        __syncthreads();
    }
    // right-looking elimination of the p pivots; column k is scaled by its
    // owner after the barrier that closes step k (no later step reads it)
    for (int k = 0; k < p; ++k) {
        T piv = F[k * LD + k];
        if (tabs(piv) < T(1e-14)) {
            if (t == 0 && a.fail) a.fail[inst] = 1;
            piv = T(1);
        }
        T const rp = T(1) / piv;
        if (t > k && t < m) {
            T const ukj = F[k * LD + t];
            for (int i = k + 1; i < m; ++i) F[i * LD + t] -= (F[i * LD + k] * rp) * ukj;
        }
        __syncthreads();
        if (t == k)
            for (int i = k + 1; i < m; ++i) F[i * LD + k] *= rp;
    }
    __syncthreads();
    T* G = front + a.f_off[s];
    for (int r = 0; r < m; ++r)
        if (t < m) G[r * m + t] = F[r * LD + t];
}

// ---------------------------------------------------- large fronts (> 128)
// Blocked right-looking partial LU in global memory, many workgroups per front:
// assembly by rows, then per panel of NB pivot columns a panel factorisation,
// a triangular solve for the U12 block and a trailing GEMM update.  All fronts
// of one (level, class) group advance together; a front whose panels are done
// simply exits.  Per-row maps (host-built): arrowhead entries sorted by row
// (arow) and the child update rows that land on each parent row (crow_c /
// crow_a), listed in child order so sums match the small-front kernels.
struct BigMaps {
    const int* bigbase;  // per front: offset of its m+1 row pointers
    const int* arow;     // row r of front f: arrowhead entries [arow[b+r], arow[b+r+1]) after f_arrow_ptr[f]
    const int* crow;     // row r: child rows [crow[b+r], crow[b+r+1])
    const int* crow_c;   // child front of each child row
    const int* crow_a;   // row index inside that child's update block
};

template <typename T>
__global__ void __launch_bounds__(256) k_big_assemble(FrontArgs a, BigMaps b, int begin, const T* jv, T* front) {
    int const s = a.level_front[begin + blockIdx.y];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s];
    int const r = blockIdx.x;
    if (r >= m) return;
    int const tid = threadIdx.x;
    T* row = front + a.f_off[s] + static_cast<long long>(r) * m;
    for (int c = tid; c < m; c += 256) row[c] = T(0);
    __syncthreads();
    int const base = b.bigbase[s];
    int const a0 = a.f_arrow_ptr[s];
    for (int q = a0 + b.arow[base + r] + tid; q < a0 + b.arow[base + r + 1]; q += 256)
        row[a.arrow_dst[q] & 0xffff] = jv[a.arrow_src[q]];
    __syncthreads();
    for (int e = b.crow[base + r]; e < b.crow[base + r + 1]; ++e) {
        int const c = b.crow_c[e], ra = b.crow_a[e];
        int const mc = a.f_m[c], pc = a.f_p[c], nu = mc - pc;
        const T* src = front + a.f_off[c] + static_cast<long long>(pc + ra) * mc + pc;
        const int* rl = a.rel + a.f_rel_off[c];
        for (int t = tid; t < nu; t += 256) row[rl[t]] += src[t];
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_big_panel(FrontArgs a, int begin, int k0, int nb, T* front) {
    int const s = a.level_front[begin + blockIdx.x];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    if (k0 >= p) return;
    int const kend = min(k0 + nb, p);
    int const tid = threadIdx.x;
    T* F = front + a.f_off[s];
    for (int k = k0; k < kend; ++k) {
        T piv = F[static_cast<long long>(k) * m + k];
        if (tabs(piv) < T(1e-14)) {
            if (tid == 0 && a.fail) a.fail[inst] = 1;
            piv = T(1);
        }
        T const rp = T(1) / piv;
        const T* uk = F + static_cast<long long>(k) * m;
        for (int i = k + 1 + tid; i < m; i += 256) {
            T* ri = F + static_cast<long long>(i) * m;
            T const l = ri[k] * rp;
            ri[k] = l;
            for (int j = k + 1; j < kend; ++j) ri[j] -= l * uk[j];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_big_trsm(FrontArgs a, int begin, int k0, int nb, T* front) {
    int const s = a.level_front[begin + blockIdx.y];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    if (k0 >= p) return;
    int const kend = min(k0 + nb, p);
    int const j = kend + blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    T* F = front + a.f_off[s];
    for (int r = k0 + 1; r < kend; ++r) {
        T acc = F[static_cast<long long>(r) * m + j];
        const T* lr = F + static_cast<long long>(r) * m;
        for (int q = k0; q < r; ++q) acc -= lr[q] * F[static_cast<long long>(q) * m + j];
        F[static_cast<long long>(r) * m + j] = acc;
    }
}

// C[kend:m, kend:m] -= L[kend:m, k0:kend] * U[k0:kend, kend:m]; 64x64 tiles,
// 256 threads with a 4x4 register tile each, the K slice (<= 32) staged in LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_big_gemm(FrontArgs a, int begin, int k0, int nb, T* front) {
    constexpr int TB = 64, KB = 32;
    __shared__ T As[KB][TB + 1];
    __shared__ T Bs[KB][TB + 1];
    int const s = a.level_front[begin + blockIdx.z];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    if (k0 >= p) return;
    int const kend = min(k0 + nb, p);
    int const kb = kend - k0;
    int const i0 = kend + blockIdx.y * TB, j0 = kend + blockIdx.x * TB;
    if (i0 >= m || j0 >= m) return;
    T* F = front + a.f_off[s];
    int const tid = threadIdx.x;
    for (int e = tid; e < KB * TB; e += 256) {
        int const kk = e / TB, x = e % TB;
        int const gi = i0 + x, gj = j0 + x;
        As[kk][x] = (kk < kb && gi < m) ? F[static_cast<long long>(gi) * m + k0 + kk] : T(0);
        Bs[kk][x] = (kk < kb && gj < m) ? F[static_cast<long long>(k0 + kk) * m + gj] : T(0);
    }
    __syncthreads();
    int const tr = (tid / 16) * 4, tc = (tid % 16) * 4;
    T acc[4][4] = {};
    for (int kk = 0; kk < kb; ++kk) {
        T av[4], bv[4];
        for (int u = 0; u < 4; ++u) av[u] = As[kk][tr + u];
        for (int u = 0; u < 4; ++u) bv[u] = Bs[kk][tc + u];
        for (int u = 0; u < 4; ++u)
            for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
    }
    for (int u = 0; u < 4; ++u) {
        int const gi = i0 + tr + u;
        if (gi >= m) continue;
        T* ci = F + static_cast<long long>(gi) * m;
        for (int v = 0; v < 4; ++v) {
            int const gj = j0 + tc + v;
            if (gj < m) ci[gj] -= acc[u][v];
        }
    }
}

// Forward solve L y = b, front by front from the leaves.  Slot vector w (m):
// y of the pivots in [0, p) and the update passed to the parent in [p, m).
template <typename T>
__global__ void __launch_bounds__(256) k_front_fwd(FrontArgs a, int begin, const T* front, T* vec,
                                                   const double* rhs, double rhs_sign) {
    int const s = a.level_front[begin + blockIdx.x];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    int const tid = threadIdx.x, bs = blockDim.x;
    const T* F = front + a.f_off[s];
    T* w = vec + a.f_voff[s];
    const int* rows = a.slot_row + a.f_slot_ptr[s];
    for (int t = tid; t < m; t += bs) w[t] = t < p ? static_cast<T>(rhs_sign * rhs[rows[t]]) : T(0);
    __syncthreads();
    for (int cq = a.f_child_ptr[s]; cq < a.f_child_ptr[s + 1]; ++cq) {
        int const c = a.child[cq];
        int const nu = a.f_m[c] - a.f_p[c];
        const T* wc = vec + a.f_voff[c] + a.f_p[c];
        const int* rl = a.rel + a.f_rel_off[c];
        for (int r = tid; r < nu; r += bs) w[rl[r]] += wc[r];
        __syncthreads();
    }
    for (int k = 0; k < p; ++k) {
        T const xk = w[k];
        for (int i = k + 1 + tid; i < m; i += bs) w[i] -= F[static_cast<long long>(i) * m + k] * xk;
        __syncthreads();
    }
}

// Backward solve U x = y, front by front from the roots; structure values come
// from the (already solved) ancestors in the global solution vector.
template <typename T>
__global__ void __launch_bounds__(256) k_front_bwd(FrontArgs a, int begin, const T* front, T* vec,
                                                   double* sol) {
    int const s = a.level_front[begin + blockIdx.x];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    int const tid = threadIdx.x, bs = blockDim.x;
    const T* F = front + a.f_off[s];
    T* w = vec + a.f_voff[s];
    const int* cols = a.slot_col + a.f_slot_ptr[s];
    for (int t = p + tid; t < m; t += bs) w[t] = static_cast<T>(sol[cols[t]]);
    __syncthreads();
    for (int k = tid; k < p; k += bs) {
        T acc = w[k];
        const T* row = F + static_cast<long long>(k) * m;
        for (int t = p; t < m; ++t) acc -= row[t] * w[t];
        w[k] = acc;
    }
    __syncthreads();
    // w[k] is only read in step k (every thread divides it itself), so the
    // result goes straight to the solution vector and w[k] stays untouched.
    for (int k = p - 1; k >= 0; --k) {
        T const xk = w[k] / F[static_cast<long long>(k) * m + k];
        for (int i = tid; i < k; i += bs) w[i] -= F[static_cast<long long>(i) * m + k] * xk;
        if (tid == 0) sol[cols[k]] = static_cast<double>(xk);
        __syncthreads();
    }
}

// Small-front forward / backward solves: one workgroup per front with the
// front staged in LDS.
template <typename T, int MM, int NT>
__global__ void __launch_bounds__(NT) k_front_fwd_lds(FrontArgs a, int begin, const T* front, T* vec,
                                                      const double* rhs, double rhs_sign) {
    constexpr int LD = MM + 1;
    __shared__ T F[MM * LD];
    __shared__ T w[MM];
    int const s = a.level_front[begin + blockIdx.x];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    int const t = threadIdx.x;
    const T* G = front + a.f_off[s];
    for (int r = 0; r < m; ++r)
        if (t < m) F[r * LD + t] = G[r * m + t];
    const int* rows = a.slot_row + a.f_slot_ptr[s];
    if (t < m) w[t] = t < p ? static_cast<T>(rhs_sign * rhs[rows[t]]) : T(0);
    __syncthreads();
    for (int cq = a.f_child_ptr[s]; cq < a.f_child_ptr[s + 1]; ++cq) {
        int const c = a.child[cq];
        int const nu = a.f_m[c] - a.f_p[c];
        const T* wc = vec + a.f_voff[c] + a.f_p[c];
        const int* rl = a.rel + a.f_rel_off[c];
        if (t < nu) w[rl[t]] += wc[t];
        __syncthreads();
    }
    for (int k = 0; k < p; ++k) {
        T const xk = w[k];
        if (t > k && t < m) w[t] -= F[t * LD + k] * xk;
        __syncthreads();
    }
    T* wout = vec + a.f_voff[s];
    if (t < m) wout[t] = w[t];
}

template <typename T, int MM, int NT>
__global__ void __launch_bounds__(NT) k_front_bwd_lds(FrontArgs a, int begin, const T* front, T* vec,
                                                      double* sol) {
    constexpr int LD = MM + 1;
    __shared__ T F[MM * LD];
    __shared__ T w[MM];
    int const s = a.level_front[begin + blockIdx.x];
    int const inst = a.f_inst[s];
    if (a.done && a.done[inst]) return;
    int const m = a.f_m[s], p = a.f_p[s];
    int const t = threadIdx.x;
    const T* G = front + a.f_off[s];
    for (int r = 0; r < p; ++r)
        if (t < m) F[r * LD + t] = G[r * m + t];
    const int* cols = a.slot_col + a.f_slot_ptr[s];
    const T* win = vec + a.f_voff[s];
    if (t < m) w[t] = t < p ? win[t] : static_cast<T>(sol[cols[t]]);
    __syncthreads();
    if (t < p) {
        T acc = w[t];
        for (int q = p; q < m; ++q) acc -= F[t * LD + q] * w[q];
        w[t] = acc;
    }
    __syncthreads();
    for (int k = p - 1; k >= 0; --k) {
        T const xk = w[k] / F[k * LD + k];
        if (t < k) w[t] -= F[t * LD + k] * xk;
        if (t == k) sol[cols[k]] = static_cast<double>(xk);
        __syncthreads();
    }
}

// ------------------------------------------------------------ update
// Damped, clamped polar update of cpu_newton_raphson.cpp:438-504, one thread
// per bus: angle step (0.9, |d| <= 0.15) for every bus but the first slack,
// then magnitude step (0.9, |d| <= 0.08, |V| in [0.5, 1.5]) for PQ buses.
struct UpdArgs {
    int n_bus;
    const int* bus_inst;
    const int* done;
    const int* fail;
    const int* ang;
    const int* mag;
    const double* dx;
    double2* v;
};
__global__ void __launch_bounds__(256) k_update(UpdArgs a) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_bus) return;
    int inst = a.bus_inst[i];
    if (a.done[inst]) return;
    bool zero = a.fail[inst] != 0;
    double2 v = a.v[i];
    int av = a.ang[i], mv = a.mag[i];
    if (av >= 0) {
        double vm = hypot(v.x, v.y);
        double ac = 0.9 * (zero ? 0.0 : a.dx[av]);
        if (fabs(ac) > 0.15) ac = 0.15 * (ac > 0 ? 1.0 : -1.0);
        double va = atan2(v.y, v.x) + ac;
        v = make_double2(vm * cos(va), vm * sin(va));
    }
    if (mv >= 0) {
        double old = hypot(v.x, v.y);
        double mc = 0.9 * (zero ? 0.0 : a.dx[mv]);
        if (fabs(mc) > 0.08) mc = 0.08 * (mc > 0 ? 1.0 : -1.0);
        double vm = old + mc;
        double va = atan2(v.y, v.x);
        vm = fmax(vm, 0.5);
        vm = fmin(vm, 1.5);
        v = make_double2(vm * cos(va), vm * sin(va));
    }
    a.v[i] = v;
}

}  // namespace gaphip
