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

// The reference's max|F| (cpu_newton_raphson.cpp:115-119) is max_element
// under `|a| < |b|`: a NaN in the first mismatch row is kept (no comparison
// with it is true), a NaN in any later row is never taken.  The kernels reduce
// with fmax, which skips NaN, and then apply the first row: `max_nonnan` is
// the NaN-free maximum, `f0` the instance's first mismatch row.  A NaN result
// fails `max < tol`, so such a network is not reported converged.
__device__ inline double reference_max(double max_nonnan, double f0) { return f0 != f0 ? fabs(f0) : max_nonnan; }

// |V| clamped to [0.5, 1.5] as std::max(v, 0.5) then std::min(v, 1.5)
// (cpu_newton_raphson.cpp:493-496): each returns its first argument unless
// the comparison holds, so a NaN magnitude stays NaN (fmax / fmin would
// replace it by the bound).
__device__ inline double ref_clamp_magnitude(double vm, double hi = 1.5) {
    vm = (vm < 0.5) ? 0.5 : vm;
    return (hi < vm) ? hi : vm;
}

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

// One Y slot: its contributions in the builder's order (segment k of code).
// The first kHead contributions' codes and parameters are all loaded before
// any is summed, so a bus row of degree d costs three dependent round trips
// instead of 1 + 2d; the sum itself runs in code order.
__device__ inline void ybus_slot(YbusArgs const& a, int k) {
#pragma clang fp contract(off)
    int const s0 = a.seg[k], s1 = a.seg[k + 1];
    int const inst = a.slot_inst[k];
    constexpr int kHead = 4;
    int c[kHead];
#pragma unroll
    for (int u = 0; u < kHead; ++u) c[u] = s0 + u < s1 ? a.code[s0 + u] : 0;
    // Per contribution: a shunt's (g, b), or a branch's (r, x, g, b).
    double4 par[kHead];
#pragma unroll
    for (int u = 0; u < kHead; ++u) {
        par[u] = make_double4(0.0, 0.0, 0.0, 0.0);
        if (s0 + u >= s1) continue;
        int const idx = c[u] >> 2;
        if ((c[u] & 3) == 0) {
            par[u].x = a.app_g[idx];
            par[u].y = a.app_b[idx];
        } else {
            par[u] = make_double4(a.br_r[idx], a.br_x[idx], a.br_g[idx], a.br_b[idx]);
        }
    }
    double re = 0.0, im = 0.0;
    auto add = [&](int code, double4 const& p) {
        int const kind = code & 3;
        if (kind == 0) {
            re += p.x;
            im += p.y;
        } else {
            double r = p.x, x = p.y;
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
                double bar = yr + p.z, bai = yi + 0.0;
                double tr = bar + 0.0, ti = bai + p.w / 2.0;
                re += tr;
                im += ti;
            }
        }
    };
#pragma unroll
    for (int u = 0; u < kHead; ++u)
        if (s0 + u < s1) add(c[u], par[u]);
    for (int s = s0 + kHead; s < s1; ++s) {
        int const cs = a.code[s], idx = cs >> 2;
        if ((cs & 3) == 0)
            add(cs, make_double4(a.app_g[idx], a.app_b[idx], 0.0, 0.0));
        else
            add(cs, make_double4(a.br_r[idx], a.br_x[idx], a.br_g[idx], a.br_b[idx]));
    }
    if (a.y_si) a.y_si[k] = make_double2(re, im);
    double zb = a.zb[inst];
    a.y_pu[k] = make_double2(re * zb, im * zb);
}

__global__ void __launch_bounds__(256) k_ybus(YbusArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.nnz) ybus_slot(a, k);
}

// Incremental update: new r / x of n branches, then the listed slots (those
// the branches touch) recomputed by the same per-slot routine, so the result
// equals a full assembly of the changed network bit for bit.
__global__ void __launch_bounds__(256) k_branch_rx_set(int n, const int* idx, const double* r, const double* x,
                                                       double* br_r, double* br_x) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    br_r[idx[t]] = r[t];
    br_x[idx[t]] = x[t];
}
__global__ void __launch_bounds__(256) k_ybus_slots(YbusArgs a, int n, const int* slots) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) ybus_slot(a, slots[t]);
}

// Caller-supplied Y-bus: y_pu = given * scale (scale = Z_base, or 1 when the
// values are already per unit), slots [begin, end).
__global__ void __launch_bounds__(256) k_yscale(int begin, int end, const double2* given, double scale,
                                                double2* y_pu, const int* slot_g, double2* y_g) {
    int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    double2 g = given[k];
    double2 const y = make_double2(g.x * scale, g.y * scale);
    if (y_pu) y_pu[k] = y;
    if (y_g) y_g[slot_g[k]] = y;  // the grid schedule's copy (its own slot order)
}

// The plan's Y from the grid schedule's copy (slot k is its slot slot_g[k]).
__global__ void __launch_bounds__(256) k_ygather(long long n, const int* slot_g, const double2* y_g, double2* y_pu) {
    long long const k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) y_pu[k] = y_g[slot_g[k]];
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
    // grid schedule only: the mismatch pass also writes each -F in the factor
    // type, per Newton equation (the factor tasks gather it), and the FP64 F
    // of each network's first equation only (the convergence test's first row)
    void* fe;
    // grid schedule with FP32 factors: FP32 copies of V and S, the inputs of
    // the FP32 Jacobian blocks, written by the mismatch pass; per bus V then
    // S (16 bytes), so a diagonal block's V and S share a cache line
    float4* vsf;
};

__device__ inline unsigned long long dbits(double x) { return __double_as_longlong(x); }

// per bus: S, (|V|, arg V) and the P / Q mismatch rows; returns this bus's max |F|
__device__ inline double mismatch_bus(MisArgs const& a, int i) {
#pragma clang fp contract(off)
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
    double m = 0.0;
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
    return m;
}

__global__ void __launch_bounds__(256) k_mismatch(MisArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = i < a.n_bus;
    int inst = live ? a.bus_inst[i] : -1;
    if (live && a.done[inst]) live = false;
    double m = live ? mismatch_bus(a, i) : 0.0;
    // wave-uniform instance: reduce in registers, one atomic per wave
    int lane = threadIdx.x & (kWave - 1);
    int inst0 = __shfl(inst, 0);
    bool uniform = __all(inst == inst0);
    if (uniform) {
        for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if (lane == 0 && inst0 >= 0 && m > 0.0) atomicMax(&a.maxf[inst0], dbits(m));
    } else if (live && m > 0.0) {
        // fabs clears the sign, so a NaN's bits lie above +inf's and win the max
        atomicMax(&a.maxf[inst], dbits(m));
    }
}

// ------------------------------------------------- convergence bookkeeping
struct CheckArgs {
    int n_inst;
    int iter;
    double tol;
    unsigned long long* maxf;  // this iteration (its reference maximum is written back)
    int* done;
    int* iters;
    int* lu_fail;
    double* final_mismatch;
    int* active;  // this iteration's counter of instances still iterating
    const double* f;        // this iteration's mismatch vector
    const int* var_off;     // per instance: its first mismatch row (n_inst + 1 entries)
};
__global__ void __launch_bounds__(256) k_check(CheckArgs a) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n_inst) return;
    if (a.done[g]) return;
    double m = __longlong_as_double(static_cast<long long>(a.maxf[g]));
    int const r0 = a.var_off[g];
    if (r0 < a.var_off[g + 1]) {
        m = reference_max(m, a.f[r0]);
        a.maxf[g] = dbits(m);  // the trajectory row holds the reference's maximum
    }
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

// dP/dtheta_j, dP/d|V_j|, dQ/dtheta_j, dQ/d|V_j| of bus i from its Y value yv,
// (|V|, arg V) of buses i and j, and S_i (used on the diagonal)
__device__ inline void jac_block_v(bool diag, double2 yv, double2 pi, double2 pj, double2 si, double& pt,
                                   double& pv, double& qt, double& qv) {
    double G = yv.x, B = yv.y;
    double Vi = pi.x < 1e-6 ? 1.0 : pi.x;
    if (diag) {
        double P = si.x, Q = si.y;
        pt = -Q - B * Vi * Vi;
        pv = (P / Vi) + G * Vi;
        qt = P - G * Vi * Vi;
        qv = (Q / Vi) - B * Vi;
    } else {
        double Vj = pj.x < 1e-6 ? 1.0 : pj.x;
        double th = pi.y - pj.y;
        double sn, cs;
        sincos(th, &sn, &cs);
        pt = Vi * Vj * (G * sn - B * cs);
        pv = Vi * (G * cs + B * sn);
        qt = -Vi * Vj * (G * cs + B * sn);
        qv = Vi * (G * sn - B * cs);
    }
}

__device__ inline void jac_block(const double2* y, const double2* s, const double2* vpol, int i, int j, int yi,
                                 double& pt, double& pv, double& qt, double& qv) {
    bool const diag = i == j;
    jac_block_v(diag, y[yi], vpol[i], diag ? make_double2(0.0, 0.0) : vpol[j], diag ? s[i] : make_double2(0.0, 0.0),
                pt, pv, qt, qv);
}

template <typename T>
__device__ inline void jacobian_item(JacArgs const& a, int t, T* jv) {
    double pt, pv, qt, qv;
    jac_block(a.y, a.s, a.vpol, a.it_i[t], a.it_j[t], a.it_y[t], pt, pv, qt, qv);
    const int* p4 = a.it_pos + 4 * t;
    if (p4[0] >= 0) jv[p4[0]] = static_cast<T>(pt);
    if (p4[1] >= 0) jv[p4[1]] = static_cast<T>(pv);
    if (p4[2] >= 0) jv[p4[2]] = static_cast<T>(qt);
    if (p4[3] >= 0) jv[p4[3]] = static_cast<T>(qv);
}

template <typename T>
__global__ void __launch_bounds__(256) k_jacobian(JacArgs a, T* jv) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_items) return;
    if (a.done[a.bus_inst[a.it_i[t]]]) return;
    jacobian_item<T>(a, t, jv);
}

// ------------------------------------------------------- multifrontal LU
// A front is a dense m x m row-major block: pivots first (p of them), then the
// structure slots shared with ancestors.  Factorisation leaves L (unit,
// below the diagonal) and U in the first p rows / columns and the Schur
// complement in the trailing block, which the parent reads during assembly.
// Arrowhead targets are packed as (row << 16) | col.  Fronts are numbered in
// (level, size class) order, so a launch covers ids [begin, begin + count).
struct alignas(16) FrontMeta {
    long long off;   // front storage offset (m*m entries)
    long long voff;  // slot-vector offset (m entries)
    int m, p;
    int child_beg, child_end;  // ChildInfo range
    int arrow_beg, arrow_end;  // arrowhead entries
    int slot_beg, inst;
};
struct alignas(16) ChildInfo {
    long long off, voff;  // the child's front / vector offsets
    int mc, pc;           // its size and pivot count
    int rel_off, id;      // its extend-add map into this front; its front id
};

struct FrontArgs {
    const FrontMeta* meta;
    const ChildInfo* cinfo;
    const int* rel;
    const int* arrow_src;
    const int* arrow_dst;
    const int* slot_row;  // per slot: index into the right-hand side
    const int* slot_col;  // per slot: index into the solution
    const int* done;      // per instance; may be null
    int* fail;            // per instance; may be null
};

template <typename T>
__device__ inline T tabs(T x) {
    return x < T(0) ? -x : x;
}

// all lanes of a group (L lanes) hold the same value; fold across the groups
// of the workgroup so loop trip counts are uniform and barriers legal
template <int L, int NT>
__device__ inline int group_max(int v) {
    for (int off = L; off < NT && off < kWave; off <<= 1) v = max(v, __shfl_xor(v, off));
    if constexpr (NT > kWave) {
        __shared__ int red[NT / kWave];
        int const w = threadIdx.x / kWave;
        if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
        __syncthreads();
        v = red[0];
        for (int k = 1; k < NT / kWave; ++k) v = max(v, red[k]);
        __syncthreads();
    }
    return v;
}

// rows [0, n) of a row-major global block (leading dimension ld) into the LDS
// tile F (leading dimension LD), lane t copying column t; CH independent loads
// are issued before their LDS stores
template <typename T, int MM, int L, int CH>
__device__ inline void load_rows(T* F, int LD, const T* G, int ld, int n, int t, bool active) {
    for (int r0 = 0; r0 < n; r0 += CH) {
        T v[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) v[u] = (active && r0 + u < n) ? G[static_cast<long long>(r0 + u) * ld + t] : T(0);
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (active && r0 + u < n) F[(r0 + u) * LD + t] = v[u];
    }
}

// F[rel[r], rel[t]] += child update block (nu x nu); the block rows are loaded
// CH at a time.  The map entry of row r comes from lane r of the group when a
// group fits in one wavefront (shuffles do not cross wavefronts), otherwise
// from memory.
template <typename T, int L, int CH>
__device__ inline void extend_add(T* F, int LD, const T* Fc, int mc, int nu, const int* rl, int t) {
    int const myrl = t < nu ? rl[t] : 0;
    for (int r0 = 0; r0 < nu; r0 += CH) {
        T v[CH];
        int rr[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            v[u] = (t < nu && r0 + u < nu) ? Fc[static_cast<long long>(r0 + u) * mc + t] : T(0);
            if constexpr (L <= kWave)
                rr[u] = __shfl(myrl, (r0 + u) & (L - 1), L);
            else
                rr[u] = (r0 + u < nu) ? rl[r0 + u] : 0;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (t < nu && r0 + u < nu) F[rr[u] * LD + myrl] += v[u];
    }
}

// Small fronts: G = NT / L fronts per workgroup, L lanes per front (L >= MM),
// each front staged in its own LDS tile (leading dimension MM + 1).  Lane t of a
// group owns column t.  Loop trip counts are made uniform over the workgroup
// (max over its fronts) so every barrier is reached by all threads; a front
// that is finished, inactive or converged only idles.
template <typename T, int MM, int L, int NT>
__global__ void __launch_bounds__(NT) k_front_factor_lds(FrontArgs a, int begin, int count, const T* jv,
                                                          T* front) {
    static_assert(L >= MM && NT % L == 0, "one lane per column");
    constexpr int G = NT / L, LD = MM + 1, CH = MM < 8 ? MM : 8;
    __shared__ T Fs[G][MM * LD];
    int const g = threadIdx.x / L, t = threadIdx.x % L;
    int const idx = blockIdx.x * G + g;
    FrontMeta fm{};
    if (idx < count) {
        fm = a.meta[begin + idx];
        if (a.done && a.done[fm.inst]) fm.m = fm.p = fm.child_end = fm.child_beg = fm.arrow_end = fm.arrow_beg = 0;
    }
    int const m = fm.m, p = fm.p;
    T* F = Fs[g];
    for (int e = t; e < m * LD; e += L) F[e] = T(0);
    __syncthreads();
    for (int q0 = fm.arrow_beg; q0 < fm.arrow_end; q0 += L) {
        int const q = q0 + t;
        if (q < fm.arrow_end) {
            int const d = a.arrow_dst[q];
            F[(d >> 16) * LD + (d & 0xffff)] = jv[a.arrow_src[q]];
        }
    }
    __syncthreads();
    int const nch = fm.child_end - fm.child_beg;
    int const nch_max = group_max<L, NT>(nch);
    for (int ci = 0; ci < nch_max; ++ci) {
        if (ci < nch) {
            ChildInfo const c = a.cinfo[fm.child_beg + ci];
            int const nu = c.mc - c.pc;
            extend_add<T, L, CH>(F, LD, front + c.off + static_cast<long long>(c.pc) * c.mc + c.pc, c.mc, nu,
                                 a.rel + c.rel_off, t);
        }
        __syncthreads();
    }
    // right-looking elimination; column k is scaled by its owner after the
    // barrier that closes step k (no later step reads it)
    int const p_max = group_max<L, NT>(p);
    for (int k = 0; k < p_max; ++k) {
        T piv = k < p ? F[k * LD + k] : T(1);
        bool const tiny = tabs(piv) < T(1e-14);
        if (tiny) {
            if (k < p && t == 0 && a.fail) a.fail[fm.inst] = 1;
            piv = T(1);
        }
        T const rp = T(1) / piv;
        if (k < p && t > k && t < m) {
            T const ukj = F[k * LD + t];
            // column k is read, column t > k written: they never overlap, so
            // the reads of later rows need not wait for earlier rows' writes
            T const* __restrict__ ck = F + k;
            T* __restrict__ ct = F + t;
#pragma unroll 4
            for (int i = k + 1; i < m; ++i) ct[i * LD] -= (ck[i * LD] * rp) * ukj;
        }
        __syncthreads();
        if (k < p && t == k)
            for (int i = k + 1; i < m; ++i) F[i * LD + k] *= rp;
    }
    __syncthreads();
    if (m > 0 && t < m) {
        T* Gm = front + fm.off;
        for (int r = 0; r < m; ++r) Gm[r * m + t] = F[r * LD + t];
    }
}

// ---------------------------------------------------- large fronts (> 128)
// Blocked right-looking partial LU in global memory, many workgroups per front:
// assembly by rows, then per panel of NB pivot columns a panel factorisation,
// a triangular solve for the U12 block and a trailing GEMM update.  All fronts
// of one (level, class) group advance together; a front whose panels are done
// simply exits.  Per-row maps (host-built): arrowhead entries sorted by row
// (arow) and the child update rows that land on each parent row (crow_ci as
// (child info index, child row) pairs), listed in child order so sums match
// the small-front kernels.
struct BigMaps {
    const int* bigbase;  // per front: offset of its m+1 row pointers
    const int* arow;     // row r of front f: arrowhead entries [arow[b+r], arow[b+r+1]) after arrow_beg
    const int* crow;     // row r: child rows [crow[b+r], crow[b+r+1])
    const int* crow_ci;  // pairs (child info index, row inside that child's update block)
};

template <typename T>
__global__ void __launch_bounds__(256) k_big_assemble(FrontArgs a, BigMaps b, int begin, const T* jv, T* front) {
    FrontMeta const fm = a.meta[begin + blockIdx.y];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m;
    int const r = blockIdx.x;
    if (r >= m) return;
    int const tid = threadIdx.x;
    T* row = front + fm.off + static_cast<long long>(r) * m;
    for (int c = tid; c < m; c += 256) row[c] = T(0);
    __syncthreads();
    int const base = b.bigbase[begin + blockIdx.y];
    for (int q = fm.arrow_beg + b.arow[base + r] + tid; q < fm.arrow_beg + b.arow[base + r + 1]; q += 256)
        row[a.arrow_dst[q] & 0xffff] = jv[a.arrow_src[q]];
    __syncthreads();
    for (int e = b.crow[base + r]; e < b.crow[base + r + 1]; ++e) {
        ChildInfo const c = a.cinfo[b.crow_ci[2 * e]];
        int const ra = b.crow_ci[2 * e + 1];
        int const nu = c.mc - c.pc;
        const T* src = front + c.off + static_cast<long long>(c.pc + ra) * c.mc + c.pc;
        const int* rl = a.rel + c.rel_off;
        for (int t = tid; t < nu; t += 256) row[rl[t]] += src[t];
        __syncthreads();
    }
}

// Two-level blocked LU of fronts above 128 slots: outer panels of kBigOuter
// pivots (groups whose largest front has at least kTwoLevelMin slots; else
// outer = inner = kBigPanel, the plain right-looking scheme), each factored
// as inner panels of kBigPanel.  Per inner panel: the
// diagonal block A11 (one wavefront per front, multipliers stored below its
// diagonal; k_big_diag), L21 = A21 U11^-1 for every row below
// (k_big_trsm_rows), U12 = L11^-1 A12 for the columns inside the outer panel
// only (k_big_trsm) and their MFMA update (k_big_gemm with a column limit).
// Then the outer panel's rows are solved for the columns to its right block
// by block (k_big_gemm_rows, k_big_trsm), and the trailing block takes one
// MFMA update with K = kBigOuter (k_big_gemm), so the trailing matrix is read
// and written once per outer panel instead of once per inner panel.
constexpr int kBigPanel = 64, kBigOuter = 256, kTwoLevelMin = 4096;

template <typename T>
__device__ inline T pivot_recip(T piv, int* fail, int inst, bool flag) {
    if (tabs(piv) < T(1e-14)) {
        if (flag && fail) fail[inst] = 1;
        piv = T(1);
    }
    return T(1) / piv;
}

// The diagonal block of one front per workgroup of one wavefront: lane t
// holds row t of the block in registers.  At step k lane k's row is final;
// every lane reads it entry by entry from lane k's registers (readlane into a
// scalar register: no LDS round trip and no barrier per step), takes its
// multiplier l_t = a_tk / a_kk from its own row and updates it:
// a_tj -= l_t u_kj, the same formula per entry as the LDS kernels.  (Through
// LDS, a single-lane write, a wave barrier and a broadcast read per step, the
// 10k-bus grid's FP64 solve took 49.6 against 46.6 ms.)
// lane k's value of x, in every lane (a scalar read of one lane's register)
__device__ inline float lane_value(float x, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}
__device__ inline double lane_value(double x, int k) {
    int const lo = __builtin_amdgcn_readlane(__double2loint(x), k);
    int const hi = __builtin_amdgcn_readlane(__double2hiint(x), k);
    return __hiloint2double(hi, lo);
}

template <typename T>
__global__ void __launch_bounds__(kWave) k_big_diag(FrontArgs a, int begin, int k0, T* front) {
    FrontMeta const fm = a.meta[begin + blockIdx.x];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kb = min(kBigPanel, p - k0);
    int const t = threadIdx.x;
    T* F = front + fm.off + static_cast<long long>(k0) * m + k0;
    T row[kBigPanel];
    const T* src = F + static_cast<long long>(min(t, kb - 1)) * m;
#pragma unroll
    for (int j = 0; j < kBigPanel; ++j) row[j] = src[min(j, kb - 1)];
#pragma unroll
    for (int k = 0; k < kBigPanel; ++k) {
        if (k < kb) {  // uniform
            T const rp = pivot_recip(lane_value(row[k], k), a.fail, fm.inst, t == 0);
            T const l = row[k] * rp;
#pragma unroll
            for (int j = k + 1; j < kBigPanel; ++j) {
                T const u = lane_value(row[j], k);  // the pivot row, read before this step's updates
                if (t > k) row[j] -= l * u;
            }
            if (t > k) row[k] = l;
        }
    }
    if (t < kb) {
        T* dst = F + static_cast<long long>(t) * m;
#pragma unroll
        for (int j = 0; j < kBigPanel; ++j)
            if (j < kb) dst[j] = row[j];
    }
}

// L21 = A21 U11^-1 for the rows below the panel: a thread per row with the
// row's panel segment in registers (unconditional, clamped loads) and U11
// transposed in LDS, so each step reads a contiguous, broadcast column.
constexpr int kTrsmRowsBlock = 256;
template <typename T>
__global__ void __launch_bounds__(kTrsmRowsBlock) k_big_trsm_rows(FrontArgs a, int begin, int k0, T* front) {
    __shared__ T Ut[kBigPanel][kBigPanel];  // Ut[k][q] = U[q][k]
    __shared__ T Rp[kBigPanel];
    FrontMeta const fm = a.meta[begin + blockIdx.y];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kb = min(kBigPanel, p - k0), kend = k0 + kb;
    int const r0 = kend + static_cast<int>(blockIdx.x) * kTrsmRowsBlock;
    if (r0 >= m) return;  // uniform: no rows for this block
    int const tid = threadIdx.x;
    T* F = front + fm.off;
    for (int e = tid; e < kb * kb; e += kTrsmRowsBlock)
        Ut[e % kb][e / kb] = F[static_cast<long long>(k0 + e / kb) * m + k0 + e % kb];
    if (tid < kb) Rp[tid] = pivot_recip(F[static_cast<long long>(k0 + tid) * m + k0 + tid], nullptr, 0, false);
    int const r = r0 + tid;
    const T* row = F + static_cast<long long>(min(r, m - 1)) * m + k0;
    T x[kBigPanel];
#pragma unroll
    for (int c = 0; c < kBigPanel; ++c) x[c] = row[min(c, kb - 1)];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBigPanel; ++k) {
        if (k < kb) {  // uniform
            T acc = x[k];
#pragma unroll
            for (int q = 0; q < k; ++q) acc -= x[q] * Ut[k][q];
            x[k] = acc * Rp[k];
        }
    }
    if (r < m) {
        T* out = F + static_cast<long long>(r) * m + k0;
#pragma unroll
        for (int c = 0; c < kBigPanel; ++c)
            if (c < kb) out[c] = x[c];
    }
}

// U12 = L11^-1 A12 for the columns right of the panel: a thread per column,
// the column's panel segment in registers, L11 in LDS.
// Columns [cb, ce): cb = kend (cbeg < 0) or min(cbeg, p); ce = m (cend < 0) or
// min(cend, p) -- the two-level factorisation solves the columns inside the
// outer panel first and the ones right of it later.
template <typename T>
__global__ void __launch_bounds__(256) k_big_trsm(FrontArgs a, int begin, int k0, int nb, int cbeg, int cend,
                                                  T* front) {
    __shared__ T Ls[kBigPanel][kBigPanel + 1];
    FrontMeta const fm = a.meta[begin + blockIdx.y];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kend = min(k0 + nb, p), kb = kend - k0;
    int const cb = cbeg < 0 ? kend : min(cbeg, p), ce = cend < 0 ? m : min(cend, p);
    if (cb + static_cast<int>(blockIdx.x) * 256 >= ce) return;  // uniform: no columns for this block
    T* F = front + fm.off;
    for (int e = threadIdx.x; e < kb * kb; e += 256)
        Ls[e / kb][e % kb] = F[static_cast<long long>(k0 + e / kb) * m + k0 + e % kb];
    __syncthreads();
    int const j = cb + blockIdx.x * 256 + threadIdx.x;
    if (j >= ce) return;
    T* cj = F + static_cast<long long>(k0) * m + j;
    T x[kBigPanel];
#pragma unroll
    for (int r = 0; r < kBigPanel; ++r) {
        T const v = cj[static_cast<long long>(min(r, kb - 1)) * m];  // unconditional, inside the panel rows
        x[r] = r < kb ? v : T(0);
    }
#pragma unroll
    for (int r = 1; r < kBigPanel; ++r) {
        if (r < kb) {
            T acc = x[r];
#pragma unroll
            for (int q = 0; q < r; ++q) acc -= Ls[r][q] * x[q];
            x[r] = acc;
        }
    }
#pragma unroll
    for (int r = 1; r < kBigPanel; ++r)
        if (r < kb) cj[static_cast<long long>(r) * m] = x[r];
}

// MFMA 16x16x4 blocks.  Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// accumulator register r holds row 4 (l / 16) + r (f32) or l / 16 + 4 r (f64) of
// column l % 16.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using V = f32x4;
    __device__ static V mma(float x, float y, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct Mfma<double> {
    using V = f64x4;
    __device__ static V mma(double x, double y, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// C[r0:r1, c0:c1] -= A[r0:r1, ka:ke] * B[ka:ke, c0:c1] inside one front: 128x128
// tiles, four wavefronts with 4x4 MFMA blocks each; A and B staged k-major in
// LDS in K chunks (A read along its rows, so both loads are coalesced).
template <typename T>
__device__ inline void gemm_tile(T* F, int m, int r0, int r1, int c0, int c1, int ka, int ke, int i0, int j0) {
    constexpr int TB = 128, KC = sizeof(T) == 4 ? 32 : 16;
    __shared__ T As[KC][TB + 4];
    __shared__ T Bs[KC][TB + 4];
    int const tid = threadIdx.x;
    using V = typename Mfma<T>::V;
    int const wave = tid / kWave, lane = tid % kWave;
    int const wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
    int const kb = ke - ka;
    // the accumulators start from C and A enters negated, so the MFMAs leave
    // C - A B; the C loads overlap the K loop
    V acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const gi = i0 + wr + u * 16 + Mfma<T>::row(lane, r), gj = j0 + wc + v * 16 + (lane & 15);
                acc[u][v][r] = (gi < r1 && gj < c1) ? F[static_cast<long long>(gi) * m + gj] : T(0);
            }
    for (int k1 = 0; k1 < kb; k1 += KC) {
        int const kc = min(KC, kb - k1);
        __syncthreads();  // the previous chunk has been consumed
        for (int e = tid; e < KC * TB; e += 256) {
            int const ar = e / KC, ak = e % KC;  // A: along the row
            int const gi = i0 + ar;
            As[ak][ar] = (ak < kc && gi < r1) ? -F[static_cast<long long>(gi) * m + ka + k1 + ak] : T(0);
            int const bx = e % TB, bk = e / TB;  // B: along the row of U12
            int const gj = j0 + bx;
            Bs[bk][bx] = (bk < kc && gj < c1) ? F[static_cast<long long>(ka + k1 + bk) * m + gj] : T(0);
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk += 4) {
            int const k = kk + (lane >> 4);
            T av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = As[k][wr + u * 16 + (lane & 15)];
#pragma unroll
            for (int v = 0; v < 4; ++v) bv[v] = Bs[k][wc + v * 16 + (lane & 15)];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = Mfma<T>::mma(av[u], bv[v], acc[u][v]);
        }
    }
    // store addresses recomputed from an opaque copy of the lane index, so the
    // load addresses are not kept live across the K loop
    int sl = lane;
    asm volatile("" : "+v"(sl));
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const gi = i0 + wr + u * 16 + Mfma<T>::row(sl, r), gj = j0 + wc + v * 16 + (sl & 15);
                if (gi < r1 && gj < c1) F[static_cast<long long>(gi) * m + gj] = acc[u][v][r];
            }
}

// Trailing update of a panel [k0, kend), kend = min(k0 + nb, p): rows [kend, m),
// columns [cb, ce) with cb = kend (cbeg < 0) or min(cbeg, p) and ce = m (cend < 0)
// or min(cend, p).  Its own copy of the tile loop (a single bound per
// direction keeps it within the register file without spills).
template <typename T>
__global__ void __launch_bounds__(256) k_big_gemm(FrontArgs a, int begin, int k0, int nb, int cbeg, int cend,
                                                  T* front) {
    constexpr int TB = 128, KC = sizeof(T) == 4 ? 32 : 16;
    __shared__ T As[KC][TB + 4];
    __shared__ T Bs[KC][TB + 4];
    FrontMeta const fm = a.meta[begin + blockIdx.z];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kend = min(k0 + nb, p), cb = cbeg < 0 ? kend : min(cbeg, p), ce = cend < 0 ? m : min(cend, p);
    int const kb = kend - k0;
    int const i0 = kend + blockIdx.y * TB, j0 = cb + blockIdx.x * TB;
    if (i0 >= m || j0 >= ce) return;
    T* F = front + fm.off;
    int const tid = threadIdx.x;
    using V = typename Mfma<T>::V;
    int const wave = tid / kWave, lane = tid % kWave;
    int const wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
    V acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const gi = i0 + wr + u * 16 + Mfma<T>::row(lane, r), gj = j0 + wc + v * 16 + (lane & 15);
                acc[u][v][r] = (gi < m && gj < ce) ? F[static_cast<long long>(gi) * m + gj] : T(0);
            }
    // K chunks double-buffered through registers: the next chunk's loads
    // (clamped addresses, unconditional; the bounds are applied when the
    // values enter LDS) are in flight while the current chunk's MFMAs run
    constexpr int PER = KC * TB / 256;
    T ra[PER], rb[PER];
    auto load_chunk = [&](int c0) {
        int const kc = min(KC, kb - c0);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int const e = tid + q * 256;
            int const ar = e / KC, ak = e % KC;  // A: along the row
            ra[q] = F[static_cast<long long>(min(i0 + ar, m - 1)) * m + k0 + c0 + min(ak, kc - 1)];
            int const bx = e % TB, bk = e / TB;  // B: along the row of U12
            rb[q] = F[static_cast<long long>(k0 + c0 + min(bk, kc - 1)) * m + min(j0 + bx, ce - 1)];
        }
    };
    load_chunk(0);
    for (int c0 = 0; c0 < kb; c0 += KC) {
        int const kc = min(KC, kb - c0);
        __syncthreads();  // the previous chunk has been consumed
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int const e = tid + q * 256;
            int const ar = e / KC, ak = e % KC;
            As[ak][ar] = (ak < kc && i0 + ar < m) ? -ra[q] : T(0);
            int const bx = e % TB, bk = e / TB;
            Bs[bk][bx] = (bk < kc && j0 + bx < ce) ? rb[q] : T(0);
        }
        __syncthreads();
        if (c0 + KC < kb) load_chunk(c0 + KC);
        for (int kk = 0; kk < kc; kk += 4) {
            int const k = kk + (lane >> 4);
            T av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = As[k][wr + u * 16 + (lane & 15)];
#pragma unroll
            for (int v = 0; v < 4; ++v) bv[v] = Bs[k][wc + v * 16 + (lane & 15)];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = Mfma<T>::mma(av[u], bv[v], acc[u][v]);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const gi = i0 + wr + u * 16 + Mfma<T>::row(lane, r), gj = j0 + wc + v * 16 + (lane & 15);
                if (gi < m && gj < ce) F[static_cast<long long>(gi) * m + gj] = acc[u][v][r];
            }
}

// Row block [k0, min(k0 + 64, p)) of an outer panel starting at j, columns
// [jend, m) with jend = min(j + nbo, p): subtracts the earlier blocks'
// contribution, L[k0.., j:k0] U[j:k0, jend..] (K = k0 - j).
template <typename T>
__global__ void __launch_bounds__(256) k_big_gemm_rows(FrontArgs a, int begin, int j, int nbo, int k0, T* front) {
    FrontMeta const fm = a.meta[begin + blockIdx.z];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const jend = min(j + nbo, p), r1 = min(k0 + kBigPanel, p);
    int const j0 = jend + blockIdx.x * 128;
    if (j0 >= m) return;
    gemm_tile<T>(front + fm.off, m, k0, r1, jend, m, j, k0, k0, j0);
}

// Forward solve L y = b (TR: U^T z = b), front by front from the leaves (large
// fronts, one workgroup each, front in global memory).  Slot vector w (m): the
// pivot results in [0, p) and the update passed to the parent in [p, m).
// The transposed pair solves J^T lambda = g on the same factors: U^T is lower
// triangular with U's diagonal, L^T unit upper; the right-hand side is then
// indexed by solution slots and the result by right-hand-side slots.
template <typename T, bool TR>
__global__ void __launch_bounds__(256) k_front_fwd(FrontArgs a, int begin, const T* front, T* vec,
                                                   const double* rhs, double rhs_sign) {
    FrontMeta const fm = a.meta[begin + blockIdx.x];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    int const tid = threadIdx.x, bs = blockDim.x;
    const T* F = front + fm.off;
    T* w = vec + fm.voff;
    const int* rows = (TR ? a.slot_col : a.slot_row) + fm.slot_beg;
    for (int t = tid; t < m; t += bs) w[t] = t < p ? static_cast<T>(rhs_sign * rhs[rows[t]]) : T(0);
    __syncthreads();
    for (int cq = fm.child_beg; cq < fm.child_end; ++cq) {
        ChildInfo const c = a.cinfo[cq];
        int const nu = c.mc - c.pc;
        const T* wc = vec + c.voff + c.pc;
        const int* rl = a.rel + c.rel_off;
        for (int r = tid; r < nu; r += bs) w[rl[r]] += wc[r];
        __syncthreads();
    }
    for (int k = 0; k < p; ++k) {
        T xk = w[k];
        if constexpr (TR) xk = xk / F[static_cast<long long>(k) * m + k];
        for (int i = k + 1 + tid; i < m; i += bs) {
            T const c = TR ? F[static_cast<long long>(k) * m + i] : F[static_cast<long long>(i) * m + k];
            w[i] -= c * xk;
        }
        __syncthreads();
        if (TR && tid == 0) w[k] = xk;
        __syncthreads();
    }
}

// Backward solve U x = y (TR: L^T x = z), front by front from the roots;
// structure values come from the (already solved) ancestors in the global
// solution vector.
template <typename T, bool TR>
__global__ void __launch_bounds__(256) k_front_bwd(FrontArgs a, int begin, const T* front, T* vec,
                                                   double* sol) {
    FrontMeta const fm = a.meta[begin + blockIdx.x];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    int const tid = threadIdx.x, bs = blockDim.x;
    const T* F = front + fm.off;
    T* w = vec + fm.voff;
    const int* cols = (TR ? a.slot_row : a.slot_col) + fm.slot_beg;
    for (int t = p + tid; t < m; t += bs) w[t] = static_cast<T>(sol[cols[t]]);
    __syncthreads();
    for (int k = tid; k < p; k += bs) {
        T acc = w[k];
        for (int t = p; t < m; ++t)
            acc -= (TR ? F[static_cast<long long>(t) * m + k] : F[static_cast<long long>(k) * m + t]) * w[t];
        w[k] = acc;
    }
    __syncthreads();
    // w[k] is only read in step k (every thread computes x_k itself), so the
    // result goes straight to the solution vector and w[k] stays untouched.
    for (int k = p - 1; k >= 0; --k) {
        T const xk = TR ? w[k] : w[k] / F[static_cast<long long>(k) * m + k];
        for (int i = tid; i < k; i += bs)
            w[i] -= (TR ? F[static_cast<long long>(k) * m + i] : F[static_cast<long long>(i) * m + k]) * xk;
        if (tid == 0) sol[cols[k]] = static_cast<double>(xk);
        __syncthreads();
    }
}

// Blocked triangular solves of large fronts (the Newton solves; the transposed
// pair keeps the kernels above).  Panels of kSolvePanel pivots: the diagonal
// block by one wavefront per front (lane t owns row k0 + t, the step values
// broadcast by shuffles), the rest of the rows by many workgroups.  Every row
// subtracts its terms one by one in the same order as the kernels above
// (ascending k forward, descending k backward).
constexpr int kSolvePanel = 64;

// w = vec: -F of the pivot slots (rhs_sign), plus the children's update vectors
template <typename T>
__global__ void __launch_bounds__(256) k_big_fwd_init(FrontArgs a, int begin, T* vec, const double* rhs,
                                                      double rhs_sign) {
    FrontMeta const fm = a.meta[begin + blockIdx.x];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    int const tid = threadIdx.x;
    T* w = vec + fm.voff;
    const int* rows = a.slot_row + fm.slot_beg;
    for (int t = tid; t < m; t += 256) w[t] = t < p ? static_cast<T>(rhs_sign * rhs[rows[t]]) : T(0);
    __syncthreads();
    for (int cq = fm.child_beg; cq < fm.child_end; ++cq) {
        ChildInfo const c = a.cinfo[cq];
        int const nu = c.mc - c.pc;
        const T* wc = vec + c.voff + c.pc;
        const int* rl = a.rel + c.rel_off;
        for (int r = tid; r < nu; r += 256) w[rl[r]] += wc[r];
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(64) k_big_fwd_diag(FrontArgs a, int begin, int k0, const T* front, T* vec) {
    FrontMeta const fm = a.meta[begin + blockIdx.x];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kb = min(kSolvePanel, p - k0), t = threadIdx.x;
    T* w = vec + fm.voff + k0;
    const T* row = front + fm.off + static_cast<long long>(k0 + t) * m + k0;
    T lr[kSolvePanel];
#pragma unroll
    for (int k = 0; k < kSolvePanel; ++k) lr[k] = (k < t && t < kb) ? row[k] : T(0);
    T wt = t < kb ? w[t] : T(0);
#pragma unroll
    for (int k = 0; k < kSolvePanel; ++k) {
        if (k < kb) {
            T const yk = lane_value(wt, k);  // lane k's value, by readlane (no LDS round trip)
            if (t > k && t < kb) wt -= lr[k] * yk;
        }
    }
    if (t < kb) w[t] = wt;
}

template <typename T>
__global__ void __launch_bounds__(256) k_big_fwd_update(FrontArgs a, int begin, int k0, const T* front, T* vec) {
    __shared__ T y[kSolvePanel];
    FrontMeta const fm = a.meta[begin + blockIdx.y];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kb = min(kSolvePanel, p - k0), kend = k0 + kb;
    if (kend + static_cast<int>(blockIdx.x) * 256 >= m) return;  // uniform
    T* w = vec + fm.voff;
    if (threadIdx.x < kb) y[threadIdx.x] = w[k0 + threadIdx.x];
    __syncthreads();
    int const i = kend + blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const T* ri = front + fm.off + static_cast<long long>(i) * m + k0;
    T acc = w[i];
    for (int k = 0; k < kb; ++k) acc -= ri[k] * y[k];
    w[i] = acc;
}

// pivot rows: w_k -= sum_t U[k][t] x_t over the structure slots t = p .. m-1;
// one wavefront per row, lanes along the row (coalesced), summed by a fixed
// butterfly so every solve adds in the same order
constexpr int kBwdInitRows = 4;  // rows (wavefronts) per workgroup
template <typename T>
__global__ void __launch_bounds__(kBwdInitRows * kWave) k_big_bwd_init(FrontArgs a, int begin, const T* front,
                                                                       T* vec, const double* sol) {
    FrontMeta const fm = a.meta[begin + blockIdx.y];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    int const lane = threadIdx.x % kWave;
    int const k = blockIdx.x * kBwdInitRows + threadIdx.x / kWave;
    if (k >= p || p == m) return;  // uniform per wavefront
    const int* cols = a.slot_col + fm.slot_beg;
    const T* rk = front + fm.off + static_cast<long long>(k) * m;
    T acc = T(0);
    for (int t = p + lane; t < m; t += kWave) acc += rk[t] * static_cast<T>(sol[cols[t]]);
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) vec[fm.voff + k] -= acc;
}

// panel [k0, kend) from its last row: x_k = w_k / U[k][k], written to the
// solution and over w_k
template <typename T>
__global__ void __launch_bounds__(64) k_big_bwd_diag(FrontArgs a, int begin, int k0, const T* front, T* vec,
                                                     double* sol) {
    FrontMeta const fm = a.meta[begin + blockIdx.x];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kb = min(kSolvePanel, p - k0), t = threadIdx.x;
    const int* cols = a.slot_col + fm.slot_beg + k0;
    T* w = vec + fm.voff + k0;
    const T* row = front + fm.off + static_cast<long long>(k0 + t) * m + k0;
    T ur[kSolvePanel];
#pragma unroll
    for (int k = 0; k < kSolvePanel; ++k) ur[k] = (k >= t && t < kb && k < kb) ? row[k] : T(0);
    T wt = t < kb ? w[t] : T(0);
    T diag = T(1);
#pragma unroll
    for (int k = 0; k < kSolvePanel; ++k)
        if (k == t) diag = ur[k];
#pragma unroll
    for (int k = kSolvePanel - 1; k >= 0; --k) {
        if (k < kb) {
            T const xk = lane_value(wt, k) / lane_value(diag, k);  // lane k's quotient, by readlane
            if (t < k) wt -= ur[k] * xk;
            if (t == k) {
                sol[cols[k]] = static_cast<double>(xk);
                w[k] = xk;
            }
        }
    }
}

// rows above the panel: w_i -= U[i][k] x_k for k = kend-1 .. k0
template <typename T>
__global__ void __launch_bounds__(256) k_big_bwd_update(FrontArgs a, int begin, int k0, const T* front, T* vec) {
    __shared__ T x[kSolvePanel];
    FrontMeta const fm = a.meta[begin + blockIdx.y];
    if (a.done && a.done[fm.inst]) return;
    int const m = fm.m, p = fm.p;
    if (k0 >= p) return;
    int const kb = min(kSolvePanel, p - k0);
    if (static_cast<int>(blockIdx.x) * 256 >= k0) return;  // uniform
    T* w = vec + fm.voff;
    if (threadIdx.x < kb) x[threadIdx.x] = w[k0 + threadIdx.x];
    __syncthreads();
    int const i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k0) return;
    const T* ri = front + fm.off + static_cast<long long>(i) * m + k0;
    T r[kSolvePanel];
#pragma unroll
    for (int k = 0; k < kSolvePanel; ++k) r[k] = ri[min(k, kb - 1)];  // issued together, in bounds
    T acc = w[i];
#pragma unroll
    for (int k = kSolvePanel - 1; k >= 0; --k)
        if (k < kb) acc -= r[k] * x[k];
    w[i] = acc;
}

// Small-front forward / backward solves, G fronts per workgroup as above
// (TR: the transposed pair, see k_front_fwd).
template <typename T, int MM, int L, int NT, bool TR>
__global__ void __launch_bounds__(NT) k_front_fwd_lds(FrontArgs a, int begin, int count, const T* front, T* vec,
                                                      const double* rhs, double rhs_sign) {
    constexpr int G = NT / L, LD = MM + 1, CH = MM < 8 ? MM : 8;
    __shared__ T Fs[G][MM * LD];
    __shared__ T ws[G][MM];
    int const g = threadIdx.x / L, t = threadIdx.x % L;
    int const idx = blockIdx.x * G + g;
    FrontMeta fm{};
    if (idx < count) {
        fm = a.meta[begin + idx];
        if (a.done && a.done[fm.inst]) fm.m = fm.p = fm.child_end = fm.child_beg = 0;
    }
    int const m = fm.m, p = fm.p;
    T* F = Fs[g];
    T* w = ws[g];
    load_rows<T, MM, L, CH>(F, LD, front + fm.off, m, TR ? p : m, t, t < m);
    const int* rows = TR ? a.slot_col : a.slot_row;
    if (t < m) w[t] = t < p ? static_cast<T>(rhs_sign * rhs[rows[fm.slot_beg + t]]) : T(0);
    __syncthreads();
    int const nch = fm.child_end - fm.child_beg;
    int const nch_max = group_max<L, NT>(nch);
    for (int ci = 0; ci < nch_max; ++ci) {
        if (ci < nch) {
            ChildInfo const c = a.cinfo[fm.child_beg + ci];
            int const nu = c.mc - c.pc;
            if (t < nu) w[a.rel[c.rel_off + t]] += vec[c.voff + c.pc + t];
        }
        __syncthreads();
    }
    // barriers are unconditional: fronts sharing a wavefront may have different p
    int const p_max = group_max<L, NT>(p);
    for (int k = 0; k < p_max; ++k) {
        bool const act = k < p;
        T xk = T(0);
        if (act) {
            xk = w[k];
            if constexpr (TR) xk = xk / F[k * LD + k];
            if (t > k && t < m) w[t] -= (TR ? F[k * LD + t] : F[t * LD + k]) * xk;
        }
        __syncthreads();
        // step k+1 reads only w[k+1..], so the owner may store x_k now
        if (TR && act && t == k) w[k] = xk;
    }
    __syncthreads();
    if (t < m) vec[fm.voff + t] = w[t];
}

template <typename T, int MM, int L, int NT, bool TR>
__global__ void __launch_bounds__(NT) k_front_bwd_lds(FrontArgs a, int begin, int count, const T* front, T* vec,
                                                      double* sol) {
    constexpr int G = NT / L, LD = MM + 1, CH = MM < 8 ? MM : 8;
    __shared__ T Fs[G][MM * LD];
    __shared__ T ws[G][MM];
    int const g = threadIdx.x / L, t = threadIdx.x % L;
    int const idx = blockIdx.x * G + g;
    FrontMeta fm{};
    if (idx < count) {
        fm = a.meta[begin + idx];
        if (a.done && a.done[fm.inst]) fm.m = fm.p = 0;
    }
    int const m = fm.m, p = fm.p;
    T* F = Fs[g];
    T* w = ws[g];
    const int* cols = (TR ? a.slot_row : a.slot_col) + fm.slot_beg;
    load_rows<T, MM, L, CH>(F, LD, front + fm.off, m, TR ? m : p, t, t < m);
    if (t < m) w[t] = t < p ? vec[fm.voff + t] : static_cast<T>(sol[cols[t]]);
    __syncthreads();
    if (t < p) {
        T acc = w[t];
        for (int q = p; q < m; ++q) acc -= (TR ? F[q * LD + t] : F[t * LD + q]) * w[q];
        w[t] = acc;
    }
    __syncthreads();
    int const p_max = group_max<L, NT>(p);
    for (int k = p_max - 1; k >= 0; --k) {
        if (k < p) {
            T const xk = TR ? w[k] : w[k] / F[k * LD + k];
            if (t < k) w[t] -= (TR ? F[k * LD + t] : F[t * LD + k]) * xk;
            if (t == k) sol[cols[k]] = static_cast<double>(xk);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ adjoint
// dL/dx at the converged state, one thread per bus, for
//   L = w_loss * sum_i Re(S_i) + w_dev * sum_PQ (|V_i| - 1)^2 + sum g_vm |V| + sum g_va arg V.
// sum_i Re(S_i) = sum_ij G_ij Re(V_i conj V_j) (B drops out of the symmetric sum),
// so its gradient in rectangular form is 2 G V; polar: d/dtheta_i =
// 2 Im(conj(V_i) (G V)_i), d/d|V_i| = 2 Re(conj(V_i) (G V)_i) / |V_i|.
struct AdjArgs {
    int n_bus;
    const int* rp;
    const int* ci;
    const double2* y;
    const double2* v;
    const double2* s;
    const int* ang;
    const int* mag;
    const int* bus_type;
    const double* g_vm;  // may be null
    const double* g_va;  // may be null
    double w_loss, w_dev;
    double* rhs;     // per Newton variable
    double* lossb;   // per bus loss contribution
};
__global__ void __launch_bounds__(256) k_adj_rhs(AdjArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_bus) return;
    double ge = 0.0, gf = 0.0;
    for (int q = a.rp[i]; q < a.rp[i + 1]; ++q) {
        double const G = a.y[q].x;
        double2 const vj = a.v[a.ci[q]];
        ge += G * vj.x;
        gf += G * vj.y;
    }
    double2 const vi = a.v[i];
    double const vm = hypot(vi.x, vi.y);
    double const dth = 2.0 * a.w_loss * (vi.x * gf - vi.y * ge) + (a.g_va ? a.g_va[i] : 0.0);
    double dvm = 2.0 * a.w_loss * (vi.x * ge + vi.y * gf) / vm + (a.g_vm ? a.g_vm[i] : 0.0);
    bool const pq = a.bus_type[i] == 0;
    if (pq) dvm += 2.0 * a.w_dev * (vm - 1.0);
    if (a.ang[i] >= 0) a.rhs[a.ang[i]] = dth;
    if (a.mag[i] >= 0) a.rhs[a.mag[i]] = dvm;
    double l = a.w_loss * a.s[i].x + (a.g_vm ? a.g_vm[i] * vm : 0.0) + (a.g_va ? a.g_va[i] * atan2(vi.y, vi.x) : 0.0);
    if (pq) l += a.w_dev * (vm - 1.0) * (vm - 1.0);
    a.lossb[i] = l;
}

// dL/d(bus p, q): F_P = P - p/S_base, so dL/dp = lambda_P / S_base (same for Q).
struct AdjBusArgs {
    int n_bus;
    const int* peq;
    const int* qeq;
    const int* bus_inst;
    const double* lam;
    double inv_base;
    double* gp;
    double* gq;
};
__global__ void __launch_bounds__(256) k_adj_bus(AdjBusArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_bus) return;
    a.gp[i] = a.peq[i] >= 0 ? a.lam[a.peq[i]] * a.inv_base : 0.0;
    a.gq[i] = a.qeq[i] >= 0 ? a.lam[a.qeq[i]] * a.inv_base : 0.0;
}

// dL/d(branch r, x) = w_loss * Re(dy) |V_f - V_t|^2 - lambda^T dF/d(r, x), with
// dy/dr = -Zb y_S^2 and dy/dx = -j Zb y_S^2 (y_S = 1 / (r + jx) in siemens),
// dS_f = V_f conj(dy (V_f - V_t)), dS_t = V_t conj(dy (V_t - V_f)).
struct AdjBrArgs {
    int n_br;
    const int* bf;
    const int* bt;
    const int* br_inst;
    const double* r;
    const double* x;
    const double* zb;
    const double2* v;
    const int* peq;
    const int* qeq;
    const double* lam;
    double w_loss;
    double* gr;
    double* gx;
};
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

__global__ void __launch_bounds__(256) k_adj_branch(AdjBrArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_br) return;
    int const f = a.bf[k], t = a.bt[k];
    double gr = 0.0, gx = 0.0;
    double const r = a.r[k], x = a.x[k];
    double const z2 = r * r + x * x;
    if (f >= 0 && z2 > 1e-12) {
        double2 const y = make_double2(r / z2, -x / z2);
        double2 y2 = cmul(y, y);
        double const zb = a.zb[a.br_inst[k]];
        double2 const dyr = make_double2(-zb * y2.x, -zb * y2.y);
        double2 const dyx = make_double2(zb * y2.y, -zb * y2.x);  // -j zb y^2
        double2 const vf = a.v[f], vt = a.v[t];
        double2 const dv = make_double2(vf.x - vt.x, vf.y - vt.y);
        double const dv2 = dv.x * dv.x + dv.y * dv.y;
        for (int which = 0; which < 2; ++which) {
            double2 const dy = which == 0 ? dyr : dyx;
            double2 const dsf = cmul(vf, cconj(cmul(dy, dv)));
            double2 const dst = cmul(vt, cconj(cmul(dy, make_double2(-dv.x, -dv.y))));
            double lt = 0.0;
            if (a.peq[f] >= 0) lt += a.lam[a.peq[f]] * dsf.x;
            if (a.qeq[f] >= 0) lt += a.lam[a.qeq[f]] * dsf.y;
            if (a.peq[t] >= 0) lt += a.lam[a.peq[t]] * dst.x;
            if (a.qeq[t] >= 0) lt += a.lam[a.qeq[t]] * dst.y;
            double const g = a.w_loss * dy.x * dv2 - lt;
            if (which == 0) gr = g; else gx = g;
        }
    }
    a.gr[k] = gr;
    a.gx[k] = gx;
}

// ------------------------------------------------------------ branch flows
// Flows of every branch at a solution: the fields BranchData declares
// (include/gap/core/types.h:233-243: p/q/s at both ends, current magnitude and
// angle at both ends) and the reference never fills.  The branch's own 2x2
// block of the Y-bus builder (cpu_admittance_matrix.cpp:121-135: series
// y = 1/(r + jx) when |z|^2 > 1e-12, (y + g1) + j b1/2 on each diagonal, -y off
// it), per unit with the network's impedance base; S = V conj(I) in VA and
// currents in A of the network's voltage base, I_base = S_base / (sqrt(3) u_base).
// Out-of-service branches (bf < 0) get zeros.  One thread per branch.
constexpr int kFlowFields = 10;  // p_from, p_to, q_from, q_to, s_from, s_to, i_from, i_from_angle, i_to, i_to_angle
struct FlowArgs {
    int n_br;
    const int* bf;
    const int* bt;
    const int* br_inst;
    const double* r;
    const double* x;
    const double* g;
    const double* b;
    const double* zb;     // per instance: u_base^2 / S_base
    const double* ibase;  // per instance: S_base / (sqrt(3) u_base)
    const double2* v;
    double base;
    double* out;          // kFlowFields per branch
};
__global__ void __launch_bounds__(256) k_branch_flow(FlowArgs a) {
#pragma clang fp contract(off)
    int const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_br) return;
    double* o = a.out + static_cast<long long>(kFlowFields) * k;
    int const f = a.bf[k], t = a.bt[k];
    if (f < 0) {
        for (int q = 0; q < kFlowFields; ++q) o[q] = 0.0;
        return;
    }
    double const r = a.r[k], x = a.x[k], z2 = r * r + x * x;
    double yr = 0.0, yi = 0.0;
    if (z2 > 1e-12) {
        yr = r / z2;
        yi = -x / z2;
    }
    int const g = a.br_inst[k];
    double const zb = a.zb[g];
    double const dr = (yr + a.g[k]) * zb, di = (yi + a.b[k] / 2.0) * zb;  // diagonal, per unit
    double const orr = -yr * zb, oi = -yi * zb;                          // off-diagonal
    double2 const vf = a.v[f], vt = a.v[t];
    double const ifr = (dr * vf.x - di * vf.y) + (orr * vt.x - oi * vt.y);
    double const ifi = (dr * vf.y + di * vf.x) + (orr * vt.y + oi * vt.x);
    double const itr = (dr * vt.x - di * vt.y) + (orr * vf.x - oi * vf.y);
    double const iti = (dr * vt.y + di * vt.x) + (orr * vf.y + oi * vf.x);
    double const pf = (vf.x * ifr + vf.y * ifi) * a.base, qf = (vf.y * ifr - vf.x * ifi) * a.base;
    double const pt = (vt.x * itr + vt.y * iti) * a.base, qt = (vt.y * itr - vt.x * iti) * a.base;
    double const ib = a.ibase[g];
    o[0] = pf;
    o[1] = pt;
    o[2] = qf;
    o[3] = qt;
    o[4] = sqrt(pf * pf + qf * qf);
    o[5] = sqrt(pt * pt + qt * qt);
    o[6] = sqrt(ifr * ifr + ifi * ifi) * ib;
    o[7] = atan2(ifi, ifr);
    o[8] = sqrt(itr * itr + iti * iti) * ib;
    o[9] = atan2(iti, itr);
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
    // grid schedule: per bus the list positions of its angle and magnitude
    // variables in xs, the solution in the factor type in right-hand-side
    // list order (-1: no such variable)
    const int2* bpos;
    const void* xs;
    // grid schedule: v holds the network's buses in the schedule's own order;
    // at the end of the solve bus i is written to v_out[perm[i]]
    const int* perm;
    double2* v_out;
};
__device__ inline void update_bus(UpdArgs const& a, int i, bool zero) {
#pragma clang fp contract(off)
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
        vm = ref_clamp_magnitude(vm);
        v = make_double2(vm * cos(va), vm * sin(va));
    }
    a.v[i] = v;
}

__global__ void __launch_bounds__(256) k_update(UpdArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_bus) return;
    int inst = a.bus_inst[i];
    if (a.done[inst]) return;
    update_bus(a, i, a.fail[inst] != 0);
}

}  // namespace gaphip
