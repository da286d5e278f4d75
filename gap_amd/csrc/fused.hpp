// Grid schedule of the batched Newton-Raphson loop: one workgroup solves one
// network of a batch from flat start to convergence inside a single launch.
//
// Per iteration the workgroup runs, with workgroup barriers between phases:
//   mismatch + max|F| (cpu_newton_raphson.cpp:115-157, 243-337) -> convergence
//   test -> multifrontal factorisation leaves -> root, with the Jacobian
//   entries (:509-765) evaluated while each front is assembled and the forward
//   substitution of -F folded into the elimination -> backward substitution
//   root -> leaves -> damped polar update (:438-504).  The per-bus and
//   per-pair arithmetic uses the rectangular forms below.
//
// A level of the network's elimination tree is a list of tasks; a task is up to
// 16, 8 or 4 fronts of at most 4, 8 or 16 slots (4, 8 or 16 lanes per front),
// or one front of at most 32 or 64 slots (the whole wavefront), handled by
// one wavefront (lane t of a front's lanes owning column t, the fronts staged
// in the wavefront's LDS region).
// Assembly is driven by flat per-task lists built once per plan, so all 64
// lanes gather the task's inputs together (two dependent loads per entry):
//   pairs : one per Jacobian 2x2 block (buses i, j, Y slot, LDS position and
//           which of the four entries exist), evaluated from Y, S and V;
//   rhs   : -F of the pivot equations into the right-hand side;
//   ext   : child Schur-complement entries, added into the front (LDS atomics,
//           list order = child order, so every solve sums in the same order);
//   vx    : child right-hand-side entries, added likewise;
//   sol   : (backward) solution values of the structure slots, gathered at
//           their pivots' list positions; each task writes its own pivots'
//           solution at their list positions, contiguous per task.
// Inside a task only wavefront-level synchronisation is needed.  Between tasks
// of different wavefronts the order is kept by dependencies, not by a barrier
// per level: a wavefront publishes in LDS the index of the next task it will
// run, and a task waits only until the wavefronts that hold its children (in
// the factorisation) or its parents (in the backward substitution) have moved
// past them.
#pragma once

namespace gaphip {

struct alignas(16) GridTask {
    int front_beg, cls_cnt;  // first front id; class << 8 | front count
    int pair_beg, pair_end;
    int rhs_beg, rhs_end;
    int ext_beg, ext_end;
    int vx_beg, vx_end;
    int sol_beg, sol_end;
    // For each residue r < kDepWaves of a task position modulo kDepWaves
    // (the factorisation counts from the network's first task, the backward
    // substitution from its last), 16 bits each: the distance to the nearest
    // task of residue r that holds a child (dep_f) or a parent (dep_b) of
    // one of this task's fronts; 0 for none.
    unsigned dep_f[2];
    unsigned dep_b[2];
};
constexpr int kDepWaves = 4;

struct FusedArgs {
    int n_inst;
    int max_it;
    double tol;
    int order_off;        // this launch covers order[order_off + blockIdx.x]
    const int* order;     // networks, largest first
    const int* bus_off;   // per network, n_inst + 1 entries
    const int* var_off;   // per network: first Newton equation / variable, n_inst + 1 entries
    const int* lvl_beg;   // per network: its levels are [lvl_beg[g], lvl_beg[g+1]) of lvl_ptr
    const int* lvl_ptr;   // per (network, level): tasks [lvl_ptr[e], lvl_ptr[e+1])
    const GridTask* tasks;
    const int4* gfront;   // per front, from its network's bases: (storage offset, vector offset, m | p << 16, list position of its first pivot)
    const int2* pairs;    // (i | j << 16 as buses of the network, LDS position | entry mask << 16)
    const void* ypair;    // per-unit Y value of each pair (the factor type's complex), in pair order (gathered once per solve)
    const void* fe;       // per Newton equation: -F (factor type), written by the mismatch pass
    // list entries index | LDS position << kListIdxBits, the index counted
    // from the network's base (netbase for fronts, vectors and the list,
    // var_off for equations)
    const unsigned* rhs;  // (equation, LDS position)
    const unsigned* ext;  // (front storage entry, LDS position)
    const unsigned* vx;   // (vector entry, LDS position)
    const unsigned* sol;  // (list position of the variable's pivot, LDS position)
    // per network: first front storage entry, first vector entry, first list
    // entry, and the local index (in the schedule's numbering) of the
    // network's first equation in the reference's numbering
    const int4* netbase;
    int pool;             // LDS elements per wavefront
    int ld4;              // leading dimension of the 64-slot class tiles (largest such front + 1)
    unsigned long long* maxf;  // rows of per-network max |F| per iteration
    int* done;
    int* iters;
    int* lu_fail;
    double* final_mis;
    void* xs;  // solution of each pivot (factor type), in right-hand-side list order
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// max over the groups of L lanes of one wavefront (every lane gets it)
template <int L>
__device__ inline int wave_group_max(int v) {
    for (int off = L; off < kWave; off <<= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// Complex of the factor type: the Jacobian blocks of the grid kernel are
// evaluated in it (FP32 factors: from FP32 copies of V, S and Y, since every
// entry is rounded to FP32 for the factorisation anyway; the mismatch, the
// update and the convergence test stay FP64).
template <typename T> struct Cplx;
template <> struct Cplx<float> { using type = float2; };
template <> struct Cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename Cplx<T>::type;

// Lane index, recomputed at each call: an empty volatile asm keeps the
// compiler from hoisting the lane-dependent masks and addresses of every task
// class out of the Newton loop, where they would stay live for the whole kernel.
__device__ inline int lane_id() {
    int l = static_cast<int>(threadIdx.x);
    asm volatile("" : "+v"(l));
    return l & (kWave - 1);
}

// Grid list entries: an index below 2^kListIdxBits and an LDS position below
// 2^kListPosBits in 32 bits.
constexpr int kListIdxBits = 19, kListPosBits = 32 - kListIdxBits;
__device__ inline int lidx(unsigned u) { return static_cast<int>(u & ((1u << kListIdxBits) - 1u)); }
__device__ inline int lpos(unsigned u) { return static_cast<int>(u >> kListIdxBits); }

// Lanes per front of each grid task class, as a shift: 4, 8 and 16 lanes
// for fronts of up to 4, 8 and 16 slots (16, 8 and 4 fronts per task), a
// whole wavefront for fronts of up to 32 and 64 slots (one front per task).
__device__ inline int grid_lane_shift(int cls) { return cls < 3 ? 2 + cls : 6; }

template <typename T>
__device__ inline void lds_add(T* p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Chunks of 64 list entries each lane keeps in registers, so that all index
// loads of a task are issued together, then all value loads; longer lists
// finish in a loop.
constexpr int kPairChunks = 1, kExtChunks = 2;

// Every load of the grid kernel is issued for every lane, with no bounds
// branch around it: a load under a divergent branch is waited for where the
// branch joins, which serialises a task's gathers.  The buffers it reads
// therefore extend past their last element (padding entries are neutral:
// index 0, LDS position -1, mask 0), and the bounds are applied where the
// values are used.
constexpr int kListPad = 4 * kWave;             // entries after each task list
constexpr int kTaskPad = 16;                    // descriptors after the last task (2 ahead at 8 wavefronts)
constexpr int kRowPad = 4;                      // Y-bus entries after the last row
constexpr long long kFrontPad = kWave * kWave;  // front storage: one 64 x 64 front
constexpr long long kVecPad = kWave;            // slot vectors and slot columns
static_assert(kListPad >= kPairChunks * kWave && kListPad >= kExtChunks * kWave && kListPad >= 4 * kWave,
              "list padding too short");

// ---------------------------------------------------------------------------
// Rectangular forms of the per-bus and per-pair arithmetic.  They equal the
// reference's polar formulas (cpu_newton_raphson.cpp:438-504, 546-557,
// 640-725) term by term and differ only in rounding; they avoid atan2, hypot
// and sin/cos of arbitrary arguments, whose library code would dominate the
// register budget of the grid kernel.

// S_i = V_i conj((Y V)_i) and the P / Q mismatch rows of buses i[0 .. B-1]
// (those past the first only where use_rest); returns their max |F|.  Each
// equation's -F is written in the factor type, by equation index (the factor
// tasks gather it); the FP64 F only for the network's first equation in the
// reference's numbering (rf in this schedule's numbering), stored at r0: the
// one row the convergence test reads back.  The B rows' loads are issued
// together, four entries per row and step; the sums keep the row order.
// B = 2 pays with FP64 factors; with FP32 factors its registers would spill
// (that kernel is held to 96).
template <typename T, int B>
__device__ inline double grid_mismatch_buses(MisArgs const& a, int const (&i)[B], bool use_rest, int r0, int rf) {
#pragma clang fp contract(off)
    double cr[B], cim[B];
    int qb[B], qe[B];
    int n = 0;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        cr[b] = cim[b] = 0.0;
        qb[b] = a.rp[i[b]];
        qe[b] = a.rp[i[b] + 1];
        n = max(n, qe[b] - qb[b]);
    }
    for (int s = 0; s < n; s += kRowPad) {
        // unconditional loads: an index past its row is clamped into the
        // kRowPad entries that follow it (the Y-bus arrays extend kRowPad
        // entries past the last row, and every column index is a bus);
        // entries past a row are not summed
        int c[B][kRowPad];
        double2 yv[B][kRowPad], vv[B][kRowPad];
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int u = 0; u < kRowPad; ++u) {
                int const q = min(qb[b] + s + u, qe[b] + kRowPad - 1);
                c[b][u] = a.ci[q];
                yv[b][u] = a.y[q];
            }
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int u = 0; u < kRowPad; ++u) vv[b][u] = a.v[c[b][u]];
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int u = 0; u < kRowPad; ++u)
                if (qb[b] + s + u < qe[b]) {
                    cr[b] += yv[b][u].x * vv[b][u].x - yv[b][u].y * vv[b][u].y;
                    cim[b] += yv[b][u].x * vv[b][u].y + yv[b][u].y * vv[b][u].x;
                }
    }
    T* fe = static_cast<T*>(a.fe);
    double m = 0.0;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        double2 const vi = a.v[i[b]];
        double2 const sp = a.sspec[i[b]];
        int const pe = a.peq[i[b]], qeq = a.qeq[i[b]];
        if (b > 0 && !use_rest) break;
        double const sr = vi.x * cr[b] + vi.y * cim[b];
        double const si = vi.y * cr[b] - vi.x * cim[b];
        // S feeds the Jacobian's diagonal blocks only: in FP64, or as its FP32
        // copy (nothing reads the FP64 S of an FP32-factor solve afterwards)
        if constexpr (sizeof(T) == 4) {
            a.vsf[i[b]] = make_float4(static_cast<float>(vi.x), static_cast<float>(vi.y), static_cast<float>(sr),
                                      static_cast<float>(si));
        } else {
            a.s[i[b]] = make_double2(sr, si);
        }
        if (pe >= 0) {
            double const fp = sr - sp.x;
            if (pe == rf) a.f[r0] = fp;
            fe[pe] = static_cast<T>(-1.0 * fp);
            m = fmax(m, fabs(fp));
        }
        if (qeq >= 0) {
            double const fq = si - sp.y;
            if (qeq == rf) a.f[r0] = fq;
            fe[qeq] = static_cast<T>(-1.0 * fq);
            m = fmax(m, fabs(fq));
        }
    }
    return m;
}

// The Jacobian 2x2 block of Y pair (i, j) from rectangular V: with
// e = V_i conj(V_j) = |V_i||V_j| (cos + j sin)(theta_i - theta_j),
//   dP/dtheta_j = G Im e - B Re e,   dQ/dtheta_j = -(G Re e + B Im e),
//   dP/d|V_j|  = (G Re e + B Im e) / |V_j|,   dQ/d|V_j| = (G Im e - B Re e) / |V_j|;
// on the diagonal the reference's S-based forms.  A magnitude below 1e-6 is
// replaced by 1 along the same direction, as the reference replaces |V| by 1.
template <typename R>
__device__ inline R jac_div(R x, R d) {
    if constexpr (sizeof(R) == 4) return x * __builtin_amdgcn_rcpf(d);  // 1 ulp; the entry is an FP32 factor input
    else return x / d;
}
template <typename R, typename C>
__device__ inline void jac_block_rect(bool diag, C yv, C vi, C vj, C si, R& pt, R& pv, R& qt, R& qv) {
    R const G = yv.x, B = yv.y;
    R const ri = sqrt(vi.x * vi.x + vi.y * vi.y);
    if (diag) {
        R const Vi = ri < R(1e-6) ? R(1) : ri;
        R const P = si.x, Q = si.y;
        pt = -Q - B * Vi * Vi;
        pv = jac_div(P, Vi) + G * Vi;
        qt = P - G * Vi * Vi;
        qv = jac_div(Q, Vi) - B * Vi;
        return;
    }
    R rj = sqrt(vj.x * vj.x + vj.y * vj.y);
    if (ri < R(1e-6)) {
        vi.x = ri > R(0) ? jac_div(vi.x, ri) : R(1);
        vi.y = ri > R(0) ? jac_div(vi.y, ri) : R(0);
    }
    if (rj < R(1e-6)) {
        vj.x = rj > R(0) ? jac_div(vj.x, rj) : R(1);
        vj.y = rj > R(0) ? jac_div(vj.y, rj) : R(0);
        rj = R(1);
    }
    R const re = vi.x * vj.x + vi.y * vj.y, im = vi.y * vj.x - vi.x * vj.y;
    R const c = G * re + B * im, d = G * im - B * re;
    pt = d;
    qt = -c;
    pv = jac_div(c, rj);
    qv = jac_div(d, rj);
}

// An FP64 constant materialised in scalar registers where it is used.  FP64
// operands other than a few inline values cannot be encoded in an
// instruction; left to itself the compiler hoists them out of the Newton
// loop into vector registers, where they stay live for the whole kernel.
__device__ inline double sconst(double c) {
    asm volatile("" : "+s"(c));
    return c;
}

// sin and cos of |x| <= 0.15 by their Taylor series (truncation below 1e-20)
__device__ inline void sincos_small(double x, double& sn, double& cs) {
    double const x2 = x * x;
    sn = x * (1.0 + x2 * (sconst(-1.0 / 6) + x2 * (sconst(1.0 / 120) + x2 * (sconst(-1.0 / 5040) +
                                                                           x2 * (sconst(1.0 / 362880) +
                                                                                 x2 * sconst(-1.0 / 39916800))))));
    cs = 1.0 + x2 * (-0.5 + x2 * (sconst(1.0 / 24) + x2 * (sconst(-1.0 / 720) + x2 * (sconst(1.0 / 40320) +
                                                                                 x2 * (sconst(-1.0 / 3628800) +
                                                                                       x2 * sconst(1.0 / 479001600))))));
}

// Damped, clamped polar update: rotate V by the angle step, then rescale to the
// clamped new magnitude.
template <typename T>
__device__ inline void grid_update_bus(UpdArgs const& a, int i, bool zero) {
    double2 v = a.v[i];
    int2 const bp = a.bpos[i];
    int const av = bp.x, mv = bp.y;
    T const* xs = static_cast<T const*>(a.xs);
    double const dxa = static_cast<double>(xs[av < 0 ? 0 : av]), dxm = static_cast<double>(xs[mv < 0 ? 0 : mv]);  // unconditional loads
    if (av >= 0) {
        double ac = sconst(0.9) * (zero ? 0.0 : dxa);
        double const amax = sconst(0.15);
        if (fabs(ac) > amax) ac = amax * (ac > 0 ? 1.0 : -1.0);
        double sn, cs;
        sincos_small(ac, sn, cs);
        v = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
    }
    if (mv >= 0) {
        double const old = sqrt(v.x * v.x + v.y * v.y);
        double mc = sconst(0.9) * (zero ? 0.0 : dxm);
        double const mmax = sconst(0.08);
        if (fabs(mc) > mmax) mc = mmax * (mc > 0 ? 1.0 : -1.0);
        double const vm = ref_clamp_magnitude(old + mc, sconst(1.5));
        // a NaN magnitude scales V to NaN, as the reference's polar rebuild does
        v = !(old <= 0.0) ? make_double2(v.x * (vm / old), v.y * (vm / old)) : make_double2(vm, 0.0);
    }
    a.v[i] = v;
}


// The loads of a task that depend only on its descriptor (the lane's front
// record and list entries, the pairs' Y values, the -F values of its
// right-hand side).  A wavefront issues them one task ahead, so a task waits
// only for its gathers of V, S and child data.
template <typename T>
struct TaskPre {
    int4 gf;                 // front record of the lane's front
    int2 pr[kPairChunks];    // pair list entries
    cplx_t<T> yv[kPairChunks]; // their per-unit Y
    unsigned ex[kExtChunks]; // extend-add entries
    unsigned rq;             // right-hand-side entry
    unsigned vv;             // child-vector entry
    unsigned so;             // (backward) solution entry
};

// Unconditional: entries past the task's lists are the next task's or padding.
template <typename T, bool BWD>
__device__ inline TaskPre<T> load_pre(FusedArgs const& a, GridTask const& tk) {
    int const lane = lane_id(), cls = tk.cls_cnt >> 8;
    int const gi = lane >> grid_lane_shift(cls);
    TaskPre<T> r;
    r.gf = a.gfront[tk.front_beg + gi];
    if constexpr (BWD) {
        r.so = a.sol[tk.sol_beg + lane];
    } else {
#pragma unroll
        for (int k = 0; k < kPairChunks; ++k) {
            r.pr[k] = a.pairs[tk.pair_beg + k * kWave + lane];
            r.yv[k] = static_cast<cplx_t<T> const*>(a.ypair)[tk.pair_beg + k * kWave + lane];
        }
#pragma unroll
        for (int k = 0; k < kExtChunks; ++k) r.ex[k] = a.ext[tk.ext_beg + k * kWave + lane];
        r.rq = a.rhs[tk.rhs_beg + lane];
        r.vv = a.vx[tk.vx_beg + lane];
    }
    return r;
}

// Lane k of each group of L lanes, in every lane of the group.  Groups of 4
// use a quad-permute DPP move (a VALU operation, no LDS round trip); larger
// groups a bpermute shuffle from lane src = group base + k.  k is a
// compile-time constant once the callers' loops are unrolled.
template <int K>
__device__ inline float quad_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, true));
}
template <int K>
__device__ inline double quad_bcast(double v) {
    int const lo = __builtin_amdgcn_mov_dpp(__double2loint(v), K * 0x55, 0xF, 0xF, true);
    int const hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), K * 0x55, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <int L, typename T>
__device__ inline T group_bcast(T v, int k, int src) {
    if constexpr (L == 4) {
        switch (k) {
            case 0: return quad_bcast<0>(v);
            case 1: return quad_bcast<1>(v);
            case 2: return quad_bcast<2>(v);
            default: return quad_bcast<3>(v);
        }
    } else {
        return __shfl(v, src);
    }
}

// Division by a pivot in the grid kernel's factorisation and solves: with
// FP32 factors the hardware reciprocal (1 ulp, one instruction against a
// dozen and two mode switches for the correctly rounded quotient), whose
// rounding the FP64 residual of the next Newton iteration absorbs like any
// other FP32 rounding of the factors; with FP64 factors the correctly
// rounded quotient.
template <typename T>
__device__ inline T pivot_recip(T d) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(d);
    else return T(1) / d;
}
template <typename T>
__device__ inline T pivot_div(T x, T d) {
    if constexpr (sizeof(T) == 4) return x * __builtin_amdgcn_rcpf(d);
    else return x / d;
}

// A factor task's gathers that do not depend on other tasks: V and S of its
// pairs' buses and the -F of its pivot equations (written by the mismatch
// pass).  They are issued before the task waits for its children.
template <typename T>
struct TaskGath {
    cplx_t<T> vi[kPairChunks], vj[kPairChunks], si[kPairChunks];
    T rv;
};
template <typename T>
__device__ inline TaskGath<T> factor_gathers(FusedArgs const& a, MisArgs const& ma, GridTask const& tk,
                                             TaskPre<T> const& pre, int b0, int r0) {
    using C = cplx_t<T>;
    constexpr int VS = sizeof(T) == 4 ? 2 : 1;  // FP32: V and S of a bus side by side
    int const lane = lane_id();
    C const* vsrc;
    C const* ssrc;
    if constexpr (sizeof(T) == 4) {
        vsrc = reinterpret_cast<C const*>(ma.vsf + b0);
        ssrc = vsrc + 1;
    } else {
        vsrc = ma.v + b0;
        ssrc = ma.s + b0;
    }
    // unconditional: a lane whose value is not used reads lane 0's element
    // instead, so its load falls into a line the wavefront fetches anyway
    TaskGath<T> r;
#pragma unroll
    for (int k = 0; k < kPairChunks; ++k) {
        bool const pok = tk.pair_beg + k * kWave + lane < tk.pair_end;
        int const li = pre.pr[k].x & 0xffff, lj = static_cast<int>(static_cast<unsigned>(pre.pr[k].x) >> 16);
        int const i0 = __builtin_amdgcn_readfirstlane(li);
        bool const diag = li == lj;
        r.vi[k] = vsrc[VS * (pok ? li : i0)];
        r.vj[k] = vsrc[VS * (pok && !diag ? lj : i0)];
        r.si[k] = ssrc[VS * (pok && diag ? li : i0)];
    }
    bool const rok = tk.rhs_beg + lane < tk.rhs_end;
    T const* fe = static_cast<T const*>(a.fe) + r0;  // the network's equations
    r.rv = fe[rok ? lidx(pre.rq) : __builtin_amdgcn_readfirstlane(lidx(pre.rq))];
    return r;
}

// Assemble, factor and forward-eliminate up to kWave / L fronts of one class.
// Stores the U rows (diagonal onwards), the Schur complement and the
// eliminated right-hand side w of every front.
template <typename T, int MM, int L>
__device__ __attribute__((always_inline)) void grid_factor_task(FusedArgs const& a, FrontArgs const& fa, MisArgs const& ma, GridTask const& tk,
                                 TaskPre<T> const& pre, TaskGath<T> const& gat, int g, int b0, int r0, T* front, T* vec,
                                 T* pool, int ld) {
    constexpr int G = kWave / L;
    int const lane = lane_id(), gi = lane / L, t = lane % L;
    int const rows = ld - 1, wbase = G * rows * ld, region = wbase + G * MM;
    int2 const* pr = pre.pr;
    unsigned const* ex = pre.ex;
    unsigned const vv = pre.vv;
    using C = cplx_t<T>;
    C const* yv = pre.yv;
    // the Jacobian inputs: V and S in the factor type, from the network's
    // first bus (the pair list holds buses of the network)
    constexpr int VS = sizeof(T) == 4 ? 2 : 1;  // FP32: V and S of a bus side by side
    C const* vsrc;
    C const* ssrc;
    if constexpr (sizeof(T) == 4) {
        vsrc = reinterpret_cast<C const*>(ma.vsf + b0);
        ssrc = vsrc + 1;
    } else {
        vsrc = ma.v + b0;
        ssrc = ma.s + b0;
    }
    C const* ypair = static_cast<C const*>(a.ypair);
    // gathers, unconditional: every list entry (the next task's and padding
    // included) holds valid indices; which entries are this task's is applied
    // where the values are used
    bool pok[kPairChunks], eok[kExtChunks];
#pragma unroll
    for (int k = 0; k < kPairChunks; ++k) pok[k] = tk.pair_beg + k * kWave + lane < tk.pair_end;
#pragma unroll
    for (int k = 0; k < kExtChunks; ++k) eok[k] = tk.ext_beg + k * kWave + lane < tk.ext_end;
    bool const rok = tk.rhs_beg + lane < tk.rhs_end;
    bool const vok = tk.vx_beg + lane < tk.vx_end;
    // a lane whose value is not used reads lane 0's element instead, so its
    // load falls into a line the wavefront fetches anyway
    C const* vi = gat.vi;
    C const* vj = gat.vj;
    C const* si = gat.si;
    T exv[kExtChunks];
#pragma unroll
    for (int k = 0; k < kExtChunks; ++k)
        exv[k] = front[eok[k] ? lidx(ex[k]) : __builtin_amdgcn_readfirstlane(lidx(ex[k]))];
    T const vxv = vec[vok ? lidx(vv) : __builtin_amdgcn_readfirstlane(lidx(vv))];
    T const* fe = static_cast<T const*>(a.fe) + r0;  // the network's equations
    T const rv = gat.rv;
    T* F = pool + gi * rows * ld;
    T* w = pool + wbase + gi * MM;
    wave_sync();  // the previous task of this wavefront has finished with the region
    {
        // 16-byte stores: the pool of a wavefront starts 16-byte aligned and
        // holds the region rounded up to whole 16-byte units
        constexpr int V = 16 / static_cast<int>(sizeof(T));
        float4* z = reinterpret_cast<float4*>(pool);
        for (int e = lane; e < (region + V - 1) / V; e += kWave) z[e] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wave_sync();
    auto put_pair = [&](bool ok, int2 const& p2, C y2, C pi2, C pj2, C s2) {
        int const base = p2.y & 0xffff, mask = p2.y >> 16;
        if (!ok || !mask) return;
        T pt, pv, qt, qv;
        jac_block_rect<T>((p2.x & 0xffff) == static_cast<int>(static_cast<unsigned>(p2.x) >> 16), y2, pi2, pj2, s2, pt,
                          pv, qt, qv);
        if (mask & 1) pool[base] = pt;
        if (mask & 2) pool[base + 1] = pv;
        if (mask & 4) pool[base + ld] = qt;
        if (mask & 8) pool[base + ld + 1] = qv;
    };
#pragma unroll
    for (int k = 0; k < kPairChunks; ++k) put_pair(pok[k], pr[k], yv[k], vi[k], vj[k], si[k]);
    // entries past the register chunks, two per lane and step: both list
    // entries are read before either gather (the list padding covers the
    // second read past the list's end; a lane past the task's end gathers
    // the step's first element)
    for (int q0 = tk.pair_beg + kPairChunks * kWave; q0 < tk.pair_end; q0 += 2 * kWave) {
        int const q = q0 + lane;
        int2 const p0 = a.pairs[q], p1 = a.pairs[q + kWave];
        C const y0 = ypair[q], y1 = ypair[q + kWave];
        bool const ok0 = q < tk.pair_end, ok1 = q + kWave < tk.pair_end;
        int const f = __builtin_amdgcn_readfirstlane(p0.x & 0xffff);
        int const i0 = ok0 ? p0.x & 0xffff : f, j0 = ok0 ? static_cast<int>(static_cast<unsigned>(p0.x) >> 16) : f;
        int const i1 = ok1 ? p1.x & 0xffff : f, j1 = ok1 ? static_cast<int>(static_cast<unsigned>(p1.x) >> 16) : f;
        C const vi0 = vsrc[VS * i0], vj0 = vsrc[VS * j0], s0 = ssrc[VS * i0];
        C const vi1 = vsrc[VS * i1], vj1 = vsrc[VS * j1], s1 = ssrc[VS * i1];
        put_pair(ok0, p0, y0, vi0, vj0, s0);
        put_pair(ok1, p1, y1, vi1, vj1, s1);
    }
    if (rok) pool[lpos(pre.rq)] = rv;
    for (int q = tk.rhs_beg + kWave + lane; q < tk.rhs_end; q += kWave) {
        unsigned const e = a.rhs[q];
        pool[lpos(e)] = fe[lidx(e)];
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < kExtChunks; ++k)
        if (eok[k]) lds_add(pool + lpos(ex[k]), exv[k]);
    // (the same for the child entries, four per lane and step)
    for (int q0 = tk.ext_beg + kExtChunks * kWave; q0 < tk.ext_end; q0 += 4 * kWave) {
        int const q = q0 + lane;
        unsigned e[4];
        T x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = a.ext[q + u * kWave];
        int const f = __builtin_amdgcn_readfirstlane(lidx(e[0]));
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = front[q + u * kWave < tk.ext_end ? lidx(e[u]) : f];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q + u * kWave < tk.ext_end) lds_add(pool + lpos(e[u]), x[u]);
    }
    if (vok) lds_add(pool + lpos(vv), vxv);
    for (int q0 = tk.vx_beg + kWave; q0 < tk.vx_end; q0 += 2 * kWave) {
        int const q = q0 + lane;
        unsigned const e0 = a.vx[q], e1 = a.vx[q + kWave];
        bool const ok0 = q < tk.vx_end, ok1 = q + kWave < tk.vx_end;
        int const f = __builtin_amdgcn_readfirstlane(lidx(e0));
        T const x0 = vec[ok0 ? lidx(e0) : f], x1 = vec[ok1 ? lidx(e1) : f];
        if (ok0) lds_add(pool + lpos(e0), x0);
        if (ok1) lds_add(pool + lpos(e1), x1);
    }
    wave_sync();
    bool const act = gi < (tk.cls_cnt & 0xff);
    int const m = act ? pre.gf.z & 0xffff : 0, p = act ? pre.gf.z >> 16 : 0;
    int const p_max = wave_group_max<L>(p);
    if constexpr (MM <= 16) {
        // dense elimination in registers: lane t holds column t (rows 0..MM-1)
        // and w[t]; step k broadcasts column k and w[k] from the group's lane k.
        // The tile is cleared past the front, so rows and columns past m are
        // zero and stay zero under the updates (no bounds masks needed).
        T col[MM];
#pragma unroll
        for (int r = 0; r < MM; ++r) col[r] = F[r * ld + t];
        T wt = w[t];
#pragma unroll
        for (int k = 0; k < MM; ++k) {
            if (k >= p_max) continue;  // uniform; keeps the loop fully unrolled
            int const src = gi * L + k;
            T lk[MM];
#pragma unroll
            for (int i = k; i < MM; ++i) lk[i] = group_bcast<L>(col[i], k, src);
            T const wk = group_bcast<L>(wt, k, src);
            bool const act = k < p;
            T piv = act ? lk[k] : T(1);
            if (tabs(piv) < T(1e-14)) {
                if (act && t == 0) fa.fail[g] = 1;
                piv = T(1);
            }
            T const rp = pivot_recip(piv);
            if (act && t > k && t < m) {
                T const ukj = col[k];
                T lt = T(0);
#pragma unroll
                for (int i = k + 1; i < MM; ++i) {
                    col[i] -= (lk[i] * rp) * ukj;
                    if (i == t) lt = lk[i];
                }
                wt -= (lt * rp) * wk;
            }
        }
        if (t < m) {
            T* Gm = front + pre.gf.x;
#pragma unroll
            for (int r = 0; r < MM; ++r)
                if (r < m && t >= (r < p ? r : p)) Gm[r * m + t] = col[r];
            vec[pre.gf.y + t] = wt;
        }
        return;
    }
    for (int k = 0; k < p_max; ++k) {
        T piv = k < p ? F[k * ld + k] : T(1);
        bool const tiny = tabs(piv) < T(1e-14);
        if (tiny) {
            if (k < p && t == 0) fa.fail[g] = 1;
            piv = T(1);
        }
        T const rp = pivot_recip(piv);
        if (k < p && t > k && t < m) {
            T const ukj = F[k * ld + t];
            // column k is read, column t > k written: they never overlap, so
            // the reads of later rows need not wait for earlier rows' writes
            T const* __restrict__ ck = F + k;
            T* __restrict__ ct = F + t;
#pragma unroll 4
            for (int i = k + 1; i < m; ++i) ct[i * ld] -= (ck[i * ld] * rp) * ukj;
            w[t] -= (F[t * ld + k] * rp) * w[k];
        }
        wave_sync();
    }
    // (the multipliers below the diagonal are not scaled: the forward
    // elimination above used them unscaled, and L is never stored)
    if (t < m) {
        T* Gm = front + pre.gf.x;
        for (int r = 0; r < m; ++r)
            if (t >= (r < p ? r : p)) Gm[r * m + t] = F[r * ld + t];
        vec[pre.gf.y + t] = w[t];
    }
}

// Backward substitution of the pivot block of up to kWave / L fronts.
template <typename T, int MM, int L>
__device__ __attribute__((always_inline)) void grid_backward_task(FusedArgs const& a, FrontArgs const& fa, GridTask const& tk, TaskPre<T> const& pre,
                                   T const* front, T const* vec, T* xsol, T* pool, int ld) {
    constexpr int G = kWave / L, CH = MM < 16 ? MM : 16;
    int const lane = lane_id(), gi = lane / L, t = lane % L;
    int const rows = ld - 1, wbase = G * rows * ld;
    unsigned const so = pre.so;
    bool const sok = tk.sol_beg + lane < tk.sol_end;
    bool const act = gi < (tk.cls_cnt & 0xff);
    int const m = act ? pre.gf.z & 0xffff : 0, p = act ? pre.gf.z >> 16 : 0;
    // unconditional loads: the front record of an inactive lane is another
    // front's or padding
    // (a lane whose value is not used reads lane 0's element: same line)
    T const sv = xsol[sok ? lidx(so) : __builtin_amdgcn_readfirstlane(lidx(so))];
    T const wv = vec[t < p ? pre.gf.y + t : __builtin_amdgcn_readfirstlane(pre.gf.y)];
    int const p_max = wave_group_max<L>(p);
    // small fronts: the U rows by columns (lane t reads column t of each pivot
    // row, consecutive along the row), transposed through the group's LDS tile
    T ucol[MM <= 16 ? MM : 1];
    if constexpr (MM <= 16) {
        int const g0 = __builtin_amdgcn_readfirstlane(pre.gf.x);
#pragma unroll
        for (int r = 0; r < MM; ++r) {
            if (r >= p_max) continue;  // uniform
            bool const use = r < p && t < m;
            ucol[r] = front[use ? pre.gf.x + r * m + t : g0];
        }
    }
    T* F = pool + gi * rows * ld;
    T* w = pool + wbase + gi * MM;
    wave_sync();  // the previous task of this wavefront has finished with the region
    if (sok) pool[lpos(so)] = sv;
    for (int q = tk.sol_beg + kWave + lane; q < tk.sol_end; q += kWave) {
        unsigned const e = a.sol[q];
        pool[lpos(e)] = xsol[lidx(e)];
    }
    if (t < p) w[t] = wv;
    T xme = T(0);  // this lane's solution value (t < p)
    if constexpr (MM <= 16) {
        // lane t < p holds row t of U; x of the structure slots and the steps'
        // x_k come from the owning lanes by shuffles
#pragma unroll
        for (int r = 0; r < MM; ++r)
            if (r < p_max && r < p && t < m) F[r * ld + t] = ucol[r];
        wave_sync();
        T urow[MM];
#pragma unroll
        for (int q = 0; q < MM; ++q) urow[q] = t < p && q < m ? F[t * ld + q] : T(0);
        T acc = t < m ? w[t] : T(0);
        T xs[MM];
#pragma unroll
        for (int q = 0; q < MM; ++q) xs[q] = group_bcast<L>(acc, q, gi * L + q);
        if (t < p) {
#pragma unroll
            for (int q = 0; q < MM; ++q)
                if (q >= p && q < m) acc -= urow[q] * xs[q];
        }
#pragma unroll
        for (int k = MM - 1; k >= 0; --k) {
            if (k < p_max) {
                T const xk = group_bcast<L>(pivot_div(acc, urow[k]), k, gi * L + k);
                if (k < p) {
                    if (t < k) acc -= urow[k] * xk;
                    if (t == k) xme = xk;
                }
            }
        }
        if (t < p) xsol[pre.gf.w + t] = xme;
        return;
    }
    load_rows<T, MM, L, CH>(F, ld, front + pre.gf.x, m, p, t, t < m);
    wave_sync();
    if (t < p) {
        T acc = w[t];
        for (int q = p; q < m; ++q) acc -= F[t * ld + q] * w[q];
        w[t] = acc;
    }
    wave_sync();
    for (int k = p_max - 1; k >= 0; --k) {
        if (k < p) {
            T const xk = pivot_div(w[k], F[k * ld + k]);
            if (t < k) w[t] -= F[t * ld + k] * xk;
            if (t == k) xme = xk;
        }
        wave_sync();
    }
    if (t < p) xsol[pre.gf.w + t] = xme;
}

template <typename T, bool BWD>
__device__ inline void grid_task(FusedArgs const& a, FrontArgs const& fa, MisArgs const& ma, int g, int b0, int r0, GridTask const& tk,
                                 TaskPre<T> const& pre, TaskGath<T> const& gat, T* front, T* vec, T* xs, T* pool) {
    switch (tk.cls_cnt >> 8) {
        case 0:
            if constexpr (BWD) grid_backward_task<T, 4, 4>(a, fa, tk, pre, front, vec, xs, pool, 5);
            else grid_factor_task<T, 4, 4>(a, fa, ma, tk, pre, gat, g, b0, r0, front, vec, pool, 5);
            break;
        case 1:
            if constexpr (BWD) grid_backward_task<T, 8, 8>(a, fa, tk, pre, front, vec, xs, pool, 9);
            else grid_factor_task<T, 8, 8>(a, fa, ma, tk, pre, gat, g, b0, r0, front, vec, pool, 9);
            break;
        case 2:
            if constexpr (BWD) grid_backward_task<T, 16, 16>(a, fa, tk, pre, front, vec, xs, pool, 17);
            else grid_factor_task<T, 16, 16>(a, fa, ma, tk, pre, gat, g, b0, r0, front, vec, pool, 17);
            break;
        case 3:
            if constexpr (BWD) grid_backward_task<T, 32, 64>(a, fa, tk, pre, front, vec, xs, pool, 33);
            else grid_factor_task<T, 32, 64>(a, fa, ma, tk, pre, gat, g, b0, r0, front, vec, pool, 33);
            break;
        default:
            if constexpr (BWD) grid_backward_task<T, 64, 64>(a, fa, tk, pre, front, vec, xs, pool, a.ld4);
            else grid_factor_task<T, 64, 64>(a, fa, ma, tk, pre, gat, g, b0, r0, front, vec, pool, a.ld4);
    }
}

// The grid schedule's Y assembly, in its own slot order: the same
// per-slot arithmetic and contribution order as ybus_slot, with the branch
// (r, x, g, b) and shunt (g, b) parameters packed in the order the slots
// first use them, so a slot's reads are one record each and consecutive
// slots read consecutive records.  Where slot_pair is given, the slot's
// value is also written, in the factor type, at its pair (at most one).
struct YbusGridArgs {
    int nnz;
    const int* seg;
    const int* code;  // index into brp (branches) or app (shunts) << 2 | kind
    const int* slot_inst;
    const double4* brp;
    const double2* app;
    const double* zb;
    double2* y;
    const int* slot_pair;
    void* ypair;
};
struct YbusPackArgs {
    int n_br, n_app;
    const int* br_src;  // per record: its branch
    const int* app_src; // per record: its shunt
    const double* br_r;
    const double* br_x;
    const double* br_g;
    const double* br_b;
    const double* app_g;
    const double* app_b;
    double4* brp;
    double2* app;
};
__global__ void __launch_bounds__(256) k_ybus_pack(YbusPackArgs a) {
    int const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.n_br) {
        int const b = a.br_src[k];
        a.brp[k] = make_double4(a.br_r[b], a.br_x[b], a.br_g[b], a.br_b[b]);
    }
    if (k < a.n_app) {
        int const s = a.app_src[k];
        a.app[k] = make_double2(a.app_g[s], a.app_b[s]);
    }
}
template <typename T>
__global__ void __launch_bounds__(256) k_ybus_grid(YbusGridArgs a) {
#pragma clang fp contract(off)
    int const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nnz) return;
    int const s0 = a.seg[k], s1 = a.seg[k + 1];
    int const inst = a.slot_inst[k];
    int const q = a.slot_pair ? a.slot_pair[k] : -1;
    // The first kHead contributions' codes and records are all loaded before
    // any is summed, so a bus row of degree d costs three dependent round
    // trips instead of 1 + 2d; the sum itself runs in code order.
    constexpr int kHead = 4;
    int c[kHead];
#pragma unroll
    for (int u = 0; u < kHead; ++u) c[u] = s0 + u < s1 ? a.code[s0 + u] : 0;
    double4 rec[kHead];
#pragma unroll
    for (int u = 0; u < kHead; ++u) {
        if (s0 + u >= s1) continue;
        if ((c[u] & 3) == 0) {
            double2 const sh = a.app[c[u] >> 2];
            rec[u] = make_double4(sh.x, sh.y, 0.0, 0.0);
        } else {
            rec[u] = a.brp[c[u] >> 2];
        }
    }
    double re = 0.0, im = 0.0;
    auto add = [&](int code, double4 const& p) {
        int const kind = code & 3;
        if (kind == 0) {
            re += p.x;
            im += p.y;
        } else {
            double const r = p.x, x = p.y;
            double const z2 = r * r + x * x;
            double yr = 0.0, yi = 0.0;
            if (z2 > 1e-12) {
                yr = r / z2;
                yi = -x / z2;
            }
            if (kind == 2) {
                re = -yr;
                im = -yi;
            } else {
                double const bar = yr + p.z, bai = yi + 0.0;
                double const tr = bar + 0.0, ti = bai + p.w / 2.0;
                re += tr;
                im += ti;
            }
        }
    };
#pragma unroll
    for (int u = 0; u < kHead; ++u)
        if (s0 + u < s1) add(c[u], rec[u]);
    for (int s = s0 + kHead; s < s1; ++s) {
        int const cs = a.code[s];
        if ((cs & 3) == 0) {
            double2 const sh = a.app[cs >> 2];
            add(cs, make_double4(sh.x, sh.y, 0.0, 0.0));
        } else {
            add(cs, a.brp[cs >> 2]);
        }
    }
    double const zb = a.zb[inst];
    double2 const y = make_double2(re * zb, im * zb);
    a.y[k] = y;
    if (a.slot_pair) {
        if (q >= 0) {
            cplx_t<T>* yp = static_cast<cplx_t<T>*>(a.ypair);
            yp[q].x = static_cast<T>(y.x);
            yp[q].y = static_cast<T>(y.y);
        }
    }
}

// Per-unit Y of every pair in pair order, in the factor type, once per solve.
template <typename T>
__global__ void __launch_bounds__(256) k_gather_ypair(int n, const int* pair_y, const double2* y, cplx_t<T>* ypair) {
    int const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        double2 const v = y[pair_y[k]];
        ypair[k].x = static_cast<T>(v.x);
        ypair[k].y = static_cast<T>(v.y);
    }
}

// Progress of the wavefronts of a workgroup through a sweep: prog[w] is the
// index of the next task wavefront w will run (its tasks are w, w + NW, ...
// of the network's task list, in order).  Tasks are in level order, children
// before parents, so in the factorisation every task's children have smaller
// indices and in the backward substitution every task's parents larger ones;
// the wavefront whose next task comes first (last) in the sweep always finds
// its dependencies done, so every wait ends.  A task at index q waits, for
// each other wavefront, until it has passed the task nearest q that holds a
// child (parent) of the task's fronts, from the per-residue distances of the
// descriptor; with 1, 2 or 4 wavefronts a wavefront runs the tasks of one or
// two residues, with 8 the task waits for every wavefront to pass the
// farthest such task.
__device__ inline void publish_progress(int* prog, int wave, int next) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this task's stores before the new index
    if (lane_id() == 0) __hip_atomic_store(prog + wave, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <bool BWD, int NW>
__device__ inline void wait_for_deps(int const* prog, int wave, GridTask const& tk, int q) {
    static_assert(kDepWaves == 4 && (NW == 1 || NW == 2 || NW == 4 || NW == 8), "wavefronts per network");
    unsigned const d01 = BWD ? tk.dep_b[0] : tk.dep_f[0], d23 = BWD ? tk.dep_b[1] : tk.dep_f[1];
    if (NW == 1 || (d01 | d23) == 0u) return;
    int const o[4] = {static_cast<int>(d01 & 0xffff), static_cast<int>(d01 >> 16), static_cast<int>(d23 & 0xffff),
                      static_cast<int>(d23 >> 16)};
    // the task index a wavefront must have passed, over residues r of mask
    auto need = [&](unsigned mask, int& dep) {
        bool any = false;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if ((mask >> r & 1) && o[r]) {
                int const d = BWD ? q + o[r] : q - o[r];
                dep = !any ? d : (BWD ? min(dep, d) : max(dep, d));
                any = true;
            }
        return any;
    };
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w == wave) continue;
        int dep = 0;
        if (!need(NW == 4 ? 1u << w : NW == 2 ? 5u << w : 15u, dep)) continue;
        for (;;) {
            int const at = __hip_atomic_load(prog + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (BWD ? at < dep : at > dep) break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // then the producers' stores are visible
}

// This wavefront's tasks of one sweep over the network's task list [lo, hi):
// ascending from q = lo + wave (factorisation) or descending from
// q = hi - 1 - wave (backward), NW apart, without a barrier between levels.
// While a task runs, the next task's descriptor-dependent loads and the
// descriptor after that are in flight (unrolled by two, so no value is copied
// while its load is outstanding), across level boundaries.  Descriptors read
// outside [lo, hi) are another network's or padding; only their static lists
// are loaded.
template <typename T, bool BWD, int NW>
__device__ inline void run_sweep(FusedArgs const& a, FrontArgs const& fa, MisArgs const& ma,
                                 const GridTask* __restrict__ tasks, int g, int b0, int r0, int wave, int lo, int hi,
                                 T* front, T* vec, T* xs, T* pool, int* prog) {
    constexpr int S = BWD ? -NW : NW;
    int q = BWD ? hi - 1 - wave : lo + wave;
    auto inside = [&](int k) { return BWD ? k >= lo : k < hi; };
    auto desc = [&](int k) { return tasks[BWD ? max(k, 0) : k]; };
    if (!inside(q)) return;
    GridTask t0 = tasks[q], t1 = desc(q + S);
    TaskPre<T> p0 = load_pre<T, BWD>(a, t0);
    auto gathers = [&](GridTask const& tk, TaskPre<T> const& pre) {
        if constexpr (BWD) return TaskGath<T>{};
        else return factor_gathers<T>(a, ma, tk, pre, b0, r0);
    };
    // (a task's own independent gathers and the next task's lists are in
    // flight while it waits for its dependencies)
    for (;;) {
        TaskGath<T> const g0 = gathers(t0, p0);
        TaskPre<T> const p1 = load_pre<T, BWD>(a, t1);
        GridTask const t2 = desc(q + 2 * S);
        wait_for_deps<BWD, NW>(prog, wave, t0, q);
        grid_task<T, BWD>(a, fa, ma, g, b0, r0, t0, p0, g0, front, vec, xs, pool);
        q += S;
        publish_progress(prog, wave, q);
        if (!inside(q)) break;
        TaskGath<T> const g1 = gathers(t1, p1);
        TaskPre<T> const p2 = load_pre<T, BWD>(a, t2);
        GridTask const t3 = desc(q + 2 * S);
        wait_for_deps<BWD, NW>(prog, wave, t1, q);
        grid_task<T, BWD>(a, fa, ma, g, b0, r0, t1, p1, g1, front, vec, xs, pool);
        q += S;
        publish_progress(prog, wave, q);
        if (!inside(q)) break;
        t0 = t2;
        t1 = t3;
        p0 = p2;
    }
}

template <int NW>
__device__ inline double block_max(double m, double* red) {
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    int const wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[wave] = m;
    __syncthreads();
    m = red[0];
    for (int k = 1; k < NW; ++k) m = fmax(m, red[k]);
    __syncthreads();
    return m;
}

// One workgroup per network; dynamic LDS = NW * pool elements of T.  The
// memory waits dominate, so resident wavefronts pay: with FP32 factors the
// kernel is held to 96 VGPRs (5 wavefronts per SIMD; a few values spill), with
// FP64 factors to 168 (3 per SIMD, its natural count).
//
// The task descriptors and level ranges (a.tasks, a.lvl_ptr) come again as
// restrict arguments: the kernel writes no memory they could alias, so the
// compiler reads them with scalar loads, which are waited for on their own
// counter.  Read as vector loads of a uniform address, each was followed by a
// wait for every vector load in flight, the next task's prefetches included.
template <typename T, int NW>
__global__ void __launch_bounds__(NW* kWave, sizeof(T) == 4 ? 5 : 3)
    k_nr_grid(FusedArgs a, MisArgs ma, UpdArgs ua, FrontArgs fa, T* front, T* vec, const GridTask* __restrict__ tasks,
              const int* __restrict__ lvl_ptr) {
    constexpr int NT = NW * kWave;
    extern __shared__ __align__(16) unsigned char grid_smem[];
    __shared__ double red[NW];
    __shared__ int prog_f[NW], prog_b[NW];  // per wavefront: next task of the factor / backward sweep
    int const g = a.order[a.order_off + blockIdx.x];
    int const b0 = a.bus_off[g], b1 = a.bus_off[g + 1];
    int const t_lo = lvl_ptr[a.lvl_beg[g]], t_hi = lvl_ptr[a.lvl_beg[g + 1]];  // the network's tasks
    int const r0 = a.var_off[g], r1 = a.var_off[g + 1];                       // and its Newton equations
    // the network's fronts, vectors and solution list (the grid lists index from them)
    int4 const nb = a.netbase[g];
    T* const front_n = front + nb.x;
    T* const vec_n = vec + nb.y;
    T* const xs_n = static_cast<T*>(a.xs) + nb.z;
    int const tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    T* pool = reinterpret_cast<T*>(grid_smem) + wave * a.pool;
    for (int it = 0; it < a.max_it; ++it) {
        // cleared before block_max's barriers, which order it before any
        // factor task of this iteration can set it
        if (tid == 0) a.lu_fail[g] = 0;
        if ((tid & (kWave - 1)) == 0) {
            prog_f[wave] = t_lo + wave;
            prog_b[wave] = t_hi - 1 - wave;
        }
        double m = 0.0;
        constexpr int MB = sizeof(T) == 8 ? 2 : 1;  // buses per thread and step
        for (int i = b0 + tid; i < b1; i += MB * NT) {
            int ib[MB];
#pragma unroll
            for (int b = 0; b < MB; ++b) ib[b] = min(i + b * NT, b1 - 1);
            m = fmax(m, grid_mismatch_buses<T, MB>(ma, ib, i + (MB - 1) * NT < b1, r0, r0 + nb.w));
        }
        m = block_max<NW>(m, red);  // its barriers also publish S and F
        if (r0 < r1) m = reference_max(m, ma.f[r0]);
        if (tid == 0) {
            a.maxf[static_cast<long long>(it) * a.n_inst + g] = dbits(m);
            a.final_mis[g] = m;
        }
        if (m < a.tol) {
            if (tid == 0) {
                a.done[g] = 1;
                a.iters[g] = it + 1;
            }
            break;
        }
        // (block_max's barriers also published the progress counters)
        run_sweep<T, false, NW>(a, fa, ma, tasks, g, b0, r0, wave, t_lo, t_hi, front_n, vec_n, xs_n, pool, prog_f);
        __syncthreads();
        run_sweep<T, true, NW>(a, fa, ma, tasks, g, b0, r0, wave, t_lo, t_hi, front_n, vec_n, xs_n, pool, prog_b);
        __syncthreads();
        bool const zero = a.lu_fail[g] != 0;
        for (int i = b0 + tid; i < b1; i += NT) grid_update_bus<T>(ua, i, zero);
        __syncthreads();
    }
    // V back in the plan's bus order (each thread's buses are the ones its
    // last update wrote)
    for (int i = b0 + tid; i < b1; i += NT) ua.v_out[ua.perm[i]] = ua.v[i];
}

}  // namespace gaphip
