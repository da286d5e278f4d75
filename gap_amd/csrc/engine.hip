// Device engine and C ABI of the gfx950 Newton-Raphson power-flow backend.
//
// A plan packs a batch of networks end to end on the device: buses, Y-bus
// slots, Newton equations/variables, Jacobian slots and multifrontal fronts
// all carry global indices, so every kernel launch serves the whole batch and
// the level schedule of the fronts is the union of the per-network schedules.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <queue>
#include <limits>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gap_hip.h"
#include "kernels.hpp"
#include "fused.hpp"
#include "evaluate.hpp"
#include "symbolic.hpp"

namespace gaphip {

thread_local std::string g_last_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, std::string const& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw Error(GAPHIP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(DBuf const&) = delete;
    DBuf& operator=(DBuf const&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        n = count;
        if (count) HIPCHK(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(std::vector<T> const& h, hipStream_t s) {
        alloc(h.size());
        if (!h.empty()) HIPCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

inline int blocks_for(long long n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

// Kernel ids of the timing hook: one per kernel (front kernels per size class).
enum Kid { K_YBUS = 0, K_FLAT, K_MISMATCH, K_CHECK, K_JAC, K_UPDATE, K_FACTOR, K_FWD = K_FACTOR + 7, K_BWD = K_FWD + 7,
           K_FUSED = K_BWD + 7, K_COUNT };
static char const* const kKernelNames[K_COUNT] = {
    "k_ybus", "k_flat", "k_mismatch", "k_check", "k_jacobian", "k_update",
    "k_front_factor_lds<4>", "k_front_factor_lds<8>", "k_front_factor_lds<16>", "k_front_factor_lds<32>",
    "k_front_factor_lds<64>", "k_front_factor_lds<128>", "k_big_factor (assemble+panel+trsm+gemm)",
    "k_front_fwd_lds<4>", "k_front_fwd_lds<8>", "k_front_fwd_lds<16>", "k_front_fwd_lds<32>",
    "k_front_fwd_lds<64>", "k_front_fwd_lds<128>", "k_front_fwd",
    "k_front_bwd_lds<4>", "k_front_bwd_lds<8>", "k_front_bwd_lds<16>", "k_front_bwd_lds<32>",
    "k_front_bwd_lds<64>", "k_front_bwd_lds<128>", "k_front_bwd", "k_nr_grid"};

Net copy_net(gaphip_net const& in) {
    if (in.n_bus <= 0) throw Error(GAPHIP_E_ARG, "network has no buses");
    if (in.n_branch < 0 || in.n_app < 0) throw Error(GAPHIP_E_ARG, "negative component count");
    auto need = [](void const* p, int n, char const* what) {
        if (n > 0 && !p) throw Error(GAPHIP_E_ARG, std::string("missing array ") + what);
    };
    need(in.bus_id, in.n_bus, "bus_id");
    need(in.u_rated, in.n_bus, "u_rated");
    need(in.bus_type, in.n_bus, "bus_type");
    need(in.u_pu, in.n_bus, "u_pu");
    need(in.p_w, in.n_bus, "p_w");
    need(in.q_var, in.n_bus, "q_var");
    need(in.br_id, in.n_branch, "br_id");
    need(in.br_from, in.n_branch, "br_from");
    need(in.br_to, in.n_branch, "br_to");
    need(in.br_status, in.n_branch, "br_status");
    need(in.br_r, in.n_branch, "br_r");
    need(in.br_x, in.n_branch, "br_x");
    need(in.app_id, in.n_app, "app_id");
    need(in.app_node, in.n_app, "app_node");
    need(in.app_status, in.n_app, "app_status");
    need(in.app_type, in.n_app, "app_type");
    need(in.app_p, in.n_app, "app_p");
    need(in.app_q, in.n_app, "app_q");
    need(in.app_g, in.n_app, "app_g");
    need(in.app_b, in.n_app, "app_b");
    Net n;
    auto cp = [](auto const* p, int k, auto& v) { v.assign(p, p + k); };
    cp(in.bus_id, in.n_bus, n.bus_id);
    cp(in.u_rated, in.n_bus, n.u_rated);
    cp(in.bus_type, in.n_bus, n.bus_type);
    cp(in.u_pu, in.n_bus, n.u_pu);
    cp(in.p_w, in.n_bus, n.p_w);
    cp(in.q_var, in.n_bus, n.q_var);
    cp(in.br_id, in.n_branch, n.br_id);
    cp(in.br_from, in.n_branch, n.br_from);
    cp(in.br_to, in.n_branch, n.br_to);
    cp(in.br_status, in.n_branch, n.br_status);
    cp(in.br_r, in.n_branch, n.br_r);
    cp(in.br_x, in.n_branch, n.br_x);
    if (in.br_g) cp(in.br_g, in.n_branch, n.br_g); else n.br_g.assign(in.n_branch, 0.0);
    if (in.br_b) cp(in.br_b, in.n_branch, n.br_b); else n.br_b.assign(in.n_branch, 0.0);
    if (in.n_app > 0) {
        cp(in.app_id, in.n_app, n.app_id);
        cp(in.app_node, in.n_app, n.app_node);
        cp(in.app_status, in.n_app, n.app_status);
        cp(in.app_type, in.n_app, n.app_type);
        cp(in.app_p, in.n_app, n.app_p);
        cp(in.app_q, in.n_app, n.app_q);
        cp(in.app_g, in.n_app, n.app_g);
        cp(in.app_b, in.n_app, n.app_b);
    }
    return n;
}

// Builds the LU input of a Newton system: nodes are the non-slack buses, a PQ
// node carries (theta <-> P, |V| <-> Q), a PV node (theta <-> P); entries are
// the structural Jacobian slots in reference CSR order.
LUInput nr_lu_input(Net const& net, YPlan const& y, NRPlan const& p, std::vector<int>& node_bus) {
    int const n = net.n_bus();
    std::vector<int> node_of(n, -1);
    node_bus.clear();
    for (int i = 0; i < n; ++i)
        if (net.bus_type[i] != kSlack) {
            node_of[i] = static_cast<int>(node_bus.size());
            node_bus.push_back(i);
        }
    LUInput in;
    in.n_nodes = static_cast<int>(node_bus.size());
    in.w.resize(in.n_nodes);
    for (int k = 0; k < in.n_nodes; ++k) in.w[k] = net.bus_type[node_bus[k]] == kPQ ? 2 : 1;
    in.adj_ptr.assign(in.n_nodes + 1, 0);
    std::vector<int> mark(in.n_nodes, -1);
    for (int k = 0; k < in.n_nodes; ++k) {
        int i = node_bus[k];
        mark[k] = k;
        for (int q = y.rp[i]; q < y.rp[i + 1]; ++q) {
            int u = node_of[y.ci[q]];
            if (u < 0 || mark[u] == k) continue;
            mark[u] = k;
            in.adj.push_back(u);
        }
        in.adj_ptr[k + 1] = static_cast<int>(in.adj.size());
    }
    // Y is pattern-symmetric (both directions of every branch), so adj is too.
    std::vector<int> eq_key(p.n_var), var_key(p.n_var);
    for (int i = 0; i < n; ++i) {
        int nd = node_of[i];
        if (p.peq[i] >= 0) eq_key[p.peq[i]] = 2 * nd;
        if (p.qeq[i] >= 0) eq_key[p.qeq[i]] = 2 * nd + 1;
        if (p.ang[i] >= 0 && nd >= 0) var_key[p.ang[i]] = 2 * nd;
        if (p.mag[i] >= 0) var_key[p.mag[i]] = 2 * nd + 1;
    }
    int const nnz = p.jrp[p.n_var];
    in.ent_row.resize(nnz);
    in.ent_col.resize(nnz);
    for (int r = 0; r < p.n_var; ++r)
        for (int q = p.jrp[r]; q < p.jrp[r + 1]; ++q) {
            in.ent_row[q] = eq_key[r];
            in.ent_col[q] = var_key[p.jci[q]];
        }
    return in;
}

// Front size classes: LDS-resident kernels for fronts of up to 4 / 8 / 16 /
// 32 / 64 / 128 slots (16, 8, 4, 2, 1, 1 fronts per workgroup), the blocked
// global-memory path above that.
constexpr int kNumClasses = 7;
constexpr int kClassMax[kNumClasses - 1] = {4, 8, 16, 32, 64, 128};
constexpr int kClassPerBlock[kNumClasses - 1] = {16, 8, 4, 2, 1, 1};  // fronts per workgroup
constexpr int kGridClasses = 5;  // classes the grid schedule handles (fronts <= 64 slots)
// fronts per grid task of each class (4, 8, 16 lanes per front; a whole
// wavefront per front of up to 32 or 64 slots)
constexpr int kGridPerTask[kGridClasses] = {16, 8, 4, 1, 1};
inline int front_class(int m) {
    for (int c = 0; c < kNumClasses - 1; ++c)
        if (m <= kClassMax[c]) return c;
    return kNumClasses - 1;
}

// bins of gaphip_plan_stats::front_hist: <= 2, 4, 8, ..., 128 slots, above
inline int front_hist_bin(int m) {
    int b = 0;
    for (int cap = 2; b < 7 && m > cap; cap <<= 1) ++b;
    return b;
}

// Device copy of a (possibly batched) multifrontal plan.  Fronts are numbered
// in (level, size class) order so a launch covers a contiguous id range.
struct DevFronts {
    // optional second stream and two events for the look-ahead of the
    // two-level large-front factorisation (null: everything on one stream)
    hipStream_t side = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    int n_fronts = 0, n_levels = 0;
    std::vector<int> level_ptr;  // host copy of the level schedule
    // large fronts: outer panel width and the front size from which the
    // two-level panels apply (gaphip_tuning::big_outer, big_two_level_min)
    int big_outer = kBigOuter, big_two_level_min = kTwoLevelMin;
    // grid task layout (gaphip_plan_options::grid_min_class, grid_pack)
    int grid_min_class = 0;
    bool grid_pack = true;
    // fronts of level l and class c: ids [cls_ptr[l*K+c], cls_ptr[l*K+c+1])
    std::vector<int> cls_ptr;
    std::vector<long long> cls_arrow, cls_factor, cls_slots;  // per (level, class) totals
    std::vector<int> cls_mmax, cls_pmax;                      // per (level, class) largest m / p
    DBuf<FrontMeta> meta;
    DBuf<ChildInfo> cinfo;
    DBuf<int> rel, arrow_src, arrow_dst, slot_row, slot_col;
    DBuf<int> bigbase, arow, crow, crow_ci;  // row maps of fronts above 128 slots
    long long front_entries = 0, vec_entries = 0, factor_entries = 0;
    int max_front = 0;
    // per instance: arrowhead entries, factor entries (p*m + (m-p)*p summed), slots
    std::vector<long long> inst_arrow, inst_factor, inst_slots;
    // grid schedule (every front <= 64 slots): per instance its levels
    // [fz_lvl_beg[g], fz_lvl_beg[g+1]) of fz_lvl_ptr, each a range of fz_tasks,
    // and the flat assembly lists the tasks index
    bool fz_ok = false;
    int fz_pool = 0, fz_ld4 = 1;
    DBuf<int> fz_order, fz_lvl_beg, fz_lvl_ptr;
    DBuf<GridTask> fz_tasks;
    DBuf<int4> fz_gfront;  // per front: storage offset, vector offset, m | p << 16, list position of its first pivot
    DBuf<int2> fz_pairs;
    DBuf<int> fz_pair_y;
    DBuf<int> fz_slot_pair;  // per Y slot (grid order): its pair, or -1 (when no slot has two)
    DBuf<double2> fz_ypair;  // Y value of each pair, refreshed at every solve
    long long fz_npairs = 0;
    // -F of each Newton equation (factor type), written by every mismatch
    // pass; host: the equation of each right-hand-side entry, read once by the plan
    DBuf<double> fz_fe;
    std::vector<int> fz_rhs_eq;
    // solution of each pivot, in right-hand-side list order (factor type;
    // doubles: room for either), written by the backward tasks; host: the
    // list position of each Newton variable, read once by the plan
    DBuf<double> fz_xs;
    std::vector<int> fz_varpos;
    DBuf<unsigned> fz_rhs, fz_ext, fz_vx, fz_sol;  // index | LDS position << kListIdxBits
    DBuf<int4> fz_netbase;  // per network: first front storage entry, first vector entry, first list entry
    long long fz_list_bytes = 0;
    std::vector<long long> fz_work;  // front slots + factor entries of fz_order's networks (descending)

    BigMaps big() const { return BigMaps{bigbase.p, arow.p, crow.p, crow_ci.p}; }

    FrontArgs args(int const* done, int* fail) const {
        FrontArgs a;
        a.meta = meta.p;
        a.cinfo = cinfo.p;
        a.rel = rel.p;
        a.arrow_src = arrow_src.p;
        a.arrow_dst = arrow_dst.p;
        a.slot_row = slot_row.p;
        a.slot_col = slot_col.p;
        a.done = done;
        a.fail = fail;
        return a;
    }
};

// Jacobian fill items of a plan (global bus / Y / J indices), the input of the
// grid schedule's pair lists.
struct GridItems {
    std::vector<int> const& it_i;
    std::vector<int> const& it_j;
    std::vector<int> const& it_y;
    std::vector<int> const& it_pos;
    std::vector<int> const& bus_off;  // per instance, n_inst + 1 entries
    std::vector<int> const& var_off;  // per instance first Newton equation / variable, n_inst + 1 entries
    long long y_nnz;                  // Y-bus entries of the plan
    // the grid schedule's own bus and equation order (built in build_plan)
    std::vector<int> const& bus_new;  // per bus: its local index in that order
    std::vector<int> const& eq_new;   // per Newton equation: its local index in that order
    std::vector<int> const& var_bus;  // per Newton variable: its bus, as a global index in that order
    std::vector<int> const& slot_new; // per Y slot: its slot in that order
};

// The grid schedule's pair list holds bus indices local to their network in
// 16 bits each, so it takes networks of up to this many buses.
constexpr int kGridMaxBuses = 0xffff;

// Grid schedule of a packed plan: tasks (runs of at most kGridPerTask fronts of
// one instance, level and class) and the flat assembly lists each task indexes.
// Within a (level, class) bucket the fronts of one instance are contiguous
// (the renumbering is stable), so a task is a contiguous id range.  Every
// index and LDS position is checked here.
void build_grid(DevFronts& d, std::vector<LUPlan const*> const& plans, GridItems const& it,
                std::vector<FrontMeta> const& meta, std::vector<ChildInfo> const& cinfo, std::vector<int> const& rel,
                std::vector<int> const& arrow_src, std::vector<int> const& arrow_dst,
                std::vector<int> const& slot_row, std::vector<int> const& slot_col, long long n_index,
                long long n_values, hipStream_t s) {
    int const G = static_cast<int>(plans.size());
    int const nf = d.n_fronts, K = kNumClasses;
    constexpr long long kIntMax = 0x7fffffffLL;
    d.fz_rhs_eq.clear();
    d.fz_varpos.clear();
    d.fz_ok = nf > 0 && d.max_front <= kClassMax[kGridClasses - 1] && d.front_entries < kIntMax &&
              d.vec_entries < kIntMax && n_index < kIntMax && n_values < kIntMax;
    for (int g = 0; g < G && d.fz_ok; ++g)
        if (it.bus_off[g + 1] - it.bus_off[g] > kGridMaxBuses) d.fz_ok = false;
    if (!d.fz_ok) return;
    // Jacobian entry -> (front, position) from the arrowhead lists
    std::vector<int> jfront(n_values, -1), jdst(n_values, 0);
    for (int n = 0; n < nf; ++n)
        for (int q = meta[n].arrow_beg; q < meta[n].arrow_end; ++q) {
            jfront[arrow_src[q]] = n;
            jdst[arrow_src[q]] = arrow_dst[q];
        }
    // one pair per Jacobian 2x2 block: its front, corner (P row, theta column) and entry mask
    long long const n_items = static_cast<long long>(it.it_i.size());
    std::vector<int> pf(n_items, -1), pcorner(n_items, 0), pmask(n_items, 0), fpair_ptr(nf + 1, 0);
    for (long long t = 0; t < n_items; ++t) {
        int f = -1, r0 = 0, c0 = 0, mask = 0;
        for (int q = 0; q < 4; ++q) {
            int const pos = it.it_pos[4 * t + q];
            if (pos < 0) continue;
            int const fr = jfront[pos];
            if (fr < 0) throw Error(GAPHIP_E_ARG, "Jacobian entry without a front");
            int const r = (jdst[pos] >> 16) - (q >> 1), c = (jdst[pos] & 0xffff) - (q & 1);
            if (f < 0) {
                f = fr, r0 = r, c0 = c;
            } else if (fr != f || r != r0 || c != c0) {
                d.fz_ok = false;  // 2x2 block split over fronts or slots: not expressible as one pair
                return;
            }
            mask |= 1 << q;
        }
        if (f < 0) continue;
        if (r0 < 0 || c0 < 0) throw Error(GAPHIP_E_ARG, "Jacobian block corner outside its front");
        pf[t] = f;
        pcorner[t] = (r0 << 16) | c0;
        pmask[t] = mask;
        fpair_ptr[f + 1]++;
    }
    for (int n = 0; n < nf; ++n) fpair_ptr[n + 1] += fpair_ptr[n];
    std::vector<int> fpair(fpair_ptr[nf]);
    {
        std::vector<int> fill(fpair_ptr.begin(), fpair_ptr.end() - 1);
        for (long long t = 0; t < n_items; ++t)
            if (pf[t] >= 0) fpair[fill[pf[t]]++] = static_cast<int>(t);
    }
    // tasks per (instance, level), classes in increasing order
    std::vector<int> lvl_beg(G + 1, 0);
    for (int g = 0; g < G; ++g) lvl_beg[g + 1] = lvl_beg[g] + plans[g]->n_levels;
    // the fronts of each (instance, level), by size class
    std::vector<std::array<std::vector<int>, kGridClasses>> byc(lvl_beg[G]);
    int m4 = 0;
    for (int l = 0; l < d.n_levels; ++l)
        for (int c = 0; c < kGridClasses; ++c) {
            size_t const b = static_cast<size_t>(l) * K + c;
            for (int n = d.cls_ptr[b]; n < d.cls_ptr[b + 1]; ++n) {
                int const g = meta[n].inst;
                if (l >= plans[g]->n_levels) throw Error(GAPHIP_E_ARG, "front level outside its network's tree");
                int const ce = std::max(c, d.grid_min_class);
                byc[lvl_beg[g] + l][ce].push_back(n);
                if (ce == kGridClasses - 1) m4 = std::max(m4, meta[n].m);
            }
        }
    d.fz_ld4 = m4 + 1;
    // the fronts of a bucket by their first pivot's bus in the schedule's bus
    // order, so a task's fronts gather nearby buses (a level's fronts are
    // independent: their order changes no arithmetic)
    {
        auto key = [&](int n) {
            FrontMeta const& fm = meta[n];
            if (fm.p <= 0) return 0x7fffffff;
            int const v = slot_col[fm.slot_beg];
            if (v < 0 || v >= n_index) throw Error(GAPHIP_E_ARG, "pivot variable out of range");
            return it.var_bus[v];
        };
        for (auto& bc : byc)
            for (auto& F : bc) std::stable_sort(F.begin(), F.end(), [&](int x, int y) { return key(x) < key(y); });
    }
    // Tasks of each (instance, level): classes from the largest down, and the
    // free lane groups of a class's last task take fronts of the smaller
    // classes (a front fits a larger class's group).  A task costs about the
    // same whatever its fill (its time is memory round trips), so fewer tasks
    // per level means fewer rounds for the level's wavefronts.  tfront lists
    // the fronts in task order; a task is a range of it.
    std::vector<int> tfront;
    tfront.reserve(nf);
    std::vector<std::vector<int2>> per(lvl_beg[G]);
    for (size_t e = 0; e < per.size(); ++e) {
        auto& bc = byc[e];
        for (int c = kGridClasses - 1; c >= 0; --c) {
            int const per_task = kGridPerTask[c];
            auto const& F = bc[c];
            size_t i = 0;
            while (i < F.size()) {
                int const beg = static_cast<int>(tfront.size());
                int cnt = 0;
                for (; i < F.size() && cnt < per_task; ++i, ++cnt) tfront.push_back(F[i]);
                for (int c2 = c - 1; d.grid_pack && c2 >= 0 && cnt < per_task; --c2) {
                    auto& F2 = bc[c2];
                    for (; !F2.empty() && cnt < per_task; ++cnt) {
                        tfront.push_back(F2.back());
                        F2.pop_back();
                    }
                }
                per[e].push_back(make_int2(beg, (c << 8) | cnt));
            }
        }
    }
    if (static_cast<int>(tfront.size()) != nf) throw Error(GAPHIP_E_ARG, "grid tasks do not cover every front once");
    std::vector<int> lvl_ptr(1, 0);
    std::vector<GridTask> tasks;
    std::vector<int2> pairs;   // (local i | local j << 16, LDS position | entry mask << 16)
    std::vector<int> pair_y;   // Y-bus slot of each pair
    std::vector<int2> rhs, ext, vx, sol;
    std::vector<int> fpos(nf, -1);               // per front: list position of its first pivot
    std::vector<int> varpos(n_index, -1);        // per Newton variable: list position of its pivot
    // The grid schedule's own storage layout: fronts and slot vectors in task
    // order, so each network's are contiguous and every list entry indexes
    // from its network's base (netbase: fronts, vectors, right-hand-side list).
    std::vector<long long> goff(nf, -1), gvoff(nf, -1);
    std::vector<int4> netbase(G, make_int4(0, 0, 0, 0));
    {
        long long fo = 0, vo = 0;
        int k = 0;
        for (int g = 0; g < G; ++g) {
            netbase[g].x = static_cast<int>(fo);
            netbase[g].y = static_cast<int>(vo);
            for (int e = lvl_beg[g]; e < lvl_beg[g + 1]; ++e)
                for (int2 const& tr : per[e])
                    for (int gi = 0; gi < (tr.y & 0xff); ++gi, ++k) {
                        int const n = tfront[tr.x + gi];
                        if (n < 0 || n >= nf || meta[n].inst != g || tr.x + gi != k)
                            throw Error(GAPHIP_E_ARG, "grid tasks out of network order");
                        goff[n] = fo;
                        gvoff[n] = vo;
                        fo += static_cast<long long>(meta[n].m) * meta[n].m;
                        vo += meta[n].m;
                    }
            if (fo - netbase[g].x >= (1LL << kListIdxBits) || vo - netbase[g].y >= (1LL << kListIdxBits) ||
                it.var_off[g + 1] - it.var_off[g] >= (1 << kListIdxBits)) {
                d.fz_ok = false;  // a network too large for the 4-byte list entries
                return;
            }
        }
        if (k != nf || fo > d.front_entries || vo > d.vec_entries) throw Error(GAPHIP_E_ARG, "grid storage layout");
    }
    int pool = 1;
    for (int g = 0; g < G; ++g) {
      netbase[g].z = static_cast<int>(rhs.size());
      netbase[g].w = it.var_off[g] < it.var_off[g + 1] ? it.eq_new[it.var_off[g]] : 0;
      for (int e = lvl_beg[g]; e < lvl_beg[g + 1]; ++e) {
        auto const& lv = per[e];
        for (int2 const& tr : lv) {
            int const c = tr.y >> 8, cnt = tr.y & 0xff, L = kClassMax[c], Gc = kGridPerTask[c];
            int const ld = c == kGridClasses - 1 ? d.fz_ld4 : L + 1, rows = ld - 1;
            int const wbase = Gc * rows * ld, region = wbase + Gc * L;
            pool = std::max(pool, region);
            GridTask tk{};
            tk.front_beg = tr.x;
            tk.cls_cnt = tr.y;
            tk.pair_beg = static_cast<int>(pairs.size());
            tk.rhs_beg = static_cast<int>(rhs.size());
            tk.ext_beg = static_cast<int>(ext.size());
            tk.vx_beg = static_cast<int>(vx.size());
            tk.sol_beg = static_cast<int>(sol.size());
            for (int gi = 0; gi < cnt; ++gi) {
                int const n = tfront[tr.x + gi];
                FrontMeta const& fm = meta[n];
                if (front_class(fm.m) > c || fm.m > rows) throw Error(GAPHIP_E_ARG, "grid task class mismatch");
                int const tile = gi * rows * ld, wpos = wbase + gi * L;
                auto in_tile = [&](int r, int col) { return r >= 0 && col >= 0 && r < fm.m && col < fm.m; };
                for (int q = fpair_ptr[n]; q < fpair_ptr[n + 1]; ++q) {
                    int const t = fpair[q];
                    int const r0 = pcorner[t] >> 16, c0 = pcorner[t] & 0xffff, mask = pmask[t];
                    for (int e = 0; e < 4; ++e)
                        if ((mask >> e & 1) && !in_tile(r0 + (e >> 1), c0 + (e & 1)))
                            throw Error(GAPHIP_E_ARG, "Jacobian pair outside its front");
                    int const b0 = it.bus_off[fm.inst];
                    if (it.it_i[t] < b0 || it.it_j[t] < b0 || it.it_i[t] >= it.bus_off[fm.inst + 1] ||
                        it.it_j[t] >= it.bus_off[fm.inst + 1])
                        throw Error(GAPHIP_E_ARG, "pair bus outside its network");
                    int const li = it.bus_new[it.it_i[t]], lj = it.bus_new[it.it_j[t]];
                    if (li < 0 || lj < 0 || li > kGridMaxBuses || lj > kGridMaxBuses || li >= it.bus_off[fm.inst + 1] - b0 ||
                        lj >= it.bus_off[fm.inst + 1] - b0)
                        throw Error(GAPHIP_E_ARG, "pair bus outside its network");
                    pairs.push_back(make_int2(li | (lj << 16), (tile + r0 * ld + c0) | (mask << 16)));
                    if (it.it_y[t] < 0 || it.it_y[t] >= it.y_nnz) throw Error(GAPHIP_E_ARG, "pair out of range");
                    pair_y.push_back(it.slot_new[it.it_y[t]]);
                }
                fpos[n] = static_cast<int>(rhs.size());
                for (int t = 0; t < fm.p; ++t) {
                    int const v = slot_col[fm.slot_beg + t];
                    if (v < 0 || v >= n_index || varpos[v] >= 0) throw Error(GAPHIP_E_ARG, "pivot variable out of range");
                    varpos[v] = static_cast<int>(rhs.size());
                    int const r = slot_row[fm.slot_beg + t];
                    if (r < it.var_off[g] || r >= it.var_off[g + 1]) throw Error(GAPHIP_E_ARG, "pivot equation out of range");
                    rhs.push_back(make_int2(it.eq_new[r], wpos + t));
                }
                // (variable index here; its pivot's list position once every task is built)
                for (int t = fm.p; t < fm.m; ++t) sol.push_back(make_int2(slot_col[fm.slot_beg + t], wpos + t));
                for (int ci = fm.child_beg; ci < fm.child_end; ++ci) {
                    ChildInfo const& ch = cinfo[ci];
                    int const nu = ch.mc - ch.pc;
                    int const* rl = rel.data() + ch.rel_off;
                    if (ch.id < 0 || ch.id >= nf || meta[ch.id].inst != g || meta[ch.id].m != ch.mc)
                        throw Error(GAPHIP_E_ARG, "child front out of range");
                    long long const cf = goff[ch.id] - netbase[g].x, cv = gvoff[ch.id] - netbase[g].y;
                    for (int r = 0; r < nu; ++r) {
                        if (rl[r] < 0 || rl[r] >= fm.m) throw Error(GAPHIP_E_ARG, "extend-add map out of range");
                        for (int c2 = 0; c2 < nu; ++c2)
                            ext.push_back(make_int2(static_cast<int>(cf + static_cast<long long>(ch.pc + r) * ch.mc +
                                                                     ch.pc + c2),
                                                    tile + rl[r] * ld + rl[c2]));
                        vx.push_back(make_int2(static_cast<int>(cv + ch.pc + r), wpos + rl[r]));
                    }
                }
            }
            tk.pair_end = static_cast<int>(pairs.size());
            tk.rhs_end = static_cast<int>(rhs.size());
            tk.ext_end = static_cast<int>(ext.size());
            tk.vx_end = static_cast<int>(vx.size());
            tk.sol_end = static_cast<int>(sol.size());
            if (static_cast<long long>(ext.size()) >= kIntMax) throw Error(GAPHIP_E_ARG, "grid schedule lists too long");
            tasks.push_back(tk);
        }
        lvl_ptr.push_back(static_cast<int>(tasks.size()));
      }
    }
    // the backward tasks gather the solution of each structure slot at its
    // pivot's list position
    for (int g = 0; g < G; ++g) {
        int const t0 = lvl_ptr[lvl_beg[g]], t1 = lvl_ptr[lvl_beg[g + 1]];
        if (t0 == t1) continue;
        for (int q = tasks[t0].sol_beg; q < tasks[t1 - 1].sol_end; ++q) {
            int2& e = sol[q];
            if (e.x < 0 || e.x >= n_index || varpos[e.x] < netbase[g].z) throw Error(GAPHIP_E_ARG, "structure variable without a pivot");
            e.x = varpos[e.x] - netbase[g].z;
        }
    }
    // Dependencies between the tasks of a network, per wavefront of a sweep
    // of kDepWaves wavefronts: for each residue r of a task position (counted
    // from the network's first task in the factorisation, from its last in
    // the backward substitution; the wavefront that runs it), the distance
    // to the nearest task of residue r holding a child (factorisation) or a
    // parent (backward) of one of the task's fronts, 0 for none.  Tasks are in
    // level order and a child's level is below its parent's, so children come
    // before a task and parents after it, which the kernel's waits rely on to
    // end.
    {
        std::vector<int> task_of(nf, -1), parent(nf, -1);
        for (size_t k = 0; k < tasks.size(); ++k)
            for (int q = tasks[k].front_beg; q < tasks[k].front_beg + (tasks[k].cls_cnt & 0xff); ++q)
                task_of[tfront[q]] = static_cast<int>(k);
        for (int n = 0; n < nf; ++n)
            for (int ci = meta[n].child_beg; ci < meta[n].child_end; ++ci) {
                if (cinfo[ci].id < 0 || cinfo[ci].id >= nf) throw Error(GAPHIP_E_ARG, "child front out of range");
                parent[cinfo[ci].id] = n;
            }
        for (int g = 0; g < G; ++g) {
            int const t0 = lvl_ptr[lvl_beg[g]], t1 = lvl_ptr[lvl_beg[g + 1]];
            for (int k = t0; k < t1; ++k) {
                GridTask& tk = tasks[k];
                unsigned of[kDepWaves] = {}, ob[kDepWaves] = {};
                auto note = [&](unsigned* o, int r, long long dist) {
                    if (dist <= 0) throw Error(GAPHIP_E_ARG, "grid task dependencies out of level order");
                    if (dist > 0xffff) throw Error(GAPHIP_E_ARG, "grid task dependency too far");
                    o[r] = o[r] ? std::min<unsigned>(o[r], static_cast<unsigned>(dist)) : static_cast<unsigned>(dist);
                };
                for (int q = tk.front_beg; q < tk.front_beg + (tk.cls_cnt & 0xff); ++q) {
                    int const n = tfront[q];
                    for (int ci = meta[n].child_beg; ci < meta[n].child_end; ++ci) {
                        int const kc = task_of[cinfo[ci].id];
                        if (kc < t0 || kc >= t1) throw Error(GAPHIP_E_ARG, "child task outside its network");
                        note(of, (kc - t0) % kDepWaves, k - kc);
                    }
                    if (parent[n] >= 0) {
                        int const kp = task_of[parent[n]];
                        if (kp < t0 || kp >= t1) throw Error(GAPHIP_E_ARG, "parent task outside its network");
                        note(ob, (t1 - 1 - kp) % kDepWaves, kp - k);
                    }
                }
                for (int h = 0; h < kDepWaves / 2; ++h) {
                    tk.dep_f[h] = of[2 * h] | (of[2 * h + 1] << 16);
                    tk.dep_b[h] = ob[2 * h] | (ob[2 * h + 1] << 16);
                }
            }
        }
    }
    // bounds of every gathered index and LDS position
    for (size_t q = 0; q < pairs.size(); ++q)
        if (pair_y[q] < 0 || pair_y[q] >= it.y_nnz || (pairs[q].y & 0xffff) + d.fz_ld4 + 1 >= (1 << 16))
            throw Error(GAPHIP_E_ARG, "pair out of range");
    // list entries index from their network's base (below 2^kListIdxBits by
    // the layout check above; the bases leave room for every network after
    // them) and hold LDS positions inside the pool
    if (pool >= (1 << kListPosBits)) throw Error(GAPHIP_E_ARG, "grid task region too large");
    auto check2 = [&](std::vector<int2> const& v, char const* what) {
        for (auto const& e : v)
            if (e.x < 0 || e.x >= (1 << kListIdxBits) || e.y < 0 || e.y >= pool)
                throw Error(GAPHIP_E_ARG, std::string(what) + " out of range");
    };
    check2(rhs, "rhs entry");
    check2(sol, "solution entry");
    check2(ext, "extend-add entry");
    check2(vx, "vector entry");
    d.fz_pool = (pool + 3) / 4 * 4;  // whole 16-byte units per wavefront (the task clears its region with 16-byte stores)
    d.fz_rhs_eq.resize(rhs.size());
    for (int g = 0; g < G; ++g) {
        int const t0 = lvl_ptr[lvl_beg[g]], t1 = lvl_ptr[lvl_beg[g + 1]];
        if (t0 < t1)
            for (int q = tasks[t0].rhs_beg; q < tasks[t1 - 1].rhs_end; ++q) d.fz_rhs_eq[q] = rhs[q].x + it.var_off[g];
    }
    d.fz_varpos = std::move(varpos);
    // largest networks first: they bound the launch
    std::vector<int> order(G);
    for (int g = 0; g < G; ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return d.inst_factor[x] + d.inst_slots[x] > d.inst_factor[y] + d.inst_slots[y];
    });
    d.fz_work.resize(G);
    for (int k = 0; k < G; ++k) d.fz_work[k] = d.inst_factor[order[k]] + d.inst_slots[order[k]];
    // compact per-front records of the grid kernel, in task order (offsets fit
    // in int: fz_ok)
    std::vector<int4> gfront(nf);
    for (int k = 0; k < nf; ++k) {
        int const n = tfront[k];
        FrontMeta const& fm = meta[n];
        int const g = fm.inst;
        long long const lo = goff[n] - netbase[g].x, lv = gvoff[n] - netbase[g].y;
        if (lo < 0 || lo + static_cast<long long>(fm.m) * fm.m > (1LL << kListIdxBits) || lv < 0 ||
            lv + fm.m > (1LL << kListIdxBits) || fm.m < 0 || fm.m > 0x7fff || fm.p < 0 || fm.p > fm.m)
            throw Error(GAPHIP_E_ARG, "front record out of range");
        if (fm.p > 0 && (fpos[n] < netbase[g].z || fpos[n] + fm.p > static_cast<long long>(rhs.size())))
            throw Error(GAPHIP_E_ARG, "front pivots outside the right-hand-side list");
        // storage, vector and list offsets from the network's bases
        gfront[k] = make_int4(static_cast<int>(lo), static_cast<int>(lv), fm.m | (fm.p << 16),
                              std::max(fpos[n] - netbase[g].z, 0));
    }
    // padding read by the grid kernel's unconditional loads (never executed)
    tasks.insert(tasks.end(), kTaskPad, GridTask{});
    gfront.insert(gfront.end(), kWave, make_int4(0, 0, 0, 0));
    pairs.insert(pairs.end(), kListPad, make_int2(0, 0));
    pair_y.insert(pair_y.end(), kListPad, 0);
    // 4-byte list entries: index | LDS position << kListIdxBits; padding 0
    auto pack = [](std::vector<int2> const& v) {
        std::vector<unsigned> u(v.size() + kListPad, 0u);
        for (size_t q = 0; q < v.size(); ++q)
            u[q] = static_cast<unsigned>(v[q].x) | (static_cast<unsigned>(v[q].y) << kListIdxBits);
        return u;
    };
    std::vector<unsigned> const rhs4 = pack(rhs), ext4 = pack(ext), vx4 = pack(vx), sol4 = pack(sol);
    d.fz_list_bytes = static_cast<long long>(tasks.size()) * sizeof(GridTask) + pairs.size() * sizeof(int2) +
                      (rhs4.size() + ext4.size() + vx4.size() + sol4.size()) * sizeof(unsigned);
    d.fz_order.upload(order, s);
    d.fz_lvl_beg.upload(lvl_beg, s);
    d.fz_lvl_ptr.upload(lvl_ptr, s);
    d.fz_tasks.upload(tasks, s);
    d.fz_gfront.upload(gfront, s);
    d.fz_pairs.upload(pairs, s);
    d.fz_pair_y.upload(pair_y, s);
    // the Y assembly writes each pair's value itself where no slot feeds two
    // pairs (the list padding past the pairs is not a pair)
    std::vector<int> slot_pair(std::max<long long>(it.y_nnz, 1), -1);
    {
        bool one = true;
        for (size_t q = 0; q + kListPad < pair_y.size() && one; ++q) {
            if (slot_pair[pair_y[q]] >= 0) one = false;
            slot_pair[pair_y[q]] = static_cast<int>(q);
        }
        if (one) d.fz_slot_pair.upload(slot_pair, s);
        else d.fz_slot_pair.release();
    }
    d.fz_npairs = static_cast<long long>(pairs.size());
    d.fz_ypair.alloc(d.fz_npairs);
    HIPCHK(hipMemsetAsync(d.fz_ypair.p, 0, sizeof(double2) * d.fz_npairs, s));  // (the padding stays zero)
    d.fz_fe.alloc(rhs.size());  // one per Newton equation; doubles: room for either factor type
    HIPCHK(hipMemsetAsync(d.fz_fe.p, 0, sizeof(double) * rhs.size(), s));
    d.fz_xs.alloc(rhs.size() + kVecPad);
    HIPCHK(hipMemsetAsync(d.fz_xs.p, 0, sizeof(double) * (rhs.size() + kVecPad), s));
    d.fz_rhs.upload(rhs4, s);
    d.fz_ext.upload(ext4, s);
    d.fz_vx.upload(vx4, s);
    d.fz_sol.upload(sol4, s);
    d.fz_netbase.upload(netbase, s);
    HIPCHK(hipStreamSynchronize(s));
}

// Concatenate per-instance LU plans and renumber the fronts by (level, class).
// key_row / key_col map a slot key of instance g to its global right-hand-side /
// solution index; ent_base[g] is the offset of instance g's entries in the
// global value array.
template <typename KeyRow, typename KeyCol>
void pack_fronts(std::vector<LUPlan const*> const& plans, std::vector<long long> const& ent_base,
                 KeyRow key_row, KeyCol key_col, long long n_index, long long n_values, DevFronts& d,
                 hipStream_t s, GridItems const* items = nullptr) {
    int const G = static_cast<int>(plans.size());
    // old (instance-concatenated) numbering
    std::vector<int> o_inst, o_local, o_base(G + 1, 0);
    int nlev = 0;
    d.max_front = 0;
    d.factor_entries = 0;
    for (int g = 0; g < G; ++g) {
        o_base[g + 1] = o_base[g] + plans[g]->n_fronts;
        nlev = std::max(nlev, plans[g]->n_levels);
        d.max_front = std::max(d.max_front, plans[g]->max_front);
        d.factor_entries += plans[g]->factor_entries;
    }
    int const nf = o_base[G];
    o_inst.resize(nf);
    o_local.resize(nf);
    for (int g = 0; g < G; ++g)
        for (int k = 0; k < plans[g]->n_fronts; ++k) {
            o_inst[o_base[g] + k] = g;
            o_local[o_base[g] + k] = k;
        }
    auto M = [&](int o) { return plans[o_inst[o]]->f_m[o_local[o]]; };
    auto Lv = [&](int o) { return plans[o_inst[o]]->f_level[o_local[o]]; };
    for (int o = 0; o < nf; ++o)
        if (M(o) > 65535) throw Error(GAPHIP_E_ARG, "front larger than 65535 slots");
    int const K = kNumClasses;
    d.n_fronts = nf;
    d.n_levels = nlev;
    d.cls_ptr.assign(static_cast<size_t>(nlev) * K + 1, 0);
    for (int o = 0; o < nf; ++o) d.cls_ptr[static_cast<size_t>(Lv(o)) * K + front_class(M(o)) + 1]++;
    for (size_t b = 0; b + 1 < d.cls_ptr.size(); ++b) d.cls_ptr[b + 1] += d.cls_ptr[b];
    d.level_ptr.assign(nlev + 1, 0);
    for (int l = 0; l <= nlev; ++l) d.level_ptr[l] = d.cls_ptr[static_cast<size_t>(l) * K];
    std::vector<int> perm(nf), inv(nf);  // new -> old, old -> new (stable within a bucket)
    {
        std::vector<int> fp(d.cls_ptr.begin(), d.cls_ptr.end() - 1);
        for (int o = 0; o < nf; ++o) {
            int n = fp[static_cast<size_t>(Lv(o)) * K + front_class(M(o))]++;
            perm[n] = o;
            inv[o] = n;
        }
    }
    std::vector<FrontMeta> meta(nf);
    std::vector<ChildInfo> cinfo;
    std::vector<int> rel_off_new(nf, 0), rel, arrow_src, arrow_dst, slot_row, slot_col;
    long long off = 0, voff = 0;
    for (int n = 0; n < nf; ++n) {
        int const o = perm[n];
        LUPlan const& P = *plans[o_inst[o]];
        int const k = o_local[o];
        FrontMeta& fm = meta[n];
        fm.m = P.f_m[k];
        fm.p = P.f_p[k];
        fm.inst = o_inst[o];
        fm.off = off;
        fm.voff = voff;
        off += static_cast<long long>(fm.m) * fm.m;
        voff += fm.m;
        fm.slot_beg = static_cast<int>(slot_row.size());
        for (int t = P.f_slot_ptr[k]; t < P.f_slot_ptr[k + 1]; ++t) {
            slot_row.push_back(key_row(fm.inst, P.slot_key[t]));
            slot_col.push_back(key_col(fm.inst, P.slot_key[t]));
        }
        // arrowhead entries in row-major order of their targets
        fm.arrow_beg = static_cast<int>(arrow_src.size());
        std::vector<std::pair<int, int>> e;
        for (int q = P.f_arrow_ptr[k]; q < P.f_arrow_ptr[k + 1]; ++q)
            e.emplace_back(P.arrow_dst[q], static_cast<int>(ent_base[fm.inst] + P.arrow_src[q]));
        std::sort(e.begin(), e.end());
        for (auto const& pr : e) {
            arrow_dst.push_back(pr.first);
            arrow_src.push_back(pr.second);
        }
        fm.arrow_end = static_cast<int>(arrow_src.size());
        // this front's extend-add map into its parent
        rel_off_new[n] = static_cast<int>(rel.size());
        for (int t = P.f_rel_off[k]; t < P.f_rel_off[k + 1]; ++t) rel.push_back(P.rel[t]);
    }
    d.front_entries = off;
    d.vec_entries = voff;
    for (int n = 0; n < nf; ++n) {
        int const o = perm[n];
        LUPlan const& P = *plans[o_inst[o]];
        int const k = o_local[o];
        FrontMeta& fm = meta[n];
        fm.child_beg = static_cast<int>(cinfo.size());
        for (int q = P.f_child_ptr[k]; q < P.f_child_ptr[k + 1]; ++q) {
            int const cn = inv[o_base[o_inst[o]] + P.child[q]];
            ChildInfo ci{};
            ci.off = meta[cn].off;
            ci.voff = meta[cn].voff;
            ci.mc = meta[cn].m;
            ci.pc = meta[cn].p;
            ci.rel_off = rel_off_new[cn];
            ci.id = cn;
            if (cn >= n) throw Error(GAPHIP_E_ARG, "child must precede its parent");
            cinfo.push_back(ci);
        }
        fm.child_end = static_cast<int>(cinfo.size());
    }
    // per (level, class) totals
    d.cls_mmax.assign(static_cast<size_t>(nlev) * K, 0);
    d.cls_pmax.assign(static_cast<size_t>(nlev) * K, 0);
    d.cls_arrow.assign(static_cast<size_t>(nlev) * K, 0);
    d.cls_factor.assign(static_cast<size_t>(nlev) * K, 0);
    d.cls_slots.assign(static_cast<size_t>(nlev) * K, 0);
    for (size_t b = 0; b + 1 < d.cls_ptr.size(); ++b)
        for (int n = d.cls_ptr[b]; n < d.cls_ptr[b + 1]; ++n) {
            long long m = meta[n].m, pp = meta[n].p;
            d.cls_arrow[b] += meta[n].arrow_end - meta[n].arrow_beg;
            d.cls_factor[b] += pp * m + (m - pp) * pp;
            d.cls_slots[b] += m;
            d.cls_mmax[b] = std::max<int>(d.cls_mmax[b], static_cast<int>(m));
            d.cls_pmax[b] = std::max<int>(d.cls_pmax[b], static_cast<int>(pp));
        }
    // every index the front kernels dereference is checked here, on the host
    for (int n = 0; n < nf; ++n) {
        FrontMeta const& fm = meta[n];
        long long m = fm.m, p = fm.p;
        if (m <= 0 || p <= 0 || p > m) throw Error(GAPHIP_E_ARG, "front with invalid size");
        for (int q = fm.arrow_beg; q < fm.arrow_end; ++q) {
            if ((arrow_dst[q] >> 16) >= m || (arrow_dst[q] & 0xffff) >= m)
                throw Error(GAPHIP_E_ARG, "arrowhead map out of range");
            if (arrow_src[q] < 0 || arrow_src[q] >= n_values) throw Error(GAPHIP_E_ARG, "arrowhead source out of range");
        }
        for (int q = fm.child_beg; q < fm.child_end; ++q) {
            ChildInfo const& ci = cinfo[q];
            for (int t = 0; t < ci.mc - ci.pc; ++t)
                if (rel[ci.rel_off + t] < 0 || rel[ci.rel_off + t] >= m)
                    throw Error(GAPHIP_E_ARG, "extend-add map out of range");
        }
    }
    for (size_t q = 0; q < slot_row.size(); ++q)
        if (slot_row[q] < 0 || slot_col[q] < 0 || slot_row[q] >= n_index || slot_col[q] >= n_index)
            throw Error(GAPHIP_E_ARG, "slot without equation/variable");
    // per-row maps of the fronts handled by the blocked kernels: row pointers
    // into the arrowhead list and into (child info index, child row) pairs
    std::vector<int> bigbase(nf, -1), arow, crow, crow_ci;
    int rows = 0, nent = 0;
    for (int n = 0; n < nf; ++n) {
        FrontMeta const& fm = meta[n];
        if (front_class(fm.m) != kNumClasses - 1) continue;
        int const m = fm.m;
        bigbase[n] = rows;
        rows += m + 1;
        std::vector<int> na(m + 1, 0), nc(m + 1, 0);
        for (int q = fm.arrow_beg; q < fm.arrow_end; ++q) na[(arrow_dst[q] >> 16) + 1]++;
        for (int q = fm.child_beg; q < fm.child_end; ++q)
            for (int t = 0; t < cinfo[q].mc - cinfo[q].pc; ++t) nc[rel[cinfo[q].rel_off + t] + 1]++;
        for (int r = 0; r < m; ++r) {
            na[r + 1] += na[r];
            nc[r + 1] += nc[r];
        }
        for (int r = 0; r <= m; ++r) {
            arow.push_back(na[r]);
            crow.push_back(nent + nc[r]);
        }
        crow_ci.resize(2 * static_cast<size_t>(nent + nc[m]));
        std::vector<int> fill(nc.begin(), nc.end() - 1);
        for (int q = fm.child_beg; q < fm.child_end; ++q)
            for (int t = 0; t < cinfo[q].mc - cinfo[q].pc; ++t) {
                int r = rel[cinfo[q].rel_off + t];
                int at = nent + fill[r]++;
                crow_ci[2 * static_cast<size_t>(at)] = q;
                crow_ci[2 * static_cast<size_t>(at) + 1] = t;
            }
        nent += nc[m];
    }
    if (arow.empty()) {
        arow.push_back(0);
        crow.push_back(0);
    }
    if (crow_ci.empty()) crow_ci.assign(2, 0);
    if (cinfo.empty()) cinfo.push_back(ChildInfo{});
    if (rel.empty()) rel.push_back(0);
    // per-instance totals (algorithmic byte counts of the grid-resident kernel)
    d.inst_arrow.assign(G, 0);
    d.inst_factor.assign(G, 0);
    d.inst_slots.assign(G, 0);
    for (int n = 0; n < nf; ++n) {
        long long m = meta[n].m, pp = meta[n].p;
        d.inst_arrow[meta[n].inst] += meta[n].arrow_end - meta[n].arrow_beg;
        d.inst_factor[meta[n].inst] += pp * m + (m - pp) * pp;
        d.inst_slots[meta[n].inst] += m;
    }
    if (items) build_grid(d, plans, *items, meta, cinfo, rel, arrow_src, arrow_dst, slot_row, slot_col, n_index,
                          n_values, s);
    d.meta.upload(meta, s);
    d.cinfo.upload(cinfo, s);
    d.rel.upload(rel, s);
    d.arrow_src.upload(arrow_src, s);
    d.arrow_dst.upload(arrow_dst, s);
    d.slot_row.upload(slot_row, s);
    slot_col.insert(slot_col.end(), kVecPad, 0);  // read past a front's last slot by the grid kernel
    d.slot_col.upload(slot_col, s);
    d.bigbase.upload(bigbase, s);
    d.arow.upload(arow, s);
    d.crow.upload(crow, s);
    d.crow_ci.upload(crow_ci, s);
    HIPCHK(hipStreamSynchronize(s));  // host vectors die at return
}

// Runs factorisation / solves over all levels of a packed plan, one launch per
// non-empty (level, size class).  `hook(kid, bytes, fn)` wraps every launch.
template <typename T>
struct FrontRunner {
    static constexpr int kPanel = kBigPanel;
    // blocked partial LU of the fronts [beg, beg+nb) of one (level, class > 128) group
    static void big_factor(DevFronts const& d, FrontArgs const& a, int beg, int nb, int mmax, int pmax,
                           T const* jv, T* front, hipStream_t s) {
        hipLaunchKernelGGL(k_big_assemble<T>, dim3(mmax, nb), dim3(256), 0, s, a, d.big(), beg, jv, front);
        // Grids are sized for the group: a front's rows below its panel start
        // at min(k0 + 64, p_f) > k0, so mmax - k0 rows (mmax - j columns right
        // of an outer panel) bound every front; blocks past a front's own
        // bounds return at once.
        // Two-level panels only where the trailing matrix is large enough for
        // its traffic to dominate (below kTwoLevelMin slots the extra row-block
        // updates and solves cost more than they save): outer width 64 is the
        // plain right-looking scheme.
        int const outer = mmax >= d.big_two_level_min ? std::max(d.big_outer, kPanel) : kPanel;
        bool const look = outer > kPanel && d.side && d.ev_a && d.ev_b;
        bool pending = false;  // a side-stream update awaits its join
        for (int j = 0; j < pmax; j += outer) {
            int const jlim = std::min(j + outer, pmax);
            for (int k0 = j; k0 < jlim; k0 += kPanel) {
                int const below = mmax - k0, inner = jlim - k0;
                hipLaunchKernelGGL(k_big_diag<T>, dim3(nb), dim3(kWave), 0, s, a, beg, k0, front);
                hipLaunchKernelGGL(k_big_trsm_rows<T>, dim3(blocks_for(below, kTrsmRowsBlock), nb),
                                   dim3(kTrsmRowsBlock), 0, s, a, beg, k0, front);
                if (inner <= kPanel) continue;  // no front has columns left inside the outer panel
                hipLaunchKernelGGL(k_big_trsm<T>, dim3(blocks_for(inner, 256), nb), dim3(256), 0, s, a, beg, k0,
                                   kPanel, -1, j + outer, front);
                hipLaunchKernelGGL(k_big_gemm<T>, dim3(blocks_for(inner, 128), blocks_for(below, 128), nb), dim3(256),
                                   0, s, a, beg, k0, kPanel, -1, j + outer, front);
            }
            int const right = mmax - j;
            // the previous panel's update of the columns right of this panel
            // ran on the side stream meanwhile; this panel's own rows need it
            if (pending) HIPCHK(hipStreamWaitEvent(s, d.ev_b, 0));
            pending = false;
            for (int k0 = j; k0 < jlim; k0 += kPanel) {
                if (k0 > j)
                    hipLaunchKernelGGL(k_big_gemm_rows<T>, dim3(blocks_for(right, 128), 1, nb), dim3(256), 0, s, a,
                                       beg, j, outer, k0, front);
                hipLaunchKernelGGL(k_big_trsm<T>, dim3(blocks_for(right, 256), nb), dim3(256), 0, s, a, beg, k0,
                                   kPanel, j + outer, -1, front);
            }
            int const tiles = blocks_for(right, 128);
            if (look && j + outer < pmax) {
                // look-ahead: the next outer panel's columns first, then the
                // columns right of it on the side stream while the next panel
                // is factored (disjoint columns: that panel touches only its own)
                hipLaunchKernelGGL(k_big_gemm<T>, dim3(blocks_for(outer, 128), tiles, nb), dim3(256), 0, s, a, beg, j,
                                   outer, -1, j + 2 * outer, front);
                HIPCHK(hipEventRecord(d.ev_a, s));
                HIPCHK(hipStreamWaitEvent(d.side, d.ev_a, 0));
                hipLaunchKernelGGL(k_big_gemm<T>, dim3(tiles, tiles, nb), dim3(256), 0, d.side, a, beg, j, outer,
                                   j + 2 * outer, -1, front);
                HIPCHK(hipEventRecord(d.ev_b, d.side));
                pending = true;
            } else {
                hipLaunchKernelGGL(k_big_gemm<T>, dim3(tiles, tiles, nb), dim3(256), 0, s, a, beg, j, outer, -1, -1,
                                   front);
            }
        }
        if (pending) HIPCHK(hipStreamWaitEvent(s, d.ev_b, 0));
    }
    // blocked forward / backward solves of the fronts [beg, beg+nb) of a large-front group
    static void big_forward(DevFronts const&, FrontArgs const& a, int beg, int nb, int mmax, int pmax, T const* front,
                            T* vec, double const* rhs, double sign, hipStream_t s) {
        hipLaunchKernelGGL(k_big_fwd_init<T>, dim3(nb), dim3(256), 0, s, a, beg, vec, rhs, sign);
        for (int k0 = 0; k0 < pmax; k0 += kSolvePanel) {
            hipLaunchKernelGGL(k_big_fwd_diag<T>, dim3(nb), dim3(kWave), 0, s, a, beg, k0, front, vec);
            // rows below a front's panel start at min(k0 + 64, p_f) > k0: mmax - k0 bounds them all
            hipLaunchKernelGGL(k_big_fwd_update<T>, dim3(blocks_for(mmax - k0, 256), nb), dim3(256), 0, s, a, beg, k0,
                               front, vec);
        }
    }
    static void big_backward(DevFronts const&, FrontArgs const& a, int beg, int nb, int pmax, T const* front, T* vec,
                             double* sol, hipStream_t s) {
        hipLaunchKernelGGL(k_big_bwd_init<T>, dim3(blocks_for(pmax, kBwdInitRows), nb), dim3(kBwdInitRows * kWave), 0,
                           s, a, beg, front, vec, sol);
        for (int k0 = ((pmax - 1) / kSolvePanel) * kSolvePanel; k0 >= 0; k0 -= kSolvePanel) {
            hipLaunchKernelGGL(k_big_bwd_diag<T>, dim3(nb), dim3(kWave), 0, s, a, beg, k0, front, vec, sol);
            if (k0 > 0)
                hipLaunchKernelGGL(k_big_bwd_update<T>, dim3(blocks_for(k0, 256), nb), dim3(256), 0, s, a, beg, k0,
                                   front, vec);
        }
    }
    template <typename Hook>
    static void factor(DevFronts const& d, FrontArgs const& a, T const* jv, T* front, hipStream_t s, int kid,
                       Hook&& hook) {
        double const w = sizeof(T);
        for (int l = 0; l < d.n_levels; ++l)
            for (int c = 0; c < kNumClasses; ++c) {
                size_t b = static_cast<size_t>(l) * kNumClasses + c;
                int const beg = d.cls_ptr[b], nb = d.cls_ptr[b + 1] - beg;
                if (!nb) continue;
                double bytes = (w + 4) * d.cls_arrow[b] + w * d.cls_factor[b];
                hook(kid + c, bytes, [&] {
                    int const gb = c < kNumClasses - 1 ? blocks_for(nb, kClassPerBlock[c]) : 0;
                    switch (c) {
                        case 0: hipLaunchKernelGGL((k_front_factor_lds<T, 4, 4, 64>), dim3(gb), dim3(64), 0, s, a, beg, nb, jv, front); break;
                        case 1: hipLaunchKernelGGL((k_front_factor_lds<T, 8, 8, 64>), dim3(gb), dim3(64), 0, s, a, beg, nb, jv, front); break;
                        case 2: hipLaunchKernelGGL((k_front_factor_lds<T, 16, 16, 64>), dim3(gb), dim3(64), 0, s, a, beg, nb, jv, front); break;
                        case 3: hipLaunchKernelGGL((k_front_factor_lds<T, 32, 32, 64>), dim3(gb), dim3(64), 0, s, a, beg, nb, jv, front); break;
                        case 4: hipLaunchKernelGGL((k_front_factor_lds<T, 64, 64, 64>), dim3(gb), dim3(64), 0, s, a, beg, nb, jv, front); break;
                        case 5: hipLaunchKernelGGL((k_front_factor_lds<T, 128, 128, 128>), dim3(gb), dim3(128), 0, s, a, beg, nb, jv, front); break;
                        default: big_factor(d, a, beg, nb, d.cls_mmax[b], d.cls_pmax[b], jv, front, s);
                    }
                });
            }
    }
    template <bool TR = false, typename Hook>
    static void forward(DevFronts const& d, FrontArgs const& a, T const* front, T* vec, double const* rhs,
                        double sign, hipStream_t s, int kid, Hook&& hook) {
        double const w = sizeof(T);
        for (int l = 0; l < d.n_levels; ++l)
            for (int c = 0; c < kNumClasses; ++c) {
                size_t b = static_cast<size_t>(l) * kNumClasses + c;
                int const beg = d.cls_ptr[b], nb = d.cls_ptr[b + 1] - beg;
                if (!nb) continue;
                double bytes = w * d.cls_factor[b] / 2 + (2 * w + 12) * d.cls_slots[b];
                hook(kid + c, bytes, [&] {
                    int const gb = c < kNumClasses - 1 ? blocks_for(nb, kClassPerBlock[c]) : 0;
                    switch (c) {
                        case 0: hipLaunchKernelGGL((k_front_fwd_lds<T, 4, 4, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, rhs, sign); break;
                        case 1: hipLaunchKernelGGL((k_front_fwd_lds<T, 8, 8, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, rhs, sign); break;
                        case 2: hipLaunchKernelGGL((k_front_fwd_lds<T, 16, 16, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, rhs, sign); break;
                        case 3: hipLaunchKernelGGL((k_front_fwd_lds<T, 32, 32, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, rhs, sign); break;
                        case 4: hipLaunchKernelGGL((k_front_fwd_lds<T, 64, 64, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, rhs, sign); break;
                        case 5: hipLaunchKernelGGL((k_front_fwd_lds<T, 128, 128, 128, TR>), dim3(gb), dim3(128), 0, s, a, beg, nb, front, vec, rhs, sign); break;
                        default:
                            if constexpr (TR) {
                                hipLaunchKernelGGL((k_front_fwd<T, TR>), dim3(nb), dim3(256), 0, s, a, beg, front, vec, rhs, sign);
                            } else {
                                big_forward(d, a, beg, nb, d.cls_mmax[b], d.cls_pmax[b], front, vec, rhs, sign, s);
                            }
                    }
                });
            }
    }
    template <bool TR = false, typename Hook>
    static void backward(DevFronts const& d, FrontArgs const& a, T const* front, T* vec, double* sol,
                         hipStream_t s, int kid, Hook&& hook) {
        double const w = sizeof(T);
        for (int l = d.n_levels - 1; l >= 0; --l)
            for (int c = 0; c < kNumClasses; ++c) {
                size_t b = static_cast<size_t>(l) * kNumClasses + c;
                int const beg = d.cls_ptr[b], nb = d.cls_ptr[b + 1] - beg;
                if (!nb) continue;
                double bytes = w * d.cls_factor[b] / 2 + (2 * w + 12) * d.cls_slots[b];
                hook(kid + c, bytes, [&] {
                    int const gb = c < kNumClasses - 1 ? blocks_for(nb, kClassPerBlock[c]) : 0;
                    switch (c) {
                        case 0: hipLaunchKernelGGL((k_front_bwd_lds<T, 4, 4, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, sol); break;
                        case 1: hipLaunchKernelGGL((k_front_bwd_lds<T, 8, 8, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, sol); break;
                        case 2: hipLaunchKernelGGL((k_front_bwd_lds<T, 16, 16, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, sol); break;
                        case 3: hipLaunchKernelGGL((k_front_bwd_lds<T, 32, 32, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, sol); break;
                        case 4: hipLaunchKernelGGL((k_front_bwd_lds<T, 64, 64, 64, TR>), dim3(gb), dim3(64), 0, s, a, beg, nb, front, vec, sol); break;
                        case 5: hipLaunchKernelGGL((k_front_bwd_lds<T, 128, 128, 128, TR>), dim3(gb), dim3(128), 0, s, a, beg, nb, front, vec, sol); break;
                        default:
                            if constexpr (TR) {
                                hipLaunchKernelGGL((k_front_bwd<T, TR>), dim3(nb), dim3(256), 0, s, a, beg, front, vec, sol);
                            } else {
                                big_backward(d, a, beg, nb, d.cls_pmax[b], front, vec, sol, s);
                            }
                    }
                });
            }
    }
};

// Launch hook that runs the launch as is.
struct PlainHook {
    template <typename F>
    void operator()(int, double, F&& fn) const {
        fn();
    }
};

// --------------------------------------------------------------- the plan
struct Profiler {
    bool on = false;
    struct Rec {
        int kid;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double ms[K_COUNT] = {};
    long long launches[K_COUNT] = {};
    double bytes[K_COUNT] = {};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    void reset() {
        for (int k = 0; k < K_COUNT; ++k) ms[k] = 0, launches[k] = 0, bytes[k] = 0;
    }
    void collect() {
        for (auto& r : recs) {
            HIPCHK(hipEventSynchronize(r.b));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
            ms[r.kid] += t;
            launches[r.kid]++;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    ~Profiler() {
        for (auto& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace gaphip

// hold of the grid schedule's second launch (see k_stagger)
constexpr double kStaggerUs = 20.0;

struct gaphip_plan {
    int n_inst = 0;
    int precision = GAPHIP_FP64;
    std::vector<gaphip::Net> nets;
    std::vector<gaphip::YPlan> yp;
    std::vector<gaphip::NRPlan> np;
    std::vector<gaphip::LUPlan> lp;
    std::vector<std::vector<int>> node_bus;
    std::vector<int> bus_off, br_off, app_off, y_off, var_off, j_off, item_off;
    long long n_bus = 0, n_br = 0, n_app = 0, y_nnz = 0, n_var = 0, j_nnz = 0, n_items = 0;
    double analysis_ms = 0;
    hipStream_t stream = nullptr;

    // device data
    gaphip::DBuf<int> bus_inst, ang, mag, peq, qeq, y_rp, y_ci, y_seg, y_code, y_slot_inst, it_i, it_j, it_y,
        it_pos, done, iters, lu_fail, active;
    gaphip::DBuf<double> flat, br_r, br_x, br_g, br_b, app_g, app_b, zb, f, dx, final_mis, jv64;
    gaphip::DBuf<float> jv32;
    gaphip::DBuf<int> br_f, br_t, br_inst;  // in-service branch endpoints (global bus) or -1
    gaphip::DBuf<double> flow_zb, flow_ib, flow_out;  // branch flows (per instance bases, results)
    gaphip::DBuf<int> bus_type_d;
    gaphip::DBuf<double> adj_lossb, adj_gp, adj_gq, adj_gr, adj_gx, adj_gvm, adj_gva;
    gaphip::DBuf<double2> v, vpol, s, sspec, y_pu, y_given;
    gaphip::DBuf<float4> vsf;  // grid schedule with FP32 factors: V and S in FP32 per bus (allocated on first use)
    std::vector<int> y_given_mode;  // per instance: 0 assemble, 1 siemens given, 2 per-unit given
    gaphip::DBuf<unsigned long long> maxf;
    gaphip::DBuf<double> front64, vec64;
    gaphip::DBuf<float> front32, vec32;
    gaphip::DevFronts fr;
    gaphip::DBuf<int> bus_off_d;  // per-instance bus offsets for the grid schedule
    gaphip::DBuf<int2> bpos;      // grid schedule: per bus, list positions of its angle / magnitude variables (-1: none)
    // grid schedule: its own bus order (per network a depth-first preorder of
    // the Y-bus graph): Y rows, slot of each plan slot, equation indices, flat
    // start, Y contribution lists, the plan's bus of each, and V, S, S_spec
    // and per-unit Y in it.  Its solves assemble only that Y; y_pu follows
    // from it when something reads it (sync_plan_y)
    gaphip::DBuf<int> g_rp, g_ci, g_slot, g_peq, g_qeq, g_old, g_seg, g_code, g_slot_inst, g_br_src, g_app_src;
    gaphip::DBuf<double4> g_brp;  // branch (r, x, g, b) records of the grid Y assembly (repacked when dirty)
    gaphip::DBuf<double2> g_app;  // shunt (g, b) records
    int g_n_brp = 0, g_n_app = 0;
    bool g_par_dirty = true;
    gaphip::DBuf<double> g_flat;
    gaphip::DBuf<double2> g_v, g_s, g_sspec, g_y;
    std::vector<int> g_old_h;
    gaphip::DBuf<int> var_off_d;  // per-instance first Newton equation (n_inst + 1)
    int schedule = GAPHIP_SCHED_AUTO;
    gaphip_tuning tuning{2, 8, -1, 1, gaphip::kBigOuter, gaphip::kTwoLevelMin};
    gaphip_plan_options opts{4, 0.10, 0, 1};

    bool grid_schedule() const { return fr.fz_ok && schedule != GAPHIP_SCHED_LEVELS; }

    // host state
    bool sspec_dirty = true;
    bool rx_dirty = false;
    double sspec_base = -1.0;
    int last_iters_run = 0;
    int maxf_rows = 0;
    std::vector<unsigned long long> maxf_host;
    bool have_solution = false;
    bool solution_converged = false;  // every network of the last solve converged
    bool inputs_changed = false;      // injections, branch data or Y-bus changed since the last solve
    bool branches_changed = false;    // branch r / x or Y-bus changed since the last solve
    bool y_assembled = false;  // y_pu holds the Y-bus of the last solve's base
    mutable bool y_from_grid = false;  // ... once copied from the grid schedule's g_y
    std::vector<std::vector<int>> br_slot_ptr, br_slot;  // per instance: Y slots each branch touches (lazy)
    gaphip::DBuf<int> upd_idx, upd_slots;
    gaphip::DBuf<double> upd_r, upd_x;
    double last_base = 100e6;
    gaphip::Profiler prof;
    int* pinned_active = nullptr;
    hipEvent_t ev = nullptr;
    hipStream_t stream2 = nullptr;  // second stream of the grid schedule (default priority)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    unsigned long long stagger_ticks = 0;
    hipGraphExec_t newton_graph = nullptr;  // captured on the first solve that needs it (tuning.level_graph)
    void drop_graph() {
        if (newton_graph) (void)hipGraphExecDestroy(newton_graph);
        newton_graph = nullptr;
    }  // hold of the grid schedule's second launch (kStaggerUs in wall-clock ticks)

    ~gaphip_plan() {
        if (newton_graph) (void)hipGraphExecDestroy(newton_graph);
        if (pinned_active) (void)hipHostFree(pinned_active);
        if (ev) (void)hipEventDestroy(ev);
        if (fr.ev_a) (void)hipEventDestroy(fr.ev_a);
        if (fr.ev_b) (void)hipEventDestroy(fr.ev_b);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gaphip {

void analyze(gaphip_plan& P, int n_inst, gaphip_net const* nets, int precision) {
    auto t0 = std::chrono::steady_clock::now();
    if (n_inst <= 0 || !nets) throw Error(GAPHIP_E_ARG, "plan needs at least one network");
    if (precision != GAPHIP_FP64 && precision != GAPHIP_MIXED) throw Error(GAPHIP_E_ARG, "unknown precision");
    P.n_inst = n_inst;
    P.precision = precision;
    P.nets.resize(n_inst);
    P.yp.resize(n_inst);
    P.np.resize(n_inst);
    P.lp.resize(n_inst);
    P.node_bus.resize(n_inst);
    for (int g = 0; g < n_inst; ++g) P.nets[g] = copy_net(nets[g]);
    // symbolic analysis, instances spread over host threads
    std::vector<std::string> errs(n_inst);
    std::vector<int> bad(n_inst, 0);
    auto work = [&](int g) {
        try {
            P.yp[g] = plan_ybus(P.nets[g]);
            if (!plan_nr(P.nets[g], P.yp[g], P.np[g], errs[g])) {
                bad[g] = 1;
                return;
            }
            LUInput in = nr_lu_input(P.nets[g], P.yp[g], P.np[g], P.node_bus[g]);
            LUOptions lo;
            lo.relax_abs = P.opts.relax_abs;
            lo.relax_frac = P.opts.relax_frac;
            P.lp[g] = plan_lu(in, lo);
        } catch (std::exception const& e) {
            errs[g] = e.what();
            bad[g] = 1;
        }
    };
    int nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nthr = std::min(nthr, n_inst);
    if (nthr <= 1) {
        for (int g = 0; g < n_inst; ++g) work(g);
    } else {
        std::vector<std::thread> th;
        std::atomic<int> next{0};
        for (int t = 0; t < nthr; ++t)
            th.emplace_back([&]() {
                for (int g = next++; g < n_inst; g = next++) work(g);
            });
        for (auto& x : th) x.join();
    }
    for (int g = 0; g < n_inst; ++g)
        if (bad[g]) throw Error(GAPHIP_E_ARG, "network " + std::to_string(g) + ": " + errs[g]);

    // offsets
    P.bus_off.assign(n_inst + 1, 0);
    P.br_off.assign(n_inst + 1, 0);
    P.app_off.assign(n_inst + 1, 0);
    P.y_off.assign(n_inst + 1, 0);
    P.var_off.assign(n_inst + 1, 0);
    P.j_off.assign(n_inst + 1, 0);
    P.item_off.assign(n_inst + 1, 0);
    for (int g = 0; g < n_inst; ++g) {
        auto const& N = P.nets[g];
        P.bus_off[g + 1] = P.bus_off[g] + N.n_bus();
        P.br_off[g + 1] = P.br_off[g] + N.n_branch();
        P.app_off[g + 1] = P.app_off[g] + N.n_app();
        P.y_off[g + 1] = P.y_off[g] + static_cast<int>(P.yp[g].ci.size());
        P.var_off[g + 1] = P.var_off[g] + P.np[g].n_var;
        P.j_off[g + 1] = P.j_off[g] + static_cast<int>(P.np[g].jci.size());
        P.item_off[g + 1] = P.item_off[g] + static_cast<int>(P.np[g].it_i.size());
    }
    P.n_bus = P.bus_off[n_inst];
    P.n_br = P.br_off[n_inst];
    P.n_app = P.app_off[n_inst];
    P.y_nnz = P.y_off[n_inst];
    P.n_var = P.var_off[n_inst];
    P.j_nnz = P.j_off[n_inst];
    P.n_items = P.item_off[n_inst];

    P.analysis_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void build_plan(gaphip_plan& P, int n_inst, gaphip_net const* nets, int precision) {
    auto t0 = std::chrono::steady_clock::now();
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
        throw Error(GAPHIP_E_DEVICE, "no HIP device visible");
    analyze(P, n_inst, nets, precision);
    HIPCHK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&P.pinned_active), sizeof(int)));
    HIPCHK(hipEventCreateWithFlags(&P.ev, hipEventDisableTiming));
    HIPCHK(hipStreamCreateWithFlags(&P.stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&P.ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&P.ev_join, hipEventDisableTiming));
    {
        int dev = 0, khz = 0;
        HIPCHK(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
        P.stagger_ticks = static_cast<unsigned long long>(kStaggerUs * khz / 1000.0);
    }
    HIPCHK(hipEventCreateWithFlags(&P.fr.ev_a, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&P.fr.ev_b, hipEventDisableTiming));
    P.fr.side = P.stream2;
    hipStream_t s = P.stream;

    std::vector<int> bus_inst(P.n_bus), ang(P.n_bus), mag(P.n_bus), peq(P.n_bus), qeq(P.n_bus);
    std::vector<double> flat(P.n_bus), br_r(P.n_br), br_x(P.n_br), br_g(P.n_br), br_b(P.n_br), app_g(P.n_app),
        app_b(P.n_app);
    std::vector<int> y_rp(P.n_bus + 1), y_ci(P.y_nnz), y_seg, y_code, y_slot_inst(P.y_nnz);
    std::vector<int> it_i(P.n_items), it_j(P.n_items), it_y(P.n_items), it_pos(4 * P.n_items);
    y_seg.reserve(P.y_nnz + 1);
    y_seg.push_back(0);
    for (int g = 0; g < n_inst; ++g) {
        auto const& N = P.nets[g];
        auto const& Y = P.yp[g];
        auto const& R = P.np[g];
        int const b0 = P.bus_off[g], v0 = P.var_off[g], y0 = P.y_off[g];
        for (int i = 0; i < N.n_bus(); ++i) {
            bus_inst[b0 + i] = g;
            ang[b0 + i] = R.ang[i] >= 0 ? v0 + R.ang[i] : -1;
            mag[b0 + i] = R.mag[i] >= 0 ? v0 + R.mag[i] : -1;
            peq[b0 + i] = R.peq[i] >= 0 ? v0 + R.peq[i] : -1;
            qeq[b0 + i] = R.qeq[i] >= 0 ? v0 + R.qeq[i] : -1;
            flat[b0 + i] = (N.bus_type[i] == kSlack || N.bus_type[i] == kPV) ? N.u_pu[i] : 0.98;
            y_rp[b0 + i] = y0 + Y.rp[i];
        }
        for (int k = 0; k < N.n_branch(); ++k) {
            br_r[P.br_off[g] + k] = N.br_r[k];
            br_x[P.br_off[g] + k] = N.br_x[k];
            br_g[P.br_off[g] + k] = N.br_g[k];
            br_b[P.br_off[g] + k] = N.br_b[k];
        }
        for (int a = 0; a < N.n_app(); ++a) {
            app_g[P.app_off[g] + a] = N.app_g[a];
            app_b[P.app_off[g] + a] = N.app_b[a];
        }
        int const cbase = static_cast<int>(y_code.size());
        for (int c : Y.code) {
            int kind = c & 3, idx = c >> 2;
            idx += kind == kContribShunt ? P.app_off[g] : P.br_off[g];
            y_code.push_back((idx << 2) | kind);
        }
        for (int k = 0; k < static_cast<int>(Y.ci.size()); ++k) {
            y_ci[y0 + k] = b0 + Y.ci[k];
            y_seg.push_back(cbase + Y.seg[k + 1]);
            y_slot_inst[y0 + k] = g;
        }
        int const t0i = P.item_off[g];
        for (int t = 0; t < static_cast<int>(R.it_i.size()); ++t) {
            it_i[t0i + t] = b0 + R.it_i[t];
            it_j[t0i + t] = b0 + R.it_j[t];
            it_y[t0i + t] = y0 + R.it_y[t];
            for (int q = 0; q < 4; ++q) {
                int pp = R.it_pos[4 * t + q];
                it_pos[4 * (t0i + t) + q] = pp >= 0 ? P.j_off[g] + pp : -1;
            }
        }
    }
    y_rp[P.n_bus] = static_cast<int>(P.y_nnz);
    // bounds of every gathered index, checked before anything reaches the device
    for (long long i = 0; i < P.n_bus; ++i) {
        if (ang[i] >= P.n_var || mag[i] >= P.n_var || peq[i] >= P.n_var || qeq[i] >= P.n_var)
            throw Error(GAPHIP_E_ARG, "Newton index out of range");
        if (y_rp[i] > y_rp[i + 1]) throw Error(GAPHIP_E_ARG, "Y row pointer not monotone");
    }
    for (long long k = 0; k < P.y_nnz; ++k)
        if (y_ci[k] < 0 || y_ci[k] >= P.n_bus) throw Error(GAPHIP_E_ARG, "Y column out of range");
    for (int c : y_code) {
        long long idx = c >> 2;
        if (((c & 3) == kContribShunt && idx >= P.n_app) || ((c & 3) != kContribShunt && idx >= P.n_br))
            throw Error(GAPHIP_E_ARG, "Y contribution out of range");
    }
    for (long long t = 0; t < P.n_items; ++t) {
        if (it_y[t] < 0 || it_y[t] >= P.y_nnz || it_i[t] < 0 || it_i[t] >= P.n_bus || it_j[t] < 0 ||
            it_j[t] >= P.n_bus)
            throw Error(GAPHIP_E_ARG, "Jacobian item out of range");
        for (int q = 0; q < 4; ++q)
            if (it_pos[4 * t + q] >= P.j_nnz) throw Error(GAPHIP_E_ARG, "Jacobian slot out of range");
    }
    // The grid schedule's own bus order (its kernel reads V, S and Y in it):
    // per network the depth-first preorder of its Y-bus graph from its first
    // bus, neighbours in row order, further components from their first bus.
    // A bus's row neighbours then mostly lie in the same or adjacent lines of
    // V.  Its equations are numbered along that order (P before Q of a bus).
    // Rows keep their entries in the plan's order, so every sum is the same.
    std::vector<int> g_old(P.n_bus), g_bus_new(P.n_bus), g_eq_new(std::max<long long>(P.n_var, 1), -1),
        g_var_bus(std::max<long long>(P.n_var, 1), -1);
    for (int g = 0; g < n_inst; ++g) {
        int const b0 = P.bus_off[g], n = P.bus_off[g + 1] - b0, v0 = P.var_off[g];
        std::vector<char> seen(n, 0);
        std::vector<int> st;
        int k = 0;
        for (int r = 0; r < n; ++r) {
            if (seen[r]) continue;
            st.push_back(r);
            while (!st.empty()) {
                int const u = st.back();
                st.pop_back();
                if (seen[u]) continue;
                seen[u] = 1;
                g_old[b0 + k] = b0 + u;
                g_bus_new[b0 + u] = k++;
                for (int q = y_rp[b0 + u + 1] - 1; q >= y_rp[b0 + u]; --q) {
                    int const c = y_ci[q] - b0;
                    if (c < 0 || c >= n) throw Error(GAPHIP_E_ARG, "Y column outside its network");
                    if (!seen[c]) st.push_back(c);
                }
            }
        }
        int e = v0;
        for (int kk = 0; kk < n; ++kk) {
            int const c = g_old[b0 + kk];
            for (int const eq : {peq[c], qeq[c]})
                if (eq >= 0) {
                    if (eq < v0 || eq >= P.var_off[g + 1] || g_eq_new[eq] >= 0) throw Error(GAPHIP_E_ARG, "equation of two buses");
                    g_eq_new[eq] = e++ - v0;
                }
            for (int const var : {ang[c], mag[c]})
                if (var >= 0) g_var_bus[var] = b0 + kk;
        }
        if (e != P.var_off[g + 1]) throw Error(GAPHIP_E_ARG, "equation without a bus");
    }
    // the mismatch inputs in that order: Y rows (entries, and each slot's
    // contributions, in the plan's order; columns renumbered), equation
    // indices, flat start; g_slot maps a plan slot to the schedule's
    std::vector<int> grp(P.n_bus + 1), gci(P.y_nnz + kRowPad, 0), gslot(std::max<long long>(P.y_nnz, 1)),
        gpeq(P.n_bus), gqeq(P.n_bus), gseg(1, 0), gcode, gslot_inst(std::max<long long>(P.y_nnz, 1));
    std::vector<double> gflat(P.n_bus);
    // (the contribution lists index packed branch / shunt records, numbered in
    // the order the slots first use them: g_br_src / g_app_src)
    std::vector<int> g_br_src, g_app_src;
    {
        gseg.reserve(P.y_nnz + 1);
        gcode.reserve(y_code.size());
        std::vector<int> br_pos(std::max<long long>(P.n_br, 1), -1), app_pos(std::max<long long>(P.n_app, 1), -1);
        int q2 = 0;
        for (long long k = 0; k < P.n_bus; ++k) {
            int const i = g_old[k], g = bus_inst[i], b0 = P.bus_off[g], v0 = P.var_off[g];
            grp[k] = q2;
            for (int q = y_rp[i]; q < y_rp[i + 1]; ++q) {
                gci[q2] = b0 + g_bus_new[y_ci[q]];
                gslot_inst[q2] = y_slot_inst[q];
                for (int t = y_seg[q]; t < y_seg[q + 1]; ++t) {
                    int const c = y_code[t], kind = c & 3, idx = c >> 2;
                    std::vector<int>& pos = kind == kContribShunt ? app_pos : br_pos;
                    std::vector<int>& src = kind == kContribShunt ? g_app_src : g_br_src;
                    if (pos[idx] < 0) {
                        pos[idx] = static_cast<int>(src.size());
                        src.push_back(idx);
                    }
                    gcode.push_back((pos[idx] << 2) | kind);
                }
                gseg.push_back(static_cast<int>(gcode.size()));
                gslot[q] = q2++;
            }
            gpeq[k] = peq[i] >= 0 ? v0 + g_eq_new[peq[i]] : -1;
            gqeq[k] = qeq[i] >= 0 ? v0 + g_eq_new[qeq[i]] : -1;
            gflat[k] = flat[i];
        }
        grp[P.n_bus] = q2;
        if (q2 != P.y_nnz || gcode.size() != y_code.size()) throw Error(GAPHIP_E_ARG, "grid Y rows");
    }
    P.bus_inst.upload(bus_inst, s);
    P.ang.upload(ang, s);
    P.mag.upload(mag, s);
    P.peq.upload(peq, s);
    P.qeq.upload(qeq, s);
    P.flat.upload(flat, s);
    P.br_r.upload(br_r, s);
    P.br_x.upload(br_x, s);
    P.br_g.upload(br_g, s);
    P.br_b.upload(br_b, s);
    P.app_g.upload(app_g, s);
    P.app_b.upload(app_b, s);
    P.y_rp.upload(y_rp, s);
    y_ci.insert(y_ci.end(), kRowPad, 0);  // the grid mismatch reads rows in steps of kRowPad entries
    P.y_ci.upload(y_ci, s);
    P.y_seg.upload(y_seg, s);
    P.y_code.upload(y_code, s);
    P.y_slot_inst.upload(y_slot_inst, s);
    P.it_i.upload(it_i, s);
    P.it_j.upload(it_j, s);
    P.it_y.upload(it_y, s);
    P.it_pos.upload(it_pos, s);
    {
        // branch endpoints for the adjoint (out-of-service / unknown ids -> -1)
        std::vector<int> bf(std::max<long long>(P.n_br, 1), -1), bt(std::max<long long>(P.n_br, 1), -1),
            bi(std::max<long long>(P.n_br, 1), 0);
        for (int g = 0; g < n_inst; ++g) {
            auto const& N = P.nets[g];
            auto const& Y = P.yp[g];
            auto lookup = [&](int id) {
                return (id >= 0 && id < static_cast<int>(Y.id2idx.size())) ? Y.id2idx[id] : -1;
            };
            for (int k = 0; k < N.n_branch(); ++k) {
                int const gk = P.br_off[g] + k;
                bi[gk] = g;
                if (!N.br_status[k]) continue;
                int f = lookup(N.br_from[k]), t = lookup(N.br_to[k]);
                if (f < 0 || t < 0) continue;
                bf[gk] = P.bus_off[g] + f;
                bt[gk] = P.bus_off[g] + t;
            }
        }
        P.br_f.upload(bf, s);
        P.br_t.upload(bt, s);
        P.br_inst.upload(bi, s);
        std::vector<int> btype(P.n_bus);
        for (int g = 0; g < n_inst; ++g)
            for (int i = 0; i < P.nets[g].n_bus(); ++i) btype[P.bus_off[g] + i] = P.nets[g].bus_type[i];
        P.bus_type_d.upload(btype, s);
    }
    HIPCHK(hipStreamSynchronize(s));

    // fronts
    std::vector<LUPlan const*> plans(n_inst);
    std::vector<long long> ent_base(n_inst);
    for (int g = 0; g < n_inst; ++g) {
        plans[g] = &P.lp[g];
        ent_base[g] = P.j_off[g];
    }
    auto key_row = [&](int g, int key) {
        int bus = P.node_bus[g][key >> 1];
        return P.var_off[g] + ((key & 1) ? P.np[g].qeq[bus] : P.np[g].peq[bus]);
    };
    auto key_col = [&](int g, int key) {
        int bus = P.node_bus[g][key >> 1];
        return P.var_off[g] + ((key & 1) ? P.np[g].mag[bus] : P.np[g].ang[bus]);
    };
    GridItems items{it_i, it_j, it_y, it_pos, P.bus_off, P.var_off, P.y_nnz, g_bus_new, g_eq_new, g_var_bus, gslot};
    pack_fronts(plans, ent_base, key_row, key_col, P.n_var, P.j_nnz, P.fr, s, &items);
    P.bus_off_d.upload(P.bus_off, s);
    P.var_off_d.upload(P.var_off, s);
    HIPCHK(hipStreamSynchronize(s));
    if (P.fr.fz_ok) {
        // every Newton equation is the pivot of exactly one right-hand-side
        // entry (the factor tasks gather its -F by equation index)
        std::vector<char> seen(std::max<long long>(P.n_var, 1), 0);
        for (int const e : P.fr.fz_rhs_eq) {
            if (e < 0 || e >= P.n_var || seen[e]) throw Error(GAPHIP_E_ARG, "right-hand-side entry out of range");
            seen[e] = 1;
        }
        for (long long e = 0; e < P.n_var; ++e)
            if (!seen[e]) throw Error(GAPHIP_E_ARG, "equation without a pivot in the grid schedule");
        std::vector<int>().swap(P.fr.fz_rhs_eq);
        // the grid update gathers each bus's corrections at its variables' list positions
        // (per bus in the schedule's order)
        auto const& vp = P.fr.fz_varpos;
        std::vector<int2> bpos(std::max<long long>(P.n_bus, 1), make_int2(-1, -1));
        for (long long k = 0; k < P.n_bus; ++k) {
            int const i = g_old[k];
            int const a = ang[i] >= 0 ? vp[ang[i]] : -1, m = mag[i] >= 0 ? vp[mag[i]] : -1;
            if ((ang[i] >= 0 && a < 0) || (mag[i] >= 0 && m < 0))
                throw Error(GAPHIP_E_ARG, "variable without a pivot in the grid schedule");
            bpos[k] = make_int2(a, m);
        }
        P.bpos.upload(bpos, s);
        P.g_rp.upload(grp, s);
        P.g_ci.upload(gci, s);
        P.g_slot.upload(gslot, s);
        P.g_peq.upload(gpeq, s);
        P.g_qeq.upload(gqeq, s);
        P.g_flat.upload(gflat, s);
        P.g_seg.upload(gseg, s);
        P.g_code.upload(gcode, s);
        P.g_n_brp = static_cast<int>(g_br_src.size());
        P.g_n_app = static_cast<int>(g_app_src.size());
        P.g_br_src.upload(g_br_src, s);
        P.g_app_src.upload(g_app_src, s);
        P.g_brp.alloc(std::max(P.g_n_brp, 1));
        P.g_app.alloc(std::max(P.g_n_app, 1));
        P.g_par_dirty = true;
        P.g_slot_inst.upload(gslot_inst, s);
        P.g_old.upload(g_old, s);
        P.g_old_h = g_old;
        P.g_v.alloc(P.n_bus);
        P.g_s.alloc(P.n_bus);
        P.g_sspec.alloc(P.n_bus);
        P.g_y.alloc(P.y_nnz + kRowPad);
        HIPCHK(hipMemsetAsync(P.g_y.p + P.y_nnz, 0, sizeof(double2) * kRowPad, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<int>().swap(P.fr.fz_varpos);
    }

    // per-bus / per-instance state and work arrays
    P.v.alloc(P.n_bus);
    P.vpol.alloc(P.n_bus);
    P.s.alloc(P.n_bus);
    P.sspec.alloc(P.n_bus);
    P.y_pu.alloc(P.y_nnz + kRowPad);
    HIPCHK(hipMemsetAsync(P.y_pu.p + P.y_nnz, 0, sizeof(double2) * kRowPad, s));
    P.zb.alloc(n_inst);
    P.done.alloc(n_inst);
    P.iters.alloc(n_inst);
    P.lu_fail.alloc(n_inst);
    P.final_mis.alloc(n_inst);
    P.f.alloc(std::max<long long>(P.n_var, 1));
    P.dx.alloc(std::max<long long>(P.n_var, 1));
    if (precision == GAPHIP_FP64) {
        P.jv64.alloc(std::max<long long>(P.j_nnz, 1));
        P.front64.alloc(P.fr.front_entries + kFrontPad);
        P.vec64.alloc(P.fr.vec_entries + kVecPad);
    } else {
        P.jv32.alloc(std::max<long long>(P.j_nnz, 1));
        P.front32.alloc(P.fr.front_entries + kFrontPad);
        P.vec32.alloc(P.fr.vec_entries + kVecPad);
    }
    // graph replay of the level schedule's Newton step: measured equal to
    // direct launches on the 10k-bus grid (46.21 against 46.23 ms per FP64
    // solve) and 0.6% slower on the 100k-bus grid, whose two-level fronts
    // overlap a side stream that the replayed graph serialises more: direct
    // launches there
    P.tuning.level_graph = P.fr.max_front < P.tuning.big_two_level_min ? 1 : 0;
    P.analysis_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// S_spec in per unit (cpu_newton_raphson.cpp:346-377): bus injection if non-zero,
// plus every in-service SOURCE / LOADGEN appliance on that bus id in appliance
// order, divided by base_power.
void upload_sspec(gaphip_plan& P, double base) {
    std::vector<double2> sp(P.n_bus);
    std::vector<double> zb(P.n_inst);
    for (int g = 0; g < P.n_inst; ++g) {
        auto const& N = P.nets[g];
        auto const& Y = P.yp[g];
        int const n = N.n_bus();
        std::vector<double> pr(n, 0.0), pi(n, 0.0);
        for (int i = 0; i < n; ++i)
            if (N.p_w[i] != 0.0 || N.q_var[i] != 0.0) {
                pr[i] = N.p_w[i];
                pi[i] = N.q_var[i];
            }
        for (int a = 0; a < N.n_app(); ++a) {
            if (N.app_status[a] != 1 || (N.app_type[a] != kSource && N.app_type[a] != kLoadgen)) continue;
            int id = N.app_node[a];
            if (id < 0 || id >= static_cast<int>(Y.id2idx.size())) continue;
            int i = Y.id2idx[id];
            if (i < 0) continue;
            pr[i] += N.app_p[a];
            pi[i] += N.app_q[a];
        }
        for (int i = 0; i < n; ++i) {
            double r = pr[i], im = pi[i];
            if (base > 0.0) {
                r /= base;
                im /= base;
            }
            sp[P.bus_off[g] + i] = make_double2(r, im);
        }
        double vb = N.u_rated[0];
        zb[g] = (vb * vb) / base;
    }
    std::vector<double2> const& c = sp;
    HIPCHK(hipMemcpyAsync(P.sspec.p, c.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice, P.stream));
    std::vector<double2> cg;
    if (P.g_sspec.p) {
        cg.resize(sp.size());
        for (size_t k = 0; k < sp.size(); ++k) cg[k] = sp[P.g_old_h[k]];
        HIPCHK(hipMemcpyAsync(P.g_sspec.p, cg.data(), cg.size() * sizeof(double2), hipMemcpyHostToDevice, P.stream));
    }
    HIPCHK(hipMemcpyAsync(P.zb.p, zb.data(), zb.size() * sizeof(double), hipMemcpyHostToDevice, P.stream));
    HIPCHK(hipStreamSynchronize(P.stream));
    P.sspec_dirty = false;
    P.sspec_base = base;
}

// Overwrite the assembled per-unit Y of instances that were given a Y-bus
// (grid: in the grid schedule's copy instead of the plan's).
void apply_given_y(gaphip_plan& P, double base, bool grid = false) {
    if (P.y_given_mode.empty()) return;
    for (int g = 0; g < P.n_inst; ++g) {
        int mode = P.y_given_mode[g];
        if (!mode) continue;
        double vb = P.nets[g].u_rated[0];
        double scale = mode == 1 ? (vb * vb) / base : 1.0;
        int b = P.y_off[g], e = P.y_off[g + 1];
        if (e > b)
            hipLaunchKernelGGL(k_yscale, dim3(blocks_for(e - b, 256)), dim3(256), 0, P.stream, b, e, P.y_given.p,
                               scale, grid ? nullptr : P.y_pu.p, grid ? P.g_slot.p : nullptr,
                               grid ? P.g_y.p : nullptr);
    }
}

// The plan's Y (y_pu) after a grid-schedule solve, which assembled only the
// schedule's copy: gathered from it the first time something reads it.
void sync_plan_y(gaphip_plan const& P) {
    if (!P.y_from_grid) return;
    if (P.y_nnz > 0)
        hipLaunchKernelGGL(k_ygather, dim3(blocks_for(P.y_nnz, 256)), dim3(256), 0, P.stream, P.y_nnz, P.g_slot.p,
                           P.g_y.p, P.y_pu.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(P.stream));
    P.y_from_grid = false;
}

// Launch helper that optionally brackets a launch with timing events.
template <typename F>
void launch(gaphip_plan& P, int kid, double bytes, F&& fn) {
    if (P.prof.on) {
        Profiler::Rec r{kid, P.prof.get(), P.prof.get()};
        HIPCHK(hipEventRecord(r.a, P.stream));
        fn();
        HIPCHK(hipEventRecord(r.b, P.stream));
        P.prof.recs.push_back(r);
        P.prof.bytes[kid] += bytes;
    } else {
        fn();
    }
    HIPCHK(hipGetLastError());
}

// The Newton step of the level schedule after the convergence check: the
// Jacobian, its factorisation, the two front solves and the damped update
// (several hundred launches per iteration on the 10k-bus grid).  Its
// arguments do not depend on the iteration or the solve options.
template <typename T>
void newton_step(gaphip_plan& P, T* jv, T* front, T* vec) {
    hipStream_t s = P.stream;
    double const w = sizeof(T);
    long long const N = P.n_bus, Z = P.y_nnz, V = P.n_var, J = P.j_nnz;
    FrontArgs fra = P.fr.args(P.done.p, P.lu_fail.p);
    JacArgs ja{static_cast<int>(P.n_items), P.it_i.p, P.it_j.p, P.it_y.p, P.it_pos.p, P.bus_inst.p, P.done.p,
               P.y_pu.p, P.s.p, P.vpol.p};
    launch(P, K_JAC, 20.0 * Z + 44.0 * N + w * J, [&] {
        hipLaunchKernelGGL(k_jacobian<T>, dim3(blocks_for(P.n_items, 256)), dim3(256), 0, s, ja, jv);
    });
    auto hook = [&](int kid, double bytes, auto&& fn) { launch(P, kid, bytes, fn); };
    FrontRunner<T>::factor(P.fr, fra, jv, front, s, K_FACTOR, hook);
    FrontRunner<T>::forward(P.fr, fra, front, vec, P.f.p, -1.0, s, K_FWD, hook);
    FrontRunner<T>::backward(P.fr, fra, front, vec, P.dx.p, s, K_BWD, hook);
    UpdArgs ua{static_cast<int>(N), P.bus_inst.p, P.done.p, P.lu_fail.p, P.ang.p, P.mag.p, P.dx.p, P.v.p};
    launch(P, K_UPDATE, 8.0 * V + 32.0 * N,
           [&] { hipLaunchKernelGGL(k_update, dim3(blocks_for(N, 256)), dim3(256), 0, s, ua); });
}

// Captures newton_step once per plan into a graph on the plan stream (the
// large fronts' look-ahead joins the side stream through events, which the
// capture follows), instantiated for replay every iteration: one launch
// instead of hundreds.  Kernels only -- no memset or host work is recorded.
template <typename T>
void capture_newton_step(gaphip_plan& P, T* jv, T* front, T* vec) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
    try {
        newton_step<T>(P, jv, front, vec);
    } catch (...) {
        (void)hipStreamEndCapture(P.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(P.stream, &g));
    hipError_t const e = hipGraphInstantiate(&P.newton_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        P.newton_graph = nullptr;
        throw Error(GAPHIP_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
}

template <typename T>
void run_solve(gaphip_plan& P, gaphip_solve_opts const& o, T* jv, T* front, T* vec) {
    hipStream_t s = P.stream;
    int const G = P.n_inst;
    int const max_it = o.max_iterations;
    P.fr.big_outer = P.tuning.big_outer;
    P.fr.big_two_level_min = P.tuning.big_two_level_min;
    if (P.sspec_dirty || P.sspec_base != o.base_power) upload_sspec(P, o.base_power);
    int rows = max_it + 1;
    if (P.maxf_rows < rows) {
        P.maxf.alloc(static_cast<size_t>(rows) * G);
        P.active.alloc(rows);
        P.maxf_rows = rows;
    }
    HIPCHK(hipMemsetAsync(P.maxf.p, 0, sizeof(unsigned long long) * rows * G, s));
    HIPCHK(hipMemsetAsync(P.active.p, 0, sizeof(int) * rows, s));
    HIPCHK(hipMemsetAsync(P.done.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.iters.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.lu_fail.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.final_mis.p, 0, sizeof(double) * G, s));
    if (P.prof.on) P.prof.reset();
    long long const N = P.n_bus, Z = P.y_nnz, V = P.n_var;

    YbusArgs ya{static_cast<int>(Z), P.y_seg.p, P.y_code.p, P.y_slot_inst.p, P.br_r.p, P.br_x.p, P.br_g.p,
                P.br_b.p, P.app_g.p, P.app_b.p, P.zb.p, P.y_pu.p, nullptr};
    launch(P, K_YBUS, 48.0 * Z, [&] { hipLaunchKernelGGL(k_ybus, dim3(blocks_for(Z, 256)), dim3(256), 0, s, ya); });
    apply_given_y(P, o.base_power);
    P.y_assembled = true;
    P.y_from_grid = false;
    FlatArgs fa{static_cast<int>(N), P.flat.p, P.v.p};
    launch(P, K_FLAT, 24.0 * N, [&] { hipLaunchKernelGGL(k_flat, dim3(blocks_for(N, 256)), dim3(256), 0, s, fa); });

    int it = 0;
    for (; it < max_it; ++it) {
        MisArgs ma{static_cast<int>(N), P.y_rp.p, P.y_ci.p, P.y_pu.p, P.v.p, P.sspec.p, P.peq.p, P.qeq.p,
                   P.bus_inst.p, P.done.p, P.s.p, P.vpol.p, P.f.p, P.maxf.p + static_cast<size_t>(it) * G};
        launch(P, K_MISMATCH, 20.0 * Z + 56.0 * N + 8.0 * V,
               [&] { hipLaunchKernelGGL(k_mismatch, dim3(blocks_for(N, 256)), dim3(256), 0, s, ma); });
        CheckArgs ca{G, it, o.tolerance, P.maxf.p + static_cast<size_t>(it) * G, P.done.p, P.iters.p,
                     P.lu_fail.p, P.final_mis.p, P.active.p + it, P.f.p, P.var_off_d.p};
        launch(P, K_CHECK, 24.0 * G, [&] { hipLaunchKernelGGL(k_check, dim3(blocks_for(G, 256)), dim3(256), 0, s, ca); });
        HIPCHK(hipMemcpyAsync(P.pinned_active, P.active.p + it, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(P.ev, s));
        HIPCHK(hipEventSynchronize(P.ev));
        if (*P.pinned_active == 0) break;
        if (P.tuning.level_graph && !P.prof.on) {
            if (!P.newton_graph) capture_newton_step<T>(P, jv, front, vec);
            HIPCHK(hipGraphLaunch(P.newton_graph, s));
        } else {
            newton_step<T>(P, jv, front, vec);
        }
    }
    P.last_iters_run = std::min(it + 1, max_it);
    HIPCHK(hipStreamSynchronize(s));
    if (P.prof.on) P.prof.collect();
    P.have_solution = true;
    P.last_base = o.base_power;
}

// Grid schedule: one launch runs every Newton iteration of every network (one
// workgroup per network, largest first).  Same prelude and the same
// per-network results (iterations, converged flag, final mismatch, max |F|
// rows) as the level-synchronous loop.
// Networks whose elimination trees hold at least tuning.grid_big_work front
// slots and factor entries run with tuning.grid_big_waves wavefronts per
// workgroup (their levels are wide), the rest with tuning.grid_waves; the two
// launches run concurrently.
//
// k_stagger holds a stream for `ticks` of the constant wall clock (one
// wavefront, bounded wait).  The second launch of a solve waits behind it, so
// the hardware dispatches the first launch's workgroups (the largest
// networks, longest first) before the second's: on the 4,096-grid batch this
// took the solve from 12.4 to 11.8 ms, since the short workgroups then fill
// the end of the launch instead of long ones started late.
__global__ void k_stagger(unsigned long long ticks) {
    unsigned long long const t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

template <typename T, int NW>
void launch_grid(gaphip_plan& P, hipStream_t st, int first, int count, FusedArgs za, MisArgs const& ma,
                 UpdArgs const& ua, FrontArgs const& fra, T* front, T* vec);

template <typename T>
void launch_grid_n(int nw, gaphip_plan& P, hipStream_t st, int first, int count, FusedArgs const& za,
                   MisArgs const& ma, UpdArgs const& ua, FrontArgs const& fra, T* front, T* vec) {
    switch (nw) {
        case 1: launch_grid<T, 1>(P, st, first, count, za, ma, ua, fra, front, vec); break;
        case 2: launch_grid<T, 2>(P, st, first, count, za, ma, ua, fra, front, vec); break;
        case 4: launch_grid<T, 4>(P, st, first, count, za, ma, ua, fra, front, vec); break;
        default: launch_grid<T, 8>(P, st, first, count, za, ma, ua, fra, front, vec);
    }
}

template <typename T, int NW>
void launch_grid(gaphip_plan& P, hipStream_t st, int first, int count, FusedArgs za, MisArgs const& ma,
                 UpdArgs const& ua, FrontArgs const& fra, T* front, T* vec) {
    if (count <= 0) return;
    size_t const shm = static_cast<size_t>(NW) * P.fr.fz_pool * sizeof(T);
    if (shm > 64 * 1024) {
        // above the default per-workgroup limit: gfx950 has 160 KiB of LDS per CU
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nr_grid<T, NW>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)) != hipSuccess)
            throw Error(GAPHIP_E_CAPACITY, "grid schedule needs " + std::to_string(shm) +
                                               " bytes of LDS per workgroup, more than the device allows");
    }
    za.order_off = first;
    hipLaunchKernelGGL((k_nr_grid<T, NW>), dim3(count), dim3(NW * kWave), shm, st, za, ma, ua, fra, front, vec,
                       za.tasks, za.lvl_ptr);
}

template <typename T>
void run_solve_fused(gaphip_plan& P, gaphip_solve_opts const& o, T* jv, T* front, T* vec) {
    hipStream_t s = P.stream;
    int const G = P.n_inst;
    int const max_it = o.max_iterations;
    if (P.sspec_dirty || P.sspec_base != o.base_power) upload_sspec(P, o.base_power);
    int rows = max_it + 1;
    if (P.maxf_rows < rows) {
        P.maxf.alloc(static_cast<size_t>(rows) * G);
        P.active.alloc(rows);
        P.maxf_rows = rows;
    }
    HIPCHK(hipMemsetAsync(P.maxf.p, 0, sizeof(unsigned long long) * rows * G, s));
    HIPCHK(hipMemsetAsync(P.done.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.iters.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.lu_fail.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.final_mis.p, 0, sizeof(double) * G, s));
    if (P.prof.on) P.prof.reset();
    long long const N = P.n_bus, Z = P.y_nnz;

    // Y and the flat start in the grid schedule's own bus order (its kernel
    // writes V back in the plan's order; with no iteration V is the flat start)
    DevFronts const& d = P.fr;
    if (P.g_par_dirty) {
        YbusPackArgs pa{P.g_n_brp, P.g_n_app, P.g_br_src.p, P.g_app_src.p, P.br_r.p, P.br_x.p, P.br_g.p, P.br_b.p,
                        P.app_g.p, P.app_b.p, P.g_brp.p, P.g_app.p};
        int const n = std::max(P.g_n_brp, P.g_n_app);
        if (n > 0) hipLaunchKernelGGL(k_ybus_pack, dim3(blocks_for(n, 256)), dim3(256), 0, s, pa);
        P.g_par_dirty = false;
    }
    // the pairs' Y values come with the assembly unless an instance's Y is
    // given (then gathered after it is applied)
    bool const given = std::any_of(P.y_given_mode.begin(), P.y_given_mode.end(), [](int m) { return m != 0; });
    bool const fused_pairs = max_it > 0 && d.fz_slot_pair.p && !given;
    YbusGridArgs ya{static_cast<int>(Z), P.g_seg.p, P.g_code.p, P.g_slot_inst.p, P.g_brp.p, P.g_app.p, P.zb.p,
                    P.g_y.p, fused_pairs ? d.fz_slot_pair.p : nullptr, d.fz_ypair.p};
    launch(P, K_YBUS, 48.0 * Z,
           [&] { hipLaunchKernelGGL(k_ybus_grid<T>, dim3(blocks_for(Z, 256)), dim3(256), 0, s, ya); });
    apply_given_y(P, o.base_power, true);
    P.y_assembled = true;
    P.y_from_grid = true;
    FlatArgs fa = max_it > 0 ? FlatArgs{static_cast<int>(N), P.g_flat.p, P.g_v.p}
                             : FlatArgs{static_cast<int>(N), P.flat.p, P.v.p};
    launch(P, K_FLAT, 24.0 * N, [&] { hipLaunchKernelGGL(k_flat, dim3(blocks_for(N, 256)), dim3(256), 0, s, fa); });

    if (max_it > 0) {
        if (!fused_pairs)
            hipLaunchKernelGGL(k_gather_ypair<T>, dim3(blocks_for(d.fz_npairs, 256)), dim3(256), 0, s,
                               static_cast<int>(d.fz_npairs), d.fz_pair_y.p, P.g_y.p,
                               reinterpret_cast<cplx_t<T>*>(d.fz_ypair.p));
        if (sizeof(T) == 4 && !P.vsf.p) {
            P.vsf.alloc(N);
        }
        FusedArgs za{G, max_it, o.tolerance, 0, d.fz_order.p, P.bus_off_d.p, P.var_off_d.p, d.fz_lvl_beg.p, d.fz_lvl_ptr.p,
                     d.fz_tasks.p, d.fz_gfront.p, d.fz_pairs.p, d.fz_ypair.p, d.fz_fe.p, d.fz_rhs.p, d.fz_ext.p, d.fz_vx.p, d.fz_sol.p, d.fz_netbase.p, d.fz_pool, d.fz_ld4,
                     P.maxf.p, P.done.p, P.iters.p, P.lu_fail.p, P.final_mis.p, d.fz_xs.p};
        MisArgs ma{static_cast<int>(N), P.g_rp.p, P.g_ci.p, P.g_y.p, P.g_v.p, P.g_sspec.p, P.g_peq.p, P.g_qeq.p,
                   nullptr, P.done.p, P.g_s.p, nullptr, P.f.p, P.maxf.p, d.fz_fe.p, P.vsf.p};
        UpdArgs ua{static_cast<int>(N), nullptr, P.done.p, P.lu_fail.p, nullptr, nullptr, nullptr, P.g_v.p,
                   P.bpos.p, d.fz_xs.p, P.g_old.p, P.v.p};
        FrontArgs fra = d.args(nullptr, P.lu_fail.p);
        gaphip_tuning const& tu = P.tuning;
        int nbig = 0;
        if (tu.grid_big_work >= 0)
            while (nbig < G && d.fz_work[nbig] >= tu.grid_big_work) ++nbig;
        launch(P, K_FUSED, 0.0, [&] {
            if (nbig > 0 && nbig < G) {
                // the big networks on the plan stream, the rest concurrently on
                // the side stream behind a short hold: the long workgroups are
                // dispatched first and the short ones fill the end of the launch
                HIPCHK(hipEventRecord(P.ev_fork, s));
                HIPCHK(hipStreamWaitEvent(P.stream2, P.ev_fork, 0));
                launch_grid_n<T>(tu.grid_big_waves, P, s, 0, nbig, za, ma, ua, fra, front, vec);
                hipLaunchKernelGGL(k_stagger, dim3(1), dim3(kWave), 0, P.stream2, P.stagger_ticks);
                launch_grid_n<T>(tu.grid_waves, P, P.stream2, nbig, G - nbig, za, ma, ua, fra, front, vec);
                HIPCHK(hipEventRecord(P.ev_join, P.stream2));
                HIPCHK(hipStreamWaitEvent(s, P.ev_join, 0));
            } else {
                launch_grid_n<T>(nbig == G ? tu.grid_big_waves : tu.grid_waves, P, s, 0, G, za, ma, ua, fra, front,
                                 vec);
            }
        });
    }
    P.last_iters_run = max_it;
    HIPCHK(hipStreamSynchronize(s));
    if (P.prof.on) {
        P.prof.collect();
        // algorithmic bytes of the launch, per network from its own sizes
        // (N buses, Z Y-bus entries, V Newton variables, J Jacobian entries,
        // L = nnz(L+U) of its fronts; w bytes per factor entry):
        //   SpMV + mismatch + max   20 Z + 56 N + 8 V
        //   Jacobian fill           20 Z + 44 N + w J
        //   numeric factorisation   (w + 4) J + w L
        //   triangular solves       (w + 4) L + 16 V
        //   update                  8 V + 32 N
        // for every full Newton iteration, plus the final mismatch-only pass
        // of a converged network
        std::vector<int> its(G), done(G);
        HIPCHK(hipMemcpy(its.data(), P.iters.p, sizeof(int) * G, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(done.data(), P.done.p, sizeof(int) * G, hipMemcpyDeviceToHost));
        double const w = sizeof(T);
        double total = 0.0;
        for (int g = 0; g < G; ++g) {
            double const n = P.bus_off[g + 1] - P.bus_off[g], z = P.y_off[g + 1] - P.y_off[g];
            double const v = P.var_off[g + 1] - P.var_off[g];
            double const J = static_cast<double>(P.fr.inst_arrow[g]), L = static_cast<double>(P.fr.inst_factor[g]);
            double const mis = 20.0 * z + 56.0 * n + 8.0 * v;
            double const full = mis + (20.0 * z + 44.0 * n + w * J) + ((w + 4.0) * J + w * L) +
                                ((w + 4.0) * L + 16.0 * v) + (8.0 * v + 32.0 * n);
            int const newton = done[g] ? its[g] - 1 : max_it;
            total += newton * full + (done[g] ? mis : 0.0);
        }
        P.prof.bytes[K_FUSED] += total;
    }
    P.have_solution = true;
    P.last_base = o.base_power;
}

// Adjoint at the converged state of the last solve: refactor J(x*) in FP64,
// solve J^T lambda = dL/dx with the transposed front solves, then the bus and
// branch gradient kernels.
void run_adjoint(gaphip_plan& P, double w_loss, double w_dev, double const* g_vm, double const* g_va,
                 double* gp, double* gq, double* gr, double* gx, double* loss_out) {
    if (!P.have_solution) throw Error(GAPHIP_E_STATE, "adjoint needs a converged solve first");
    if (P.inputs_changed)
        throw Error(GAPHIP_E_STATE, "plan inputs changed since the last solve: solve again before the adjoint");
    if (!P.solution_converged)
        throw Error(GAPHIP_E_STATE, "adjoint needs a solve in which every network converged");
    hipStream_t s = P.stream;
    int const G = P.n_inst;
    long long const N = P.n_bus;
    if (!P.jv64.p) P.jv64.alloc(std::max<long long>(P.j_nnz, 1));
    if (!P.front64.p) P.front64.alloc(P.fr.front_entries + kFrontPad);
    if (!P.vec64.p) P.vec64.alloc(P.fr.vec_entries + kVecPad);
    P.adj_lossb.alloc(N);
    P.adj_gp.alloc(N);
    P.adj_gq.alloc(N);
    P.adj_gr.alloc(std::max<long long>(P.n_br, 1));
    P.adj_gx.alloc(std::max<long long>(P.n_br, 1));
    double const* dgvm = nullptr;
    double const* dgva = nullptr;
    if (g_vm) {
        P.adj_gvm.alloc(N);
        HIPCHK(hipMemcpyAsync(P.adj_gvm.p, g_vm, sizeof(double) * N, hipMemcpyHostToDevice, s));
        dgvm = P.adj_gvm.p;
    }
    if (g_va) {
        P.adj_gva.alloc(N);
        HIPCHK(hipMemcpyAsync(P.adj_gva.p, g_va, sizeof(double) * N, hipMemcpyHostToDevice, s));
        dgva = P.adj_gva.p;
    }
    HIPCHK(hipMemsetAsync(P.done.p, 0, sizeof(int) * G, s));
    HIPCHK(hipMemsetAsync(P.lu_fail.p, 0, sizeof(int) * G, s));
    if (P.maxf_rows < 1) {
        P.maxf.alloc(G);
        P.active.alloc(1);
        P.maxf_rows = 1;
    }
    // S, |V|, arg V at the solution, then J(x*) and its FP64 factors
    sync_plan_y(P);
    MisArgs ma{static_cast<int>(N), P.y_rp.p, P.y_ci.p, P.y_pu.p, P.v.p, P.sspec.p, P.peq.p, P.qeq.p,
               P.bus_inst.p, P.done.p, P.s.p, P.vpol.p, P.f.p, P.maxf.p};
    hipLaunchKernelGGL(k_mismatch, dim3(blocks_for(N, 256)), dim3(256), 0, s, ma);
    JacArgs ja{static_cast<int>(P.n_items), P.it_i.p, P.it_j.p, P.it_y.p, P.it_pos.p, P.bus_inst.p, P.done.p,
               P.y_pu.p, P.s.p, P.vpol.p};
    hipLaunchKernelGGL(k_jacobian<double>, dim3(blocks_for(P.n_items, 256)), dim3(256), 0, s, ja, P.jv64.p);
    FrontArgs fra = P.fr.args(P.done.p, P.lu_fail.p);
    FrontRunner<double>::factor(P.fr, fra, P.jv64.p, P.front64.p, s, 0, PlainHook{});
    // dL/dx into f (per variable), loss contributions per bus
    HIPCHK(hipMemsetAsync(P.f.p, 0, sizeof(double) * std::max<long long>(P.n_var, 1), s));
    AdjArgs aa{static_cast<int>(N), P.y_rp.p, P.y_ci.p, P.y_pu.p, P.v.p, P.s.p, P.ang.p, P.mag.p, P.bus_type_d.p,
               dgvm, dgva, w_loss, w_dev, P.f.p, P.adj_lossb.p};
    hipLaunchKernelGGL(k_adj_rhs, dim3(blocks_for(N, 256)), dim3(256), 0, s, aa);
    // J^T lambda = dL/dx: U^T then L^T; lambda lands in dx, indexed by equation
    FrontRunner<double>::forward<true>(P.fr, fra, P.front64.p, P.vec64.p, P.f.p, 1.0, s, 0, PlainHook{});
    FrontRunner<double>::backward<true>(P.fr, fra, P.front64.p, P.vec64.p, P.dx.p, s, 0, PlainHook{});
    AdjBusArgs ba{static_cast<int>(N), P.peq.p, P.qeq.p, P.bus_inst.p, P.dx.p, 1.0 / P.last_base, P.adj_gp.p,
                  P.adj_gq.p};
    hipLaunchKernelGGL(k_adj_bus, dim3(blocks_for(N, 256)), dim3(256), 0, s, ba);
    if (P.n_br > 0) {
        AdjBrArgs bra{static_cast<int>(P.n_br), P.br_f.p, P.br_t.p, P.br_inst.p, P.br_r.p, P.br_x.p, P.zb.p, P.v.p,
                      P.peq.p, P.qeq.p, P.dx.p, w_loss, P.adj_gr.p, P.adj_gx.p};
        hipLaunchKernelGGL(k_adj_branch, dim3(blocks_for(P.n_br, 256)), dim3(256), 0, s, bra);
    }
    HIPCHK(hipGetLastError());
    std::vector<int> fail(G);
    HIPCHK(hipMemcpyAsync(fail.data(), P.lu_fail.p, sizeof(int) * G, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int g = 0; g < G; ++g)
        if (fail[g]) throw Error(GAPHIP_E_SINGULAR, "Jacobian at the solution of network " + std::to_string(g) + " is singular");
    if (gp) HIPCHK(hipMemcpy(gp, P.adj_gp.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (gq) HIPCHK(hipMemcpy(gq, P.adj_gq.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (gr && P.n_br) HIPCHK(hipMemcpy(gr, P.adj_gr.p, sizeof(double) * P.n_br, hipMemcpyDeviceToHost));
    if (gx && P.n_br) HIPCHK(hipMemcpy(gx, P.adj_gx.p, sizeof(double) * P.n_br, hipMemcpyDeviceToHost));
    if (loss_out) {
        std::vector<double> lb(N);
        HIPCHK(hipMemcpy(lb.data(), P.adj_lossb.p, sizeof(double) * N, hipMemcpyDeviceToHost));
        for (int g = 0; g < G; ++g) {
            double acc = 0.0;
            for (int i = P.bus_off[g]; i < P.bus_off[g + 1]; ++i) acc += lb[i];
            loss_out[g] = acc;
        }
    }
}

// New r1 / x1 for n branches of instance g, applied to the plan's branch
// arrays and, when its Y-bus is assembled, to exactly the Y slots those
// branches contribute to (same per-slot routine as the full assembly).
void run_update_branches(gaphip_plan& P, int g, int n, const int* branch, const double* r, const double* x) {
    if (g < 0 || g >= P.n_inst || n < 0 || (n > 0 && (!branch || !r || !x))) throw Error(GAPHIP_E_ARG, "bad argument");
    if (n == 0) return;
    auto& N = P.nets[g];
    int const m = N.n_branch();
    for (int t = 0; t < n; ++t)
        if (branch[t] < 0 || branch[t] >= m) throw Error(GAPHIP_E_ARG, "branch index out of range");
    P.inputs_changed = P.branches_changed = P.g_par_dirty = true;
    if (P.br_slot_ptr.empty()) {
        P.br_slot_ptr.resize(P.n_inst);
        P.br_slot.resize(P.n_inst);
    }
    if (P.br_slot_ptr[g].empty()) {
        auto const& Y = P.yp[g];
        std::vector<int> ptr(m + 1, 0);
        int const slots = static_cast<int>(Y.ci.size());
        for (int k = 0; k < slots; ++k)
            for (int q = Y.seg[k]; q < Y.seg[k + 1]; ++q)
                if ((Y.code[q] & 3) != kContribShunt) ptr[(Y.code[q] >> 2) + 1]++;
        for (int b = 0; b < m; ++b) ptr[b + 1] += ptr[b];
        std::vector<int> sl(ptr[m]), fill(ptr.begin(), ptr.end() - 1);
        for (int k = 0; k < slots; ++k)
            for (int q = Y.seg[k]; q < Y.seg[k + 1]; ++q)
                if ((Y.code[q] & 3) != kContribShunt) sl[fill[Y.code[q] >> 2]++] = P.y_off[g] + k;
        P.br_slot_ptr[g] = std::move(ptr);
        P.br_slot[g] = std::move(sl);
    }
    std::vector<int> gidx(n), slots;
    for (int t = 0; t < n; ++t) {
        N.br_r[branch[t]] = r[t];
        N.br_x[branch[t]] = x[t];
        gidx[t] = static_cast<int>(P.br_off[g]) + branch[t];
        for (int q = P.br_slot_ptr[g][branch[t]]; q < P.br_slot_ptr[g][branch[t] + 1]; ++q) slots.push_back(P.br_slot[g][q]);
    }
    std::sort(slots.begin(), slots.end());
    slots.erase(std::unique(slots.begin(), slots.end()), slots.end());
    hipStream_t s = P.stream;
    std::vector<double> rv(r, r + n), xv(x, x + n);  // alive until the synchronisation below
    P.upd_idx.upload(gidx, s);
    P.upd_r.upload(rv, s);
    P.upd_x.upload(xv, s);
    hipLaunchKernelGGL(k_branch_rx_set, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, P.upd_idx.p, P.upd_r.p, P.upd_x.p,
                       P.br_r.p, P.br_x.p);
    bool const given = !P.y_given_mode.empty() && P.y_given_mode[g];
    if (P.y_assembled && !given && !slots.empty()) {
        sync_plan_y(P);
        P.upd_slots.upload(slots, s);
        YbusArgs ya{static_cast<int>(P.y_nnz), P.y_seg.p, P.y_code.p, P.y_slot_inst.p, P.br_r.p, P.br_x.p, P.br_g.p,
                    P.br_b.p, P.app_g.p, P.app_b.p, P.zb.p, P.y_pu.p, nullptr};
        int const ns = static_cast<int>(slots.size());
        hipLaunchKernelGGL(k_ybus_slots, dim3(blocks_for(ns, 256)), dim3(256), 0, s, ya, ns, P.upd_slots.p);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
}

// The plan's Y-bus of instance g (per unit, instance-local CSR).
void run_get_ybus(gaphip_plan const& P, int g, int* rp, int* ci, double* val_ri, int cap, int* nnz_out) {
    if (g < 0 || g >= P.n_inst) throw Error(GAPHIP_E_ARG, "bad instance");
    auto const& Y = P.yp[g];
    int const nnz = static_cast<int>(Y.ci.size());
    if (nnz_out) *nnz_out = nnz;
    if (!P.y_assembled) throw Error(GAPHIP_E_STATE, "the plan's Y-bus is assembled by the first solve");
    if (cap < nnz) throw Error(GAPHIP_E_CAPACITY, "output holds " + std::to_string(cap) + " entries, Y has " +
                                                      std::to_string(nnz));
    if (rp) std::copy(Y.rp.begin(), Y.rp.end(), rp);
    if (ci) std::copy(Y.ci.begin(), Y.ci.end(), ci);
    sync_plan_y(P);
    if (val_ri && nnz)
        HIPCHK(hipMemcpy(val_ri, P.y_pu.p + P.y_off[g], sizeof(double2) * nnz, hipMemcpyDeviceToHost));
}

// Branch flows at the last solution, kFlowFields per branch into host `out`.
void run_branch_flows(gaphip_plan& P, double* out) {
    if (!P.have_solution) throw Error(GAPHIP_E_STATE, "branch flows need a solve first");
    if (P.branches_changed)
        throw Error(GAPHIP_E_STATE, "branch data changed since the last solve: solve again before the flows");
    if (!out) throw Error(GAPHIP_E_ARG, "null output");
    if (P.n_br == 0) return;
    hipStream_t s = P.stream;
    int const G = P.n_inst;
    std::vector<double> zb(G), ib(G);
    for (int g = 0; g < G; ++g) {
        double const ub = P.nets[g].u_rated[0];
        zb[g] = ub * ub / P.last_base;
        ib[g] = P.last_base / (std::sqrt(3.0) * ub);
    }
    P.flow_zb.upload(zb, s);
    P.flow_ib.upload(ib, s);
    P.flow_out.alloc(static_cast<size_t>(kFlowFields) * P.n_br);
    FlowArgs fa{static_cast<int>(P.n_br), P.br_f.p, P.br_t.p, P.br_inst.p, P.br_r.p, P.br_x.p, P.br_g.p, P.br_b.p,
                P.flow_zb.p, P.flow_ib.p, P.v.p, P.last_base, P.flow_out.p};
    hipLaunchKernelGGL(k_branch_flow, dim3(blocks_for(P.n_br, 256)), dim3(256), 0, s, fa);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, P.flow_out.p, sizeof(double) * kFlowFields * P.n_br, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

}  // namespace gaphip

using namespace gaphip;

#define GAPHIP_GUARD(...)                                     \
    try {                                                     \
        __VA_ARGS__;                                          \
        return GAPHIP_OK;                                     \
    } catch (gaphip::Error const& e) {                        \
        gaphip::g_last_error = e.what();                      \
        return e.code;                                        \
    } catch (std::bad_alloc const&) {                         \
        gaphip::g_last_error = "host out of memory";          \
        return GAPHIP_E_DEVICE;                               \
    } catch (std::exception const& e) {                       \
        gaphip::g_last_error = e.what();                      \
        return GAPHIP_E_ARG;                                  \
    }

extern "C" {

int gaphip_abi_version(void) { return GAPHIP_ABI_VERSION; }
const char* gaphip_last_error(void) { return g_last_error.c_str(); }

int gaphip_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return std::strstr(prop.gcnArchName, "gfx950") != nullptr ? 1 : 0;
}

int gaphip_plan_create_opt(int n_inst, const gaphip_net* nets, int precision, const gaphip_plan_options* options,
                           gaphip_plan** out) {
    GAPHIP_GUARD({
        if (!out) throw Error(GAPHIP_E_ARG, "out is null");
        *out = nullptr;
        auto p = std::make_unique<gaphip_plan>();
        if (options) {
            if (options->relax_abs < 0 || !(options->relax_frac >= 0.0) || options->grid_min_class < 0 ||
                options->grid_min_class >= kGridClasses)
                throw Error(GAPHIP_E_ARG, "plan options out of range");
            p->opts = *options;
        }
        p->fr.grid_min_class = p->opts.grid_min_class;
        p->fr.grid_pack = p->opts.grid_pack != 0;
        build_plan(*p, n_inst, nets, precision);
        // default launch shape (measured on the 4,096-grid batch): with FP32
        // factors the biggest networks get 4 wavefronts; the FP64 kernel holds
        // more registers per wavefront and runs best with 2 throughout
        if (precision == GAPHIP_MIXED) {
            p->tuning.grid_waves = 2;
            p->tuning.grid_big_waves = 4;
            p->tuning.grid_big_work = 40000;
        }
        *out = p.release();
    })
}

int gaphip_plan_create(int n_inst, const gaphip_net* nets, int precision, gaphip_plan** out) {
    return gaphip_plan_create_opt(n_inst, nets, precision, nullptr, out);
}

int gaphip_analyze(int n_inst, const gaphip_net* nets, gaphip_plan_stats* o) {
    GAPHIP_GUARD({
        if (!o) throw Error(GAPHIP_E_ARG, "out is null");
        gaphip_plan P;
        analyze(P, n_inst, nets, GAPHIP_FP64);
        *o = gaphip_plan_stats{};
        o->n_inst = P.n_inst;
        o->n_bus = P.n_bus;
        o->n_branch = P.n_br;
        o->ybus_nnz = P.y_nnz;
        o->n_var = P.n_var;
        o->jac_nnz = P.j_nnz;
        for (auto const& L : P.lp) {
            for (int m : L.f_m) o->front_hist[front_hist_bin(m)]++;
            o->n_fronts += L.n_fronts;
            o->n_levels = std::max<long long>(o->n_levels, L.n_levels);
            o->front_entries += L.front_entries;
            o->factor_entries += L.factor_entries;
            o->max_front = std::max<long long>(o->max_front, L.max_front);
        }
        o->analysis_ms = P.analysis_ms;
    })
}

int gaphip_plan_destroy(gaphip_plan* plan) {
    GAPHIP_GUARD({ delete plan; })
}

int gaphip_plan_stats_get(const gaphip_plan* P, gaphip_plan_stats* o) {
    GAPHIP_GUARD({
        if (!P || !o) throw Error(GAPHIP_E_ARG, "null argument");
        o->n_inst = P->n_inst;
        o->n_bus = P->n_bus;
        o->n_branch = P->n_br;
        o->ybus_nnz = P->y_nnz;
        o->n_var = P->n_var;
        o->jac_nnz = P->j_nnz;
        o->n_fronts = P->fr.n_fronts;
        o->n_levels = P->fr.n_levels;
        o->front_entries = P->fr.front_entries;
        o->factor_entries = P->fr.factor_entries;
        o->max_front = P->fr.max_front;
        o->analysis_ms = P->analysis_ms;
        for (auto const& L : P->lp)
            for (int m : L.f_m) o->front_hist[front_hist_bin(m)]++;
    })
}

int gaphip_plan_set_injections(gaphip_plan* P, int inst, const double* p_w, const double* q_var) {
    GAPHIP_GUARD({
        if (!P || inst < 0 || inst >= P->n_inst || !p_w || !q_var) throw Error(GAPHIP_E_ARG, "bad argument");
        auto& N = P->nets[inst];
        N.p_w.assign(p_w, p_w + N.n_bus());
        N.q_var.assign(q_var, q_var + N.n_bus());
        P->sspec_dirty = true;
        P->inputs_changed = true;
    })
}

int gaphip_plan_set_branch_rx(gaphip_plan* P, int inst, const double* r, const double* x) {
    GAPHIP_GUARD({
        if (!P || inst < 0 || inst >= P->n_inst || !r || !x) throw Error(GAPHIP_E_ARG, "bad argument");
        auto& N = P->nets[inst];
        int const m = N.n_branch();
        N.br_r.assign(r, r + m);
        N.br_x.assign(x, x + m);
        P->inputs_changed = P->branches_changed = P->g_par_dirty = true;
        if (m) {
            HIPCHK(hipMemcpyAsync(P->br_r.p + P->br_off[inst], r, sizeof(double) * m, hipMemcpyHostToDevice,
                                  P->stream));
            HIPCHK(hipMemcpyAsync(P->br_x.p + P->br_off[inst], x, sizeof(double) * m, hipMemcpyHostToDevice,
                                  P->stream));
            HIPCHK(hipStreamSynchronize(P->stream));
        }
    })
}

int gaphip_plan_set_ybus(gaphip_plan* P, int inst, const int* rp, const int* ci, const double* v_ri, int nnz,
                         int per_unit) {
    GAPHIP_GUARD({
        if (!P || inst < 0 || inst >= P->n_inst) throw Error(GAPHIP_E_ARG, "bad argument");
        if (P->y_given_mode.empty()) P->y_given_mode.assign(P->n_inst, 0);
        P->inputs_changed = P->branches_changed = P->g_par_dirty = true;
        if (!rp) {
            P->y_given_mode[inst] = 0;
            return GAPHIP_OK;
        }
        auto const& Y = P->yp[inst];
        int const n = Y.n;
        if (!ci || !v_ri || nnz != static_cast<int>(Y.ci.size()) || !std::equal(rp, rp + n + 1, Y.rp.begin()) ||
            !std::equal(ci, ci + nnz, Y.ci.begin()))
            throw Error(GAPHIP_E_ARG, "given Y-bus pattern differs from the network's admittance pattern");
        if (!P->y_given.p) P->y_given.alloc(std::max<long long>(P->y_nnz, 1));
        if (nnz)
            HIPCHK(hipMemcpyAsync(P->y_given.p + P->y_off[inst], v_ri, sizeof(double2) * nnz, hipMemcpyHostToDevice,
                                  P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
        P->y_given_mode[inst] = per_unit ? 2 : 1;
    })
}

int gaphip_plan_set_schedule(gaphip_plan* P, int schedule, int* effective) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        if (schedule != GAPHIP_SCHED_AUTO && schedule != GAPHIP_SCHED_LEVELS && schedule != GAPHIP_SCHED_GRID)
            throw Error(GAPHIP_E_ARG, "unknown schedule");
        if (schedule == GAPHIP_SCHED_GRID && !P->fr.fz_ok)
            throw Error(GAPHIP_E_ARG, "the grid schedule needs every front to have at most 64 slots (largest: " +
                                          std::to_string(P->fr.max_front) +
                                          ") and every network at most 65,535 buses and fewer than 2^19 front "
                                          "storage and slot-vector entries");
        P->schedule = schedule;
        if (effective) *effective = P->grid_schedule() ? GAPHIP_SCHED_GRID : GAPHIP_SCHED_LEVELS;
    })
}

int gaphip_plan_set_graph(gaphip_plan* P, int on) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        P->tuning.level_graph = on != 0;
        P->drop_graph();
    })
}

int gaphip_plan_get_tuning(const gaphip_plan* P, gaphip_tuning* out) {
    GAPHIP_GUARD({
        if (!P || !out) throw Error(GAPHIP_E_ARG, "null argument");
        *out = P->tuning;
    })
}

int gaphip_plan_set_tuning(gaphip_plan* P, const gaphip_tuning* t) {
    GAPHIP_GUARD({
        if (!P || !t) throw Error(GAPHIP_E_ARG, "null argument");
        auto ok = [](int w) { return w == 1 || w == 2 || w == 4 || w == 8; };
        if (!ok(t->grid_waves) || !ok(t->grid_big_waves))
            throw Error(GAPHIP_E_ARG, "wavefronts per workgroup must be 1, 2, 4 or 8");
        if (t->big_outer < kBigPanel || t->big_outer > 1024 || t->big_outer % kBigPanel != 0)
            throw Error(GAPHIP_E_ARG, "big_outer must be a multiple of 64 between 64 and 1024");
        if (t->big_two_level_min < 0 || (t->level_graph != 0 && t->level_graph != 1))
            throw Error(GAPHIP_E_ARG, "big_two_level_min must be >= 0 and level_graph 0 or 1");
        P->tuning = *t;
        P->drop_graph();  // the captured step holds the panel widths
    })
}

int gaphip_solve(gaphip_plan* P, const gaphip_solve_opts* o, double* v_ri, int* iterations, int* converged,
                 double* final_mismatch) {
    GAPHIP_GUARD({
        if (!P || !o) throw Error(GAPHIP_E_ARG, "null argument");
        if (o->max_iterations < 0) throw Error(GAPHIP_E_ARG, "max_iterations must be >= 0");
        if (!(o->base_power > 0.0)) throw Error(GAPHIP_E_ARG, "base_power must be > 0");
        bool const grid = P->grid_schedule();
        if (P->precision == GAPHIP_FP64) {
            if (grid) run_solve_fused<double>(*P, *o, P->jv64.p, P->front64.p, P->vec64.p);
            else run_solve<double>(*P, *o, P->jv64.p, P->front64.p, P->vec64.p);
        } else {
            if (grid) run_solve_fused<float>(*P, *o, P->jv32.p, P->front32.p, P->vec32.p);
            else run_solve<float>(*P, *o, P->jv32.p, P->front32.p, P->vec32.p);
        }
        int const G = P->n_inst;
        std::vector<int> done(G), its(G);
        HIPCHK(hipMemcpy(done.data(), P->done.p, sizeof(int) * G, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(its.data(), P->iters.p, sizeof(int) * G, hipMemcpyDeviceToHost));
        for (int g = 0; g < G; ++g)
            if (!done[g]) its[g] = o->max_iterations;
        P->solution_converged = std::all_of(done.begin(), done.end(), [](int d) { return d != 0; });
        P->inputs_changed = P->branches_changed = false;
        if (iterations) std::copy(its.begin(), its.end(), iterations);
        if (converged) std::copy(done.begin(), done.end(), converged);
        if (final_mismatch) HIPCHK(hipMemcpy(final_mismatch, P->final_mis.p, sizeof(double) * G, hipMemcpyDeviceToHost));
        if (v_ri) HIPCHK(hipMemcpy(v_ri, P->v.p, sizeof(double2) * P->n_bus, hipMemcpyDeviceToHost));
        P->maxf_host.clear();
    })
}

int gaphip_trajectory(const gaphip_plan* P, int inst, double* out, int cap, int* n_out) {
    GAPHIP_GUARD({
        if (!P || inst < 0 || inst >= P->n_inst || !n_out) throw Error(GAPHIP_E_ARG, "bad argument");
        if (!P->have_solution) throw Error(GAPHIP_E_STATE, "no solve has run");
        int const G = P->n_inst, rows = P->last_iters_run;
        std::vector<unsigned long long> h(static_cast<size_t>(rows) * G);
        if (rows) HIPCHK(hipMemcpy(h.data(), P->maxf.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
        std::vector<int> its(G), done(G);
        HIPCHK(hipMemcpy(its.data(), P->iters.p, sizeof(int) * G, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(done.data(), P->done.p, sizeof(int) * G, hipMemcpyDeviceToHost));
        int n = done[inst] ? its[inst] : rows;
        *n_out = n;
        for (int k = 0; k < n && k < cap; ++k) {
            long long b = static_cast<long long>(h[static_cast<size_t>(k) * G + inst]);
            double d;
            std::memcpy(&d, &b, sizeof d);
            out[k] = d;
        }
    })
}

int gaphip_profile_enable(gaphip_plan* P, int on) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        P->prof.on = on != 0;
        P->prof.reset();
    })
}

int gaphip_profile_get(const gaphip_plan* P, int k, double* ms, long long* launches, double* bytes) {
    GAPHIP_GUARD({
        if (!P || k < 0 || k >= K_COUNT) throw Error(GAPHIP_E_ARG, "bad kernel id");
        if (ms) *ms = P->prof.ms[k];
        if (launches) *launches = P->prof.launches[k];
        if (bytes) *bytes = P->prof.bytes[k];
    })
}

const char* gaphip_kernel_name(int k) { return (k >= 0 && k < K_COUNT) ? kKernelNames[k] : ""; }
int gaphip_kernel_count(void) { return K_COUNT; }

}  // extern "C"

// ------------------------------------------------------------ inspection
namespace gaphip {

// Device memory of gaphip_power_mismatches_noalloc: one static array per GPU,
// sized for the largest evaluation it accepts (the EvalLayout of its limits,
// about 8.4 MB), held until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns on each device (the array's
// name is a host-side stand-in); that address is looked up once per device, on
// the entry point's first call there, and kept.  Calls on one device are
// serialised by g_scratch_call; every call overwrites the previous contents.
constexpr size_t kNoAllocScratchBytes =
    EvalLayout(GAPHIP_NOALLOC_MAX_BUS, GAPHIP_NOALLOC_MAX_BUS, GAPHIP_NOALLOC_MAX_NNZ, GAPHIP_NOALLOC_MAX_APP).total;
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kScratchDevices] = {};
std::mutex g_scratch_call[kScratchDevices];

}  // namespace gaphip

static __device__ __align__(256) unsigned char scratch[gaphip::kNoAllocScratchBytes];

namespace gaphip {

// Device address of `scratch` on device `dev` (current), looked up once.
unsigned char* scratch_address(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(GAPHIP_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_dev[dev];
}

// Argument checks of the mismatch entry points (no copy of the network: the
// no-allocation variant must not allocate); sets *m_out to the row count.
void check_eval_args(gaphip_net const* net, double const* v_ri, int y_rows, int const* rp, int const* ci,
                     double const* y_ri, int y_nnz, double base, double const* f, int cap, int* m_out) {
    if (!net || !m_out) throw Error(GAPHIP_E_ARG, "null argument");
    if (net->n_bus <= 0) throw Error(GAPHIP_E_ARG, "network has no buses");
    if (net->n_app < 0) throw Error(GAPHIP_E_ARG, "negative appliance count");
    if (!v_ri || !net->bus_id || !net->bus_type || !net->p_w || !net->q_var)
        throw Error(GAPHIP_E_ARG, "missing bus array or voltages");
    if (net->n_app > 0 && (!net->app_node || !net->app_status || !net->app_type || !net->app_p || !net->app_q))
        throw Error(GAPHIP_E_ARG, "missing appliance array");
    if (!(base > 0.0)) throw Error(GAPHIP_E_ARG, "base_power must be > 0");
    if (y_nnz < 0 || y_rows < 0) throw Error(GAPHIP_E_ARG, "negative admittance size");
    if (y_nnz > 0 && ((y_rows > 0 && !rp) || !ci || !y_ri)) throw Error(GAPHIP_E_ARG, "missing admittance array");
    int m = 0;
    for (int i = 0; i < net->n_bus; ++i) m += net->bus_type[i] == kSlack ? 0 : (net->bus_type[i] == kPQ ? 2 : 1);
    *m_out = m;
    if (m > cap) throw Error(GAPHIP_E_CAPACITY, "mismatch capacity too small");
    if (m > 0 && !f) throw Error(GAPHIP_E_ARG, "null output");
}

// Body shared by the two mismatch entry points: stages the inputs into the
// device region laid out by L, runs the row scan and the per-bus evaluation on
// stream s, copies F out.  rows = min(y_rows, n_bus).
void run_power_mismatches(gaphip_net const& net, double const* v_ri, int rows, int const* rp, int const* ci,
                          double const* y_ri, int y_nnz, double base, double* f, unsigned char* region,
                          EvalLayout const& L, hipStream_t s) {
    int const n = net.n_bus, na = net.n_app;
    auto at = [&](size_t off) { return region + off; };
    auto put = [&](size_t off, void const* src, size_t bytes) {
        if (bytes) HIPCHK(hipMemcpyAsync(at(off), src, bytes, hipMemcpyHostToDevice, s));
    };
    bool const have_y = y_nnz > 0 && rows > 0;
    if (have_y) {
        put(L.rp, rp, sizeof(int) * (static_cast<size_t>(rows) + 1));
        put(L.ci, ci, sizeof(int) * static_cast<size_t>(y_nnz));
        put(L.y, y_ri, sizeof(double2) * static_cast<size_t>(y_nnz));
    }
    put(L.v, v_ri, sizeof(double2) * n);
    put(L.bus_id, net.bus_id, sizeof(int) * n);
    put(L.bus_type, net.bus_type, sizeof(int) * n);
    put(L.p_w, net.p_w, sizeof(double) * n);
    put(L.q_var, net.q_var, sizeof(double) * n);
    if (na > 0) {
        put(L.app_node, net.app_node, sizeof(int) * na);
        put(L.app_status, net.app_status, sizeof(int) * na);
        put(L.app_type, net.app_type, sizeof(int) * na);
        put(L.app_p, net.app_p, sizeof(double) * na);
        put(L.app_q, net.app_q, sizeof(double) * na);
    }
    auto I = [&](size_t off) { return reinterpret_cast<int*>(at(off)); };
    auto D = [&](size_t off) { return reinterpret_cast<double*>(at(off)); };
    EvalArgs a{n, have_y ? rows : 0, have_y ? y_nnz : 0, na, base,
               I(L.rp), I(L.ci), reinterpret_cast<double2 const*>(at(L.y)), reinterpret_cast<double2 const*>(at(L.v)),
               I(L.bus_id), I(L.bus_type), D(L.p_w), D(L.q_var),
               I(L.app_node), I(L.app_status), I(L.app_type), D(L.app_p), D(L.app_q),
               I(L.peq), I(L.qeq), D(L.f), I(L.err)};
    hipLaunchKernelGGL(k_eval_rows, dim3(1), dim3(kEvalScanThreads), 0, s, a);
    hipLaunchKernelGGL(k_eval_mismatch, dim3(blocks_for(n, 256)), dim3(256), 0, s, a);
    HIPCHK(hipGetLastError());
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, at(L.err), sizeof(int), hipMemcpyDeviceToHost, s));
    int m = 0;
    for (int i = 0; i < n; ++i) m += net.bus_type[i] == kSlack ? 0 : (net.bus_type[i] == kPQ ? 2 : 1);
    if (m) HIPCHK(hipMemcpyAsync(f, at(L.f), sizeof(double) * m, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err) throw Error(GAPHIP_E_ARG, "negative admittance row pointer or column index");
}

// Evaluates Y, S, F and (optionally) J of one instance at the given voltages.
void evaluate_instance(gaphip_plan& P, int inst, double const* v_ri, double base, bool want_j) {
    hipStream_t s = P.stream;
    if (P.sspec_dirty || P.sspec_base != base) upload_sspec(P, base);
    if (P.maxf_rows < 1) {
        P.maxf.alloc(P.n_inst);
        P.active.alloc(1);
        P.maxf_rows = 1;
    }
    int const G = P.n_inst;
    std::vector<int> done(G, 1);
    done[inst] = 0;
    HIPCHK(hipMemcpyAsync(P.done.p, done.data(), sizeof(int) * G, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(P.maxf.p, 0, sizeof(unsigned long long) * G, s));
    int const nb = P.nets[inst].n_bus();
    HIPCHK(hipMemcpyAsync(P.v.p + P.bus_off[inst], v_ri, sizeof(double2) * nb, hipMemcpyHostToDevice, s));
    YbusArgs ya{static_cast<int>(P.y_nnz), P.y_seg.p, P.y_code.p, P.y_slot_inst.p, P.br_r.p, P.br_x.p, P.br_g.p,
                P.br_b.p, P.app_g.p, P.app_b.p, P.zb.p, P.y_pu.p, nullptr};
    hipLaunchKernelGGL(k_ybus, dim3(blocks_for(P.y_nnz, 256)), dim3(256), 0, s, ya);
    apply_given_y(P, base);
    P.y_from_grid = false;
    MisArgs ma{static_cast<int>(P.n_bus), P.y_rp.p, P.y_ci.p, P.y_pu.p, P.v.p, P.sspec.p, P.peq.p, P.qeq.p,
               P.bus_inst.p, P.done.p, P.s.p, P.vpol.p, P.f.p, P.maxf.p};
    hipLaunchKernelGGL(k_mismatch, dim3(blocks_for(P.n_bus, 256)), dim3(256), 0, s, ma);
    if (want_j) {
        if (!P.jv64.p) P.jv64.alloc(std::max<long long>(P.j_nnz, 1));
        JacArgs ja{static_cast<int>(P.n_items), P.it_i.p, P.it_j.p, P.it_y.p, P.it_pos.p, P.bus_inst.p, P.done.p,
                   P.y_pu.p, P.s.p, P.vpol.p};
        hipLaunchKernelGGL(k_jacobian<double>, dim3(blocks_for(P.n_items, 256)), dim3(256), 0, s, ja, P.jv64.p);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    P.have_solution = false;
}

}  // namespace gaphip

extern "C" {

int gaphip_jacobian(gaphip_plan* P, int inst, const double* v_ri, double base_power, int* n_var, int* row_ptr,
                    int* col_idx, double* values, int cap, int* nnz_out) {
    GAPHIP_GUARD({
        if (!P || inst < 0 || inst >= P->n_inst || !v_ri || !n_var || !nnz_out)
            throw Error(GAPHIP_E_ARG, "bad argument");
        if (!(base_power > 0.0)) throw Error(GAPHIP_E_ARG, "base_power must be > 0");
        evaluate_instance(*P, inst, v_ri, base_power, true);
        auto const& R = P->np[inst];
        int const nz = static_cast<int>(R.jci.size());
        std::vector<double> jv(nz);
        if (nz)
            HIPCHK(hipMemcpy(jv.data(), P->jv64.p + P->j_off[inst], sizeof(double) * nz, hipMemcpyDeviceToHost));
        // reference compaction: keep |J| > 1e-12 (cpu_newton_raphson.cpp:752-761)
        int k = 0;
        *n_var = R.n_var;
        if (row_ptr) row_ptr[0] = 0;
        for (int r = 0; r < R.n_var; ++r) {
            for (int q = R.jrp[r]; q < R.jrp[r + 1]; ++q) {
                if (!(std::fabs(jv[q]) > 1e-12)) continue;
                if (k < cap && col_idx && values) {
                    col_idx[k] = R.jci[q];
                    values[k] = jv[q];
                }
                ++k;
            }
            if (row_ptr) row_ptr[r + 1] = k;
        }
        *nnz_out = k;
        if (k > cap) throw Error(GAPHIP_E_CAPACITY, "Jacobian capacity too small");
    })
}

int gaphip_mismatches(gaphip_plan* P, int inst, const double* v_ri, double base_power, double* f, int cap,
                      int* m_out) {
    GAPHIP_GUARD({
        if (!P || inst < 0 || inst >= P->n_inst || !v_ri || !m_out) throw Error(GAPHIP_E_ARG, "bad argument");
        if (!(base_power > 0.0)) throw Error(GAPHIP_E_ARG, "base_power must be > 0");
        evaluate_instance(*P, inst, v_ri, base_power, false);
        int const m = P->np[inst].n_var;
        *m_out = m;
        if (m > cap) throw Error(GAPHIP_E_CAPACITY, "mismatch capacity too small");
        if (m) HIPCHK(hipMemcpy(f, P->f.p + P->var_off[inst], sizeof(double) * m, hipMemcpyDeviceToHost));
    })
}

int gaphip_power_mismatches(const gaphip_net* net, const double* v_ri, int y_rows, const int* y_row_ptr,
                            const int* y_col_idx, const double* y_ri, int y_nnz, double base_power, double* f,
                            int cap, int* m_out) {
    GAPHIP_GUARD({
        check_eval_args(net, v_ri, y_rows, y_row_ptr, y_col_idx, y_ri, y_nnz, base_power, f, cap, m_out);
        int dev_count = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
            throw Error(GAPHIP_E_DEVICE, "no HIP device visible");
        int const rows = std::min(y_rows, net->n_bus);
        EvalLayout const L(net->n_bus, rows, y_nnz, net->n_app);
        DBuf<unsigned char> region;
        region.alloc(L.total);
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        run_power_mismatches(*net, v_ri, rows, y_row_ptr, y_col_idx, y_ri, y_nnz, base_power, f, region.p, L, s);
    })
}

int gaphip_power_mismatches_noalloc(const gaphip_net* net, const double* v_ri, int y_rows, const int* y_row_ptr,
                                    const int* y_col_idx, const double* y_ri, int y_nnz, double base_power, double* f,
                                    int cap, int* m_out) {
    GAPHIP_GUARD({
        check_eval_args(net, v_ri, y_rows, y_row_ptr, y_col_idx, y_ri, y_nnz, base_power, f, cap, m_out);
        if (net->n_bus > GAPHIP_NOALLOC_MAX_BUS || y_nnz > GAPHIP_NOALLOC_MAX_NNZ || net->n_app > GAPHIP_NOALLOC_MAX_APP)
            throw Error(GAPHIP_E_CAPACITY,
                        "gaphip_power_mismatches_noalloc: network above the limits of its static device array "
                        "(at most 32768 buses, 262144 admittance entries, 32768 appliances)");
        int dev_count = 0, dev = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
            throw Error(GAPHIP_E_DEVICE, "no HIP device visible");
        HIPCHK(hipGetDevice(&dev));
        if (dev < 0 || dev >= kScratchDevices) throw Error(GAPHIP_E_DEVICE, "device ordinal above 63");
        unsigned char* const region = scratch_address(dev);
        int const rows = std::min(y_rows, net->n_bus);
        EvalLayout const L(net->n_bus, rows, y_nnz, net->n_app);
        if (L.total > kNoAllocScratchBytes) throw Error(GAPHIP_E_CAPACITY, "evaluation larger than the static array");
        std::lock_guard<std::mutex> lk(g_scratch_call[dev]);
        run_power_mismatches(*net, v_ri, rows, y_row_ptr, y_col_idx, y_ri, y_nnz, base_power, f, region, L, nullptr);
    })
}

int gaphip_ybus_build(const gaphip_net* net, int* row_ptr, int* col_idx, double* values_ri, int cap, int* nnz_out) {
    GAPHIP_GUARD({
        if (!net || !row_ptr || !nnz_out) throw Error(GAPHIP_E_ARG, "null argument");
        int dev_count = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
            throw Error(GAPHIP_E_DEVICE, "no HIP device visible");
        Net N = copy_net(*net);
        YPlan Y = plan_ybus(N);
        int const nnz = static_cast<int>(Y.ci.size());
        *nnz_out = nnz;
        std::copy(Y.rp.begin(), Y.rp.end(), row_ptr);
        if (nnz > cap) throw Error(GAPHIP_E_CAPACITY, "Y-bus capacity too small");
        std::copy(Y.ci.begin(), Y.ci.end(), col_idx);
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        DBuf<int> seg, code, slot_inst;
        DBuf<double> r, x, g, b, ag, ab, zb;
        DBuf<double2> ypu, ysi;
        seg.upload(Y.seg, s);
        code.upload(Y.code, s);
        slot_inst.upload(std::vector<int>(nnz, 0), s);
        r.upload(N.br_r, s);
        x.upload(N.br_x, s);
        g.upload(N.br_g, s);
        b.upload(N.br_b, s);
        ag.upload(N.app_g.empty() ? std::vector<double>{0.0} : N.app_g, s);
        ab.upload(N.app_b.empty() ? std::vector<double>{0.0} : N.app_b, s);
        zb.upload(std::vector<double>{1.0}, s);
        ypu.alloc(std::max(nnz, 1));
        ysi.alloc(std::max(nnz, 1));
        YbusArgs ya{nnz, seg.p, code.p, slot_inst.p, r.p, x.p, g.p, b.p, ag.p, ab.p, zb.p, ypu.p, ysi.p};
        if (nnz) hipLaunchKernelGGL(k_ybus, dim3(blocks_for(nnz, 256)), dim3(256), 0, s, ya);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        if (nnz) HIPCHK(hipMemcpy(values_ri, ysi.p, sizeof(double2) * nnz, hipMemcpyDeviceToHost));
    })
}

}  // extern "C"

// ------------------------------------------------------------ general LU
struct gaphip_lu {
    int n = 0;
    bool factorized = false;
    std::vector<int> pat_rp, pat_ci;  // pattern the symbolic plan was built for (permuted, duplicate-free)
    std::vector<int> row_of;          // row j of the factored matrix is row row_of[j] of the caller's
    std::vector<std::complex<double>> row_scale;  // per caller row: scaling (magnitude and phase)
    std::vector<double> col_scale;    // per column
    std::vector<int> uniq_of;         // CSR entry -> unique (i, j) index
    int n_uniq = 0;
    gaphip::LUPlan plan;
    gaphip::DevFronts fr;
    gaphip::DBuf<double> vals, front, vec, rhs, sol;
    gaphip::DBuf<int> fail;
    hipStream_t stream = nullptr;
    ~gaphip_lu() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gaphip {

// Maximum-product transversal with scaling (the approach of Duff and Koster's
// MC64): for every column j a distinct row i = row_of[j] such that the product
// of the |a_ij| is maximal, as a minimum-cost perfect matching with costs
// c_ij = log(max_k |a_kj|) - log |a_ij| >= 0, found by one shortest augmenting
// path per column (Dijkstra on the reduced costs c_ij - u_j - v_i >= 0, with
// the dual variables updated after each path).  At the optimum every matched
// entry has reduced cost 0 and every other a non-negative one, so the row
// scales exp(v_i) and column scales exp(u_j) / max_k |a_kj| make the matched
// entries 1 in magnitude and every other entry at most 1.  Zero entries are
// not edges.  Returns false when some column cannot be matched (the matrix is
// structurally singular).  rp / ci / mag: duplicate-free CSR of |A|.
bool max_product_matching(int n, std::vector<int> const& rp, std::vector<int> const& ci,
                          std::vector<double> const& mag, std::vector<int>& row_of, std::vector<double>& row_scale,
                          std::vector<double>& col_scale) {
    // columns as lists of (row, cost)
    std::vector<double> amax(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q) amax[ci[q]] = std::max(amax[ci[q]], mag[q]);
    std::vector<int> cp(n + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q)
            if (mag[q] > 0.0) cp[ci[q] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    std::vector<int> cr(cp[n]), fill(cp.begin(), cp.end() - 1);
    std::vector<double> cc(cp[n]);
    for (int i = 0; i < n; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            int const j = ci[q];
            if (!(mag[q] > 0.0)) continue;
            cr[fill[j]] = i;
            cc[fill[j]++] = std::log(amax[j]) - std::log(mag[q]);
        }
    for (int j = 0; j < n; ++j)
        if (cp[j] == cp[j + 1]) return false;
    std::vector<double> u(n, 0.0), v(n, 0.0);  // column and row duals (all costs are >= 0)
    std::vector<int> mate_row(n, -1), mate_col(n, -1);
    for (int j = 0; j < n; ++j)  // greedy start on zero-cost entries
        for (int q = cp[j]; q < cp[j + 1]; ++q)
            if (cc[q] == 0.0 && mate_row[cr[q]] < 0) {
                mate_row[cr[q]] = j;
                mate_col[j] = cr[q];
                break;
            }
    double const inf = std::numeric_limits<double>::infinity();
    std::vector<double> dist(n, inf);
    std::vector<int> pred(n, -1);
    std::vector<char> done(n, 0);
    std::vector<int> touched, final_rows;
    using Item = std::pair<double, int>;
    for (int j0 = 0; j0 < n; ++j0) {
        if (mate_col[j0] >= 0) continue;
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
        auto relax = [&](int j, double base) {
            for (int q = cp[j]; q < cp[j + 1]; ++q) {
                int const k = cr[q];
                if (done[k]) continue;
                double const nd = base + std::max(0.0, cc[q] - u[j] - v[k]);
                if (nd < dist[k]) {
                    if (dist[k] == inf) touched.push_back(k);
                    dist[k] = nd;
                    pred[k] = j;
                    heap.push({nd, k});
                }
            }
        };
        relax(j0, 0.0);
        int sink = -1;
        double dstar = 0.0;
        while (!heap.empty()) {
            auto const [d, i] = heap.top();
            heap.pop();
            if (done[i] || d > dist[i]) continue;
            done[i] = 1;
            final_rows.push_back(i);
            if (mate_row[i] < 0) {
                sink = i;
                dstar = d;
                break;
            }
            relax(mate_row[i], d);
        }
        if (sink < 0) return false;
        u[j0] += dstar;
        for (int i : final_rows) {
            double const delta = dstar - dist[i];
            if (i != sink) u[mate_row[i]] += delta;
            v[i] -= delta;
        }
        for (int i = sink;;) {
            int const j = pred[i];
            int const prev = mate_col[j];
            mate_col[j] = i;
            mate_row[i] = j;
            if (j == j0) break;
            i = prev;
        }
        for (int k : touched) {
            dist[k] = inf;
            pred[k] = -1;
            done[k] = 0;
        }
        touched.clear();
        final_rows.clear();
    }
    row_of = mate_col;
    row_scale.resize(n);
    col_scale.resize(n);
    for (int i = 0; i < n; ++i) row_scale[i] = std::exp(v[i]);
    for (int j = 0; j < n; ++j) col_scale[j] = std::exp(u[j]) / amax[j];
    return true;
}

// Complex matrix as a real system of twice the size: node i carries (re, im),
// entry a+ib of (i, j) becomes [[a, -b], [b, a]] on unknowns (2i, 2i+1) x (2j, 2j+1).
void lu_symbolic(gaphip_lu& L, int n, int const* rp, int const* ci) {
    LUInput in;
    in.n_nodes = n;
    in.w.assign(n, 2);
    std::vector<std::vector<int>> nb(n);
    L.uniq_of.assign(rp[n], -1);
    std::vector<int> last(n, -1);
    std::vector<int> urow, ucol;
    // unique (i, j) in CSR order; duplicates fold onto the first occurrence
    std::vector<int> first_at(n, -1);
    for (int i = 0; i < n; ++i) {
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            int j = ci[q];
            if (last[j] == i) {
                L.uniq_of[q] = first_at[j];
                continue;
            }
            last[j] = i;
            first_at[j] = static_cast<int>(urow.size());
            L.uniq_of[q] = first_at[j];
            urow.push_back(i);
            ucol.push_back(j);
            if (i != j) {
                nb[i].push_back(j);
                nb[j].push_back(i);
            }
        }
    }
    L.n_uniq = static_cast<int>(urow.size());
    in.adj_ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        std::sort(nb[i].begin(), nb[i].end());
        nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
        in.adj_ptr[i + 1] = in.adj_ptr[i] + static_cast<int>(nb[i].size());
        in.adj.insert(in.adj.end(), nb[i].begin(), nb[i].end());
    }
    for (int u = 0; u < L.n_uniq; ++u)
        for (int t = 0; t < 4; ++t) {
            in.ent_row.push_back(2 * urow[u] + (t >> 1));
            in.ent_col.push_back(2 * ucol[u] + (t & 1));
        }
    L.plan = plan_lu(in);
    std::vector<LUPlan const*> plans{&L.plan};
    std::vector<long long> base{0};
    auto ident = [](int, int key) { return key; };
    pack_fronts(plans, base, ident, ident, 2LL * n, 4LL * L.n_uniq, L.fr, L.stream);
    L.front.alloc(std::max<long long>(L.fr.front_entries, 1));
    L.vec.alloc(std::max<long long>(L.fr.vec_entries, 1));
    L.rhs.alloc(2 * n);
    L.sol.alloc(2 * n);
    L.fail.alloc(1);
    L.pat_rp.assign(rp, rp + n + 1);
    L.pat_ci.assign(ci, ci + rp[n]);
    L.n = n;
}

}  // namespace gaphip

extern "C" {

int gaphip_lu_create(gaphip_lu** out) {
    GAPHIP_GUARD({
        if (!out) throw Error(GAPHIP_E_ARG, "out is null");
        auto L = std::make_unique<gaphip_lu>();
        int dev_count = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
            throw Error(GAPHIP_E_DEVICE, "no HIP device visible");
        HIPCHK(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
        *out = L.release();
    })
}

int gaphip_lu_destroy(gaphip_lu* L) {
    GAPHIP_GUARD({ delete L; })
}

int gaphip_lu_is_factorized(const gaphip_lu* L) { return (L && L->factorized) ? 1 : 0; }

int gaphip_lu_factorize(gaphip_lu* L, int n_rows, int n_cols, const int* rp, const int* ci, const double* v_ri,
                        int nnz) {
    GAPHIP_GUARD({
        if (!L) throw Error(GAPHIP_E_ARG, "null LU handle");
        L->factorized = false;
        if (n_rows != n_cols) throw Error(GAPHIP_E_ARG, "Matrix must be square for LU factorization");
        if (nnz <= 0 || !v_ri || !rp || !ci || n_rows <= 0)
            throw Error(GAPHIP_E_ARG, "Matrix has no non-zero elements");
        int const n = n_rows;
        if (rp[0] != 0 || rp[n] != nnz) throw Error(GAPHIP_E_ARG, "row_ptr inconsistent with nnz");
        for (int i = 0; i < n; ++i)
            if (rp[i + 1] < rp[i]) throw Error(GAPHIP_E_ARG, "row_ptr not monotone");
        for (int q = 0; q < nnz; ++q)
            if (ci[q] < 0 || ci[q] >= n) throw Error(GAPHIP_E_ARG, "column index out of range");
        // duplicates summed onto their first occurrence in the row
        std::vector<int> urp(n + 1, 0), uci;
        std::vector<std::complex<double>> uval;
        {
            std::vector<int> at(n, -1);
            for (int i = 0; i < n; ++i) {
                for (int q = rp[i]; q < rp[i + 1]; ++q) {
                    int const j = ci[q];
                    std::complex<double> const a(v_ri[2 * q], v_ri[2 * q + 1]);
                    if (at[j] >= urp[i]) {
                        uval[at[j]] += a;
                        continue;
                    }
                    at[j] = static_cast<int>(uci.size());
                    uci.push_back(j);
                    uval.push_back(a);
                }
                urp[i + 1] = static_cast<int>(uci.size());
            }
        }
        // row permutation and scalings that put entries of magnitude 1 on the
        // diagonal with every other entry at most 1 (the reference pivots by
        // column maximum instead, cpu_lu_solver.cpp:469-498; the factorisation
        // below has static pivots)
        std::vector<double> mag(uci.size());
        for (size_t q = 0; q < uci.size(); ++q) mag[q] = std::abs(uval[q]);
        std::vector<double> rsc, csc;
        if (!max_product_matching(n, urp, uci, mag, L->row_of, rsc, csc))
            throw Error(GAPHIP_E_SINGULAR, "matrix is structurally singular (no row for some column)");
        // permuted, scaled matrix: row j = scaled row row_of[j]; each row also
        // takes the phase that makes its diagonal entry real and positive
        L->row_scale.assign(n, 0.0);
        L->col_scale = csc;
        std::vector<int> prp(n + 1, 0), pci;
        std::vector<double> pv;
        pci.reserve(uci.size());
        pv.reserve(2 * uci.size());
        for (int j = 0; j < n; ++j) {
            int const i = L->row_of[j];
            std::complex<double> diag(0.0);
            for (int q = urp[i]; q < urp[i + 1]; ++q)
                if (uci[q] == j) diag = uval[q] * (rsc[i] * csc[j]);
            std::complex<double> const phase = std::abs(diag) > 0.0 ? std::conj(diag) / std::abs(diag) : 1.0;
            L->row_scale[i] = rsc[i] * phase;
            for (int q = urp[i]; q < urp[i + 1]; ++q) {
                std::complex<double> const a = uval[q] * L->row_scale[i] * csc[uci[q]];
                pci.push_back(uci[q]);
                pv.push_back(a.real());
                pv.push_back(a.imag());
            }
            prp[j + 1] = static_cast<int>(pci.size());
        }
        int const pnz = prp[n];
        bool same = L->n == n && static_cast<int>(L->pat_ci.size()) == pnz &&
                    std::equal(prp.begin(), prp.end(), L->pat_rp.begin()) &&
                    std::equal(pci.begin(), pci.end(), L->pat_ci.begin());
        if (!same) lu_symbolic(*L, n, prp.data(), pci.data());
        // expanded real values in unique-entry order (no duplicates remain)
        std::vector<double> ex(4 * static_cast<size_t>(L->n_uniq));
        for (int q = 0; q < pnz; ++q) {
            int const u = L->uniq_of[q];
            ex[4 * u + 0] = pv[2 * q];
            ex[4 * u + 1] = -pv[2 * q + 1];
            ex[4 * u + 2] = pv[2 * q + 1];
            ex[4 * u + 3] = pv[2 * q];
        }
        L->vals.upload(ex, L->stream);
        HIPCHK(hipMemsetAsync(L->fail.p, 0, sizeof(int), L->stream));
        FrontArgs fa = L->fr.args(nullptr, L->fail.p);
        FrontRunner<double>::factor(L->fr, fa, L->vals.p, L->front.p, L->stream, 0, PlainHook{});
        HIPCHK(hipGetLastError());
        int failed = 0;
        HIPCHK(hipMemcpyAsync(&failed, L->fail.p, sizeof(int), hipMemcpyDeviceToHost, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
        if (failed) throw Error(GAPHIP_E_SINGULAR, "numerically zero pivot (|pivot| < 1e-14)");
        L->factorized = true;
    })
}

int gaphip_lu_solve(gaphip_lu* L, const double* rhs_ri, double* x_ri, int n) {
    GAPHIP_GUARD({
        if (!L) throw Error(GAPHIP_E_ARG, "null LU handle");
        if (!L->factorized) throw Error(GAPHIP_E_STATE, "Matrix not factorized. Call factorize() first.");
        if (n != L->n) throw Error(GAPHIP_E_ARG, "RHS size does not match matrix dimension");
        if (!rhs_ri || !x_ri) throw Error(GAPHIP_E_ARG, "null vector");
        // A x = b  <=>  (P Dr A Dc) y = P Dr b,  x = Dc y
        std::vector<double> b(2 * static_cast<size_t>(n));
        for (int j = 0; j < n; ++j) {
            int const i = L->row_of[j];
            std::complex<double> const v = std::complex<double>(rhs_ri[2 * i], rhs_ri[2 * i + 1]) * L->row_scale[i];
            b[2 * j] = v.real();
            b[2 * j + 1] = v.imag();
        }
        HIPCHK(hipMemcpyAsync(L->rhs.p, b.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, L->stream));
        FrontArgs fa = L->fr.args(nullptr, L->fail.p);
        FrontRunner<double>::forward(L->fr, fa, L->front.p, L->vec.p, L->rhs.p, 1.0, L->stream, 0, PlainHook{});
        FrontRunner<double>::backward(L->fr, fa, L->front.p, L->vec.p, L->sol.p, L->stream, 0, PlainHook{});
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(b.data(), L->sol.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
        for (int c = 0; c < n; ++c) {
            x_ri[2 * c] = b[2 * c] * L->col_scale[c];
            x_ri[2 * c + 1] = b[2 * c + 1] * L->col_scale[c];
        }
    })
}

}  // extern "C"

extern "C" int gaphip_plan_update_branches(gaphip_plan* P, int inst, int n, const int* branch, const double* r,
                                           const double* x) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        run_update_branches(*P, inst, n, branch, r, x);
    })
}

extern "C" int gaphip_plan_get_ybus(const gaphip_plan* P, int inst, int* row_ptr, int* col_idx, double* values_ri,
                                    int cap, int* nnz) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        run_get_ybus(*P, inst, row_ptr, col_idx, values_ri, cap, nnz);
    })
}

extern "C" int gaphip_branch_flows(gaphip_plan* P, double* out) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        run_branch_flows(*P, out);
    })
}

extern "C" int gaphip_adjoint(gaphip_plan* P, double w_loss, double w_dev, const double* g_vm, const double* g_va,
                              double* grad_p, double* grad_q, double* grad_r, double* grad_x, double* loss_out) {
    GAPHIP_GUARD({
        if (!P) throw Error(GAPHIP_E_ARG, "null plan");
        run_adjoint(*P, w_loss, w_dev, g_vm, g_va, grad_p, grad_q, grad_r, grad_x, loss_out);
    })
}
