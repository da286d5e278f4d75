// Host-side symbolic analysis for the GPU Newton-Raphson path (C++20).
//
// Everything here depends only on topology and bus types, so it runs once per
// network and is reused by every solve:
//   * YPlan  : CSR pattern of the Y-bus and, per CSR slot, the ordered list of
//              admittance contributions (the reference accumulation order of
//              src/admittance/cpu/cpu_admittance_matrix.cpp:86-154), so the GPU
//              assembly is a segmented scatter with bit-identical sums;
//   * NRPlan : Newton variable/equation numbering (cpu_newton_raphson.cpp:603-627),
//              the structural Jacobian pattern in the reference layout, and the
//              per-(i,j) work items of the fused Jacobian fill;
//   * LUPlan : multifrontal factorisation plan on a node graph with one or two
//              unknowns per node: minimum-degree order, elimination tree,
//              fundamental supernodes, fronts, level schedule and all scatter maps.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gaphip {

struct Net {
    std::vector<int> bus_id, bus_type, br_id, br_from, br_to, br_status, app_id, app_node, app_status,
        app_type;
    std::vector<double> u_rated, u_pu, p_w, q_var, br_r, br_x, br_g, br_b, app_p, app_q, app_g, app_b;
    int n_bus() const { return static_cast<int>(bus_id.size()); }
    int n_branch() const { return static_cast<int>(br_id.size()); }
    int n_app() const { return static_cast<int>(app_id.size()); }
};

enum : int { kPQ = 0, kPV = 1, kSlack = 2 };
enum : int { kSource = 0, kLoadgen = 1, kShunt = 2 };

// Contribution codes of the Y assembly: (index << 2) | kind.
enum : int { kContribShunt = 0, kContribBranchDiag = 1, kContribBranchOff = 2 };

struct YPlan {
    int n = 0;
    std::vector<int> rp, ci;   // CSR pattern, duplicates kept in branch order
    std::vector<int> seg;      // nnz+1: contributions of slot k are code[seg[k] .. seg[k+1])
    std::vector<int> code;
    std::vector<int> id2idx;   // dense bus-id table (cpu_admittance_matrix.cpp:56-73)
};
YPlan plan_ybus(Net const& net);

struct NRPlan {
    int n_bus = 0, slack = -1, n_var = 0, n_pq = 0;
    std::vector<int> ang, mag;  // per bus: variable index or -1
    std::vector<int> peq, qeq;  // per bus: equation index or -1
    std::vector<int> jrp, jci;  // structural Jacobian pattern (reference layout)
    // Jacobian fill items, one per unique (i, j) of Y with i non-slack
    std::vector<int> it_i, it_j, it_y;  // it_y: Y slot used for G/B (last duplicate)
    std::vector<int> it_pos;            // 4 per item: J slots of (P,th) (P,V) (Q,th) (Q,V), or -1
    std::vector<int> ydiag;             // per bus: Y slot of the (last) diagonal entry or -1
};
// Returns false with a message when the network is not supported on the GPU path.
bool plan_nr(Net const& net, YPlan const& y, NRPlan& out, std::string& err);

// Node graph with w[node] in {1,2} unknowns; unknown key = 2*node + sub.
struct LUInput {
    int n_nodes = 0;
    std::vector<int> adj_ptr, adj;  // symmetric, unique, no self loops
    std::vector<int> w;
    std::vector<int> ent_row, ent_col;  // keys of each matrix entry, in source order
};

struct LUPlan {
    int n_nodes = 0, n_unk = 0, n_fronts = 0, n_levels = 0;
    std::vector<int> order;                // elimination position -> node
    std::vector<int> f_m, f_p, f_level, f_parent;
    std::vector<int64_t> f_off;            // front storage offset (m*m entries)
    std::vector<int64_t> f_voff;           // vector slot offset (m entries)
    std::vector<int> f_slot_ptr, slot_key; // slots of each front (pivots first)
    std::vector<int> f_child_ptr, child;   // children of each front
    std::vector<int> f_rel_off;            // per front: offset of its map into the parent front
    std::vector<int> rel;
    std::vector<int> f_arrow_ptr, arrow_src, arrow_dst;  // entry k -> (row << 16) | col in its front
    std::vector<int> level_ptr, level_front;             // fronts grouped by level
    int64_t front_entries = 0, vec_entries = 0, factor_entries = 0;
    int max_front = 0;
};
// Supernode relaxation: a child supernode is merged into its parent when it
// is contiguous with it in the postorder and the merge adds at most
// max(relax_abs, relax_frac * merged entries) explicit zeros (in node units).
struct LUOptions {
    int relax_abs = 4;
    double relax_frac = 0.10;
};
LUPlan plan_lu(LUInput const& in, LUOptions const& opt = LUOptions{});

// Minimum-degree elimination order of an undirected graph (no self loops).
std::vector<int> min_degree_order(int n, std::vector<int> const& adj_ptr, std::vector<int> const& adj);

}  // namespace gaphip
