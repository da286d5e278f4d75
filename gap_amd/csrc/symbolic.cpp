// Host symbolic analysis; see symbolic.hpp for what each plan holds.
#include "symbolic.hpp"

#include <algorithm>
#include <numeric>
#include <queue>
#include <stdexcept>

namespace gaphip {

// ------------------------------------------------------------------ Y-bus
// Triplet order of cpu_admittance_matrix.cpp:100-154: two off-diagonal
// triplets per in-service branch in branch order, then one diagonal triplet per
// bus, stable-sorted by (row, col).  The diagonal of bus i accumulates its shunt
// appliances (appliance order, :88-97) and then every branch touching it
// (branch order, :134-135).
YPlan plan_ybus(Net const& net) {
    YPlan y;
    int const n = net.n_bus();
    y.n = n;
    int max_id = -1;
    for (int id : net.bus_id) max_id = std::max(max_id, id);
    y.id2idx.assign(static_cast<size_t>(std::max(max_id + 1, 0)), -1);
    for (int i = 0; i < n; ++i) {
        int id = net.bus_id[i];
        if (id >= 0 && id <= max_id) y.id2idx[id] = i;
    }
    auto lookup = [&](int id) { return (id >= 0 && id <= max_id) ? y.id2idx[id] : -1; };

    // diagonal contribution lists (counting pass, then fill)
    std::vector<int> dcount(n, 0);
    for (int a = 0; a < net.n_app(); ++a) {
        if (net.app_type[a] != kShunt) continue;
        int b = lookup(net.app_node[a]);
        if (b >= 0 && b < n) dcount[b]++;
    }
    struct Trip {
        int row, col, code;
    };
    std::vector<Trip> trips;
    trips.reserve(static_cast<size_t>(2 * net.n_branch() + n));
    std::vector<int> bfrom(net.n_branch(), -1), bto(net.n_branch(), -1);
    for (int k = 0; k < net.n_branch(); ++k) {
        if (!net.br_status[k]) continue;
        int f = lookup(net.br_from[k]), t = lookup(net.br_to[k]);
        if (f < 0 || f >= n || t < 0 || t >= n) continue;
        bfrom[k] = f;
        bto[k] = t;
        dcount[f]++;
        dcount[t]++;
        trips.push_back({f, t, (k << 2) | kContribBranchOff});
        trips.push_back({t, f, (k << 2) | kContribBranchOff});
    }
    std::vector<int> dptr(n + 1, 0);
    for (int i = 0; i < n; ++i) dptr[i + 1] = dptr[i] + dcount[i];
    std::vector<int> dcode(dptr[n]);
    std::vector<int> fillp(dptr.begin(), dptr.end() - 1);
    for (int a = 0; a < net.n_app(); ++a) {
        if (net.app_type[a] != kShunt) continue;
        int b = lookup(net.app_node[a]);
        if (b >= 0 && b < n) dcode[fillp[b]++] = (a << 2) | kContribShunt;
    }
    for (int k = 0; k < net.n_branch(); ++k) {
        if (bfrom[k] < 0) continue;
        dcode[fillp[bfrom[k]]++] = (k << 2) | kContribBranchDiag;
        dcode[fillp[bto[k]]++] = (k << 2) | kContribBranchDiag;
    }
    // diagonal triplets carry -(i+1) as a marker for "bus i's accumulated list"
    for (int i = 0; i < n; ++i) trips.push_back({i, i, -(i + 1)});
    std::stable_sort(trips.begin(), trips.end(), [](Trip const& a, Trip const& b) {
        if (a.row != b.row) return a.row < b.row;
        return a.col < b.col;
    });
    int const nnz = static_cast<int>(trips.size());
    y.rp.assign(n + 1, 0);
    y.ci.resize(nnz);
    y.seg.resize(nnz + 1);
    y.seg[0] = 0;
    for (int k = 0; k < nnz; ++k) {
        y.rp[trips[k].row + 1]++;
        y.ci[k] = trips[k].col;
        if (trips[k].code >= 0) {
            y.code.push_back(trips[k].code);
        } else {
            int i = -trips[k].code - 1;
            for (int p = dptr[i]; p < dptr[i + 1]; ++p) y.code.push_back(dcode[p]);
        }
        y.seg[k + 1] = static_cast<int>(y.code.size());
    }
    for (int i = 0; i < n; ++i) y.rp[i + 1] += y.rp[i];
    return y;
}

// --------------------------------------------------------- Newton numbering
bool plan_nr(Net const& net, YPlan const& y, NRPlan& p, std::string& err) {
    int const n = net.n_bus();
    p = NRPlan{};
    p.n_bus = n;
    int n_slack = 0;
    for (int i = 0; i < n; ++i) {
        int t = net.bus_type[i];
        if (t != kPQ && t != kPV && t != kSlack) {
            err = "bus " + std::to_string(net.bus_id[i]) + " has unknown bus_type " + std::to_string(t);
            return false;
        }
        if (t == kSlack) {
            if (p.slack < 0) p.slack = i;
            ++n_slack;
        }
        if (t == kPQ) ++p.n_pq;
    }
    if (n_slack > 1) {
        // the reference gives angle variables to every bus but the first slack while
        // skipping equations of all slack buses (cpu_newton_raphson.cpp:530-541, 615-627),
        // which is not a square pairing; the GPU path refuses it.
        err = "more than one SLACK bus is not supported by the GPU path";
        return false;
    }
    p.ang.assign(n, -1);
    p.mag.assign(n, -1);
    p.peq.assign(n, -1);
    p.qeq.assign(n, -1);
    int v = 0;
    for (int i = 0; i < n; ++i)
        if (i != p.slack) p.ang[i] = v++;
    for (int i = 0; i < n; ++i)
        if (net.bus_type[i] == kPQ) p.mag[i] = v++;
    int e = 0;
    for (int i = 0; i < n; ++i) {
        if (net.bus_type[i] == kSlack) continue;
        p.peq[i] = e++;
        if (net.bus_type[i] == kPQ) p.qeq[i] = e++;
    }
    p.n_var = v;
    if (e != v) {
        err = "inconsistent Newton variable count";
        return false;
    }
    p.ydiag.assign(n, -1);
    // unique columns per row with their last duplicate slot
    std::vector<int> ucol, uslot;
    p.jrp.assign(p.n_var + 1, 0);
    struct RowItems {
        int first, count;
    };
    std::vector<int> row_len(p.n_var, 0);
    // first pass: count J entries per row
    for (int i = 0; i < n; ++i) {
        for (int q = y.rp[i]; q < y.rp[i + 1]; ++q)
            if (y.ci[q] == i) p.ydiag[i] = q;
        if (net.bus_type[i] == kSlack) continue;
        int na = 0, nm = 0;
        for (int q = y.rp[i]; q < y.rp[i + 1]; ++q) {
            int j = y.ci[q];
            if (q + 1 < y.rp[i + 1] && y.ci[q + 1] == j) continue;  // not the last duplicate
            if (p.ang[j] >= 0) ++na;
            if (p.mag[j] >= 0) ++nm;
        }
        row_len[p.peq[i]] = na + nm;
        if (p.qeq[i] >= 0) row_len[p.qeq[i]] = na + nm;
    }
    for (int r = 0; r < p.n_var; ++r) p.jrp[r + 1] = p.jrp[r] + row_len[r];
    p.jci.resize(p.jrp[p.n_var]);
    for (int i = 0; i < n; ++i) {
        if (net.bus_type[i] == kSlack) continue;
        int na = 0;
        for (int q = y.rp[i]; q < y.rp[i + 1]; ++q) {
            int j = y.ci[q];
            if (q + 1 < y.rp[i + 1] && y.ci[q + 1] == j) continue;
            if (p.ang[j] >= 0) ++na;
        }
        int ra = 0, rm = 0;
        for (int q = y.rp[i]; q < y.rp[i + 1]; ++q) {
            int j = y.ci[q];
            if (q + 1 < y.rp[i + 1] && y.ci[q + 1] == j) continue;
            int pos4[4] = {-1, -1, -1, -1};
            if (p.ang[j] >= 0) {
                pos4[0] = p.jrp[p.peq[i]] + ra;
                p.jci[pos4[0]] = p.ang[j];
                if (p.qeq[i] >= 0) {
                    pos4[2] = p.jrp[p.qeq[i]] + ra;
                    p.jci[pos4[2]] = p.ang[j];
                }
                ++ra;
            }
            if (p.mag[j] >= 0) {
                pos4[1] = p.jrp[p.peq[i]] + na + rm;
                p.jci[pos4[1]] = p.mag[j];
                if (p.qeq[i] >= 0) {
                    pos4[3] = p.jrp[p.qeq[i]] + na + rm;
                    p.jci[pos4[3]] = p.mag[j];
                }
                ++rm;
            }
            if (pos4[0] < 0 && pos4[1] < 0) continue;
            p.it_i.push_back(i);
            p.it_j.push_back(j);
            p.it_y.push_back(q);
            for (int t = 0; t < 4; ++t) p.it_pos.push_back(pos4[t]);
        }
    }
    return true;
}

// --------------------------------------------------------- minimum degree
// Explicit elimination graph with a lazy (degree, node) heap; ties go to the
// lower node index so the order is deterministic.  A remaining graph of at
// least kDenseMin nodes whose minimum degree reaches half its size is ordered
// directly (it becomes one dense front either way).
constexpr int kDenseMin = 512;

std::vector<int> min_degree_order(int n, std::vector<int> const& adj_ptr, std::vector<int> const& adj) {
    std::vector<std::vector<int>> g(n);
    for (int i = 0; i < n; ++i) g[i].assign(adj.begin() + adj_ptr[i], adj.begin() + adj_ptr[i + 1]);
    std::vector<char> gone(n, 0);
    std::vector<int> mark(n, -1), order;
    order.reserve(n);
    using Ent = std::pair<int, int>;
    std::priority_queue<Ent, std::vector<Ent>, std::greater<Ent>> heap;
    for (int i = 0; i < n; ++i) heap.push({static_cast<int>(g[i].size()), i});
    int stamp = 0, left = n;
    while (!heap.empty()) {
        auto [d, v] = heap.top();
        heap.pop();
        if (gone[v] || d != static_cast<int>(g[v].size())) continue;
        if (d == left - 1 && left > 2) {
            // the remaining graph is a clique: every node has degree left - 1,
            // and eliminating one keeps the rest a clique with equal degrees,
            // so the order continues by ascending index (the tie rule)
            for (int u = 0; u < n; ++u)
                if (!gone[u]) order.push_back(u);
            break;
        }
        if (left >= kDenseMin && 2LL * d >= left - 1) {
            // at least half dense: eliminating it fills it in after a few
            // steps whatever the order, while each step here would cost
            // O(degree^2); finish by (current degree, index)
            std::vector<std::pair<int, int>> rest;
            for (int u = 0; u < n; ++u)
                if (!gone[u]) rest.push_back({static_cast<int>(g[u].size()), u});
            std::sort(rest.begin(), rest.end());
            for (auto const& e : rest) order.push_back(e.second);
            break;
        }
        gone[v] = 1;
        --left;
        order.push_back(v);
        std::vector<int> nb;
        nb.swap(g[v]);
        for (int u : nb) {
            auto& gu = g[u];
            gu.erase(std::remove(gu.begin(), gu.end(), v), gu.end());
            ++stamp;
            for (int x : gu) mark[x] = stamp;
            mark[u] = stamp;
            for (int x : nb)
                if (mark[x] != stamp) {
                    gu.push_back(x);
                    mark[x] = stamp;
                }
        }
        for (int u : nb) heap.push({static_cast<int>(g[u].size()), u});
    }
    return order;
}

// ------------------------------------------------------------ multifrontal
LUPlan plan_lu(LUInput const& in, LUOptions const& opt) {
    LUPlan P;
    int const n = in.n_nodes;
    P.n_nodes = n;
    P.n_unk = 0;
    for (int w : in.w) P.n_unk += w;
    std::vector<int> order = min_degree_order(n, in.adj_ptr, in.adj);
    std::vector<int> pos(n);
    for (int k = 0; k < n; ++k) pos[order[k]] = k;

    // symbolic factorisation in position space: S[k] = rows of L below k
    std::vector<std::vector<int>> S(n), kids(n);
    std::vector<int> parent(n, -1), mark(n, -1);
    for (int k = 0; k < n; ++k) {
        std::vector<int>& s = S[k];
        int v = order[k];
        mark[k] = k;
        for (int q = in.adj_ptr[v]; q < in.adj_ptr[v + 1]; ++q) {
            int r = pos[in.adj[q]];
            if (r > k && mark[r] != k) {
                mark[r] = k;
                s.push_back(r);
            }
        }
        for (int c : kids[k])
            for (int r : S[c])
                if (r > k && mark[r] != k) {
                    mark[r] = k;
                    s.push_back(r);
                }
        std::sort(s.begin(), s.end());
        if (!s.empty()) {
            parent[k] = s[0];
            kids[s[0]].push_back(k);
        }
    }
    // postorder the elimination forest so supernodes are contiguous
    std::vector<int> post;
    post.reserve(n);
    {
        std::vector<int> stk, it;
        for (int r = 0; r < n; ++r) {
            if (parent[r] >= 0) continue;
            stk.assign(1, r);
            it.assign(1, 0);
            while (!stk.empty()) {
                int x = stk.back();
                int& c = it.back();
                if (c < static_cast<int>(kids[x].size())) {
                    int ch = kids[x][c++];
                    stk.push_back(ch);
                    it.push_back(0);
                } else {
                    post.push_back(x);
                    stk.pop_back();
                    it.pop_back();
                }
            }
        }
    }
    std::vector<int> newpos(n);
    for (int k = 0; k < n; ++k) newpos[post[k]] = k;
    std::vector<std::vector<int>> S2(n);
    std::vector<int> par2(n, -1), ord2(n);
    for (int k = 0; k < n; ++k) {
        int o = post[k];
        ord2[k] = order[o];
        par2[k] = parent[o] >= 0 ? newpos[parent[o]] : -1;
        S2[k].reserve(S[o].size());
        for (int r : S[o]) S2[k].push_back(newpos[r]);
        std::sort(S2[k].begin(), S2[k].end());
    }
    S.swap(S2);
    parent.swap(par2);
    order.swap(ord2);
    for (int k = 0; k < n; ++k) pos[order[k]] = k;
    P.order = order;
    std::vector<int> nkids(n, 0);
    for (int k = 0; k < n; ++k)
        if (parent[k] >= 0) nkids[parent[k]]++;

    // fundamental supernodes
    std::vector<int> sn_first;
    for (int k = 0; k < n; ++k) {
        bool join = k > 0 && parent[k - 1] == k && nkids[k] == 1 && S[k - 1].size() == S[k].size() + 1;
        if (!join) sn_first.push_back(k);
    }
    // relaxed amalgamation, bottom-up in postorder: absorb the supernode that
    // ends right before this one when it is a child and adds few zeros
    {
        int const n0 = static_cast<int>(sn_first.size());
        std::vector<int> first = sn_first, last(n0);
        for (int s = 0; s < n0; ++s) last[s] = (s + 1 < n0 ? sn_first[s + 1] : n) - 1;
        std::vector<int> out_first;
        std::vector<long long> out_zeros;  // explicit zeros already inside each merged supernode
        std::vector<int> out_last;
        for (int s = 0; s < n0; ++s) {
            int f = first[s], l = last[s];
            long long zeros = 0;
            while (!out_first.empty()) {
                int cf = out_first.back(), cl = out_last.back();
                if (parent[cl] < f || parent[cl] > l || cl + 1 != f) break;
                long long nc = cl - cf + 1, np = l - f + 1;
                long long sp = static_cast<long long>(S[l].size());
                long long sc = static_cast<long long>(S[cl].size());
                long long add = nc * (np + sp - sc);
                long long tot = zeros + out_zeros.back() + add;
                long long mrg = nc + np;
                long long entries = mrg * (mrg + 1) / 2 + mrg * sp;
                if (tot > std::max<long long>(opt.relax_abs, static_cast<long long>(opt.relax_frac * entries))) break;
                f = cf;
                zeros = tot;
                out_first.pop_back();
                out_last.pop_back();
                out_zeros.pop_back();
            }
            out_first.push_back(f);
            out_last.push_back(l);
            out_zeros.push_back(zeros);
        }
        sn_first = out_first;
    }
    std::vector<int> sn_of(n);
    for (int s = 0, k = 0; s < static_cast<int>(sn_first.size()); ++s) {
        int end = s + 1 < static_cast<int>(sn_first.size()) ? sn_first[s + 1] : n;
        for (; k < end; ++k) sn_of[k] = s;
    }
    int const ns = static_cast<int>(sn_first.size());
    sn_first.push_back(n);
    P.n_fronts = ns;
    P.f_m.assign(ns, 0);
    P.f_p.assign(ns, 0);
    P.f_level.assign(ns, 0);
    P.f_parent.assign(ns, -1);
    P.f_slot_ptr.assign(ns + 1, 0);
    P.f_off.assign(ns + 1, 0);
    P.f_voff.assign(ns + 1, 0);
    // slots: pivot nodes then structure nodes, each node expanded to its unknowns
    for (int s = 0; s < ns; ++s) {
        int k0 = sn_first[s], k1 = sn_first[s + 1] - 1;
        int pcount = 0;
        for (int k = k0; k <= k1; ++k) {
            int node = order[k];
            for (int sub = 0; sub < in.w[node]; ++sub) P.slot_key.push_back(2 * node + sub);
            pcount += in.w[node];
        }
        int m = pcount;
        for (int r : S[k1]) {
            int node = order[r];
            for (int sub = 0; sub < in.w[node]; ++sub) P.slot_key.push_back(2 * node + sub);
            m += in.w[node];
        }
        P.f_p[s] = pcount;
        P.f_m[s] = m;
        P.f_slot_ptr[s + 1] = static_cast<int>(P.slot_key.size());
        P.f_off[s + 1] = P.f_off[s] + static_cast<int64_t>(m) * m;
        P.f_voff[s + 1] = P.f_voff[s] + m;
        P.factor_entries += static_cast<int64_t>(pcount) * m + static_cast<int64_t>(m - pcount) * pcount;
        P.max_front = std::max(P.max_front, m);
        if (parent[k1] >= 0) P.f_parent[s] = sn_of[parent[k1]];
    }
    P.front_entries = P.f_off[ns];
    P.vec_entries = P.f_voff[ns];
    // children and levels (children precede parents in postorder)
    std::vector<int> ccount(ns + 1, 0);
    for (int s = 0; s < ns; ++s)
        if (P.f_parent[s] >= 0) ccount[P.f_parent[s] + 1]++;
    P.f_child_ptr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) P.f_child_ptr[s + 1] = P.f_child_ptr[s] + ccount[s + 1];
    P.child.resize(P.f_child_ptr[ns]);
    {
        std::vector<int> fp(P.f_child_ptr.begin(), P.f_child_ptr.end() - 1);
        for (int s = 0; s < ns; ++s)
            if (P.f_parent[s] >= 0) P.child[fp[P.f_parent[s]]++] = s;
    }
    P.n_levels = 0;
    for (int s = 0; s < ns; ++s) {
        int lv = 0;
        for (int q = P.f_child_ptr[s]; q < P.f_child_ptr[s + 1]; ++q)
            lv = std::max(lv, P.f_level[P.child[q]] + 1);
        P.f_level[s] = lv;
        P.n_levels = std::max(P.n_levels, lv + 1);
    }
    P.level_ptr.assign(P.n_levels + 1, 0);
    for (int s = 0; s < ns; ++s) P.level_ptr[P.f_level[s] + 1]++;
    for (int l = 0; l < P.n_levels; ++l) P.level_ptr[l + 1] += P.level_ptr[l];
    P.level_front.resize(ns);
    {
        std::vector<int> fp(P.level_ptr.begin(), P.level_ptr.end() - 1);
        for (int s = 0; s < ns; ++s) P.level_front[fp[P.f_level[s]]++] = s;
    }
    // slot position of a node inside a front: first slot of that node
    std::vector<int> slot_at(n, -1), stamp_of(n, -1);
    auto map_front = [&](int s) {
        for (int t = P.f_slot_ptr[s]; t < P.f_slot_ptr[s + 1]; ++t) {
            int node = P.slot_key[t] >> 1;
            if (stamp_of[node] != s) {
                stamp_of[node] = s;
                slot_at[node] = t - P.f_slot_ptr[s];
            }
        }
    };
    // extend-add maps: structure slots of each front -> slots of its parent
    P.f_rel_off.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) P.f_rel_off[s + 1] = P.f_rel_off[s] + (P.f_parent[s] >= 0 ? P.f_m[s] - P.f_p[s] : 0);
    P.rel.resize(P.f_rel_off[ns]);
    for (int s = 0; s < ns; ++s) {
        int par = P.f_parent[s];
        if (par < 0) continue;
        map_front(par);
        int base = P.f_slot_ptr[s];
        for (int t = P.f_p[s]; t < P.f_m[s]; ++t) {
            int key = P.slot_key[base + t];
            int node = key >> 1, sub = key & 1;
            if (stamp_of[node] != par) throw std::logic_error("front structure not contained in parent");
            P.rel[P.f_rel_off[s] + (t - P.f_p[s])] = slot_at[node] + sub;
        }
    }
    // arrowhead maps: each matrix entry belongs to the front owning the earlier node
    int const ne = static_cast<int>(in.ent_row.size());
    std::vector<int> owner(ne);
    P.f_arrow_ptr.assign(ns + 1, 0);
    for (int e = 0; e < ne; ++e) {
        int a = pos[in.ent_row[e] >> 1], b = pos[in.ent_col[e] >> 1];
        owner[e] = sn_of[std::min(a, b)];
        P.f_arrow_ptr[owner[e] + 1]++;
    }
    for (int s = 0; s < ns; ++s) P.f_arrow_ptr[s + 1] += P.f_arrow_ptr[s];
    P.arrow_src.resize(ne);
    P.arrow_dst.resize(ne);
    {
        std::vector<int> fp(P.f_arrow_ptr.begin(), P.f_arrow_ptr.end() - 1);
        for (int e = 0; e < ne; ++e) P.arrow_src[fp[owner[e]]++] = e;
    }
    for (int s = 0; s < ns; ++s) {
        map_front(s);
        int m = P.f_m[s];
        for (int q = P.f_arrow_ptr[s]; q < P.f_arrow_ptr[s + 1]; ++q) {
            int e = P.arrow_src[q];
            int nr = in.ent_row[e] >> 1, nc = in.ent_col[e] >> 1;
            if (stamp_of[nr] != s || stamp_of[nc] != s) throw std::logic_error("entry outside its front");
            int r = slot_at[nr] + (in.ent_row[e] & 1);
            int c = slot_at[nc] + (in.ent_col[e] & 1);
            if (m > 65535) throw std::length_error("front larger than 65535 slots");
            P.arrow_dst[q] = (r << 16) | c;
        }
    }
    return P;
}

}  // namespace gaphip
