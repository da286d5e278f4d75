/*
 * pforacle.c — FP64 CPU restatement of the reference Newton-Raphson power flow.
 * TEST INFRASTRUCTURE ONLY (see pforacle.h for the reference line map).
 *
 * Arithmetic is written operation-for-operation in the order the reference
 * evaluates it (std::complex<double> sums and products expanded by hand), so
 * the Y-bus is bit-identical and the Newton trajectory agrees to rounding.
 * Build: gcc -O2 -std=c99 (no -ffast-math, no FMA contraction on x86-64 baseline).
 */
#include "pforacle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ utils */

typedef struct {
    int n, nnz;
    int* rp;
    int* ci;
    double* re;
    double* im;
} ccsr; /* complex CSR */

static void* xmalloc(size_t n) {
    void* p = malloc(n ? n : 1);
    if (!p) abort();
    return p;
}
static void* xcalloc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s ? s : 1);
    if (!p) abort();
    return p;
}
static void ccsr_free(ccsr* a) {
    free(a->rp);
    free(a->ci);
    free(a->re);
    free(a->im);
    memset(a, 0, sizeof(*a));
}

/* --------------------------------------------- Y-bus (cpu_admittance_matrix.cpp:44-202) */

typedef struct {
    int row, col, seq;
    double re, im;
} trip;

static int trip_cmp(const void* a, const void* b) {
    const trip* x = (const trip*)a;
    const trip* y = (const trip*)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq); /* stable: push order */
}

static void build_ybus(const pf_net* net, ccsr* y) {
    int n = net->n_bus;
    int min_id = n, max_id = -1; /* cpu_admittance_matrix.cpp:56-62 */
    for (int i = 0; i < n; ++i) {
        if (net->bus_id[i] < min_id) min_id = net->bus_id[i];
        if (net->bus_id[i] > max_id) max_id = net->bus_id[i];
    }
    (void)min_id;
    int* id2idx = (int*)xmalloc(sizeof(int) * (size_t)(max_id + 1 > 0 ? max_id + 1 : 1));
    for (int k = 0; k <= max_id; ++k) id2idx[k] = -1;
    for (int i = 0; i < n; ++i) {
        int id = net->bus_id[i];
        if (id >= 0 && id <= max_id) id2idx[id] = i;
    }
    double* dre = (double*)xcalloc((size_t)n, sizeof(double));
    double* dim = (double*)xcalloc((size_t)n, sizeof(double));
    /* shunt appliances first, in appliance order, status not consulted (:88-97) */
    for (int a = 0; a < net->n_app; ++a) {
        if (net->app_type[a] != PF_SHUNT) continue;
        int node = net->app_node[a];
        if (node >= 0 && node <= max_id) {
            int b = id2idx[node];
            if (b >= 0 && b < n) {
                dre[b] += net->app_g[a];
                dim[b] += net->app_b[a];
            }
        }
    }
    int cap = 2 * net->n_branch + n;
    trip* t = (trip*)xmalloc(sizeof(trip) * (size_t)cap);
    int nt = 0;
    for (int k = 0; k < net->n_branch; ++k) {
        if (!net->br_status[k]) continue;
        int f = -1, to = -1;
        if (net->br_from[k] >= 0 && net->br_from[k] <= max_id) f = id2idx[net->br_from[k]];
        if (net->br_to[k] >= 0 && net->br_to[k] <= max_id) to = id2idx[net->br_to[k]];
        if (f < 0 || f >= n || to < 0 || to >= n) continue;
        double r = net->br_r[k], x = net->br_x[k];
        double z2 = r * r + x * x; /* :121 */
        double yr = 0.0, yi = 0.0;
        if (z2 > 1e-12) {
            yr = r / z2;
            yi = -x / z2;
        }
        double g1 = net->br_g ? net->br_g[k] : 0.0;
        double b1 = net->br_b ? net->br_b[k] : 0.0;
        /* (y + (g1, 0)) + (0, b1/2), each a std::complex sum (:130-135) */
        double bar = yr + g1, bai = yi + 0.0;
        double tr = bar + 0.0, ti = bai + b1 / 2.0;
        dre[f] += tr;
        dim[f] += ti;
        dre[to] += tr;
        dim[to] += ti;
        t[nt].row = f; t[nt].col = to; t[nt].re = -yr; t[nt].im = -yi; t[nt].seq = nt; ++nt;
        t[nt].row = to; t[nt].col = f; t[nt].re = -yr; t[nt].im = -yi; t[nt].seq = nt; ++nt;
    }
    for (int i = 0; i < n; ++i) {
        t[nt].row = i; t[nt].col = i; t[nt].re = dre[i]; t[nt].im = dim[i]; t[nt].seq = nt; ++nt;
    }
    qsort(t, (size_t)nt, sizeof(trip), trip_cmp);
    y->n = n;
    y->nnz = nt;
    y->rp = (int*)xcalloc((size_t)n + 1, sizeof(int));
    y->ci = (int*)xmalloc(sizeof(int) * (size_t)nt);
    y->re = (double*)xmalloc(sizeof(double) * (size_t)nt);
    y->im = (double*)xmalloc(sizeof(double) * (size_t)nt);
    for (int k = 0; k < nt; ++k) {
        y->rp[t[k].row + 1]++;
        y->ci[k] = t[k].col;
        y->re[k] = t[k].re;
        y->im[k] = t[k].im;
    }
    for (int i = 0; i < n; ++i) y->rp[i + 1] += y->rp[i];
    free(t);
    free(dre);
    free(dim);
    free(id2idx);
}

int pfo_build_ybus(const pf_net* net, int* row_ptr, int* col_idx, double* val_ri, int cap) {
    ccsr y;
    build_ybus(net, &y);
    if (y.nnz > cap) {
        ccsr_free(&y);
        return -1;
    }
    memcpy(row_ptr, y.rp, sizeof(int) * (size_t)(y.n + 1));
    memcpy(col_idx, y.ci, sizeof(int) * (size_t)y.nnz);
    for (int k = 0; k < y.nnz; ++k) {
        val_ri[2 * k] = y.re[k];
        val_ri[2 * k + 1] = y.im[k];
    }
    int nnz = y.nnz;
    ccsr_free(&y);
    return nnz;
}

/* ----------------------------------- S = V conj(Y V)  (cpu_newton_raphson.cpp:296-337) */

static void power_injections(const ccsr* y, const double* vr, const double* vi, double* sr,
                             double* si) {
    for (int i = 0; i < y->n; ++i) {
        double cr = 0.0, cim = 0.0;
        for (int p = y->rp[i]; p < y->rp[i + 1]; ++p) {
            int k = y->ci[p];
            double a = y->re[p], b = y->im[p], c = vr[k], d = vi[k];
            double pr = a * c - b * d, pi = a * d + b * c; /* complex product */
            cr += pr;
            cim += pi;
        }
        /* V * conj(I): (a + ib)(c - id) */
        double a = vr[i], b = vi[i], c = cr, d = -cim;
        sr[i] = a * c - b * d;
        si[i] = a * d + b * c;
    }
}

/* ------------------------- S_spec per bus (cpu_newton_raphson.cpp:346-377), pu if sb > 0 */

static void specified_power(const pf_net* net, double sb, double* pr, double* pi) {
    int n = net->n_bus;
    for (int i = 0; i < n; ++i) {
        pr[i] = 0.0;
        pi[i] = 0.0;
        if (net->p_w[i] != 0.0 || net->q_var[i] != 0.0) {
            pr[i] = net->p_w[i];
            pi[i] = net->q_var[i];
        }
    }
    /* appliances are summed in appliance order for every bus with a matching id */
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < net->n_app; ++a) {
            if (net->app_node[a] == net->bus_id[i] && net->app_status[a] == 1 &&
                (net->app_type[a] == PF_SOURCE || net->app_type[a] == PF_LOADGEN)) {
                pr[i] += net->app_p[a];
                pi[i] += net->app_q[a];
            }
        }
    }
    if (sb > 0.0) {
        for (int i = 0; i < n; ++i) {
            pr[i] /= sb;
            pi[i] /= sb;
        }
    }
}

/* F over non-slack buses, P then Q (if PQ) (cpu_newton_raphson.cpp:269-289); returns length */
static int mismatch(const pf_net* net, const double* sr, const double* si, const double* pr,
                    const double* pi, double* f) {
    int m = 0;
    for (int i = 0; i < net->n_bus; ++i) {
        if (net->bus_type[i] == PF_SLACK) continue;
        f[m++] = sr[i] - pr[i];
        if (net->bus_type[i] == PF_PQ) f[m++] = si[i] - pi[i];
    }
    return m;
}

/* ------------------------------------------------- Jacobian (cpu_newton_raphson.cpp:509-765) */

typedef struct {
    int n, nnz, cap;
    int* rp;
    int* ci;
    double* v;
} rcsr;

static void rcsr_push(rcsr* j, int col, double val) {
    if (j->nnz == j->cap) {
        j->cap = j->cap ? 2 * j->cap : 64;
        j->ci = (int*)realloc(j->ci, sizeof(int) * (size_t)j->cap);
        j->v = (double*)realloc(j->v, sizeof(double) * (size_t)j->cap);
        if (!j->ci || !j->v) abort();
    }
    j->ci[j->nnz] = col;
    j->v[j->nnz] = val;
    j->nnz++;
}

static int count_vars(const pf_net* net) { /* :530-541 */
    int ns = 0, npq = 0;
    for (int i = 0; i < net->n_bus; ++i) {
        ns += net->bus_type[i] == PF_SLACK;
        npq += net->bus_type[i] == PF_PQ;
    }
    return (net->n_bus - ns) + npq;
}

static int first_slack(const pf_net* net) {
    for (int i = 0; i < net->n_bus; ++i)
        if (net->bus_type[i] == PF_SLACK) return i;
    return -1;
}

/* The sparse equivalent of the dense fill + compaction.  Off-pattern entries of
 * the dense matrix evaluate to +-0 and are dropped by |J| > 1e-12, so only Y
 * pattern columns are visited; G/B take the last duplicate of each column. */
static void jacobian(const pf_net* net, const ccsr* y, const double* vr, const double* vi,
                     const double* sr, const double* si, rcsr* J) {
    int n = net->n_bus;
    int nv = count_vars(net);
    J->n = nv;
    J->nnz = 0;
    J->rp = (int*)xcalloc((size_t)nv + 1, sizeof(int));
    if (nv == 0) return;
    if (y->nnz == 0) { /* identity (:562-568) */
        for (int i = 0; i < nv; ++i) {
            rcsr_push(J, i, 1.0);
            J->rp[i + 1] = J->nnz;
        }
        return;
    }
    double* vm = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* va = (double*)xmalloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        vm[i] = hypot(vr[i], vi[i]);
        va[i] = atan2(vi[i], vr[i]);
        if (vm[i] < 1e-6) vm[i] = 1.0;
    }
    int slack = first_slack(net);
    int* ang = (int*)xmalloc(sizeof(int) * (size_t)n);
    int* mag = (int*)xmalloc(sizeof(int) * (size_t)n);
    int vix = 0;
    for (int i = 0; i < n; ++i) ang[i] = (i != slack) ? vix++ : -1;
    for (int i = 0; i < n; ++i) mag[i] = (net->bus_type[i] == PF_PQ) ? vix++ : -1;

    /* per-row scratch: unique pattern columns (ascending) with last-duplicate G, B */
    int* cols = (int*)xmalloc(sizeof(int) * (size_t)(n + 1));
    double* G = (double*)xmalloc(sizeof(double) * (size_t)(n + 1));
    double* B = (double*)xmalloc(sizeof(double) * (size_t)(n + 1));
    int eq = 0;
    for (int i = 0; i < n; ++i) {
        if (net->bus_type[i] == PF_SLACK) continue;
        int nc = 0;
        double gii = 0.0, bii = 0.0;
        for (int p = y->rp[i]; p < y->rp[i + 1]; ++p) {
            int j = y->ci[p];
            if (j >= n) continue;
            if (nc > 0 && cols[nc - 1] == j) {
                G[nc - 1] = y->re[p];
                B[nc - 1] = y->im[p];
            } else {
                cols[nc] = j;
                G[nc] = y->re[p];
                B[nc] = y->im[p];
                ++nc;
            }
            if (j == i) {
                gii = y->re[p];
                bii = y->im[p];
            }
        }
        double Vi = vm[i], ti = va[i], Pi = sr[i], Qi = si[i];
        for (int pass = 0; pass < 2; ++pass) {
            int is_q = pass == 1;
            if (is_q && net->bus_type[i] != PF_PQ) break;
            if (eq >= nv) break;
            /* angle columns, then magnitude columns, each ascending in bus order */
            for (int part = 0; part < 2; ++part) {
                for (int c = 0; c < nc; ++c) {
                    int j = cols[c];
                    int var = part == 0 ? ang[j] : mag[j];
                    if (var < 0 || var >= nv) continue;
                    double val;
                    if (part == 0) {
                        if (i == j) {
                            val = is_q ? Pi - gii * Vi * Vi : -Qi - bii * Vi * Vi;
                        } else {
                            double Vj = vm[j], tij = ti - va[j];
                            double s = sin(tij), co = cos(tij);
                            val = is_q ? -Vi * Vj * (G[c] * co + B[c] * s)
                                       : Vi * Vj * (G[c] * s - B[c] * co);
                        }
                    } else {
                        if (i == j) {
                            val = is_q ? (Qi / Vi) - bii * Vi : (Pi / Vi) + gii * Vi;
                        } else {
                            double tij = ti - va[j];
                            double s = sin(tij), co = cos(tij);
                            val = is_q ? Vi * (G[c] * s - B[c] * co) : Vi * (G[c] * co + B[c] * s);
                        }
                    }
                    if (fabs(val) > 1e-12) rcsr_push(J, var, val);
                }
            }
            J->rp[eq + 1] = J->nnz;
            ++eq;
        }
    }
    for (; eq < nv; ++eq) J->rp[eq + 1] = J->nnz;
    free(cols);
    free(G);
    free(B);
    free(vm);
    free(va);
    free(ang);
    free(mag);
}

/* ------------------------------------------------------------------ sparse LU
 * Left-looking LU with threshold partial pivoting (prefer the diagonal when it
 * is within 0.1 of the column maximum) and a minimum-degree column order on the
 * symmetrised pattern.  Complex arithmetic, as the reference solves J dx = -F as
 * a complex system with zero imaginary parts (cpu_newton_raphson.cpp:402-431). */

typedef struct {
    int n;
    int *Lp, *Li, *Up, *Ui, *pinv, *q;
    double *Lr, *Li_, *Ur, *Ui_; /* real / imaginary parts */
    int lcap, ucap;
} lufac;

static void lu_free(lufac* f) {
    free(f->Lp); free(f->Li); free(f->Up); free(f->Ui); free(f->pinv); free(f->q);
    free(f->Lr); free(f->Li_); free(f->Ur); free(f->Ui_);
    memset(f, 0, sizeof(*f));
}

/* minimum degree on an undirected graph given as CSR adjacency (no self loops) */
typedef struct {
    int deg, node;
} hent;
static void heap_push(hent* h, int* hn, hent e) {
    int i = (*hn)++;
    h[i] = e;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (h[p].deg < h[i].deg || (h[p].deg == h[i].deg && h[p].node < h[i].node)) break;
        hent t = h[p]; h[p] = h[i]; h[i] = t;
        i = p;
    }
}
static hent heap_pop(hent* h, int* hn) {
    hent top = h[0];
    h[0] = h[--(*hn)];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        if (l < *hn && (h[l].deg < h[m].deg || (h[l].deg == h[m].deg && h[l].node < h[m].node))) m = l;
        if (r < *hn && (h[r].deg < h[m].deg || (h[r].deg == h[m].deg && h[r].node < h[m].node))) m = r;
        if (m == i) break;
        hent t = h[m]; h[m] = h[i]; h[i] = t;
        i = m;
    }
    return top;
}

static void min_degree(int n, const int* ap, const int* ai, int* order) {
    int** adj = (int**)xmalloc(sizeof(int*) * (size_t)(n ? n : 1));
    int* sz = (int*)xcalloc((size_t)n, sizeof(int));
    int* cp = (int*)xcalloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) {
        cp[i] = (ap[i + 1] - ap[i]) + 4;
        adj[i] = (int*)xmalloc(sizeof(int) * (size_t)cp[i]);
        for (int p = ap[i]; p < ap[i + 1]; ++p) adj[i][sz[i]++] = ai[p];
    }
    char* gone = (char*)xcalloc((size_t)n, 1);
    int* mark = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) mark[i] = -1;
    size_t hcap = (size_t)n * 4 + 16;
    hent* h = (hent*)xmalloc(sizeof(hent) * hcap);
    int hn = 0;
    for (int i = 0; i < n; ++i) heap_push(h, &hn, (hent){sz[i], i});
    int k = 0, stamp = -1;
    while (k < n) {
        hent e = heap_pop(h, &hn);
        int v = e.node;
        if (gone[v] || e.deg != sz[v]) continue;
        gone[v] = 1;
        order[k++] = v;
        int* nb = adj[v];
        int dv = sz[v];
        for (int a = 0; a < dv; ++a) {
            int u = nb[a];
            /* drop v from u, then connect u to the other neighbours of v */
            int w = 0;
            for (int p = 0; p < sz[u]; ++p)
                if (adj[u][p] != v) adj[u][w++] = adj[u][p];
            sz[u] = w;
            ++stamp; /* fresh stamp per (v, u) pair: stale marks must not alias */
            for (int p = 0; p < sz[u]; ++p) mark[adj[u][p]] = stamp;
            mark[u] = stamp;
            for (int b = 0; b < dv; ++b) {
                int x = nb[b];
                if (mark[x] == stamp) continue;
                if (sz[u] == cp[u]) {
                    cp[u] *= 2;
                    adj[u] = (int*)realloc(adj[u], sizeof(int) * (size_t)cp[u]);
                    if (!adj[u]) abort();
                }
                adj[u][sz[u]++] = x;
                mark[x] = stamp;
            }
        }
        for (int a = 0; a < dv; ++a) {
            int u = nb[a];
            if ((size_t)hn == hcap) {
                hcap *= 2;
                h = (hent*)realloc(h, sizeof(hent) * hcap);
                if (!h) abort();
            }
            heap_push(h, &hn, (hent){sz[u], u});
        }
    }
    for (int i = 0; i < n; ++i) free(adj[i]);
    free(adj); free(sz); free(cp); free(gone); free(mark); free(h);
}

/* column order: minimum degree on pattern(A + A^T) */
static void lu_order(int n, const int* rp, const int* ci, int* q) {
    /* count symmetric adjacency without diagonal */
    int* cnt = (int*)xcalloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (ci[p] != i) {
                cnt[i + 1]++;
                cnt[ci[p] + 1]++;
            }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    int* adj = (int*)xmalloc(sizeof(int) * (size_t)(cnt[n] ? cnt[n] : 1));
    int* pos = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) pos[i] = cnt[i];
    for (int i = 0; i < n; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (ci[p] != i) {
                adj[pos[i]++] = ci[p];
                adj[pos[ci[p]]++] = i;
            }
    /* deduplicate each list */
    int* mark = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) mark[i] = -1;
    int* ap = (int*)xcalloc((size_t)n + 1, sizeof(int));
    int w = 0;
    for (int i = 0; i < n; ++i) {
        ap[i] = w;
        for (int p = cnt[i]; p < cnt[i + 1]; ++p) {
            int j = adj[p];
            if (mark[j] == i) continue;
            mark[j] = i;
            adj[w++] = j;
        }
    }
    ap[n] = w;
    min_degree(n, ap, adj, q);
    free(cnt); free(adj); free(pos); free(mark); free(ap);
}

static void lgrow(lufac* f, int need) {
    if (need <= f->lcap) return;
    while (f->lcap < need) f->lcap = f->lcap ? 2 * f->lcap : 1024;
    f->Li = (int*)realloc(f->Li, sizeof(int) * (size_t)f->lcap);
    f->Lr = (double*)realloc(f->Lr, sizeof(double) * (size_t)f->lcap);
    f->Li_ = (double*)realloc(f->Li_, sizeof(double) * (size_t)f->lcap);
    if (!f->Li || !f->Lr || !f->Li_) abort();
}
static void ugrow(lufac* f, int need) {
    if (need <= f->ucap) return;
    while (f->ucap < need) f->ucap = f->ucap ? 2 * f->ucap : 1024;
    f->Ui = (int*)realloc(f->Ui, sizeof(int) * (size_t)f->ucap);
    f->Ur = (double*)realloc(f->Ur, sizeof(double) * (size_t)f->ucap);
    f->Ui_ = (double*)realloc(f->Ui_, sizeof(double) * (size_t)f->ucap);
    if (!f->Ui || !f->Ur || !f->Ui_) abort();
}

/* Factor A (CSR, complex) with column order q: returns 0, or 1 when a pivot
 * column is numerically empty (|pivot| < 1e-14, the reference's pivot_tolerance_,
 * cpu_lu_solver.cpp:65). */
static int lu_factor(int n, const int* rp, const int* ci, const double* ar, const double* ai,
                     lufac* f) {
    memset(f, 0, sizeof(*f));
    f->n = n;
    /* CSC copy */
    int nnz = rp[n];
    int* cp = (int*)xcalloc((size_t)n + 1, sizeof(int));
    int* ri = (int*)xmalloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    double* cr = (double*)xmalloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    double* cim = (double*)xmalloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    for (int p = 0; p < nnz; ++p) cp[ci[p] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    int* pos = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int j = 0; j < n; ++j) pos[j] = cp[j];
    for (int i = 0; i < n; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            int d = pos[ci[p]]++;
            ri[d] = i;
            cr[d] = ar[p];
            cim[d] = ai[p];
        }
    f->q = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    lu_order(n, rp, ci, f->q);
    f->pinv = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) f->pinv[i] = -1;
    f->Lp = (int*)xcalloc((size_t)n + 1, sizeof(int));
    f->Up = (int*)xcalloc((size_t)n + 1, sizeof(int));
    lgrow(f, 4 * nnz + n);
    ugrow(f, 4 * nnz + n);
    double* xr = (double*)xcalloc((size_t)n, sizeof(double));
    double* xi = (double*)xcalloc((size_t)n, sizeof(double));
    int* stack = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    int* pstk = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    int* reach = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    int* mark = (int*)xmalloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) mark[i] = -1;
    int lnz = 0, unz = 0, rc = 0;
    for (int k = 0; k < n; ++k) {
        int col = f->q[k];
        f->Lp[k] = lnz;
        f->Up[k] = unz;
        lgrow(f, lnz + n);
        ugrow(f, unz + n);
        /* reach of A(:,col) in the graph of L: DFS, topological order in reach[top..n-1] */
        int top = n;
        for (int p = cp[col]; p < cp[col + 1]; ++p) {
            int s = ri[p];
            if (mark[s] == k) continue;
            int head = 0;
            stack[0] = s;
            while (head >= 0) {
                int j = stack[head];
                int J = f->pinv[j];
                if (mark[j] != k) {
                    mark[j] = k;
                    pstk[head] = (J < 0) ? 0 : f->Lp[J] + 1; /* skip unit diagonal */
                }
                int done = 1;
                int pend = (J < 0) ? 0 : (J < k ? f->Lp[J + 1] : lnz);
                for (int p2 = pstk[head]; p2 < pend; ++p2) {
                    int i = f->Li[p2];
                    if (mark[i] == k) continue;
                    pstk[head] = p2 + 1;
                    stack[++head] = i;
                    done = 0;
                    break;
                }
                if (done) {
                    --head;
                    reach[--top] = j;
                }
            }
        }
        /* numeric: x = L \ A(:,col) */
        for (int p = top; p < n; ++p) xr[reach[p]] = xi[reach[p]] = 0.0;
        for (int p = cp[col]; p < cp[col + 1]; ++p) {
            xr[ri[p]] = cr[p];
            xi[ri[p]] = cim[p];
        }
        for (int px = top; px < n; ++px) {
            int j = reach[px];
            int J = f->pinv[j];
            if (J < 0) continue;
            double a = xr[j], b = xi[j];
            for (int p = f->Lp[J] + 1; p < f->Lp[J + 1]; ++p) {
                int i = f->Li[p];
                double c = f->Lr[p], d = f->Li_[p];
                xr[i] -= c * a - d * b;
                xi[i] -= c * b + d * a;
            }
        }
        /* pivot choice among non-pivotal rows */
        int ipiv = -1;
        double amax = -1.0;
        for (int px = top; px < n; ++px) {
            int i = reach[px];
            if (f->pinv[i] < 0) {
                double t = hypot(xr[i], xi[i]);
                if (t > amax) {
                    amax = t;
                    ipiv = i;
                }
            } else {
                f->Ui[unz] = f->pinv[i];
                f->Ur[unz] = xr[i];
                f->Ui_[unz] = xi[i];
                ++unz;
            }
        }
        if (ipiv < 0 || amax < 1e-14) {
            rc = 1;
            break;
        }
        if (f->pinv[col] < 0 && mark[col] == k && hypot(xr[col], xi[col]) >= 0.1 * amax) ipiv = col;
        double pr = xr[ipiv], pim = xi[ipiv];
        f->Ui[unz] = k;
        f->Ur[unz] = pr;
        f->Ui_[unz] = pim;
        ++unz;
        f->pinv[ipiv] = k;
        f->Li[lnz] = ipiv;
        f->Lr[lnz] = 1.0;
        f->Li_[lnz] = 0.0;
        ++lnz;
        double den = pr * pr + pim * pim;
        for (int px = top; px < n; ++px) {
            int i = reach[px];
            if (f->pinv[i] < 0) {
                /* x / pivot */
                double a = xr[i], b = xi[i];
                f->Li[lnz] = i;
                f->Lr[lnz] = (a * pr + b * pim) / den;
                f->Li_[lnz] = (b * pr - a * pim) / den;
                ++lnz;
            }
            xr[i] = xi[i] = 0.0;
        }
    }
    if (rc == 0) {
        f->Lp[n] = lnz;
        f->Up[n] = unz;
        for (int p = 0; p < lnz; ++p) f->Li[p] = f->pinv[f->Li[p]];
    }
    free(cp); free(ri); free(cr); free(cim); free(pos);
    free(xr); free(xi); free(stack); free(pstk); free(reach); free(mark);
    return rc;
}

/* x = A^{-1} b with A Q = P^T L U */
static void lu_solve_fac(const lufac* f, const double* br, const double* bi, double* xr, double* xi) {
    int n = f->n;
    double* yr = (double*)xmalloc(sizeof(double) * (size_t)(n ? n : 1));
    double* yi = (double*)xmalloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; ++i) {
        yr[f->pinv[i]] = br[i];
        yi[f->pinv[i]] = bi[i];
    }
    for (int j = 0; j < n; ++j) { /* L y = P b, unit diagonal first in each column */
        double a = yr[j], b = yi[j];
        for (int p = f->Lp[j] + 1; p < f->Lp[j + 1]; ++p) {
            int i = f->Li[p];
            double c = f->Lr[p], d = f->Li_[p];
            yr[i] -= c * a - d * b;
            yi[i] -= c * b + d * a;
        }
    }
    for (int j = n - 1; j >= 0; --j) { /* U w = y, diagonal last in each column */
        int pd = f->Up[j + 1] - 1;
        double c = f->Ur[pd], d = f->Ui_[pd];
        double den = c * c + d * d;
        double a = yr[j], b = yi[j];
        double wr = (a * c + b * d) / den, wi = (b * c - a * d) / den;
        yr[j] = wr;
        yi[j] = wi;
        for (int p = f->Up[j]; p < pd; ++p) {
            int i = f->Ui[p];
            double e = f->Ur[p], g = f->Ui_[p];
            yr[i] -= e * wr - g * wi;
            yi[i] -= e * wi + g * wr;
        }
    }
    for (int k = 0; k < n; ++k) {
        xr[f->q[k]] = yr[k];
        xi[f->q[k]] = yi[k];
    }
    free(yr);
    free(yi);
}

int pfo_lu_solve(int n, const int* row_ptr, const int* col_idx, const double* val_ri, int nnz,
                 const double* rhs_ri, double* x_ri) {
    if (n <= 0 || nnz <= 0) return 1; /* cpu_lu_solver.cpp:100-108 */
    double* ar = (double*)xmalloc(sizeof(double) * (size_t)nnz);
    double* ai = (double*)xmalloc(sizeof(double) * (size_t)nnz);
    for (int k = 0; k < nnz; ++k) {
        ar[k] = val_ri[2 * k];
        ai[k] = val_ri[2 * k + 1];
    }
    lufac f;
    int rc = lu_factor(n, row_ptr, col_idx, ar, ai, &f);
    if (rc == 0) {
        double* br = (double*)xmalloc(sizeof(double) * (size_t)n);
        double* bi = (double*)xmalloc(sizeof(double) * (size_t)n);
        double* xr = (double*)xmalloc(sizeof(double) * (size_t)n);
        double* xi = (double*)xmalloc(sizeof(double) * (size_t)n);
        for (int i = 0; i < n; ++i) {
            br[i] = rhs_ri[2 * i];
            bi[i] = rhs_ri[2 * i + 1];
        }
        lu_solve_fac(&f, br, bi, xr, xi);
        for (int i = 0; i < n; ++i) {
            x_ri[2 * i] = xr[i];
            x_ri[2 * i + 1] = xi[i];
        }
        free(br); free(bi); free(xr); free(xi);
    }
    lu_free(&f);
    free(ar);
    free(ai);
    return rc;
}

/* ------------------------------------------ voltage update (cpu_newton_raphson.cpp:438-504) */

static void update_voltages(const pf_net* net, double* vr, double* vi, const double* corr, int m) {
    if (m == 0) return;
    int n = net->n_bus;
    int slack = first_slack(net);
    const double damping = 0.9;
    int c = 0;
    for (int i = 0; i < n; ++i) {
        if (i != slack && c < m) {
            double vm = hypot(vr[i], vi[i]);
            double ac = damping * corr[c];
            if (fabs(ac) > 0.15) ac = 0.15 * (ac > 0 ? 1.0 : -1.0);
            double va = atan2(vi[i], vr[i]) + ac;
            vr[i] = vm * cos(va);
            vi[i] = vm * sin(va);
            ++c;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (net->bus_type[i] == PF_PQ && c < m) {
            double old = hypot(vr[i], vi[i]);
            double mc = damping * corr[c];
            if (fabs(mc) > 0.08) mc = 0.08 * (mc > 0 ? 1.0 : -1.0);
            double vm = old + mc;
            double va = atan2(vi[i], vr[i]);
            if (vm < 0.5) vm = 0.5;
            if (vm > 1.5) vm = 1.5;
            vr[i] = vm * cos(va);
            vi[i] = vm * sin(va);
            ++c;
        }
    }
}

/* ------------------------------------------------------------------ driver */

static int* eq_to_var = NULL; /* equation row -> variable index, set by permute_rows_to_vars */

/* Row permutation putting P_b at theta_b's index and Q_b at |V|_b's; returns 0
 * when the counts do not line up (several slack buses), leaving J unpermuted. */
static int permute_rows_to_vars(const pf_net* net, const rcsr* J, rcsr* P) {
    int n = net->n_bus, ns = 0;
    for (int i = 0; i < n; ++i) ns += net->bus_type[i] == PF_SLACK;
    if (ns != 1) return 0;
    int nv = J->n;
    free(eq_to_var);
    eq_to_var = (int*)xmalloc(sizeof(int) * (size_t)(nv ? nv : 1));
    int slack = first_slack(net);
    int a = 0, eq = 0, npq = 0;
    for (int i = 0; i < n; ++i) npq += net->bus_type[i] == PF_PQ;
    int nang = n - 1;
    int mg = nang;
    for (int i = 0; i < n; ++i) {
        if (i == slack) continue;
        eq_to_var[eq++] = a++;
        if (net->bus_type[i] == PF_PQ) eq_to_var[eq++] = mg++;
    }
    if (eq != nv || mg != nang + npq) return 0;
    P->n = nv;
    P->nnz = J->nnz;
    P->cap = J->nnz;
    P->rp = (int*)xcalloc((size_t)nv + 1, sizeof(int));
    P->ci = (int*)xmalloc(sizeof(int) * (size_t)(J->nnz ? J->nnz : 1));
    P->v = (double*)xmalloc(sizeof(double) * (size_t)(J->nnz ? J->nnz : 1));
    for (int r = 0; r < nv; ++r) P->rp[eq_to_var[r] + 1] = J->rp[r + 1] - J->rp[r];
    for (int r = 0; r < nv; ++r) P->rp[r + 1] += P->rp[r];
    for (int r = 0; r < nv; ++r) {
        int d = P->rp[eq_to_var[r]];
        for (int p = J->rp[r]; p < J->rp[r + 1]; ++p, ++d) {
            P->ci[d] = J->ci[p];
            P->v[d] = J->v[p];
        }
    }
    return 1;
}

static void scale_y(ccsr* y, double zb) {
    for (int k = 0; k < y->nnz; ++k) {
        y->re[k] = y->re[k] * zb;
        y->im[k] = y->im[k] * zb;
    }
}

int pfo_solve(const pf_net* net, double tol, int max_it, double base_power, double* v_out,
              int* iterations, int* converged, double* final_mismatch, double* traj, int traj_cap,
              int* n_traj, int j_call, int* j_n, int* j_row_ptr, int* j_col_idx, double* j_val,
              int j_cap, int* j_nnz) {
    int n = net->n_bus;
    if (n <= 0) return -1;
    ccsr y;
    build_ybus(net, &y);
    double vb = net->u_rated[0];
    double zb = (vb * vb) / base_power; /* cpu_newton_raphson.cpp:46-47 */
    scale_y(&y, zb);
    double* vr = (double*)xcalloc((size_t)n, sizeof(double));
    double* vi = (double*)xcalloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i) /* flat start (:77-88) */
        vr[i] = (net->bus_type[i] == PF_SLACK || net->bus_type[i] == PF_PV) ? net->u_pu[i] : 0.98;
    double *sr = (double*)xmalloc(sizeof(double) * (size_t)n), *si = (double*)xmalloc(sizeof(double) * (size_t)n);
    double *pr = (double*)xmalloc(sizeof(double) * (size_t)n), *pi = (double*)xmalloc(sizeof(double) * (size_t)n);
    specified_power(net, base_power, pr, pi);
    double* f = (double*)xmalloc(sizeof(double) * (size_t)(2 * n + 1));
    double* rr = (double*)xmalloc(sizeof(double) * (size_t)(2 * n + 1));
    double* ri = (double*)xmalloc(sizeof(double) * (size_t)(2 * n + 1));
    double* xr = (double*)xmalloc(sizeof(double) * (size_t)(2 * n + 1));
    double* xi = (double*)xmalloc(sizeof(double) * (size_t)(2 * n + 1));
    int conv = 0, its = 0, nt = 0, rc = 0, calls = 0;
    double fm = 0.0;
    if (j_n) *j_n = 0;
    if (j_nnz) *j_nnz = 0;
    for (int it = 0; it < max_it; ++it) {
        power_injections(&y, vr, vi, sr, si);
        int m = mismatch(net, sr, si, pr, pi, f);
        /* |F| of the first maximum under max_element's `|a| < |b|` test
         * (:115-120): a NaN in row 0 stays the maximum (no comparison with
         * it is true), a NaN in a later row is never taken */
        double mx = m > 0 ? fabs(f[0]) : 0.0;
        for (int k = 1; k < m; ++k)
            if (mx < fabs(f[k])) mx = fabs(f[k]);
        fm = mx;
        if (traj && nt < traj_cap) traj[nt] = mx;
        ++nt;
        if (mx < tol) {
            conv = 1;
            its = it + 1;
            break;
        }
        rcsr J;
        memset(&J, 0, sizeof(J));
        jacobian(net, &y, vr, vi, sr, si, &J);
        if (calls == j_call && j_n) {
            *j_n = J.n;
            *j_nnz = J.nnz;
            if (J.nnz > j_cap) {
                rc = 1;
            } else {
                memcpy(j_row_ptr, J.rp, sizeof(int) * (size_t)(J.n + 1));
                memcpy(j_col_idx, J.ci, sizeof(int) * (size_t)J.nnz);
                memcpy(j_val, J.v, sizeof(double) * (size_t)J.nnz);
            }
        }
        ++calls;
        int nv = J.n;
        for (int k = 0; k < m; ++k) xr[k] = 0.0;
        if (J.nnz > 0 && nv == m) {
            double* ai = (double*)xcalloc((size_t)J.nnz, sizeof(double));
            lufac fac;
            /* Row order of J is (P_b, Q_b) per bus while its column order is all
             * angles then all magnitudes; moving row P_b to the index of column
             * theta_b and Q_b to |V|_b gives the same solution with a structurally
             * symmetric matrix, so the symmetric ordering and the diagonal-preferring
             * pivot keep fill low. */
            rcsr P;
            int perm_ok = permute_rows_to_vars(net, &J, &P);
            const rcsr* A = perm_ok ? &P : &J;
            if (lu_factor(nv, A->rp, A->ci, A->v, ai, &fac) == 0) {
                for (int k = 0; k < m; ++k) {
                    int dst = perm_ok ? eq_to_var[k] : k;
                    rr[dst] = -f[k];
                    ri[dst] = 0.0;
                }
                lu_solve_fac(&fac, rr, ri, xr, xi);
            }
            if (perm_ok) {
                free(P.rp);
                free(P.ci);
                free(P.v);
            }
            lu_free(&fac);
            free(ai);
        }
        free(J.rp);
        free(J.ci);
        free(J.v);
        update_voltages(net, vr, vi, xr, m);
    }
    if (!conv) its = max_it;
    for (int i = 0; i < n; ++i) {
        v_out[2 * i] = vr[i];
        v_out[2 * i + 1] = vi[i];
    }
    *iterations = its;
    *converged = conv;
    *final_mismatch = fm;
    if (n_traj) *n_traj = nt;
    ccsr_free(&y);
    free(vr); free(vi); free(sr); free(si); free(pr); free(pi);
    free(f); free(rr); free(ri); free(xr); free(xi);
    return rc;
}

int pfo_mismatches(const pf_net* net, const double* v_ri, const double* y_ri, const int* y_row_ptr,
                   const int* y_col_idx, int y_nnz, double* out, int cap) {
    int n = net->n_bus;
    ccsr y;
    y.n = n;
    y.nnz = y_nnz;
    y.rp = (int*)xcalloc((size_t)n + 1, sizeof(int));
    y.ci = (int*)xmalloc(sizeof(int) * (size_t)(y_nnz ? y_nnz : 1));
    y.re = (double*)xmalloc(sizeof(double) * (size_t)(y_nnz ? y_nnz : 1));
    y.im = (double*)xmalloc(sizeof(double) * (size_t)(y_nnz ? y_nnz : 1));
    if (y_nnz > 0) {
        memcpy(y.rp, y_row_ptr, sizeof(int) * (size_t)(n + 1));
        memcpy(y.ci, y_col_idx, sizeof(int) * (size_t)y_nnz);
        for (int k = 0; k < y_nnz; ++k) {
            y.re[k] = y_ri[2 * k];
            y.im[k] = y_ri[2 * k + 1];
        }
    }
    double* vr = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* vi = (double*)xmalloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        vr[i] = v_ri[2 * i];
        vi[i] = v_ri[2 * i + 1];
    }
    double *sr = (double*)xcalloc((size_t)n, sizeof(double)), *si = (double*)xcalloc((size_t)n, sizeof(double));
    double *pr = (double*)xmalloc(sizeof(double) * (size_t)n), *pi = (double*)xmalloc(sizeof(double) * (size_t)n);
    if (y_nnz > 0) power_injections(&y, vr, vi, sr, si); /* empty Y -> zero S (:301-305) */
    const double sb = 100e6;
    specified_power(net, sb, pr, pi);
    double* f = (double*)xmalloc(sizeof(double) * (size_t)(2 * n + 1));
    int m = mismatch(net, sr, si, pr, pi, f);
    int rc = m;
    if (m > cap) {
        rc = -1;
    } else {
        for (int k = 0; k < m; ++k) out[k] = f[k] * sb;
    }
    ccsr_free(&y);
    free(vr); free(vi); free(sr); free(si); free(pr); free(pi); free(f);
    return rc;
}

int pfo_jacobian(const pf_net* net, const double* v_ri, const double* ypu_ri, const int* y_row_ptr,
                 const int* y_col_idx, int* j_n, int* j_row_ptr, int* j_col_idx, double* j_val,
                 int j_cap) {
    int n = net->n_bus;
    ccsr y;
    y.n = n;
    y.nnz = y_row_ptr[n];
    y.rp = (int*)y_row_ptr;
    y.ci = (int*)y_col_idx;
    y.re = (double*)xmalloc(sizeof(double) * (size_t)(y.nnz ? y.nnz : 1));
    y.im = (double*)xmalloc(sizeof(double) * (size_t)(y.nnz ? y.nnz : 1));
    for (int k = 0; k < y.nnz; ++k) {
        y.re[k] = ypu_ri[2 * k];
        y.im[k] = ypu_ri[2 * k + 1];
    }
    double* vr = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* vi = (double*)xmalloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        vr[i] = v_ri[2 * i];
        vi[i] = v_ri[2 * i + 1];
    }
    double *sr = (double*)xmalloc(sizeof(double) * (size_t)n), *si = (double*)xmalloc(sizeof(double) * (size_t)n);
    power_injections(&y, vr, vi, sr, si);
    rcsr J;
    memset(&J, 0, sizeof(J));
    jacobian(net, &y, vr, vi, sr, si, &J);
    int rc = J.nnz;
    *j_n = J.n;
    if (J.nnz > j_cap) {
        rc = -1;
    } else {
        memcpy(j_row_ptr, J.rp, sizeof(int) * (size_t)(J.n + 1));
        memcpy(j_col_idx, J.ci, sizeof(int) * (size_t)J.nnz);
        memcpy(j_val, J.v, sizeof(double) * (size_t)J.nnz);
    }
    free(J.rp); free(J.ci); free(J.v);
    free(y.re); free(y.im); free(vr); free(vi); free(sr); free(si);
    return rc;
}
