// Host-side analysis for the batched GPU Newton-Raphson engine (see analysis.hpp).
#include "analysis.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <string>

namespace pfgpu {

// ---------------------------------------------------------------------------
// Jacobian pattern + scatter map.
// Layout follows the reference exactly so that get_J() matches it:
//   distributed slack (BaseNRAlgo.tpp:396-555): row/col 0 = reference-slack
//   active-power equation / slack_absorbed unknown; columns 1.. = Va of pvpq
//   then Vm of pq; rows 1.. = P of pvpq then Q of pq. Column 0 holds the
//   slack weights (w[slack0] at row 0, w[pvpq[k]] at row k+1 when |w|>1e-7).
//   single slack (BaseNRSingleSlackAlgo.tpp:247-...): same without row/col 0.
// The value of J entry t is Re/Im of dS_dVa or dS_dVm at Ybus entry src[t]
// (BaseNRAlgo.tpp:562-635: J11 = dS_dVa.real, J21 = dS_dVa.imag,
// J12 = dS_dVm.real, J22 = dS_dVm.imag, slack row = real parts).
// ---------------------------------------------------------------------------
JPattern build_jacobian_pattern(const YbusPattern& Y, const int* slack_ids, int nslack, const double* slack_w,
                                const int* pv, int npv, const int* pq, int npq, bool single_slack) {
    if (nslack < 1) throw std::runtime_error("pfgpu: at least one slack bus is required");
    JPattern J;
    const int n = Y.n;
    J.n_bus = n;
    J.off = single_slack ? 0 : 1;
    J.slack0 = slack_ids[0];
    if (!single_slack)
        for (int s = 1; s < nslack; ++s) J.my_pv.push_back(slack_ids[s]);  // BaseAlgo.hpp:201-216
    J.my_pv.insert(J.my_pv.end(), pv, pv + npv);
    J.pq.assign(pq, pq + npq);
    J.pvpq = J.my_pv;
    J.pvpq.insert(J.pvpq.end(), pq, pq + npq);
    const int npvpq = (int)J.pvpq.size();
    std::vector<int> pvpq_inv(n, -1), pq_inv(n, -1);
    for (int k = 0; k < npvpq; ++k) {
        const int b = J.pvpq[k];
        if (b < 0 || b >= n) throw std::runtime_error("pfgpu: pv/pq bus index out of range");
        if (pvpq_inv[b] >= 0) throw std::runtime_error("pfgpu: a bus is listed twice in pv/pq/slack");
        pvpq_inv[b] = k;
    }
    for (int k = 0; k < npq; ++k) pq_inv[J.pq[k]] = k;
    if (J.slack0 < 0 || J.slack0 >= n) throw std::runtime_error("pfgpu: slack bus index out of range");
    if (pvpq_inv[J.slack0] >= 0) throw std::runtime_error("pfgpu: the reference slack bus is also pv/pq");
    const int off = J.off;
    J.N = npvpq + npq + off;

    struct E { int r; int src; uint8_t sel; double c; };
    std::vector<std::vector<E>> cols(J.N);
    auto add_bus_column = [&](int jc, int bus, uint8_t sel_re, uint8_t sel_im) {
        for (int q = Y.p[bus]; q < Y.p[bus + 1]; ++q) {
            const int r = Y.i[q];
            if (pvpq_inv[r] >= 0) cols[jc].push_back({pvpq_inv[r] + off, q, sel_re, 0.0});
            if (pq_inv[r] >= 0) cols[jc].push_back({pq_inv[r] + npvpq + off, q, sel_im, 0.0});
            if (off && r == J.slack0) cols[jc].push_back({0, q, sel_re, 0.0});
        }
    };
    for (int k = 0; k < npvpq; ++k) add_bus_column(k + off, J.pvpq[k], SEL_DVA_RE, SEL_DVA_IM);
    for (int k = 0; k < npq; ++k) add_bus_column(k + npvpq + off, J.pq[k], SEL_DVM_RE, SEL_DVM_IM);
    if (off) {
        cols[0].push_back({0, -1, SEL_CONST, slack_w[J.slack0]});
        for (int k = 0; k < npvpq; ++k) {
            const double w = slack_w[J.pvpq[k]];
            if (std::fabs(w) > 1e-7) cols[0].push_back({k + 1, -1, SEL_CONST, w});  // _tol_equal_float
        }
    }
    J.Jp.assign(J.N + 1, 0);
    for (int c = 0; c < J.N; ++c) {
        auto& v = cols[c];
        std::sort(v.begin(), v.end(), [](const E& a, const E& b) { return a.r < b.r; });
        for (size_t t = 1; t < v.size(); ++t)
            if (v[t].r == v[t - 1].r) throw std::runtime_error("pfgpu: duplicate Ybus entry");
        J.Jp[c + 1] = J.Jp[c] + (int)v.size();
        for (const E& e : v) {
            J.Ji.push_back(e.r);
            J.src.push_back(e.src);
            J.sel.push_back(e.sel);
            J.cval.push_back(e.c);
        }
    }
    return J;
}

// ---------------------------------------------------------------------------
// Ordering: minimum degree on the bus graph. Each bus with unknowns is one
// block (Va, then Vm for pq buses; the slack_absorbed unknown rides with the
// reference slack bus). The degree of a block is the number of unknowns in its
// neighbouring blocks of the current elimination graph; ties go to the lowest
// bus id. Blocks are expanded to unknowns in elimination order, so the factor
// keeps the 2x2 bus structure of the power-flow Jacobian.
// ---------------------------------------------------------------------------
static std::vector<int> bus_block_order(const JPattern& J) {
    const int N = J.N, n = J.n_bus, off = J.off;
    const int npvpq = (int)J.pvpq.size(), npq = (int)J.pq.size();
    std::vector<int> blk(N);
    if (off) blk[0] = J.slack0;
    for (int k = 0; k < npvpq; ++k) blk[k + off] = J.pvpq[k];
    for (int k = 0; k < npq; ++k) blk[k + npvpq + off] = J.pq[k];
    std::vector<std::vector<int>> members(n);
    for (int u = 0; u < N; ++u) members[blk[u]].push_back(u);  // ascending J index: Va before Vm
    std::vector<int> w(n, 0);
    for (int b = 0; b < n; ++b) w[b] = (int)members[b].size();
    std::vector<std::vector<int>> adj(n);
    for (int c = 0; c < N; ++c)
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) {
            const int a = blk[J.row_of(J.Ji[q])], b = blk[c];
            if (a == b) continue;
            adj[a].push_back(b);
            adj[b].push_back(a);
        }
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<long> deg(n, 0);
    using QE = std::pair<long, int>;
    std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
    for (int b = 0; b < n; ++b) {
        if (!w[b]) continue;
        for (int u : adj[b]) deg[b] += w[u];
        pq.push({deg[b], b});
    }
    std::vector<char> done(n, 0);
    std::vector<int> order;
    std::vector<int> merged;
    while (!pq.empty()) {
        auto [d, b] = pq.top();
        pq.pop();
        if (done[b] || d != deg[b]) continue;  // stale entry
        done[b] = 1;
        for (int u : members[b]) order.push_back(u);
        const std::vector<int> nb = adj[b];
        for (int u : nb) {
            merged.clear();
            std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(), std::back_inserter(merged));
            auto& au = adj[u];
            au.clear();
            long du = 0;
            for (int x : merged)
                if (x != u && x != b) { au.push_back(x); du += w[x]; }
            if (du != deg[u]) { deg[u] = du; pq.push({du, u}); }
        }
        adj[b].clear();
        adj[b].shrink_to_fit();
    }
    if ((int)order.size() != N) throw std::runtime_error("pfgpu: ordering lost unknowns");
    return order;
}

static inline int find_sorted(const std::vector<int>& v, int b, int e, int x) {
    auto it = std::lower_bound(v.begin() + b, v.begin() + e, x);
    if (it == v.begin() + e || *it != x) return -1;
    return (int)(it - v.begin());
}

// ---------------------------------------------------------------------------
// Symbolic LU on the symmetrised pattern with static diagonal pivots:
// S_k = struct(A(k+1:,k)) U  union over etree children c of (S_c \ {k}),
// parent(k) = min S_k (George & Liu's symbolic Cholesky on A + A^T).
// ---------------------------------------------------------------------------
std::vector<int> column_order(const JPattern& J) { return bus_block_order(J); }

SymbolicLU symbolic_lu(const JPattern& J, int max_tail, const std::vector<int>* perm) {
    SymbolicLU S;
    const int N = J.N;
    S.N = N;
    if (perm) {
        if ((int)perm->size() != N) throw std::runtime_error("pfgpu: elimination order of the wrong size");
        S.perm = *perm;
    } else {
        S.perm = bus_block_order(J);
    }
    S.pinv.assign(N, -1);
    for (int k = 0; k < N; ++k) S.pinv[S.perm[k]] = k;
    for (int k = 0; k < N; ++k)
        if (S.pinv[k] < 0) throw std::runtime_error("pfgpu: the elimination order is not a permutation");
    std::vector<std::vector<int>> low(N);
    for (int c = 0; c < N; ++c)
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) {
            const int a = S.pinv[J.row_of(J.Ji[q])], b = S.pinv[c];
            if (a == b) continue;
            low[std::min(a, b)].push_back(std::max(a, b));
        }
    std::vector<std::vector<int>> children(N);
    S.parent.assign(N, -1);
    S.cp.assign(N + 1, 0);
    std::vector<std::vector<int>> arms(N);
    std::vector<int> mark(N, -1);
    for (int k = 0; k < N; ++k) {
        std::vector<int>& s = arms[k];
        mark[k] = k;
        for (int r : low[k]) if (mark[r] != k) { mark[r] = k; s.push_back(r); }
        for (int c : children[k])
            for (int r : arms[c]) if (mark[r] != k) { mark[r] = k; s.push_back(r); }
        std::sort(s.begin(), s.end());
        if (!s.empty()) { S.parent[k] = s[0]; children[s[0]].push_back(k); }
        S.cp[k + 1] = S.cp[k] + (int)s.size();
    }
    S.ci.reserve(S.cp[N]);
    for (int k = 0; k < N; ++k) S.ci.insert(S.ci.end(), arms[k].begin(), arms[k].end());
    S.nnzL = S.cp[N];
    S.off_L = N;
    S.off_U = S.off_L + S.nnzL;
    S.off_y = S.off_U + S.nnzL;
    S.nvals = S.off_y + N;
    // J entry -> value slot
    S.jslot.resize(J.Ji.size());
    for (int c = 0; c < N; ++c)
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) {
            const int a = S.pinv[J.row_of(J.Ji[q])], b = S.pinv[c];
            int slot;
            if (a == b) slot = a;
            else if (a > b) slot = (int)S.off_L + find_sorted(S.ci, S.cp[b], S.cp[b + 1], a);
            else slot = (int)S.off_U + find_sorted(S.ci, S.cp[a], S.cp[a + 1], b);
            S.jslot[q] = slot;
        }
    double fl = 0;
    for (int k = 0; k < N; ++k) {
        const double c = S.cp[k + 1] - S.cp[k];
        fl += 2.0 * c * c + c + 1;  // rank-1 update of the arm block, scaling, reciprocal
    }
    S.flops = fl;
    // dense tail = the top chain of the elimination tree (parent(k) == k+1 up to the root)
    int k0 = N - 1;
    while (k0 > 0 && S.parent[k0 - 1] == k0) --k0;
    int m = N - k0;
    if (m > max_tail) m = max_tail;
    if (m < 8) m = 0;
    S.m = m;
    S.k0 = N - m;
    return S;
}

// ---------------------------------------------------------------------------
// Schedule. Value slot of the target (a, b) of a pair through pivot k:
//   a == b -> diag a ; a > b -> L(a, b) in arm b ; a < b -> U(a, b) in arm a.
// Tasks are entries; a task's level is 1 + the max level of what it reads.
// ---------------------------------------------------------------------------
Schedule build_schedule(const SymbolicLU& S) {
    Schedule sch;
    const int N = S.N;
    const int64_t nv = S.nvals;
    auto slot_of = [&](int a, int b) -> int {
        if (a == b) return a;
        if (a > b) return (int)S.off_L + find_sorted(S.ci, S.cp[b], S.cp[b + 1], a);
        return (int)S.off_U + find_sorted(S.ci, S.cp[a], S.cp[a + 1], b);
    };
    const int k0 = S.k0;
    // count pairs per target (pivots of the dense tail are applied by the dense phase)
    std::vector<int> cnt(nv + 1, 0);
    for (int k = 0; k < k0; ++k) {
        const int b0 = S.cp[k], b1 = S.cp[k + 1];
        for (int qa = b0; qa < b1; ++qa) {
            const int a = S.ci[qa];
            for (int qb = b0; qb < b1; ++qb) cnt[slot_of(a, S.ci[qb])]++;
            cnt[S.off_y + a]++;
        }
    }
    std::vector<int> start(nv + 1, 0);
    for (int64_t t = 0; t < nv; ++t) start[t + 1] = start[t] + cnt[t];
    const int64_t npairs = start[nv];
    std::vector<int> pl(npairs), pu(npairs), pd(npairs);
    std::vector<int> fillp(start.begin(), start.end() - 1);
    for (int k = 0; k < k0; ++k) {
        const int b0 = S.cp[k], b1 = S.cp[k + 1];
        for (int qa = b0; qa < b1; ++qa) {
            const int a = S.ci[qa];
            const int lslot = (int)S.off_L + qa;
            for (int qb = b0; qb < b1; ++qb) {
                const int t = slot_of(a, S.ci[qb]);
                const int p = fillp[t]++;
                pl[p] = lslot; pu[p] = (int)S.off_U + qb; pd[p] = k;
            }
            const int t = (int)S.off_y + a;
            const int p = fillp[t]++;
            pl[p] = lslot; pu[p] = (int)S.off_y + k; pd[p] = k;
        }
    }
    // levels in a topological order: column by column, then y of that column
    std::vector<int> level(nv, 0);
    auto eval_level = [&](int t) {
        int lv = 0;
        for (int p = start[t]; p < start[t + 1]; ++p)
            lv = std::max({lv, level[pl[p]] + 1, level[pu[p]] + 1, level[pd[p]] + 1});
        level[t] = lv;
    };
    for (int j = 0; j < N; ++j) {
        eval_level(j);
        for (int q = S.cp[j]; q < S.cp[j + 1]; ++q) {
            eval_level((int)S.off_L + q);
            eval_level((int)S.off_U + q);
        }
        eval_level((int)S.off_y + j);
    }
    int nlev = 0;
    for (int64_t t = 0; t < nv; ++t) nlev = std::max(nlev, level[t] + 1);
    // bucket tasks by level (stable: slot order inside a level)
    std::vector<int> lc(nlev + 1, 0);
    for (int64_t t = 0; t < nv; ++t) lc[level[t] + 1]++;
    for (int l = 0; l < nlev; ++l) lc[l + 1] += lc[l];
    sch.lvl_ptr = lc;
    std::vector<int> order(nv);
    {
        std::vector<int> f(lc.begin(), lc.end() - 1);
        for (int64_t t = 0; t < nv; ++t) order[f[level[t]]++] = (int)t;
    }
    sch.task_out.resize(nv);
    sch.task_pbeg.resize(nv);
    sch.task_pend.resize(nv);
    sch.task_diag.resize(nv);
    sch.pair_l.resize(npairs);
    sch.pair_u.resize(npairs);
    sch.pair_d.resize(npairs);
    int64_t pp = 0;
    for (int64_t i = 0; i < nv; ++i) {
        const int t = order[i];
        sch.task_out[i] = t;
        sch.task_diag[i] = (t < k0) ? 1 : 0;  // tail pivots are inverted by the dense phase
        sch.task_pbeg[i] = (int)pp;
        for (int p = start[t]; p < start[t + 1]; ++p, ++pp) {
            sch.pair_l[pp] = pl[p];
            sch.pair_u[pp] = pu[p];
            sch.pair_d[pp] = pd[p];
        }
        sch.task_pend[i] = (int)pp;
        sch.max_pairs_per_task = std::max(sch.max_pairs_per_task, start[t + 1] - start[t]);
    }
    sch.npairs = npairs;
    // backward levels (columns of the dense tail are solved by the dense phase first)
    std::vector<int> bl(N, -1);  // tail columns: solved before the first level
    int nb = 0;
    for (int j = k0 - 1; j >= 0; --j) {
        int lv = 0;
        for (int q = S.cp[j]; q < S.cp[j + 1]; ++q) lv = std::max(lv, bl[S.ci[q]] + 1);
        bl[j] = lv;
        nb = std::max(nb, lv + 1);
    }
    std::vector<int> bc(nb + 1, 0);
    for (int j = 0; j < k0; ++j) bc[bl[j] + 1]++;
    for (int l = 0; l < nb; ++l) bc[l + 1] += bc[l];
    sch.blvl_ptr = bc;
    sch.bcol.resize(k0);
    {
        std::vector<int> f(bc.begin(), bc.end() - 1);
        for (int j = 0; j < k0; ++j) sch.bcol[f[bl[j]]++] = j;
    }
    return sch;
}

std::string describe(const JPattern& J, const SymbolicLU& S, const Schedule& sch) {
    std::ostringstream os;
    const int nlev = (int)sch.lvl_ptr.size() - 1;
    int maxw = 0;
    for (int l = 0; l < nlev; ++l) maxw = std::max(maxw, sch.lvl_ptr[l + 1] - sch.lvl_ptr[l]);
    int etree_h = 0;
    {
        std::vector<int> h(S.N, 0);
        for (int k = 0; k < S.N; ++k)
            if (S.parent[k] >= 0) h[S.parent[k]] = std::max(h[S.parent[k]], h[k] + 1);
        for (int k = 0; k < S.N; ++k) etree_h = std::max(etree_h, h[k] + 1);
    }
    os << "N=" << J.N << " nnzJ=" << J.Ji.size() << " nnzL=" << S.nnzL << " nnzLU=" << (S.N + 2 * S.nnzL)
       << " slots=" << S.nvals << " flops=" << S.flops << " etree_h=" << etree_h << " levels=" << nlev
       << " max_level_width=" << maxw << " pairs=" << sch.npairs << " max_pairs=" << sch.max_pairs_per_task
       << " back_levels=" << (sch.blvl_ptr.size() - 1);
    return os.str();
}

}  // namespace pfgpu

namespace pfgpu {

void recycle_slots(DevicePlan& P);
void build_jfill_table(DevicePlan& P);

DevicePlan build_device_plan(const YbusPattern& Y, const JPattern& J, const SymbolicLU& S, const Schedule& sch) {
    DevicePlan P;
    const int n = Y.n, N = J.N;
    P.n = n; P.N = N; P.off = J.off; P.nnzY = Y.p[n]; P.nJ = (int)J.Ji.size(); P.slack0 = J.slack0;
    P.nvals = (int)S.nvals;
    P.flops = S.flops;
    P.nnzLU = S.N + 2 * S.nnzL;
    const int nlev = (int)sch.lvl_ptr.size() - 1;
    const int nblev = (int)sch.blvl_ptr.size() - 1;
    P.nlev = nlev; P.nblev = nblev;

    // --- slot hotness and renumbering ----------------------------------
    std::vector<int> hot(S.nvals, -1);
    for (int l = 0; l < nlev; ++l)
        for (int t = sch.lvl_ptr[l]; t < sch.lvl_ptr[l + 1]; ++t) {
            hot[sch.task_out[t]] = std::max(hot[sch.task_out[t]], l);
            for (int p = sch.task_pbeg[t]; p < sch.task_pend[t]; ++p) {
                hot[sch.pair_l[p]] = std::max(hot[sch.pair_l[p]], l);
                hot[sch.pair_u[p]] = std::max(hot[sch.pair_u[p]], l);
                hot[sch.pair_d[p]] = std::max(hot[sch.pair_d[p]], l);
            }
        }
    // backward reads U, x, 1/d; make it count as later than every factor level
    for (int l = 0; l < nblev; ++l)
        for (int q = sch.blvl_ptr[l]; q < sch.blvl_ptr[l + 1]; ++q) {
            const int j = sch.bcol[q];
            const int lv = nlev + l;
            hot[S.off_y + j] = std::max(hot[S.off_y + j], lv);
            hot[j] = std::max(hot[j], lv);
            for (int a = S.cp[j]; a < S.cp[j + 1]; ++a) hot[S.off_U + a] = std::max(hot[S.off_U + a], lv);
        }
    // dense tail first: tile [0, m*(m+1)) row-major with row stride m + 1, its vector as
    // column m; every other slot follows by decreasing hotness
    const int m = S.m, k0 = S.k0;
    P.m = m;
    std::vector<int> colof(S.nnzL);
    for (int k = 0; k < N; ++k)
        for (int q = S.cp[k]; q < S.cp[k + 1]; ++q) colof[q] = k;
    std::vector<int> nw(S.nvals, -1);
    for (int j = k0; j < N; ++j) {
        nw[j] = (j - k0) * (m + 2);
        nw[S.off_y + j] = (j - k0) * (m + 1) + m;
        for (int q = S.cp[j]; q < S.cp[j + 1]; ++q) {
            const int r = S.ci[q];  // > j >= k0
            nw[S.off_L + q] = (r - k0) * (m + 1) + (j - k0);
            nw[S.off_U + q] = (j - k0) * (m + 1) + (r - k0);
        }
    }
    std::vector<int> ord;
    for (int i = 0; i < (int)S.nvals; ++i)
        if (nw[i] < 0) ord.push_back(i);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return hot[a] > hot[b]; });
    const int tail_slots = m * m + m;
    for (size_t i = 0; i < ord.size(); ++i) nw[ord[i]] = tail_slots + (int)i;
    P.nvals = tail_slots + (int)ord.size();
    auto ys = [&](int k) { return nw[S.off_y + k]; };  // slot of y/x for permuted index k

    // --- Ybus row view ----------------------------------------------------
    P.yr_ptr.assign(n + 1, 0);
    for (int c = 0; c < n; ++c)
        for (int q = Y.p[c]; q < Y.p[c + 1]; ++q) P.yr_ptr[Y.i[q] + 1]++;
    for (int r = 0; r < n; ++r) P.yr_ptr[r + 1] += P.yr_ptr[r];
    P.yr_col.resize(P.nnzY); P.yr_pos.resize(P.nnzY);
    {
        std::vector<int> f(P.yr_ptr.begin(), P.yr_ptr.end() - 1);
        for (int c = 0; c < n; ++c)
            for (int q = Y.p[c]; q < Y.p[c + 1]; ++q) {
                const int k = f[Y.i[q]]++;
                P.yr_col[k] = c; P.yr_pos[k] = q;
            }
    }
    // --- F rows and unknowns --------------------------------------------
    const int npvpq = (int)J.pvpq.size(), npq = (int)J.pq.size(), off = J.off;
    // the F entry of J row r is the right-hand side of factored row row_of(r)
    P.fslot_p.assign(n, -1); P.fslot_q.assign(n, -1);
    if (off) P.fslot_p[J.slack0] = ys(S.pinv[J.row_of(0)]);
    for (int k = 0; k < npvpq; ++k) P.fslot_p[J.pvpq[k]] = ys(S.pinv[J.row_of(k + off)]);
    for (int k = 0; k < npq; ++k) P.fslot_q[J.pq[k]] = ys(S.pinv[J.row_of(k + npvpq + off)]);
    P.ux_bus.resize(N); P.ux_kind.resize(N); P.ux_slot.resize(N);
    if (off) { P.ux_bus[0] = J.slack0; P.ux_kind[0] = 0; P.ux_slot[0] = ys(S.pinv[0]); }
    for (int k = 0; k < npvpq; ++k) { P.ux_bus[k + off] = J.pvpq[k]; P.ux_kind[k + off] = 1; P.ux_slot[k + off] = ys(S.pinv[k + off]); }
    for (int k = 0; k < npq; ++k) {
        const int j = k + npvpq + off;
        P.ux_bus[j] = J.pq[k]; P.ux_kind[j] = 2; P.ux_slot[j] = ys(S.pinv[j]);
    }
    // --- J entries ----------------------------------------------------------
    // bus of a J row / column: rows/cols 'off..off+npvpq' -> pvpq, then pq, 0 -> slack0
    auto jbus = [&](int jr) -> int {
        if (off && jr == 0) return J.slack0;
        const int k = jr - off;
        return k < npvpq ? J.pvpq[k] : J.pq[k - npvpq];
    };
    P.j_ypos.resize(P.nJ); P.j_rbus.resize(P.nJ); P.j_cbus.resize(P.nJ); P.j_sel.resize(P.nJ); P.j_cval.resize(P.nJ);
    P.j_slot.resize(P.nJ);
    std::vector<int> npairs_of(S.nvals, 0);
    for (int l = 0; l < nlev; ++l)
        for (int t = sch.lvl_ptr[l]; t < sch.lvl_ptr[l + 1]; ++t)
            npairs_of[sch.task_out[t]] = sch.task_pend[t] - sch.task_pbeg[t];
    std::vector<char> is_jslot(S.nvals, 0);
    for (int c = 0; c < N; ++c)
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) {
            P.j_ypos[q] = J.src[q];
            P.j_sel[q] = J.sel[q];
            P.j_cval[q] = J.cval[q];
            if (J.src[q] >= 0) {
                P.j_rbus[q] = Y.i[J.src[q]];
                // column bus of the Ybus entry: the bus whose CSC column holds src
                P.j_cbus[q] = (int)(std::upper_bound(Y.p.begin(), Y.p.end(), J.src[q]) - Y.p.begin()) - 1;
            } else {
                P.j_rbus[q] = jbus(J.Ji[q]);
                P.j_cbus[q] = jbus(c);
            }
            const int slot = S.jslot[q];
            is_jslot[slot] = 1;
            const bool recip = slot < k0 && npairs_of[slot] == 0;  // pivot with nothing to eliminate
            P.j_slot[q] = nw[slot] | (recip ? (1 << 30) : 0);
        }
    {
        std::vector<char> jfinal(P.nvals, 0);
        for (int64_t sl = 0; sl < S.off_y; ++sl)
            if (is_jslot[sl]) jfinal[nw[sl]] = 1;
        std::vector<char> isy(P.nvals, 0);
        for (int k = 0; k < N; ++k) isy[nw[S.off_y + k]] = 1;
        for (int sl = 0; sl < P.nvals; ++sl)
            if (!jfinal[sl] && !isy[sl]) P.zero_slots.push_back(sl);
    }
    // --- factor tasks (tasks without pairs are covered by the J fill / mismatch) ---------
    std::vector<int> lvl(1, 0);
    for (int l = 0; l < nlev; ++l) {
        int mp = 0;
        for (int t = sch.lvl_ptr[l]; t < sch.lvl_ptr[l + 1]; ++t) {
            const int np = sch.task_pend[t] - sch.task_pbeg[t];
            if (np == 0) continue;
            const int o = sch.task_out[t];
            const int kind = sch.task_diag[t] ? 1 : (o >= S.off_y ? 2 : 0);
            P.tasks.insert(P.tasks.end(), {nw[o], sch.task_pbeg[t], sch.task_pend[t], kind});
            mp = std::max(mp, np);
        }
        const int cnt = (int)P.tasks.size() / 4;
        if (cnt == lvl.back()) continue;  // empty level after pruning
        lvl.push_back(cnt);
        P.f_maxp.push_back(mp);
    }
    P.f_lvl = lvl;
    P.nlev = (int)lvl.size() - 1;
    P.pairs.resize(4 * sch.npairs);
    for (int64_t p = 0; p < sch.npairs; ++p) {
        P.pairs[4 * p] = nw[sch.pair_l[p]];
        P.pairs[4 * p + 1] = nw[sch.pair_u[p]];
        P.pairs[4 * p + 2] = nw[sch.pair_d[p]];
        P.pairs[4 * p + 3] = 0;
    }
    // --- backward tasks ---------------------------------------------------------
    P.b_lvl = sch.blvl_ptr;
    P.b_maxp.assign(nblev, 0);
    P.btasks.resize(4 * (size_t)k0);
    int bp = 0;
    for (int l = 0; l < nblev; ++l)
        for (int q = sch.blvl_ptr[l]; q < sch.blvl_ptr[l + 1]; ++q) {
            const int j = sch.bcol[q];
            P.btasks[4 * q] = ys(j);
            P.btasks[4 * q + 1] = bp;
            for (int a = S.cp[j]; a < S.cp[j + 1]; ++a) {
                P.bpairs.push_back(nw[S.off_U + a]);
                P.bpairs.push_back(ys(S.ci[a]));
                ++bp;
            }
            P.btasks[4 * q + 2] = bp;
            P.btasks[4 * q + 3] = nw[j];
            P.b_maxp[l] = std::max(P.b_maxp[l], S.cp[j + 1] - S.cp[j]);
        }
    recycle_slots(P);
    build_jfill_table(P);
    return P;
}

// Groups the J entries by the Ybus entry they are computed from (after the slots are final).
void build_jfill_table(DevicePlan& P) {
    std::vector<int> ent(P.nnzY, -1);
    for (int q = 0; q < P.nJ; ++q) {
        const int sel = P.j_sel[q];
        if (sel == SEL_CONST) {
            P.jc_slot.push_back(P.j_slot[q]);
            P.jc_jidx.push_back(q);
            P.jc_val.push_back(P.j_cval[q]);
            continue;
        }
        const int yp = P.j_ypos[q];
        if (ent[yp] < 0) {
            ent[yp] = (int)P.ye_pos.size();
            P.ye_pos.push_back(yp);
            P.ye_rc.push_back(P.j_rbus[q]);
            P.ye_rc.push_back(P.j_cbus[q]);
            P.ye_slot.insert(P.ye_slot.end(), {-1, -1, -1, -1});
            P.ye_jidx.insert(P.ye_jidx.end(), {-1, -1, -1, -1});
        }
        const int e = ent[yp];
        if (P.ye_rc[2 * e] != P.j_rbus[q] || P.ye_rc[2 * e + 1] != P.j_cbus[q])
            throw std::runtime_error("pfgpu: J entries of one Ybus entry disagree on its buses");
        if (P.ye_slot[4 * e + sel] != -1) throw std::runtime_error("pfgpu: two J entries of one Ybus entry and selector");
        P.ye_slot[4 * e + sel] = P.j_slot[q];
        P.ye_jidx[4 * e + sel] = q;
    }
    // diagonal entries first (their J values need Ibus of their bus), in bus order
    const int ne = (int)P.ye_pos.size();
    std::vector<int> ord(ne);
    for (int e = 0; e < ne; ++e) ord[e] = e;
    auto diag = [&](int e) { return P.ye_rc[2 * e] == P.ye_rc[2 * e + 1]; };
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        if (diag(x) != diag(y)) return diag(x);
        return diag(x) && P.ye_rc[2 * x] < P.ye_rc[2 * y];
    });
    std::vector<int> pos(ne), rc(2 * ne), sl(4 * ne), jx(4 * ne);
    for (int i = 0; i < ne; ++i) {
        const int e = ord[i];
        pos[i] = P.ye_pos[e];
        rc[2 * i] = P.ye_rc[2 * e];
        rc[2 * i + 1] = P.ye_rc[2 * e + 1];
        for (int k = 0; k < 4; ++k) {
            sl[4 * i + k] = P.ye_slot[4 * e + k];
            jx[4 * i + k] = P.ye_jidx[4 * e + k];
        }
    }
    P.ye_pos.swap(pos);
    P.ye_rc.swap(rc);
    P.ye_slot.swap(sl);
    P.ye_jidx.swap(jx);
    P.nyd = 0;
    P.ye_diag_of.assign(P.n, -1);
    while (P.nyd < ne && P.ye_rc[2 * P.nyd] == P.ye_rc[2 * P.nyd + 1]) {
        P.ye_diag_of[P.ye_rc[2 * P.nyd]] = P.nyd;
        ++P.nyd;
    }
}

// ---------------------------------------------------------------------------
// Slot recycling. Time axis of one NR iteration: 0 = J fill, l + 1 = factor level
// l, nlev + 1 = dense tail, nlev + 2 + l = backward level l, nlev + 2 + nblev =
// update. A value lives from its first write to its last read. J entries are
// written at 0; a fill-in entry (neither a J entry nor in the dense tile) is
// assigned by its own task (kTaskAssign) instead of being zeroed at 0, so it lives from
// its task's level only. The tile and the right-hand side / solution slots (written
// by the mismatch of the previous iteration, read by the update) keep their slots;
// every other value takes a physical slot freed by a value whose last read came at
// an earlier time (greedy interval colouring, optimal for interval graphs).
// Physical slots are then ordered tile first, then the right-hand side / solution slots,
// then by decreasing accesses per iteration.
// ---------------------------------------------------------------------------
void recycle_slots(DevicePlan& P) {
    // one extra slot: the shared zero (structural zeros, and the operand of padding items)
    const int NV = P.nvals + 1, tile = P.m * P.m + P.m;
    const int zslot = P.nvals;
    const int nl = P.nlev, nb = P.nblev, T_END = nl + 2 + nb;
    std::vector<char> fixed(NV, 0), isj(NV, 0);
    for (int s = 0; s < tile; ++s) fixed[s] = 1;
    for (int b = 0; b < P.n; ++b) {
        if (P.fslot_p[b] >= 0) fixed[P.fslot_p[b]] = 1;
        if (P.fslot_q[b] >= 0) fixed[P.fslot_q[b]] = 1;
    }
    for (int x : P.ux_slot) fixed[x] = 1;
    for (int q = 0; q < P.nJ; ++q) isj[P.j_slot[q] & ((1 << 30) - 1)] = 1;
    // fill-in entries outside the tile: their task assigns them. Entries of the
    // symmetrised pattern that are neither J entries nor written by any task are
    // structural zeros: they all share one slot, zeroed by the J fill.
    std::vector<char> fill(NV, 0), written(NV, 0);
    for (size_t t = 0; t < P.tasks.size() / 4; ++t) written[P.tasks[4 * t]] = 1;
    std::vector<char> zero(NV, 0);
    for (int s = 0; s < NV; ++s) {
        fill[s] = !fixed[s] && !isj[s] && written[s];
        zero[s] = !fixed[s] && !isj[s] && !written[s];
    }
    for (size_t t = 0; t < P.tasks.size() / 4; ++t)
        if (fill[P.tasks[4 * t]]) P.tasks[4 * t + 3] |= kTaskAssign;
    {
        std::vector<int> z;
        for (int s : P.zero_slots)
            if (!fill[s] && !zero[s]) z.push_back(s);
        z.push_back(zslot);
        P.zero_slots.swap(z);
    }
    fixed[zslot] = 1;
    // intervals
    std::vector<int> start(NV, 1 << 30), end(NV, -1), hot(NV, -1);
    std::vector<long long> nacc(NV, 0);  // reads and writes of each value per iteration
    auto acc = [&](int s, int t, bool w) {
        end[s] = std::max(end[s], t);
        if (w) start[s] = std::min(start[s], t);
        if (t < T_END) hot[s] = std::max(hot[s], t);
        ++nacc[s];
    };
    for (int q = 0; q < P.nJ; ++q) acc(P.j_slot[q] & ((1 << 30) - 1), 0, true);
    for (int s : P.zero_slots) acc(s, 0, true);
    for (int l = 0; l < nl; ++l)
        for (int t = P.f_lvl[l]; t < P.f_lvl[l + 1]; ++t) {
            acc(P.tasks[4 * t], l + 1, true);
            for (int q = P.tasks[4 * t + 1]; q < P.tasks[4 * t + 2]; ++q)
                for (int k = 0; k < 3; ++k) acc(P.pairs[4 * q + k], l + 1, false);
        }
    for (int l = 0; l < nb; ++l)
        for (int t = P.b_lvl[l]; t < P.b_lvl[l + 1]; ++t) {
            const int tt = nl + 2 + l;
            acc(P.btasks[4 * t], tt, true);
            acc(P.btasks[4 * t + 3], tt, false);
            for (int q = P.btasks[4 * t + 1]; q < P.btasks[4 * t + 2]; ++q) {
                acc(P.bpairs[2 * q], tt, false);
                acc(P.bpairs[2 * q + 1], tt, false);
            }
        }
    for (int s = 0; s < NV; ++s) {
        if (zero[s] && s != zslot) continue;  // aliases of the shared zero slot
        if (fixed[s]) { start[s] = -1; end[s] = T_END; }
        if (end[s] < 0) throw std::runtime_error("pfgpu: value slot never used");
        if (start[s] > end[s])
            throw std::runtime_error("pfgpu: value slot read before it is written (slot " + std::to_string(s) + " start " +
                                     std::to_string(start[s]) + " end " + std::to_string(end[s]) + " fill " +
                                     std::to_string((int)fill[s]) + " j " + std::to_string((int)isj[s]) + ")");
    }
    // colour the non-fixed values
    std::vector<int> phys(NV, -1);
    int nphys = 0;
    for (int s = 0; s < NV; ++s)
        if (fixed[s]) phys[s] = nphys++;  // placeholder ids, reordered below
    for (int s = 0; s < NV; ++s)
        if (zero[s] && s != zslot) phys[s] = phys[zslot];
    std::vector<int> ord;
    for (int s = 0; s < NV; ++s)
        if (!fixed[s] && !zero[s]) ord.push_back(s);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return start[a] < start[b]; });
    using FE = std::pair<int, int>;  // (last end, physical)
    std::priority_queue<FE, std::vector<FE>, std::greater<FE>> freeq;
    for (int s : ord) {
        if (!freeq.empty() && freeq.top().first < start[s]) {
            const int ph = freeq.top().second;
            freeq.pop();
            phys[s] = ph;
        } else {
            phys[s] = nphys++;
        }
        freeq.push({end[s], phys[s]});
    }
    // final order: tile first, then the right-hand side / solution slots, then by decreasing
    // accesses per iteration (ties: decreasing hotness), so that a launch whose LDS holds only the
    // first slots sends the fewest accesses to the spilled ones. The solution slots come first
    // because the backward levels write and read them: one of them spilled puts a memory round
    // trip between two of a level's barriers.
    std::vector<int> phot(nphys, -1);
    std::vector<long long> pacc(nphys, 0);
    for (int s = 0; s < NV; ++s) {
        phot[phys[s]] = std::max(phot[phys[s]], hot[s]);
        pacc[phys[s]] += nacc[s];
    }
    for (int s = 0; s < NV; ++s)  // the zero slot's aliases are read-only zeros at all times
        if (zero[s]) { start[s] = -1; end[s] = T_END; }
    std::vector<int> pord, fin(nphys, -1);
    for (int s = 0; s < tile; ++s) fin[phys[s]] = s;
    for (int ph = 0; ph < nphys; ++ph)
        if (fin[ph] < 0) pord.push_back(ph);
    std::vector<char> prhs(nphys, 0);
    for (int b = 0; b < P.n; ++b) {
        if (P.fslot_p[b] >= 0) prhs[phys[P.fslot_p[b]]] = 1;
        if (P.fslot_q[b] >= 0) prhs[phys[P.fslot_q[b]]] = 1;
    }
    for (int x : P.ux_slot) prhs[phys[x]] = 1;
    std::stable_sort(pord.begin(), pord.end(), [&](int a, int b) {
        if (prhs[a] != prhs[b]) return prhs[a] > prhs[b];
        return pacc[a] != pacc[b] ? pacc[a] > pacc[b] : phot[a] > phot[b];
    });
    for (size_t i = 0; i < pord.size(); ++i) fin[pord[i]] = tile + (int)i;
    std::vector<int> mp(NV);
    for (int s = 0; s < NV; ++s) mp[s] = fin[phys[s]];
    // self-check: the values sharing a physical slot have disjoint lifetimes
    {
        std::vector<std::vector<std::pair<int, int>>> by(nphys);
        for (int s = 0; s < NV; ++s)
            if (!zero[s] || s == zslot) by[mp[s]].push_back({start[s], end[s]});
        for (auto& v : by) {
            std::sort(v.begin(), v.end());
            for (size_t i = 1; i < v.size(); ++i)
                if (v[i].first <= v[i - 1].second) throw std::runtime_error("pfgpu: slot recycling overlap");
        }
    }
    // rewrite every slot reference
    for (int& x : P.j_slot) x = mp[x & ((1 << 30) - 1)] | (x & (1 << 30));
    for (int& x : P.fslot_p) if (x >= 0) x = mp[x];
    for (int& x : P.fslot_q) if (x >= 0) x = mp[x];
    for (int& x : P.ux_slot) x = mp[x];
    for (int& x : P.zero_slots) x = mp[x];
    for (size_t t = 0; t < P.tasks.size() / 4; ++t) P.tasks[4 * t] = mp[P.tasks[4 * t]];
    for (size_t q = 0; q < P.pairs.size() / 4; ++q)
        for (int k = 0; k < 3; ++k) P.pairs[4 * q + k] = mp[P.pairs[4 * q + k]];
    for (size_t t = 0; t < P.btasks.size() / 4; ++t) {
        P.btasks[4 * t] = mp[P.btasks[4 * t]];
        P.btasks[4 * t + 3] = mp[P.btasks[4 * t + 3]];
    }
    for (int& x : P.bpairs) x = mp[x];
    P.nvals = nphys;
    P.zslot = mp[zslot];
}

std::vector<int> level_teams(const std::vector<int>& lvl_ptr, const std::vector<int>& maxp, int G, int nslot) {
    const int nlev = (int)lvl_ptr.size() - 1;
    std::vector<int> team(nlev, 1);
    const int tmax = std::max(1, 64 / G);
    for (int l = 0; l < nlev; ++l) {
        const int ntask = lvl_ptr[l + 1] - lvl_ptr[l];
        const int mp = std::max(1, maxp[l]);
        double best = 1e300;
        for (int t = 1; t <= tmax && t <= nslot; t *= 2) {
            const int conc = std::max(1, nslot / t);                 // tasks in flight
            const double rounds = std::ceil((double)ntask / conc);
            const double cost = rounds * (std::ceil((double)mp / t) * 3.0 + 2.0 * std::log2((double)t) + 4.0);
            if (cost < best - 1e-9) { best = cost; team[l] = t; }
        }
    }
    return team;
}

namespace {

// Placement of one level's tasks on the lanes of one instance.
struct LevelPlacement {
    int K = 0;                  // items per lane
    std::vector<int> lane0;     // per task: first lane of its team
    std::vector<int> lt;        // per task: log2 of its team size
    std::vector<int> k0;        // per task: first item index in its team's lanes
};

int ceil_div(int a, int b) { return (a + b - 1) / b; }
int log2_ceil(int x) {
    int l = 0;
    while ((1 << l) < x) ++l;
    return l;
}

// For a target of K items per lane, each task takes the smallest power-of-two
// team T with ceil(np / T) <= K; the tasks of one team size are packed onto
// teams of that size (largest first, onto the least loaded team that still has
// room), and the team-size classes are laid out from lane 0 in decreasing size,
// so every team is aligned to its size. Returns false if the lanes do not suffice.
bool place_level_k(const std::vector<int>& np, int K, int tmax_log, int nslot, LevelPlacement& pl,
                   double& cost) {
    const int nt = (int)np.size();
    std::vector<std::vector<int>> cls(tmax_log + 1);
    for (int i = 0; i < nt; ++i) {
        const int need = std::max(1, ceil_div(std::max(np[i], 1), K));
        const int l = log2_ceil(need);
        if (l > tmax_log) return false;
        cls[l].push_back(i);
    }
    pl.K = K;
    pl.lane0.assign(nt, 0);
    pl.lt.assign(nt, 0);
    pl.k0.assign(nt, 0);
    int lane = 0;
    double worst = 0;
    for (int l = tmax_log; l >= 0; --l) {
        if (cls[l].empty()) continue;
        const int T = 1 << l;
        auto items = [&](int i) { return std::max(1, ceil_div(np[i], T)); };
        std::vector<int>& v = cls[l];
        std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return items(x) > items(y); });
        // teams of this class: (load, tasks), least loaded first among those with room
        std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<>> heap;
        std::vector<int> load, ntask;
        for (int i : v) {
            const int c = items(i);
            int t;
            if (!heap.empty() && heap.top().first + c <= K) {
                t = heap.top().second;
                heap.pop();
            } else {
                t = (int)load.size();
                load.push_back(0);
                ntask.push_back(0);
            }
            pl.lane0[i] = lane + t * T;
            pl.lt[i] = l;
            pl.k0[i] = load[t];
            load[t] += c;
            ntask[t] += 1;
            heap.push({load[t], t});
        }
        for (int t = 0; t < (int)load.size(); ++t) worst = std::max(worst, ntask[t] * (2.0 * l + 4.0));
        lane += (int)load.size() * T;
        if (lane > nslot) return false;
    }
    // per lane: items (one pipelined load and the LDS reads of one pair each) plus the task
    // ends of its team (a log2(T)-step shuffle reduction and the leader's read-modify-write)
    cost = K + worst;
    return true;
}

LevelPlacement choose_placement(const std::vector<int>& np, int G, int nslot) {
    const int tmax = std::max(1, std::min(64 / G, nslot));
    const int tmax_log = log2_ceil(tmax + 1) - 1;  // floor(log2(tmax))
    long total = 0;
    int maxnp = 1;
    for (int x : np) {
        total += std::max(x, 1);
        maxnp = std::max(maxnp, x);
    }
    const int k_lo = std::max({1, (int)((total + nslot - 1) / nslot), ceil_div(maxnp, 1 << tmax_log)});
    LevelPlacement best, pl;
    double best_cost = 1e300;
    // candidate K: k_lo, then growing by ~12% up to the all-singleton layout
    const int k_hi = std::max(k_lo, maxnp) * 2 + (int)np.size();
    for (int K = k_lo; K <= k_hi; K = std::max(K + 1, (int)(K * 1.125))) {
        double cost;
        if (!place_level_k(np, K, tmax_log, nslot, pl, cost)) continue;
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = pl;
        }
        if (K > 4 * best_cost) break;
    }
    if (best.K == 0) throw std::runtime_error("lane streams: no placement fits the lanes");
    return best;
}

// Chunked placement of one level (lean encoding): a lane's items form c chunks of Lc
// items, a task takes one chunk of a team of T = 2^lt lanes (the smallest power of two
// with ceil(np / T) <= Lc). Teams are laid out in decreasing size, column after column of
// nslot lanes, so that each is aligned to its size. Lc minimises an estimate of the level's
// time in item units: the issue work of a lane (c chunks of Lc items plus an end each: a
// reduction of lt steps, the read of the old value and the leader's write) times the waves
// that share a SIMD, plus the latency of the item loads past the kHead items per lane that
// are loaded during the previous level (two items per round trip).
struct ChunkPlacement {
    int Lc = 0, c = 0, U = 0, maxlt = 0;
    std::vector<int> col, lane0, lt;  // per task: chunk column, first lane, log2 team size
};

ChunkPlacement place_chunked(const std::vector<int>& np, int G, int nslot, int lane_round, int kHead) {
    const int tmax = std::max(1, std::min(64 / G, nslot));
    const int tmax_log = log2_ceil(tmax + 1) - 1;  // floor(log2(tmax))
    const int nt = (int)np.size();
    int maxnp = 1;
    for (int x : np) maxnp = std::max(maxnp, x);
    const int lc_lo = std::max(1, ceil_div(maxnp, 1 << tmax_log));
    ChunkPlacement best;
    double best_cost = 1e300;
    std::vector<int> lt(nt);
    for (int Lc = lc_lo; Lc <= std::max(lc_lo, maxnp); Lc = Lc < 64 ? Lc + 1 : Lc + Lc / 8) {
        long total = 0;
        int maxlt = 0;
        for (int i = 0; i < nt; ++i) {
            lt[i] = log2_ceil(std::max(1, ceil_div(std::max(np[i], 1), Lc)));
            total += 1L << lt[i];
            maxlt = std::max(maxlt, lt[i]);
        }
        const int c = (int)((total + nslot - 1) / nslot);
        const int U = c == 1 ? std::min(nslot, (int)((total + lane_round - 1) / lane_round * lane_round)) : nslot;
        const double simd_load = std::max(1.0, std::ceil((double)U * G / 64.0) / 4.0);
        const int K = c * Lc;
        const double cost = simd_load * (K + c * (3.0 + 1.0 * maxlt)) + 12.0 * std::max(0, (K - kHead + 1) / 2);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best.Lc = Lc;
            best.c = c;
            best.U = U;
            best.maxlt = maxlt;
            best.lt = lt;
        }
    }
    std::vector<int> ord(nt);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return best.lt[x] > best.lt[y]; });
    best.col.assign(nt, 0);
    best.lane0.assign(nt, 0);
    long p = 0;
    for (int i : ord) {
        best.col[i] = (int)(p / nslot);
        best.lane0[i] = (int)(p % nslot);
        p += 1L << best.lt[i];
    }
    return best;
}

// Lean layout of the levels of a task list: K = c * Lc items per lane, lanes [0, U), item k
// of lane s at base + k * U + s. emit(task, pair index or -1, end, log2 team size, item)
// fills one item pre-set to pad.
template <class NPairs, class Emit>
void layout_chunked(const std::vector<int>& lvl_ptr, int G, int nslot, NPairs npairs_of, Emit emit,
                    const int (&pad)[4], int lane_round, int k_head, std::vector<int>& lv, std::vector<int>& lcs,
                    std::vector<int>& stream) {
    const int nlev = (int)lvl_ptr.size() - 1;
    lv.assign(4 * (size_t)std::max(nlev, 0), 0);
    lcs.assign(std::max(nlev, 0), 0);
    size_t base = stream.size() / 4;
    for (int l = 0; l < nlev; ++l) {
        const int t0 = lvl_ptr[l], nt = lvl_ptr[l + 1] - t0;
        std::vector<int> np(nt);
        for (int i = 0; i < nt; ++i) np[i] = npairs_of(t0 + i);
        const ChunkPlacement pl = place_chunked(np, G, nslot, lane_round, k_head);
        const int K = pl.c * pl.Lc, U = pl.U;
        lv[4 * l] = (int)base;
        lv[4 * l + 1] = K;
        lv[4 * l + 2] = 1 << pl.maxlt;
        lv[4 * l + 3] = U;
        lcs[l] = pl.Lc;
        if (K >= (1 << 16) || pl.Lc >= (1 << 15)) throw std::runtime_error("lane streams: level too long");
        if (base + (size_t)K * U >= (size_t)1 << 31) throw std::runtime_error("lane stream too large");
        const size_t at = stream.size();
        stream.resize(4 * (base + (size_t)K * U));
        for (size_t q = at; q < stream.size(); q += 4)
            for (int w = 0; w < 4; ++w) stream[q + w] = pad[w];
        for (int i = 0; i < nt; ++i) {
            const int T = 1 << pl.lt[i];
            for (int tp = 0; tp < T; ++tp)
                for (int jj = 0; jj < pl.Lc; ++jj) {
                    const int pi = jj * T + tp;
                    const size_t k = (size_t)pl.col[i] * pl.Lc + jj;
                    int* it = &stream[4 * (base + k * U + pl.lane0[i] + tp)];
                    emit(t0 + i, pi < np[i] ? pi : -1, jj == pl.Lc - 1, pl.lt[i], it);
                }
        }
        base += (size_t)K * U;
    }
}

// Lays out the levels [lvl_ptr[l], lvl_ptr[l+1]) of a task list. Level l uses lanes
// [0, U) and K items per lane, stored lane-minor at base + k * U + s. emit(task, pair
// index within the task or -1 for none, end (is the task's last item), log2 team size,
// item) fills one 4-int item (pre-set to {-1, -1, -1, 0}); the lane's last item of the
// level then gets kStreamLast in its end word.
template <class NPairs, class Emit>
void layout_streams(const std::vector<int>& lvl_ptr, int G, int nslot, NPairs npairs_of, Emit emit, int end_word,
                    const int (&pad)[4], int lane_round, int k_round, bool flag_last, std::vector<int>& lv,
                    std::vector<int>& stream) {
    const int nlev = (int)lvl_ptr.size() - 1;
    lv.assign(4 * (size_t)std::max(nlev, 0), 0);
    size_t base = stream.size() / 4;  // items already in the stream (a leading padding item)
    for (int l = 0; l < nlev; ++l) {
        const int t0 = lvl_ptr[l], nt = lvl_ptr[l + 1] - t0;
        std::vector<int> np(nt);
        for (int i = 0; i < nt; ++i) np[i] = npairs_of(t0 + i);
        const LevelPlacement pl = choose_placement(np, G, nslot);
        const int K = (pl.K + k_round - 1) / k_round * k_round;
        int maxlt = 0, U = 0;
        for (int i = 0; i < nt; ++i) {
            maxlt = std::max(maxlt, pl.lt[i]);
            U = std::max(U, pl.lane0[i] + (1 << pl.lt[i]));
        }
        U = std::min(nslot, (U + lane_round - 1) / lane_round * lane_round);
        lv[4 * l] = (int)base;
        lv[4 * l + 1] = K;
        lv[4 * l + 2] = 1 << maxlt;
        lv[4 * l + 3] = U;
        if (base + (size_t)K * U >= (size_t)1 << 31) throw std::runtime_error("lane stream too large");
        const size_t at = stream.size();
        stream.resize(4 * (base + (size_t)K * U));
        for (size_t q = at; q < stream.size(); q += 4)
            for (int w = 0; w < 4; ++w) stream[q + w] = pad[w];
        std::vector<int> count(U, 0);
        for (int i = 0; i < nt; ++i) {
            const int T = 1 << pl.lt[i];
            const int c = std::max(1, ceil_div(np[i], T));
            for (int tp = 0; tp < T; ++tp) {
                const int lane = pl.lane0[i] + tp;
                for (int j = 0; j < c; ++j) {
                    const int pi = j * T + tp;
                    const size_t k = (size_t)pl.k0[i] + j;
                    int* it = &stream[4 * (base + k * U + lane)];
                    emit(t0 + i, pi < np[i] ? pi : -1, j == c - 1, pl.lt[i], it);
                }
                count[lane] = std::max(count[lane], pl.k0[i] + c);
            }
        }
        if (flag_last)
            for (int lane = 0; lane < U; ++lane) {
                if (count[lane] == 0) throw std::runtime_error("lane streams: idle lane inside the used range");
                stream[4 * (base + (size_t)(count[lane] - 1) * U + lane) + end_word] |= kStreamLast;
            }
        base += (size_t)K * U;
    }
}

}  // namespace

LaneStreams build_lane_streams(const DevicePlan& P, int G, int nslot, const LeanFormat* lean) {
    if (G < 1 || nslot < 1) throw std::invalid_argument("build_lane_streams: bad launch shape");
    if (P.nvals >= kStreamSlotMask) throw std::invalid_argument("build_lane_streams: too many value slots");
    const int sc = lean ? lean->elem_bytes : 1;  // slot -> stored value
    if (lean && (int64_t)P.nvals * sc >= kStreamSlotMask) throw std::invalid_argument("lean streams: offsets too large");
    const int zb = P.zslot * sc;
    const int round = lean ? std::max(1, lean->lane_round) : 1;
    LaneStreams L;
    L.G = G;
    L.nslot = nslot;
    const int fpad[4] = {lean ? zb : -1, lean ? zb : -1, lean ? zb : -1, 0};
    const int bpad[4] = {lean ? zb : -1, lean ? zb : -1, 0, lean ? zb : -1};
    // lean streams start with one padding item (item 0), read in place of items past a level's end
    if (lean) {
        L.fs.assign(fpad, fpad + 4);
        L.bs.assign(bpad, bpad + 4);
    }
    auto lay = [&](const std::vector<int>& lvl_ptr, auto npairs_of, auto emit, int end_word, const int(&pad)[4],
                   std::vector<int>& lv, std::vector<int>& lcs, std::vector<int>& stream) {
        if (lean)
            layout_chunked(lvl_ptr, G, nslot, npairs_of, emit, pad, round, std::max(0, lean->head_items), lv, lcs, stream);
        else
            layout_streams(lvl_ptr, G, nslot, npairs_of, emit, end_word, pad, 1, 1, true, lv, stream);
    };
    lay(
        P.f_lvl, [&](int t) { return P.tasks[4 * t + 2] - P.tasks[4 * t + 1]; },
        [&](int t, int pi, bool last, int lt, int* it) {
            if (pi >= 0) {
                const int p = P.tasks[4 * t + 1] + pi;
                it[0] = P.pairs[4 * p] * sc;
                it[1] = P.pairs[4 * p + 1] * sc;
                it[2] = P.pairs[4 * p + 2] * sc;
            }
            if (last)
                it[3] = P.tasks[4 * t] * sc | (lt << kStreamTeamShift) | ((P.tasks[4 * t + 3] & 3) == 1 ? kStreamRecip : 0) |
                        ((P.tasks[4 * t + 3] & kTaskAssign) ? kStreamAssign : 0) | kStreamEnd;
        },
        3, fpad, L.f_lv, L.f_lc, L.fs);
    lay(
        P.b_lvl, [&](int t) { return P.btasks[4 * t + 2] - P.btasks[4 * t + 1]; },
        [&](int t, int pi, bool last, int lt, int* it) {
            if (pi >= 0) {
                const int p = P.btasks[4 * t + 1] + pi;
                it[0] = P.bpairs[2 * p] * sc;
                it[1] = P.bpairs[2 * p + 1] * sc;
            }
            if (last) {
                it[2] = P.btasks[4 * t] * sc | (lt << kStreamTeamShift) | kStreamEnd;
                it[3] = P.btasks[4 * t + 3] * sc;
            }
        },
        2, bpad, L.b_lv, L.b_lc, L.bs);
    // lean streams end with nslot padding items: a lane reads item base + k * U + s of a level
    // without clamping s to the level's lanes (the waves of lanes s >= U skip the level), and
    // the levels past the last one read at the end of the stream
    if (lean)
        for (int i = 0; i < nslot; ++i) {
            L.fs.insert(L.fs.end(), fpad, fpad + 4);
            L.bs.insert(L.bs.end(), bpad, bpad + 4);
        }
    return L;
}

// ---------------------------------------------------------------------------
// Host re-pivot: threshold partial pivoting on J's numbers (see analysis.hpp).
// Left-looking LU of J(:, perm): column k is the sparse triangular solve L x = J(:, perm[k]) over
// the columns its nonzeros reach (depth-first search, reverse post-order), then the pivot row is
// chosen among the rows not pivotal yet. Only L is kept: the pivots are what this is for.
// ---------------------------------------------------------------------------
std::vector<int> threshold_pivot_rows(const JPattern& J, const std::vector<double>& Jx, const std::vector<int>& perm,
                                      double tau) {
    const int N = J.N;
    if ((int)Jx.size() != (int)J.Ji.size() || (int)perm.size() != N)
        throw std::runtime_error("pfgpu: re-pivot inputs of the wrong size");
    std::vector<int> step_of_row(N, -1), pivrow(N, -1);  // row -> step it pivots, step -> row
    std::vector<int> Lp(1, 0), Li;                          // L by steps, original row indices
    std::vector<double> Lx;
    std::vector<double> x(N, 0.0);
    // inz: row is a nonzero of the current column; vis: step reached by the current search
    std::vector<int> inz(N, -1), vis(N, -1), nz, topo, stack, pos;
    nz.reserve(N);
    for (int k = 0; k < N; ++k) {
        const int c = perm[k];
        nz.clear();
        topo.clear();
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) {
            const int r = J.Ji[q];
            x[r] += Jx[q];
            if (inz[r] != k) { inz[r] = k; nz.push_back(r); }
        }
        // the steps the column reaches: depth-first from the pivotal rows of the column through
        // the graph of L, each step finished (appended to topo) after every step it reaches
        const size_t n0 = nz.size();
        for (size_t t = 0; t < n0; ++t) {
            const int j0 = step_of_row[nz[t]];
            if (j0 < 0 || vis[j0] == k) continue;
            vis[j0] = k;
            stack.assign(1, j0);
            pos.assign(1, Lp[j0]);
            while (!stack.empty()) {
                const int j = stack.back();
                bool pushed = false;
                while (pos.back() < Lp[j + 1]) {
                    const int r = Li[pos.back()++];
                    if (inz[r] != k) { inz[r] = k; nz.push_back(r); }
                    const int js = step_of_row[r];
                    if (js >= 0 && vis[js] != k) {
                        vis[js] = k;
                        stack.push_back(js);
                        pos.push_back(Lp[js]);
                        pushed = true;
                        break;
                    }
                }
                if (!pushed) {
                    topo.push_back(j);
                    stack.pop_back();
                    pos.pop_back();
                }
            }
        }
        // the steps in reverse post-order: x -= L(:, j) * x[pivrow[j]]
        for (auto it = topo.rbegin(); it != topo.rend(); ++it) {
            const int j = *it;
            const double xj = x[pivrow[j]];
            if (xj == 0.0) continue;
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) x[Li[p]] -= Lx[p] * xj;
        }
        // pivot: the diagonal row c while it holds at least tau times the largest candidate
        int best = -1;
        double amax = 0.0;
        for (int r : nz)
            if (step_of_row[r] < 0 && std::fabs(x[r]) > amax) { amax = std::fabs(x[r]); best = r; }
        if (best < 0 || !(amax > 0.0) || !std::isfinite(amax)) {
            for (int r : nz) x[r] = 0.0;
            return {};
        }
        if (step_of_row[c] < 0 && x[c] != 0.0 && std::fabs(x[c]) >= tau * amax) best = c;
        const double piv = x[best];
        step_of_row[best] = k;
        pivrow[k] = best;
        for (int r : nz) {
            if (step_of_row[r] < 0 && x[r] != 0.0) {
                Li.push_back(r);
                Lx.push_back(x[r] / piv);
            }
            x[r] = 0.0;
        }
        Lp.push_back((int)Li.size());
    }
    std::vector<int> rowpos(N, -1);
    for (int k = 0; k < N; ++k) rowpos[pivrow[k]] = perm[k];
    return rowpos;
}

}  // namespace pfgpu
