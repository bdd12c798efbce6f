// Copyright (c) 2026. Licensed under MPL-2.0 (same terms as the reference it restates).
//
// ORACLE — test infrastructure only.
// A CPU FP64 restatement of lightsim2grid's Newton-Raphson power-flow path,
// used as the parity checker for the GPU engine and as the CPU baseline.
// Nothing in the product links or calls this file; only tests/, the smoke
// check and the CPU-baseline leg of the benchmark load it (through ctypes).
//
// What it restates (reference file:line, paths relative to the reference root):
//   * BaseNRAlgo<KLULinearSolver>::compute_pf       src/powerflow_algorithm/BaseNRAlgo.tpp:15-186
//   * BaseNRSingleSlackAlgo::compute_pf             src/powerflow_algorithm/BaseNRSingleSlackAlgo.tpp:12-183
//   * BaseAlgo::_evaluate_Fx (both forms)           src/powerflow_algorithm/BaseAlgo.cpp:30-118
//   * BaseAlgo::_check_for_convergence              src/powerflow_algorithm/BaseAlgo.cpp:120-129
//   * BaseAlgo::retrieve_pv_with_slack              src/powerflow_algorithm/BaseAlgo.hpp:201-216
//   * BaseNRAlgo::_dSbus_dV                          src/powerflow_algorithm/BaseNRAlgo.tpp:202-260
//   * J pattern (distributed slack)                  src/powerflow_algorithm/BaseNRAlgo.tpp:396-555
//   * J pattern (single slack)                       src/powerflow_algorithm/BaseNRSingleSlackAlgo.tpp:247-...
//   * value map + known-pattern fill                 src/powerflow_algorithm/BaseNRAlgo.tpp:562-688
//   * KLULinearSolver initialize/refactor/solve     src/linear_solvers/KLUSolver.cpp:24-69
//   * ContingencyAnalysis::compute loop              src/batch_algorithm/ContingencyAnalysis.cpp:166-233
//   * ContingencyAnalysis::check_invertible          src/batch_algorithm/ContingencyAnalysis.cpp:14-45
//   * TimeSeries::compute_Vs step loop               src/batch_algorithm/TimeSeries.cpp:80-103
//
// The linear solver follows the published KLU algorithm (Davis & Palamadai
// Natarajan, ACM TOMS 37(3), 2010, "Algorithm 907"): a fill-reducing symmetric
// ordering of A+A^T (here: exact minimum degree on the elimination graph,
// written for this oracle), a left-looking Gilbert-Peierls factorisation with
// threshold partial pivoting that prefers the diagonal, a refactor that reuses
// the pivot sequence and the L/U patterns, and an in-place solve. BTF is off
// and the pivot tolerance is 0 because the reference value-initialises
// klu_common (KLUSolver.cpp:18,29) instead of calling klu_defaults.
//
// Build: see oracle/Makefile (g++ -O3 -march=native -shared -fPIC, -pthread).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <queue>
#include <set>
#include <stdexcept>
#include <thread>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

// Error codes: identical numbering to the reference's ErrorType (src/Utils.hpp:47-56).
enum Err : int {
    NoError = 0, SingularMatrix = 1, TooManyIterations = 2, InifiniteValue = 3,
    SolverAnalyze = 4, SolverFactor = 5, SolverReFactor = 6, SolverSolve = 7,
    NotInitError = 8, LicenseError = 9
};

struct CscC {  // complex CSC (Ybus)
    int n = 0;
    const int* p = nullptr;
    const int* i = nullptr;
    const cplx* x = nullptr;
};

// ---------------------------------------------------------------------------
// Minimum-degree ordering on the symmetric pattern of A + A^T.
// Exact minimum degree on the explicit elimination graph; ties go to the
// lowest index. Quality is what matters here (fill decides the LU cost); the
// explicit graph keeps the code short and is fast on low-fill grid Jacobians.
// ---------------------------------------------------------------------------
static std::vector<int> min_degree_order(int n, const std::vector<int>& Ap, const std::vector<int>& Ai) {
    std::vector<std::vector<int>> adj(n);
    for (int j = 0; j < n; ++j)
        for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
            int i = Ai[q];
            if (i == j) continue;
            adj[i].push_back(j);
            adj[j].push_back(i);
        }
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::set<std::pair<int, int>> heap;
    std::vector<int> deg(n);
    for (int v = 0; v < n; ++v) { deg[v] = (int)adj[v].size(); heap.insert({deg[v], v}); }
    std::vector<char> dead(n, 0);
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> tmp;
    while (!heap.empty()) {
        int v = heap.begin()->second;
        heap.erase(heap.begin());
        dead[v] = 1;
        order.push_back(v);
        const std::vector<int> nb = adj[v];  // alive neighbours only (kept clean below)
        for (int u : nb) {
            // adj[u] <- (adj[u] U nb) \ {u, v}
            tmp.clear();
            std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(), std::back_inserter(tmp));
            auto& au = adj[u];
            au.clear();
            for (int w : tmp) if (w != u && w != v) au.push_back(w);
            if ((int)au.size() != deg[u]) {
                heap.erase({deg[u], u});
                deg[u] = (int)au.size();
                heap.insert({deg[u], u});
            }
        }
        adj[v].clear();
    }
    return order;
}

// ---------------------------------------------------------------------------
// KLU-contract sparse LU: analyze (ordering) -> factor (partial pivoting) ->
// refactor (same pivots, same patterns) -> solve (in place).
// Matrix A is CSC (n x n). We factor P*A*Q = L*U, Q = column order from the
// ordering, P = row pivots chosen during factor (initial row order = Q).
// ---------------------------------------------------------------------------
class KluLikeLU {
public:
    double pivot_tol = 0.0;  // zeroed klu_common => tol 0 (diagonal preferred whenever nonzero)

    int analyze(int n, const std::vector<int>& Ap, const std::vector<int>& Ai) {
        n_ = n;
        Q_ = min_degree_order(n, Ap, Ai);
        return NoError;
    }

    // Left-looking Gilbert-Peierls with threshold partial pivoting.
    int factor(const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax) {
        const int n = n_;
        // Work in the "ordered" index space: row r of the permuted matrix is
        // original row Qinv-mapped; we keep a row permutation pinv (orig row -> pivot step).
        std::vector<int> qinv(n);
        for (int k = 0; k < n; ++k) qinv[Q_[k]] = k;
        pinv_.assign(n, -1);
        P_.assign(n, -1);
        Lp_.assign(1, 0); Li_.clear(); Lx_.clear();
        Up_.assign(1, 0); Ui_.clear(); Ux_.clear();
        std::vector<double> x(n, 0.0);
        std::vector<int> xi(2 * n), pstack(n), mark(n, -1);
        for (int k = 0; k < n; ++k) {
            const int col = Q_[k];
            // symbolic: reach of A(:,col) in graph of L (columns indexed by pivot step)
            int top = n;
            for (int q = Ap[col]; q < Ap[col + 1]; ++q) {
                int r = Ai[q];
                if (mark[r] == k) continue;
                top = dfs(r, k, top, xi, pstack, mark);
            }
            // numeric sparse triangular solve
            for (int t = top; t < n; ++t) x[xi[t]] = 0.0;
            for (int q = Ap[col]; q < Ap[col + 1]; ++q) x[Ai[q]] += Ax[q];
            for (int t = top; t < n; ++t) {
                int r = xi[t];
                int s = pinv_[r];
                if (s < 0) continue;  // not yet pivotal -> part of L candidate
                double xr = x[r];
                for (int q = Lp_[s]; q < Lp_[s + 1]; ++q) x[Li_[q]] -= Lx_[q] * xr;
            }
            // pivot choice: prefer the "diagonal" = original row Q_[k]
            int ipiv = -1; double amax = -1.0;
            for (int t = top; t < n; ++t) {
                int r = xi[t];
                if (pinv_[r] < 0) { double a = std::fabs(x[r]); if (a > amax) { amax = a; ipiv = r; } }
            }
            if (ipiv < 0 || amax <= 0.0) return SingularMatrix;
            const int diag = Q_[k];
            if (pinv_[diag] < 0 && mark[diag] == k) {
                double ad = std::fabs(x[diag]);
                if (ad > 0.0 && ad >= pivot_tol * amax) ipiv = diag;
            }
            const double piv = x[ipiv];
            // store U(:,k): pivotal rows (in step index) and the pivot itself last
            for (int t = top; t < n; ++t) {
                int r = xi[t];
                int s = pinv_[r];
                if (s >= 0) { Ui_.push_back(s); Ux_.push_back(x[r]); }
            }
            Ui_.push_back(k); Ux_.push_back(piv);
            Up_.push_back((int)Ui_.size());
            pinv_[ipiv] = k; P_[k] = ipiv;
            // store L(:,k) (original row ids, scaled)
            for (int t = top; t < n; ++t) {
                int r = xi[t];
                if (pinv_[r] < 0) { Li_.push_back(r); Lx_.push_back(x[r] / piv); }
            }
            Lp_.push_back((int)Li_.size());
            for (int t = top; t < n; ++t) x[xi[t]] = 0.0;
        }
        // Convert L row ids to pivot-step ids for the solve and refactor.
        for (auto& r : Li_) r = pinv_[r];
        factored_ = true;
        return NoError;
    }

    // Static-pivot refactor: numeric values on the stored patterns (klu_refactor).
    int refactor(const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax) {
        if (!factored_) return NotInitError;
        const int n = n_;
        std::vector<double> x(n, 0.0);  // indexed by pivot step
        for (int k = 0; k < n; ++k) {
            const int col = Q_[k];
            for (int q = Ap[col]; q < Ap[col + 1]; ++q) x[pinv_[Ai[q]]] += Ax[q];
            // U(:,k) rows are stored in topological order of the reach (last = pivot)
            const int u0 = Up_[k], u1 = Up_[k + 1] - 1;
            for (int q = u0; q < u1; ++q) {
                const int s = Ui_[q];
                const double xs = x[s];
                Ux_[q] = xs;
                for (int l = Lp_[s]; l < Lp_[s + 1]; ++l) x[Li_[l]] -= Lx_[l] * xs;
            }
            const double piv = x[k];
            Ux_[u1] = piv;
            if (piv == 0.0 || !std::isfinite(piv)) { clear_x(x, k); return SolverReFactor; }
            for (int l = Lp_[k]; l < Lp_[k + 1]; ++l) Lx_[l] = x[Li_[l]] / piv;
            clear_x(x, k);
        }
        return NoError;
    }

    // Solve A x = b in place.
    int solve(std::vector<double>& b) const {
        if (!factored_) return NotInitError;
        const int n = n_;
        std::vector<double> y(n);
        for (int k = 0; k < n; ++k) y[k] = b[P_[k]];
        for (int k = 0; k < n; ++k) {
            const double yk = y[k];
            for (int l = Lp_[k]; l < Lp_[k + 1]; ++l) y[Li_[l]] -= Lx_[l] * yk;
        }
        for (int k = n - 1; k >= 0; --k) {
            const int u1 = Up_[k + 1] - 1;
            const double xk = y[k] / Ux_[u1];
            y[k] = xk;
            for (int q = Up_[k]; q < u1; ++q) y[Ui_[q]] -= Ux_[q] * xk;
        }
        for (int k = 0; k < n; ++k) b[Q_[k]] = y[k];
        return NoError;
    }

    void reset() { factored_ = false; n_ = 0; }
    long nnz_lu() const { return (long)Lx_.size() + (long)Ux_.size(); }
    bool factored() const { return factored_; }

    // accessors for tests (L unit-lower without the diagonal, U with diagonal last per column)
    const std::vector<int>& P() const { return P_; }
    const std::vector<int>& Q() const { return Q_; }
    const std::vector<int>& Lp() const { return Lp_; }
    const std::vector<int>& Li() const { return Li_; }
    const std::vector<double>& Lx() const { return Lx_; }
    const std::vector<int>& Up() const { return Up_; }
    const std::vector<int>& Ui() const { return Ui_; }
    const std::vector<double>& Ux() const { return Ux_; }

private:
    void clear_x(std::vector<double>& x, int k) const {
        for (int q = Up_[k]; q < Up_[k + 1]; ++q) x[Ui_[q]] = 0.0;
        for (int l = Lp_[k]; l < Lp_[k + 1]; ++l) x[Li_[l]] = 0.0;
    }
    // iterative DFS from original row r through the graph of L (columns = pivot steps)
    int dfs(int r0, int k, int top, std::vector<int>& xi, std::vector<int>& pstack, std::vector<int>& mark) {
        int head = 0;
        xi[0] = r0;
        while (head >= 0) {
            int r = xi[head];
            int s = pinv_[r];
            if (mark[r] != k) {
                mark[r] = k;
                pstack[head] = (s < 0) ? 0 : Lp_[s];
            }
            bool done = true;
            if (s >= 0) {
                int end = Lp_[s + 1];
                for (int q = pstack[head]; q < end; ++q) {
                    int rr = Li_[q];  // original row ids during factor
                    if (mark[rr] == k) continue;
                    pstack[head] = q + 1;
                    xi[++head] = rr;
                    done = false;
                    break;
                }
            }
            if (done) {
                --head;
                xi[--top] = r;
            }
        }
        return top;
    }

    int n_ = 0;
    bool factored_ = false;
    std::vector<int> Q_, P_, pinv_;
    std::vector<int> Lp_, Li_, Up_, Ui_;
    std::vector<double> Lx_, Ux_;
};

// ---------------------------------------------------------------------------
// Newton-Raphson solver (distributed slack, BaseNRAlgo; or single slack,
// BaseNRSingleSlackAlgo). Keeps J pattern/value map and the LU across calls
// until reset(), like the reference does under tell_none_changed.
// ---------------------------------------------------------------------------
struct Timers {
    double Fx = 0, solve = 0, refactor = 0, initialize = 0, check = 0, dSbus = 0, fillJ = 0, VaVm = 0, pre_proc = 0, total = 0;
};

class NRSolver {
public:
    explicit NRSolver(bool single_slack) : single_(single_slack) {}

    void reset() {
        Jp_.clear(); Ji_.clear(); Jx_.clear();
        vmap_.clear();
        lu_.reset();
        need_factorize_ = true;
        nJ_ = -1;
    }

    bool compute_pf(const CscC& Y, std::vector<cplx>& V, const cplx* Sbus,
                    const int* slack_ids, int nslack, const double* slack_w,
                    const int* pv, int npv_in, const int* pq, int npq_in,
                    int max_iter, double tol) {
        using clk = std::chrono::steady_clock;
        auto t_all = clk::now();
        tm_ = Timers();
        err_ = NoError;
        const int n = Y.n;
        if ((int)V.size() != n) throw std::runtime_error("oracle compute_pf: V size mismatch");
        npq_cur_ = npq_in;
        // my_pv = slack_ids[1:] ++ pv (BaseAlgo.hpp:201-216); single slack uses pv as is
        std::vector<int> my_pv;
        if (!single_) for (int s = 1; s < nslack; ++s) my_pv.push_back(slack_ids[s]);
        for (int s = 0; s < npv_in; ++s) my_pv.push_back(pv[s]);
        const int npv = (int)my_pv.size(), npq = npq_in;
        std::vector<int> pvpq(my_pv);
        pvpq.insert(pvpq.end(), pq, pq + npq);
        const int npvpq = (int)pvpq.size();
        std::vector<int> pvpq_inv(n, -1), pq_inv(n, -1);
        for (int k = 0; k < npvpq; ++k) pvpq_inv[pvpq[k]] = k;
        for (int k = 0; k < npq; ++k) pq_inv[pq[k]] = k;
        const int slack0 = slack_ids[0];
        double slack_absorbed = 0.0;
        for (int b = 0; b < n; ++b) slack_absorbed += Sbus[b].real();
        Vc_ = V;
        Vm_.resize(n); Va_.resize(n);
        for (int b = 0; b < n; ++b) { Vm_[b] = std::abs(Vc_[b]); Va_[b] = std::arg(Vc_[b]); }

        std::vector<double> F;
        eval_F(Y, Vc_, Sbus, slack0, slack_absorbed, slack_w, my_pv, pq, F);
        bool conv = inf_norm(F) < tol;
        nr_iter_ = 0;
        bool res = true;
        const int off = single_ ? 0 : 1;
        while (!conv && nr_iter_ < max_iter) {
            ++nr_iter_;
            fill_J(Y, slack0, slack_w, pq, pvpq, pq_inv, pvpq_inv);
            auto t0 = clk::now();
            if (need_factorize_) {
                int e = lu_.analyze(nJ_, Jp_, Ji_);
                if (e == NoError) e = lu_.factor(Jp_, Ji_, Jx_);
                need_factorize_ = false;
                tm_.initialize += sec(clk::now() - t0);
                if (e != NoError) { err_ = (e == SingularMatrix) ? SolverFactor : e; res = false; break; }
            } else {
                int e = lu_.refactor(Jp_, Ji_, Jx_);
                tm_.refactor += sec(clk::now() - t0);
                if (e != NoError) { err_ = SolverReFactor; res = false; break; }
            }
            t0 = clk::now();
            int e = lu_.solve(F);
            tm_.solve += sec(clk::now() - t0);
            if (e != NoError) { err_ = SolverSolve; res = false; break; }
            t0 = clk::now();
            if (single_) {
                // BaseNRSingleSlackAlgo.tpp:131-139 re-derives Vm/Va from V before the update
                for (int b = 0; b < n; ++b) { Vm_[b] = std::abs(Vc_[b]); Va_[b] = std::arg(Vc_[b]); }
            } else {
                slack_absorbed -= F[0];
            }
            for (int k = 0; k < npv; ++k) Va_[my_pv[k]] -= F[off + k];
            for (int k = 0; k < npq; ++k) {
                Va_[pq[k]] -= F[off + npv + k];
                Vm_[pq[k]] -= F[off + npv + npq + k];
            }
            double vmin = 1e300;
            for (int b = 0; b < n; ++b) {
                Vc_[b] = Vm_[b] * cplx(std::cos(Va_[b]), std::sin(Va_[b]));
                vmin = std::min(vmin, Vm_[b]);
            }
            if (vmin < 0.0)
                for (int b = 0; b < n; ++b) { Vm_[b] = std::abs(Vc_[b]); Va_[b] = std::arg(Vc_[b]); }
            tm_.VaVm += sec(clk::now() - t0);
            eval_F(Y, Vc_, Sbus, slack0, slack_absorbed, slack_w, my_pv, pq, F);
            bool finite = true;
            for (double f : F) if (!std::isfinite(f)) { finite = false; break; }
            if (!finite) { err_ = InifiniteValue; break; }
            conv = inf_norm(F) < tol;
        }
        if (!conv) { if (err_ == NoError) err_ = TooManyIterations; res = false; }
        for (int b = 0; b < n; ++b) { Vm_[b] = std::abs(Vc_[b]); Va_[b] = std::arg(Vc_[b]); }
        converged_ = res;
        tm_.total = sec(clk::now() - t_all);
        return res;
    }

    const std::vector<cplx>& V() const { return Vc_; }
    const std::vector<double>& Va() const { return Va_; }
    const std::vector<double>& Vm() const { return Vm_; }
    int nb_iter() const { return nr_iter_; }
    int error() const { return err_; }
    bool converged() const { return converged_; }
    const Timers& timers() const { return tm_; }
    int nJ() const { return nJ_; }
    const std::vector<int>& Jp() const { return Jp_; }
    const std::vector<int>& Ji() const { return Ji_; }
    const std::vector<double>& Jx() const { return Jx_; }
    const KluLikeLU& lu() const { return lu_; }

private:
    static double sec(std::chrono::steady_clock::duration d) { return std::chrono::duration<double>(d).count(); }
    static double inf_norm(const std::vector<double>& F) {
        double m = 0.0;
        for (double f : F) m = std::max(m, std::fabs(f));
        return m;
    }

    // BaseAlgo::_evaluate_Fx (BaseAlgo.cpp:30-54 single, :56-118 distributed)
    void eval_F(const CscC& Y, const std::vector<cplx>& V, const cplx* S, int slack0, double slack_absorbed,
                const double* w, const std::vector<int>& pv, const int* pq, std::vector<double>& F) {
        auto t0 = std::chrono::steady_clock::now();
        const int n = Y.n;
        Ibus_.assign(n, cplx(0, 0));
        for (int c = 0; c < n; ++c)
            for (int q = Y.p[c]; q < Y.p[c + 1]; ++q) Ibus_[Y.i[q]] += Y.x[q] * V[c];
        mis_.resize(n);
        for (int b = 0; b < n; ++b) {
            mis_[b] = V[b] * std::conj(Ibus_[b]) - S[b];
            if (!single_) mis_[b] += slack_absorbed * w[b];
        }
        const int npv = (int)pv.size();
        const int npq = npq_cur_;
        const int off = single_ ? 0 : 1;
        F.assign(off + npv + 2 * npq, 0.0);
        for (int k = 0; k < npv; ++k) F[off + k] = mis_[pv[k]].real();
        for (int k = 0; k < npq; ++k) {
            F[off + npv + k] = mis_[pq[k]].real();
            F[off + npv + npq + k] = mis_[pq[k]].imag();
        }
        if (!single_) F[0] = mis_[slack0].real();
        tm_.Fx += sec(std::chrono::steady_clock::now() - t0);
    }

    // BaseNRAlgo::_dSbus_dV (BaseNRAlgo.tpp:202-260), values in Ybus CSC order
    void dSbus_dV(const CscC& Y, const std::vector<cplx>& V) {
        auto t0 = std::chrono::steady_clock::now();
        const int n = Y.n;
        const int nnz = Y.p[n];
        dVm_.resize(nnz); dVa_.resize(nnz);
        std::vector<cplx> Vnorm(n), Ib(n, cplx(0, 0));
        for (int b = 0; b < n; ++b) Vnorm[b] = V[b] / std::abs(V[b]);
        for (int c = 0; c < n; ++c)
            for (int q = Y.p[c]; q < Y.p[c + 1]; ++q) Ib[Y.i[q]] += Y.x[q] * V[c];
        for (int c = 0; c < n; ++c)
            for (int q = Y.p[c]; q < Y.p[c + 1]; ++q) {
                const int r = Y.i[q];
                cplx dvm = std::conj(Y.x[q] * Vnorm[c]) * V[r];
                cplx dva = Y.x[q] * V[c];
                if (r == c) {
                    dvm += std::conj(Ib[r]) * Vnorm[r];
                    dva -= Ib[r];
                }
                dva = std::conj(-dva) * (cplx(0, 1) * V[r]);
                dVm_[q] = dvm; dVa_[q] = dva;
            }
        tm_.dSbus += sec(std::chrono::steady_clock::now() - t0);
    }

    static int find_in_col(const CscC& Y, int r, int c) {
        for (int q = Y.p[c]; q < Y.p[c + 1]; ++q) if (Y.i[q] == r) return q;
        return -1;
    }

    // Pattern build (first call) + value map + known-pattern fill.
    // Distributed-slack layout: BaseNRAlgo.tpp:396-555 (pattern) and :562-635 (map).
    // Single-slack layout: BaseNRSingleSlackAlgo.tpp (same blocks without row/col 0).
    void fill_J(const CscC& Y, int slack0, const double* w, const int* pq,
                const std::vector<int>& pvpq, const std::vector<int>& pq_inv, const std::vector<int>& pvpq_inv) {
        dSbus_dV(Y, Vc_);
        auto t0 = std::chrono::steady_clock::now();
        const int n = Y.n;
        const int npvpq = (int)pvpq.size();
        const int npq = npq_cur_;
        const int off = single_ ? 0 : 1;
        const int size_j = npvpq + npq + off;
        if (nJ_ != size_j) {
            // triplets (row, col, ybus-index, selector) ; selector: 0 dVa.re 1 dVa.im 2 dVm.re 3 dVm.im, -1 const
            struct T { int r, c, k, sel; double v; };
            std::vector<T> trip;
            for (int cj = 0; cj < npvpq; ++cj) {
                const int bc = pvpq[cj];
                for (int q = Y.p[bc]; q < Y.p[bc + 1]; ++q) {
                    const int br = Y.i[q];
                    if (pvpq_inv[br] >= 0) trip.push_back({pvpq_inv[br] + off, cj + off, q, 0, 0});
                }
                for (int q = Y.p[bc]; q < Y.p[bc + 1]; ++q) {
                    const int br = Y.i[q];
                    if (pq_inv[br] >= 0) trip.push_back({pq_inv[br] + npvpq + off, cj + off, q, 1, 0});
                }
            }
            for (int cj = 0; cj < npq; ++cj) {
                const int bc = pq[cj];
                for (int q = Y.p[bc]; q < Y.p[bc + 1]; ++q) {
                    const int br = Y.i[q];
                    if (pvpq_inv[br] >= 0) trip.push_back({pvpq_inv[br] + off, cj + npvpq + off, q, 2, 0});
                }
                for (int q = Y.p[bc]; q < Y.p[bc + 1]; ++q) {
                    const int br = Y.i[q];
                    if (pq_inv[br] >= 0) trip.push_back({pq_inv[br] + npvpq + off, cj + npvpq + off, q, 3, 0});
                }
            }
            if (!single_) {
                // slack row 0: dS_dVa[slack0, pvpq].real and dS_dVm[slack0, pq].real
                for (int b = 0; b < n; ++b) {
                    if (pvpq_inv[b] < 0) continue;
                    int q = find_in_col(Y, slack0, b);
                    if (q >= 0) trip.push_back({0, pvpq_inv[b] + 1, q, 0, 0});
                }
                for (int b = 0; b < n; ++b) {
                    if (pq_inv[b] < 0) continue;
                    int q = find_in_col(Y, slack0, b);
                    if (q >= 0) trip.push_back({0, pq_inv[b] + npvpq + 1, q, 2, 0});
                }
                // column 0: slack weights (constant)
                trip.push_back({0, 0, -1, -1, w[slack0]});
                for (int k = 0; k < npvpq; ++k) {
                    const double sw = w[pvpq[k]];
                    if (std::fabs(sw) > 1e-7) trip.push_back({k + 1, 0, -1, -1, sw});
                }
            }
            // CSC assembly (duplicates cannot occur: each (row, col) comes from one Ybus entry)
            std::stable_sort(trip.begin(), trip.end(), [](const T& a, const T& b) {
                return a.c != b.c ? a.c < b.c : a.r < b.r;
            });
            nJ_ = size_j;
            Jp_.assign(size_j + 1, 0);
            Ji_.resize(trip.size()); Jx_.assign(trip.size(), 0.0);
            vmap_.resize(trip.size());
            for (size_t t = 0; t < trip.size(); ++t) {
                Jp_[trip[t].c + 1]++;
                Ji_[t] = trip[t].r;
                vmap_[t] = {trip[t].k, trip[t].sel};
                if (trip[t].sel < 0) Jx_[t] = trip[t].v;
            }
            for (int c = 0; c < size_j; ++c) Jp_[c + 1] += Jp_[c];
        }
        // known-pattern fill (BaseNRAlgo.tpp:637-688)
        for (size_t t = 0; t < vmap_.size(); ++t) {
            const int k = vmap_[t].first, sel = vmap_[t].second;
            if (sel < 0) continue;
            const cplx z = (sel < 2) ? dVa_[k] : dVm_[k];
            Jx_[t] = (sel & 1) ? z.imag() : z.real();
        }
        tm_.fillJ += sec(std::chrono::steady_clock::now() - t0);
    }

    bool single_;
    bool need_factorize_ = true;
    int nJ_ = -1;
    int npq_cur_ = 0;
    std::vector<int> Jp_, Ji_;
    std::vector<double> Jx_;
    std::vector<std::pair<int, int>> vmap_;
    std::vector<cplx> dVm_, dVa_, Ibus_, mis_, Vc_;
    std::vector<double> Vm_, Va_;
    KluLikeLU lu_;
    int nr_iter_ = 0;
    int err_ = NoError;
    bool converged_ = false;
    Timers tm_;
};

// BFS connectivity with |y| > 1e-8 (ContingencyAnalysis.cpp:14-45)
static bool check_invertible(const CscC& Y, const std::vector<cplx>& Yx) {
    const int n = Y.n;
    std::vector<char> visited(n, 0), added(n, 0);
    std::queue<int> nb;
    int col = 0;
    while (true) {
        visited[col] = 1;
        for (int q = Y.p[col]; q < Y.p[col + 1]; ++q) {
            const int r = Y.i[q];
            if (!visited[r] && !added[r] && std::abs(Yx[q]) > 1e-8) { nb.push(r); added[r] = 1; }
        }
        if (nb.empty()) break;
        col = nb.front(); nb.pop();
    }
    for (char v : visited) if (!v) return false;
    return true;
}

}  // namespace oracle

using namespace oracle;

// ---------------------------------------------------------------------------
// C API (ctypes). Complex arrays are interleaved (re, im) doubles.
// ---------------------------------------------------------------------------
extern "C" {

void* oracle_solver_create(int single_slack) { return new NRSolver(single_slack != 0); }
void oracle_solver_destroy(void* h) { delete static_cast<NRSolver*>(h); }
void oracle_solver_reset(void* h) { static_cast<NRSolver*>(h)->reset(); }

// Returns 1 if converged. V is in/out.
int oracle_solver_compute_pf(void* h, int n, const int* Yp, const int* Yi, const double* Yx,
                             double* V, const double* Sbus, const int* slack_ids, int nslack,
                             const double* slack_w, const int* pv, int npv, const int* pq, int npq,
                             int max_iter, double tol) {
    NRSolver* s = static_cast<NRSolver*>(h);
    CscC Y{n, Yp, Yi, reinterpret_cast<const cplx*>(Yx)};
    std::vector<cplx> Vv(reinterpret_cast<cplx*>(V), reinterpret_cast<cplx*>(V) + n);
    bool ok = s->compute_pf(Y, Vv, reinterpret_cast<const cplx*>(Sbus), slack_ids, nslack, slack_w, pv, npv, pq, npq,
                            max_iter, tol);
    const auto& Vo = s->V();
    for (int b = 0; b < n; ++b) { V[2 * b] = Vo[b].real(); V[2 * b + 1] = Vo[b].imag(); }
    return ok ? 1 : 0;
}

int oracle_solver_nb_iter(void* h) { return static_cast<NRSolver*>(h)->nb_iter(); }
int oracle_solver_error(void* h) { return static_cast<NRSolver*>(h)->error(); }
int oracle_solver_J_dim(void* h) { return static_cast<NRSolver*>(h)->nJ(); }
int oracle_solver_J_nnz(void* h) { auto s = static_cast<NRSolver*>(h); return s->nJ() < 0 ? 0 : (int)s->Ji().size(); }
void oracle_solver_get_J(void* h, int* Jp, int* Ji, double* Jx) {
    auto s = static_cast<NRSolver*>(h);
    std::copy(s->Jp().begin(), s->Jp().end(), Jp);
    std::copy(s->Ji().begin(), s->Ji().end(), Ji);
    std::copy(s->Jx().begin(), s->Jx().end(), Jx);
}
long oracle_solver_nnz_lu(void* h) { return static_cast<NRSolver*>(h)->lu().nnz_lu(); }
void oracle_solver_get_timers(void* h, double* out10) {
    const auto& t = static_cast<NRSolver*>(h)->timers();
    double v[10] = {t.Fx, t.solve, t.refactor, t.initialize, t.check, t.dSbus, t.fillJ, t.VaVm, t.pre_proc, t.total};
    std::memcpy(out10, v, sizeof(v));
}

// Dense export of the current LU for reconstruction tests: fills P (n), Q (n),
// L (n*n, row-major, unit diagonal implied) and U (n*n). Only for small n.
int oracle_solver_get_lu_dense(void* h, int* P, int* Q, double* L, double* U) {
    auto s = static_cast<NRSolver*>(h);
    const auto& lu = s->lu();
    if (!lu.factored()) return -1;
    const int n = s->nJ();
    std::copy(lu.P().begin(), lu.P().end(), P);
    std::copy(lu.Q().begin(), lu.Q().end(), Q);
    std::fill(L, L + (size_t)n * n, 0.0);
    std::fill(U, U + (size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        L[(size_t)k * n + k] = 1.0;
        for (int q = lu.Lp()[k]; q < lu.Lp()[k + 1]; ++q) L[(size_t)lu.Li()[q] * n + k] = lu.Lx()[q];
        for (int q = lu.Up()[k]; q < lu.Up()[k + 1]; ++q) U[(size_t)lu.Ui()[q] * n + k] = lu.Ux()[q];
    }
    return n;
}

// Stand-alone KLU-contract LU on a real CSC matrix: factor then solve b in place.
// reps > 1 repeats refactor+solve (timing). Returns an ErrorType code.
int oracle_lu_solve(int n, const int* Ap, const int* Ai, const double* Ax, double* b, int reps) {
    KluLikeLU lu;
    std::vector<int> p(Ap, Ap + n + 1), i(Ai, Ai + Ap[n]);
    std::vector<double> x(Ax, Ax + Ap[n]);
    int e = lu.analyze(n, p, i);
    if (e) return e;
    e = lu.factor(p, i, x);
    if (e) return e;
    for (int r = 1; r < reps; ++r) { e = lu.refactor(p, i, x); if (e) return e; }
    std::vector<double> bb(b, b + n);
    e = lu.solve(bb);
    std::copy(bb.begin(), bb.end(), b);
    return e;
}

// DC power flow, BaseDCAlgo<KLU>::compute_pf (src/powerflow_algorithm/BaseDCAlgo.tpp:15-179):
// my_pv = slack_ids[1:] ++ pv (BaseAlgo.hpp:201-216); the buses in neither my_pv nor pq are the
// slack buses (BaseAlgo.cpp:143-181); the others are numbered in bus order (fill_mat_bus_id,
// :196-205); B = Re(Ybus) and P = Re(Sbus) without the slack rows / columns (:207-233, :93-101);
// B theta = P is factored (analyze + factor) and solved; theta must be finite with |theta|inf <
// 1e6, else SolverSolve (:131-147); Va = theta (0 at the slack buses) + arg(V[first slack]),
// Vm = |V|, V = Vm exp(j Va) (:154-177). V is in/out; returns the ErrorType code (0 converged).
int oracle_dc(int n, const int* Yp, const int* Yi, const double* Yx, double* V, const double* Sbus,
              const int* slack_ids, int nslack, const int* pv, int npv, const int* pq, int npq) {
    std::vector<char> in_pv_pq(n, 0);
    for (int s = 1; s < nslack; ++s) in_pv_pq[slack_ids[s]] = 1;
    for (int k = 0; k < npv; ++k) in_pv_pq[pv[k]] = 1;
    for (int k = 0; k < npq; ++k) in_pv_pq[pq[k]] = 1;
    std::vector<int> id(n, -1), slack;
    int m = 0;
    for (int b = 0; b < n; ++b) {
        if (in_pv_pq[b]) id[b] = m++;
        else slack.push_back(b);
    }
    if (slack.empty()) return SolverAnalyze;
    std::vector<int> Bp(m + 1, 0), Bi;
    std::vector<double> Bx;
    for (int c = 0; c < n; ++c) {
        if (id[c] < 0) continue;
        for (int q = Yp[c]; q < Yp[c + 1]; ++q) {
            if (id[Yi[q]] < 0) continue;
            Bi.push_back(id[Yi[q]]);
            Bx.push_back(Yx[2 * q]);
        }
        Bp[id[c] + 1] = (int)Bi.size();
    }
    std::vector<double> theta(m);
    for (int b = 0; b < n; ++b)
        if (id[b] >= 0) theta[id[b]] = Sbus[2 * b];
    KluLikeLU lu;
    int e = lu.analyze(m, Bp, Bi);
    if (e == NoError) e = lu.factor(Bp, Bi, Bx);
    if (e != NoError) return e == SingularMatrix ? SolverFactor : e;
    if (lu.solve(theta) != NoError) return SolverSolve;
    double tmax = 0.0;
    bool finite = true;
    for (double t : theta) {
        finite = finite && std::isfinite(t);
        tmax = std::max(tmax, std::fabs(t));
    }
    if (!finite || tmax >= 1e6) return SolverSolve;
    const double va_s = std::atan2(V[2 * slack[0] + 1], V[2 * slack[0]]);
    for (int b = 0; b < n; ++b) {
        const double vm = std::hypot(V[2 * b], V[2 * b + 1]);
        const double va = (id[b] >= 0 ? theta[id[b]] : 0.0) + va_s;
        V[2 * b] = vm * std::cos(va);
        V[2 * b + 1] = vm * std::sin(va);
    }
    return NoError;
}

// Fast-decoupled power flow (BaseFDPFAlgo<KLU, XB_BX>::compute_pf, BaseFDPFAlgo.tpp:12-160,
// has_converged / evaluate_mismatch BaseFDPFAlgo.hpp:64-164). grid_Bp / grid_Bpp are the n x n
// matrices of GridModel::fillBp_Bpp (CSC); the solver keeps Bp = grid_Bp[pvpq, pvpq] and
// Bpp = grid_Bpp[pq, pq] (fill_sparse_matrices, BaseFDPFAlgo.tpp:164-214) with my_pv = slack_ids[1:]
// followed by pv (retrieve_pv_with_slack, BaseAlgo.hpp:201-216), factors both once and alternates
// the P half-step (Va[pvpq] -= Bp^-1 P) and the Q half-step (Vm[pq] -= Bpp^-1 Q), P and Q being the
// mismatch divided by Vm, with a convergence test after each half. slack_absorbed stays Re(sum Sbus).
// Returns the ErrorType; V in/out, *iters = P half-steps done.
int oracle_fdpf(int n, const int* Yp, const int* Yi, const double* Yx, double* V, const double* Sbus,
                const int* slack_ids, int nslack, const double* slack_w, const int* pv, int npv, const int* pq, int npq,
                const int* Bp_p, const int* Bp_i, const double* Bp_x, const int* Bpp_p, const int* Bpp_i,
                const double* Bpp_x, int max_iter, double tol, int* iters) {
    std::vector<int> pvpq;
    for (int s = 1; s < nslack; ++s) pvpq.push_back(slack_ids[s]);
    for (int k = 0; k < npv; ++k) pvpq.push_back(pv[k]);
    const int npvpq_ = (int)pvpq.size();
    for (int k = 0; k < npq; ++k) pvpq.push_back(pq[k]);
    const int nvv = (int)pvpq.size();
    std::vector<int> pvpq_inv(n, -1), pq_inv(n, -1);
    for (int k = 0; k < nvv; ++k) pvpq_inv[pvpq[k]] = k;
    for (int k = 0; k < npq; ++k) pq_inv[pq[k]] = k;
    // sub-matrices in CSC with sorted rows, duplicates summed (setFromTriplets)
    auto sub = [&](const int* Ap, const int* Ai, const double* Ax, const std::vector<int>& inv, int m,
                   std::vector<int>& p, std::vector<int>& i, std::vector<double>& x) {
        std::vector<std::vector<std::pair<int, double>>> cols(m);
        for (int c = 0; c < n; ++c) {
            if (inv[c] < 0) continue;
            for (int q = Ap[c]; q < Ap[c + 1]; ++q)
                if (inv[Ai[q]] >= 0) cols[inv[c]].push_back({inv[Ai[q]], Ax[q]});
        }
        p.assign(m + 1, 0);
        i.clear();
        x.clear();
        for (int c = 0; c < m; ++c) {
            auto& v = cols[c];
            std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            for (size_t k = 0; k < v.size(); ++k) {
                if (!i.empty() && (int)i.size() > p[c] && i.back() == v[k].first) x.back() += v[k].second;
                else { i.push_back(v[k].first); x.push_back(v[k].second); }
            }
            p[c + 1] = (int)i.size();
        }
    };
    std::vector<int> Pp, Pi, Qp, Qi;
    std::vector<double> Px, Qx;
    sub(Bp_p, Bp_i, Bp_x, pvpq_inv, nvv, Pp, Pi, Px);
    sub(Bpp_p, Bpp_i, Bpp_x, pq_inv, npq, Qp, Qi, Qx);
    (void)npvpq_;
    const cplx* Yc = reinterpret_cast<const cplx*>(Yx);
    const cplx* S = reinterpret_cast<const cplx*>(Sbus);
    std::vector<cplx> Vc(reinterpret_cast<cplx*>(V), reinterpret_cast<cplx*>(V) + n);
    std::vector<double> Vm(n), Va(n);
    for (int b = 0; b < n; ++b) { Vm[b] = std::abs(Vc[b]); Va[b] = std::arg(Vc[b]); }
    double slack_absorbed = 0.0;
    for (int b = 0; b < n; ++b) slack_absorbed += S[b].real();
    std::vector<double> p(nvv), q(npq);
    int err = NoError;
    // mis = (V conj(Ybus V) - Sbus + slack_absorbed w) / Vm; P = mis[pvpq].re, Q = mis[pq].im
    auto mismatch = [&]() -> bool {
        std::vector<cplx> I(n, cplx(0, 0));
        for (int c = 0; c < n; ++c)
            for (int k = Yp[c]; k < Yp[c + 1]; ++k) I[Yi[k]] += Yc[k] * Vc[c];
        std::vector<cplx> mis(n);
        bool finite = true;
        for (int b = 0; b < n; ++b) {
            mis[b] = (Vc[b] * std::conj(I[b]) - S[b] + slack_absorbed * slack_w[b]) / Vm[b];
            finite = finite && std::isfinite(mis[b].real()) && std::isfinite(mis[b].imag());
        }
        if (!finite) return false;
        for (int k = 0; k < nvv; ++k) p[k] = mis[pvpq[k]].real();
        for (int k = 0; k < npq; ++k) q[k] = mis[pq[k]].imag();
        return true;
    };
    auto check = [&]() -> bool {
        double np = 0.0, nq = 0.0;
        for (double v : p) np = std::max(np, std::fabs(v));
        for (double v : q) nq = std::max(nq, std::fabs(v));
        return np < tol && nq < tol;
    };
    // has_converged: V = Vm exp(j Va) (tmp_va), Vm and Va taken back from V, then the mismatch
    auto has_converged = [&](const std::vector<cplx>& tmp_va) -> bool {
        for (int b = 0; b < n; ++b) {
            Vc[b] = Vm[b] * tmp_va[b];
            Vm[b] = std::abs(Vc[b]);
            Va[b] = std::arg(Vc[b]);
        }
        if (!mismatch()) {
            err = InifiniteValue;
            return false;
        }
        return check();
    };
    bool converged = mismatch() && check();
    int it = 0;
    KluLikeLU lu_p, lu_q;
    bool factored = false;
    while (!converged && it < max_iter) {
        ++it;
        if (!factored) {
            int e = lu_p.analyze(nvv, Pp, Pi);
            if (e == NoError) e = lu_p.factor(Pp, Pi, Px);
            if (e == NoError) e = lu_q.analyze(npq, Qp, Qi);
            if (e == NoError) e = lu_q.factor(Qp, Qi, Qx);
            if (e != NoError) {
                err = e == SingularMatrix ? SolverFactor : e;
                break;
            }
            factored = true;
        }
        std::vector<double> x = p;
        if (lu_p.solve(x) != NoError) { err = SolverSolve; break; }
        for (int k = 0; k < nvv; ++k) Va[pvpq[k]] -= x[k];
        std::vector<cplx> tmp_va(n);
        for (int b = 0; b < n; ++b) tmp_va[b] = cplx(std::cos(Va[b]), std::sin(Va[b]));
        if (has_converged(tmp_va)) { converged = true; break; }
        if (lu_q.solve(q) != NoError) { err = SolverSolve; break; }
        for (int k = 0; k < npq; ++k) Vm[pq[k]] -= q[k];
        converged = has_converged(tmp_va);
    }
    if (!converged && err == NoError) err = TooManyIterations;
    for (int b = 0; b < n; ++b) { V[2 * b] = Vc[b].real(); V[2 * b + 1] = Vc[b].imag(); }
    *iters = it;
    return err;
}

// Contingency sweep (ContingencyAnalysis::compute loop). Each contingency c
// subtracts its Ybus deltas dv[dptr[c]:dptr[c+1]] at Ybus value positions
// dpos[...], checks connectivity, solves from Vinit and restores Ybus.
// The base case is solved first (analyze+factor), then every contingency is a
// refactor. With nthreads > 1, contingencies are split in contiguous chunks,
// one solver per thread, each thread solving the base case first like an
// independent ContingencyAnalysis object would. Outputs per contingency:
// V (ncont x n, zero rows when not converged / not connected), iters, err,
// conv (1/0), connected (1/0). Returns wall seconds.
double oracle_contingency(int n, const int* Yp, const int* Yi, const double* Yx_base,
                          const double* Sbus, const int* slack_ids, int nslack, const double* slack_w,
                          const int* pv, int npv, const int* pq, int npq,
                          const double* Vinit, int init_from_n, int single_slack,
                          int ncont, const int* dptr, const int* dpos, const double* dv,
                          int max_iter, double tol, int nthreads,
                          double* Vout, int* iters, int* errs, int* convs, int* connected) {
    auto t0 = std::chrono::steady_clock::now();
    if (nthreads < 1) nthreads = 1;
    const int nnz = Yp[n];
    auto worker = [&](int c0, int c1) {
        NRSolver s(single_slack != 0);
        std::vector<cplx> Yx(reinterpret_cast<const cplx*>(Yx_base), reinterpret_cast<const cplx*>(Yx_base) + nnz);
        CscC Y{n, Yp, Yi, Yx.data()};
        std::vector<cplx> Vb(reinterpret_cast<const cplx*>(Vinit), reinterpret_cast<const cplx*>(Vinit) + n);
        std::vector<cplx> V0 = Vb;
        bool base_conv = s.compute_pf(Y, V0, reinterpret_cast<const cplx*>(Sbus), slack_ids, nslack, slack_w,
                                      pv, npv, pq, npq, max_iter, tol);
        if (init_from_n) Vb = s.V();
        for (int c = c0; c < c1; ++c) {
            iters[c] = 0; errs[c] = NoError; convs[c] = 0; connected[c] = 0;
            std::fill(Vout + 2 * (size_t)c * n, Vout + 2 * (size_t)(c + 1) * n, 0.0);
            if (!base_conv) continue;
            for (int d = dptr[c]; d < dptr[c + 1]; ++d) Yx[dpos[d]] -= cplx(dv[2 * d], dv[2 * d + 1]);
            bool inv = check_invertible(Y, Yx);
            connected[c] = inv ? 1 : 0;
            if (inv) {
                std::vector<cplx> V = Vb;
                bool conv = s.compute_pf(Y, V, reinterpret_cast<const cplx*>(Sbus), slack_ids, nslack, slack_w,
                                         pv, npv, pq, npq, max_iter, tol);
                iters[c] = s.nb_iter(); errs[c] = s.error(); convs[c] = conv ? 1 : 0;
                if (conv) {
                    const auto& Vo = s.V();
                    for (int b = 0; b < n; ++b) {
                        Vout[2 * ((size_t)c * n + b)] = Vo[b].real();
                        Vout[2 * ((size_t)c * n + b) + 1] = Vo[b].imag();
                    }
                }
            }
            for (int d = dptr[c]; d < dptr[c + 1]; ++d) Yx[dpos[d]] += cplx(dv[2 * d], dv[2 * d + 1]);
        }
    };
    if (nthreads == 1) {
        worker(0, ncont);
    } else {
        std::vector<std::thread> th;
        const int chunk = (ncont + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; ++t) {
            int c0 = t * chunk, c1 = std::min(ncont, c0 + chunk);
            if (c0 >= c1) break;
            th.emplace_back(worker, c0, c1);
        }
        for (auto& t : th) t.join();
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Time series (TimeSeries::compute_Vs step loop): base case from Vinit, then
// each step warm-started from the previous step's V; stops at the first
// divergence (status 0). Chains: the T steps are split in nchains contiguous
// chunks solved by independent threads, each chunk head warm-started from the
// base-case V (nchains = 1 is the reference's exact semantics).
// Returns wall seconds; status_out = 1 if every step converged.
double oracle_time_series(int n, const int* Yp, const int* Yi, const double* Yx,
                          int T, const double* Sbuses, const int* slack_ids, int nslack, const double* slack_w,
                          const int* pv, int npv, const int* pq, int npq,
                          const double* Vinit, int single_slack, int max_iter, double tol, int nchains,
                          int nthreads, double* Vout, int* iters, int* errs, int* convs, int* status_out) {
    auto t0 = std::chrono::steady_clock::now();
    if (nchains < 1) nchains = 1;
    CscC Y{n, Yp, Yi, reinterpret_cast<const cplx*>(Yx)};
    std::atomic<int> first_fail(T);
    auto worker = [&](int s0, int s1) {
        NRSolver s(single_slack != 0);
        std::vector<cplx> V(reinterpret_cast<const cplx*>(Vinit), reinterpret_cast<const cplx*>(Vinit) + n);
        // The caller passes Vinit (the reference's optional base-case V); the
        // first step of a chain runs analyze+factor, later steps refactor.
        for (int t = s0; t < s1; ++t) {
            const cplx* S = reinterpret_cast<const cplx*>(Sbuses) + (size_t)t * n;
            bool conv = s.compute_pf(Y, V, S, slack_ids, nslack, slack_w, pv, npv, pq, npq, max_iter, tol);
            iters[t] = s.nb_iter(); errs[t] = s.error(); convs[t] = conv ? 1 : 0;
            if (!conv) {
                int ff = first_fail.load();
                while (t < ff && !first_fail.compare_exchange_weak(ff, t)) {}
                for (int u = t; u < s1; ++u)
                    std::fill(Vout + 2 * (size_t)u * n, Vout + 2 * (size_t)(u + 1) * n, 0.0);
                return;
            }
            V = s.V();
            for (int b = 0; b < n; ++b) {
                Vout[2 * ((size_t)t * n + b)] = V[b].real();
                Vout[2 * ((size_t)t * n + b) + 1] = V[b].imag();
            }
        }
    };
    if (nchains == 1) {
        worker(0, T);
    } else {
        // chains of ceil(T / nchains) steps, handed out to nthreads threads (each chain keeps
        // its own solver object: a fresh factorisation at its head, as the reference's threaded use)
        const int chunk = (T + nchains - 1) / nchains;
        const int nth = std::max(1, std::min(nthreads > 0 ? nthreads : nchains, nchains));
        std::atomic<int> next(0);
        std::vector<std::thread> th;
        for (int k = 0; k < nth; ++k)
            th.emplace_back([&]() {
                for (int c = next++; c < nchains; c = next++) {
                    const int s0 = c * chunk, s1 = std::min(T, s0 + chunk);
                    if (s0 < s1) worker(s0, s1);
                }
            });
        for (auto& t : th) t.join();
    }
    const int ff = first_fail.load();
    for (int t = ff; t < T; ++t) {
        std::fill(Vout + 2 * (size_t)t * n, Vout + 2 * (size_t)(t + 1) * n, 0.0);
        if (t > ff) { convs[t] = 0; }
    }
    *status_out = (ff == T) ? 1 : 0;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
