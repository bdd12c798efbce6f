// Host-side mirror of lightsim2grid's solver surface (see host.hpp).
#include "host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <sstream>

namespace ls2g {

namespace {
using clk = std::chrono::steady_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }
constexpr double kTolEqualFloat = 1e-7;  // BaseConstants.cpp:22
}  // namespace

std::string solver_type_name(SolverType t) {
    static const char* names[] = {
        "SparseLU", "KLU", "GaussSeidel", "DC", "GaussSeidelSynch", "NICSLU", "SparseLUSingleSlack",
        "KLUSingleSlack", "NICSLUSingleSlack", "KLUDC", "NICSLUDC", "CKTSO", "CKTSOSingleSlack", "CKTSODC",
        "FDPF_XB_SparseLU", "FDPF_BX_SparseLU", "FDPF_XB_KLU", "FDPF_BX_KLU", "FDPF_XB_NICSLU", "FDPF_BX_NICSLU",
        "FDPF_XB_CKTSO", "FDPF_BX_CKTSO", "AMDGPU", "AMDGPUSingleSlack", "AMDGPUMixed", "AMDGPUMixedSingleSlack",
        "AMDGPUDC", "AMDGPUFDPF_XB", "AMDGPUFDPF_BX"};
    const int k = (int)t;
    return (k >= 0 && k < (int)(sizeof(names) / sizeof(names[0]))) ? names[k] : "Unknown";
}

bool is_gpu_type(SolverType t) {
    return t == SolverType::AMDGPU || t == SolverType::AMDGPUSingleSlack || t == SolverType::AMDGPUMixed ||
           t == SolverType::AMDGPUMixedSingleSlack || t == SolverType::AMDGPUDC || t == SolverType::AMDGPUFDPF_XB ||
           t == SolverType::AMDGPUFDPF_BX;
}
int fdpf_kind(SolverType t) {
    if (t == SolverType::AMDGPUFDPF_XB) return 1 + (int)FDPFMethod::XB;
    if (t == SolverType::AMDGPUFDPF_BX) return 1 + (int)FDPFMethod::BX;
    return 0;
}
// ChooseSolver::is_dc (ChooseSolver.hpp:88-97)
bool is_dc_type(SolverType t) {
    return t == SolverType::DC || t == SolverType::KLUDC || t == SolverType::NICSLUDC || t == SolverType::CKTSODC ||
           t == SolverType::AMDGPUDC;
}

void check_rc(int rc, const char* what) {
    if (rc != 0) {
        std::ostringstream os;
        os << what << ": pfgpu error " << rc << " (" << pfgpu_last_error() << ")";
        throw std::runtime_error(os.str());
    }
}

pfgpu_ctx* engine_ctx() {
    static std::once_flag once;
    static pfgpu_ctx* ctx = nullptr;
    static std::string err;
    std::call_once(once, [] {
        if (pfgpu_ctx_create(0, 0, &ctx) != 0) {
            err = pfgpu_last_error();
            ctx = nullptr;
        }
    });
    if (!ctx) throw std::runtime_error("AMDGPU solver: no usable GPU (" + err + ")");
    return ctx;
}

void* DevMem::get(size_t bytes) {
    if (bytes > n_ || !p_) {
        release();
        void* q = nullptr;
        check_rc(pfgpu_dev_alloc(engine_ctx(), std::max<size_t>(bytes, 1), &q), "pfgpu_dev_alloc");
        p_ = q;
        n_ = std::max<size_t>(bytes, 1);
    }
    return p_;
}
void DevMem::release() {
    if (p_) pfgpu_dev_free(engine_ctx(), p_);
    p_ = nullptr;
    n_ = 0;
}

// ---------------------------------------------------------------------------
// GpuNRSolver
// ---------------------------------------------------------------------------
GpuNRSolver::~GpuNRSolver() {
    if (grid_) pfgpu_grid_destroy(grid_);
}

void GpuNRSolver::reset() {
    if (grid_) pfgpu_grid_destroy(grid_);
    grid_ = nullptr;
    key_p_.clear();
    V_.clear();
    Jx_.clear();
    nr_iter_ = 0;
    err_ = ErrorType::NotInitError;
}

RealVect GpuNRSolver::get_Va() const {
    RealVect r(V_.size());
    for (size_t b = 0; b < V_.size(); ++b) r[b] = std::arg(V_[b]);
    return r;
}
RealVect GpuNRSolver::get_Vm() const {
    RealVect r(V_.size());
    for (size_t b = 0; b < V_.size(); ++b) r[b] = std::abs(V_[b]);
    return r;
}

void GpuNRSolver::ensure_grid(const CscC& Y, const IntVect& slack_ids, const RealVect& w, const IntVect& pv,
                              const IntVect& pq) {
    if (fdpf_ && (Bp_.ncol != Y.ncol || Bpp_.ncol != Y.ncol))
        throw std::runtime_error("BaseFDPFAlgo::compute_pf: B' and B'' are built by GridModel::ac_pf (fillBp_Bpp); "
                                 "this solver needs a GridModel");
    auto same = [](const CscR& a, const CscR& b) { return a.p == b.p && a.i == b.i && a.x == b.x; };
    if (grid_ && Y.p == key_p_ && Y.i == key_i_ && slack_ids == key_slack_ && pv == key_pv_ && pq == key_pq_ &&
        w == key_w_ && (!fdpf_ || (same(Bp_, key_Bp_) && same(Bpp_, key_Bpp_))))
        return;
    auto t0 = clk::now();
    if (grid_) pfgpu_grid_destroy(grid_);
    grid_ = nullptr;
    pfgpu_grid* g = nullptr;
    if (fdpf_)
        check_rc(pfgpu_grid_load_fdpf(engine_ctx(), Y.ncol, Y.p.data(), Y.i.data(), slack_ids.data(),
                                      (int)slack_ids.size(), w.data(), pv.data(), (int)pv.size(), pq.data(),
                                      (int)pq.size(), Bp_.p.data(), Bp_.i.data(), Bp_.x.data(), Bpp_.p.data(),
                                      Bpp_.i.data(), Bpp_.x.data(), &g),
                 "pfgpu_grid_load_fdpf");
    else if (dc_)
        check_rc(pfgpu_grid_load_dc(engine_ctx(), Y.ncol, Y.p.data(), Y.i.data(), slack_ids.data(),
                                    (int)slack_ids.size(), pv.data(), (int)pv.size(), pq.data(), (int)pq.size(), &g),
                 "pfgpu_grid_load_dc");
    else
        check_rc(pfgpu_grid_load(engine_ctx(), Y.ncol, Y.p.data(), Y.i.data(), slack_ids.data(), (int)slack_ids.size(),
                                 w.data(), pv.data(), (int)pv.size(), pq.data(), (int)pq.size(), single_ ? 1 : 0, &g),
                 "pfgpu_grid_load");
    grid_ = g;
    check_rc(pfgpu_grid_get_info(grid_, &info_), "pfgpu_grid_get_info");
    key_p_ = Y.p; key_i_ = Y.i; key_slack_ = slack_ids; key_pv_ = pv; key_pq_ = pq; key_w_ = w;
    key_Bp_ = Bp_;
    key_Bpp_ = Bpp_;
    timer_initialize_ += since(t0);
}

static void check_sizes(const CscC& Y, size_t nV, size_t nS, const char* who) {
    if (Y.nrow != Y.ncol || (int)nS != Y.nrow || (int)nS != Y.ncol) {
        std::ostringstream os;
        os << who << ": Size of the Sbus should be the same as the size of Ybus. Currently: Sbus  (" << nS
           << ") and Ybus (" << Y.nrow << ", " << Y.ncol << ").";
        throw std::runtime_error(os.str());
    }
    if ((int)nV != Y.nrow) {
        std::ostringstream os;
        os << who << ": Size of V (init voltages) should be the same as the size of Ybus. Currently: V  (" << nV
           << ") and Ybus (" << Y.nrow << ", " << Y.ncol << ").";
        throw std::runtime_error(os.str());
    }
}

bool GpuNRSolver::compute_pf(const CscC& Ybus, CplxVect& V, const CplxVect& Sbus, const IntVect& slack_ids,
                             const RealVect& slack_weights, const IntVect& pv, const IntVect& pq, int max_iter,
                             double tol) {
    check_sizes(Ybus, V.size(), Sbus.size(), single_ ? "BaseNRSingleSlackAlgo::compute_pf" : "BaseNRAlgo::compute_pf");
    auto t0 = clk::now();
    timer_Fx_ = timer_solve_ = timer_check_ = timer_total_ = timer_initialize_ = timer_pre_proc_ = 0;
    timer_refactor_ = timer_fillJ_ = timer_VaVm_ = 0;
    ensure_grid(Ybus, slack_ids, slack_weights, pv, pq);
    const int n = Ybus.ncol;
    V_ = V;
    int it = 0, st = 0;
    Jx_.assign(info_.nnz_jac, 0.0);
    pfgpu_stats stats{};
    check_rc(pfgpu_solve_batch(engine_ctx(), grid_, 1, reinterpret_cast<const double*>(Ybus.x.data()), 0, nullptr,
                               nullptr, nullptr, reinterpret_cast<const double*>(Sbus.data()), 0,
                               reinterpret_cast<double*>(V_.data()), n, max_iter, tol,
                               precision_, 0, &it, &st, Jx_.data(), &stats),
             "pfgpu_solve_batch");
    nr_iter_ = it;
    err_ = (ErrorType)st;
    timer_fillJ_ = stats.phase_s[0];
    timer_refactor_ = stats.phase_s[1] + stats.phase_s[2];
    timer_solve_ = stats.phase_s[3];
    timer_VaVm_ = stats.phase_s[4];
    timer_Fx_ = stats.phase_s[5];
    timer_pre_proc_ = stats.h2d_s;
    timer_total_ = since(t0);
    return err_ == ErrorType::NoError;
}

void GpuNRSolver::solve_batch(const CscC& Ybus, const IntVect& slack_ids, const RealVect& w, const IntVect& pv,
                              const IntVect& pq, int ninst, const IntVect* dptr, const IntVect* dpos,
                              const CplxVect* dval, const CplxVect& Sbus, bool sbus_per_instance, const CplxVect& V0,
                              bool v0_per_instance, int max_iter, double tol, CplxVect& Vout, IntVect& iters,
                              IntVect& status, pfgpu_stats* st) {
    ensure_grid(Ybus, slack_ids, w, pv, pq);
    const int n = Ybus.ncol;
    Vout.assign((size_t)ninst * n, cplx(0, 0));
    if (v0_per_instance) std::copy(V0.begin(), V0.end(), Vout.begin());
    else std::copy(V0.begin(), V0.begin() + n, Vout.begin());
    iters.assign(ninst, 0);
    status.assign(ninst, 0);
    if (ninst == 0) return;
    check_rc(pfgpu_solve_batch(engine_ctx(), grid_, ninst, reinterpret_cast<const double*>(Ybus.x.data()), 0,
                               dptr ? dptr->data() : nullptr, dpos ? dpos->data() : nullptr,
                               dval ? reinterpret_cast<const double*>(dval->data()) : nullptr,
                               reinterpret_cast<const double*>(Sbus.data()), sbus_per_instance ? n : 0,
                               reinterpret_cast<double*>(Vout.data()), v0_per_instance ? n : 0, max_iter, tol,
                               precision_, 0, iters.data(), status.data(), nullptr, st),
             "pfgpu_solve_batch");
}

void GpuNRSolver::solve_batch_device(const CscC& Ybus, const IntVect& slack_ids, const RealVect& w, const IntVect& pv,
                                     const IntVect& pq, int ninst, const double* d_Yx, const int* d_dptr,
                                     const int* d_dpos, const double* d_dval, const double* d_S,
                                     bool sbus_per_instance, const double* d_V0, bool v0_per_instance, double* d_V,
                                     int max_iter, double tol, int* d_iters, int* d_status, pfgpu_stats* st) {
    ensure_grid(Ybus, slack_ids, w, pv, pq);
    if (ninst == 0) return;
    const int n = Ybus.ncol;
    check_rc(pfgpu_solve_batch_device(engine_ctx(), grid_, ninst, d_Yx, 0, d_dptr, d_dpos, d_dval, d_S,
                                      sbus_per_instance ? n : 0, d_V0, v0_per_instance ? n : 0, d_V, max_iter, tol,
                                      precision_, 0, d_iters, d_status, st),
             "pfgpu_solve_batch_device");
}

void GpuNRSolver::solve_chains_device(const CscC& Ybus, const IntVect& slack_ids, const RealVect& w, const IntVect& pv,
                                      const IntVect& pq, int nrows, int chain_len, const double* d_Yx,
                                      const double* d_S, const double* d_V0, double* d_V, int max_iter, double tol,
                                      int* d_iters, int* d_status, pfgpu_stats* st) {
    ensure_grid(Ybus, slack_ids, w, pv, pq);
    if (nrows == 0) return;
    check_rc(pfgpu_solve_chains_device(engine_ctx(), grid_, nrows, chain_len, d_Yx, 0, d_S, d_V0, 0, d_V, max_iter, tol,
                                       precision_, 0, d_iters, d_status, st),
             "pfgpu_solve_chains_device");
}

CscR GpuNRSolver::get_J() const {
    CscR J;
    if (!grid_) return J;
    J.nrow = J.ncol = info_.n_jac;
    J.p.resize(info_.n_jac + 1);
    J.i.resize(info_.nnz_jac);
    check_rc(pfgpu_grid_get_J_pattern(grid_, J.p.data(), J.i.data()), "pfgpu_grid_get_J_pattern");
    J.x = Jx_;
    if ((int)J.x.size() != info_.nnz_jac) J.x.assign(info_.nnz_jac, 0.0);
    return J;
}

// ---------------------------------------------------------------------------
// ChooseSolver
// ---------------------------------------------------------------------------
ChooseSolver::ChooseSolver() : type_(SolverType::AMDGPU) {
    s_[0] = std::make_unique<GpuNRSolver>(false, PFGPU_FP64);
    s_[1] = std::make_unique<GpuNRSolver>(true, PFGPU_FP64);
    s_[2] = std::make_unique<GpuNRSolver>(false, PFGPU_FP32);
    s_[3] = std::make_unique<GpuNRSolver>(true, PFGPU_FP32);
    s_[4] = std::make_unique<GpuNRSolver>(true, PFGPU_FP64, true);
    s_[5] = std::make_unique<GpuNRSolver>(true, PFGPU_FP64, false, 1 + (int)FDPFMethod::XB);
    s_[6] = std::make_unique<GpuNRSolver>(true, PFGPU_FP64, false, 1 + (int)FDPFMethod::BX);
}

std::vector<SolverType> ChooseSolver::available_solvers() const {
    return {SolverType::AMDGPU, SolverType::AMDGPUSingleSlack, SolverType::AMDGPUMixed,
            SolverType::AMDGPUMixedSingleSlack, SolverType::AMDGPUDC, SolverType::AMDGPUFDPF_XB,
            SolverType::AMDGPUFDPF_BX};
}

void ChooseSolver::change_solver(SolverType t) {
    if (t == type_) return;
    if (!is_gpu_type(t)) {
        // ChooseSolver.hpp:155-192: asking for a solver that is not compiled in raises
        std::ostringstream os;
        os << "Impossible to use the solver " << solver_type_name(t)
           << ": it is not available in this build (available: AMDGPU, AMDGPUSingleSlack, AMDGPUMixed, "
              "AMDGPUMixedSingleSlack, AMDGPUDC, AMDGPUFDPF_XB, AMDGPUFDPF_BX).";
        throw std::runtime_error(os.str());
    }
    current().reset();
    type_ = t;
    current().reset();
}

GpuNRSolver& ChooseSolver::current() { return *s_[(int)type_ - (int)SolverType::AMDGPU]; }
const GpuNRSolver& ChooseSolver::current() const { return *s_[(int)type_ - (int)SolverType::AMDGPU]; }

// ---------------------------------------------------------------------------
// GridModel
// ---------------------------------------------------------------------------
void GridModel::init_bus(int n_sub, int n_busbar_per_sub, const RealVect& bus_vn_kv, int nb_line, int nb_trafo) {
    const int nb = n_sub * std::max(1, n_busbar_per_sub);
    if ((int)bus_vn_kv.size() != nb && (int)bus_vn_kv.size() != n_sub)
        throw std::runtime_error("GridModel::init_bus: bus_vn_kv must have n_sub or n_sub * n_busbar_per_sub entries");
    bus_vn_kv_.assign(nb, 0.0);
    for (int b = 0; b < nb; ++b) bus_vn_kv_[b] = bus_vn_kv[b % bus_vn_kv.size()];
    (void)nb_line;
    if (nb_trafo != 0) throw std::runtime_error("GridModel::init_bus: transformers are not part of this build");
}

void GridModel::init_powerlines(const RealVect& r, const RealVect& x, const CplxVect& h, const IntVect& from,
                                const IntVect& to) {
    const size_t m = r.size();
    if (x.size() != m || h.size() != m || from.size() != m || to.size() != m)
        throw std::runtime_error("GridModel::init_powerlines: inconsistent sizes");
    for (size_t k = 0; k < m; ++k)
        if (from[k] < 0 || from[k] >= total_bus() || to[k] < 0 || to[k] >= total_bus())
            throw std::runtime_error("GridModel::init_powerlines: bus id out of range");
    line_r_ = r; line_x_ = x; line_h_ = h; line_from_ = from; line_to_ = to;
    line_status_.assign(m, true);
}

void GridModel::init_generators(const RealVect& p, const RealVect& v, const RealVect& min_q, const RealVect& max_q,
                                const IntVect& bus) {
    const size_t m = p.size();
    if (v.size() != m || bus.size() != m || min_q.size() != m || max_q.size() != m)
        throw std::runtime_error("GridModel::init_generators: inconsistent sizes");
    gen_p_ = p; gen_v_ = v; gen_bus_ = bus;
    gen_slack_.assign(m, false);
    gen_slack_w_.assign(m, 0.0);
}

void GridModel::init_loads(const RealVect& p, const RealVect& q, const IntVect& bus) {
    if (q.size() != p.size() || bus.size() != p.size())
        throw std::runtime_error("GridModel::init_loads: inconsistent sizes");
    load_p_ = p; load_q_ = q; load_bus_ = bus;
}

void GridModel::init_sgens(const RealVect& p, const RealVect& q, const RealVect&, const RealVect&, const RealVect&,
                           const RealVect&, const IntVect& bus) {
    if (q.size() != p.size() || bus.size() != p.size())
        throw std::runtime_error("GridModel::init_sgens: inconsistent sizes");
    sgen_p_ = p; sgen_q_ = q; sgen_bus_ = bus;
}

void GridModel::add_gen_slackbus(int gen_id, double weight) {
    if (gen_id < 0 || gen_id >= (int)gen_bus_.size()) throw std::runtime_error("GridModel::add_gen_slackbus: bad id");
    if (weight <= 0.0) throw std::runtime_error("GridModel::add_gen_slackbus: weight must be > 0");
    gen_slack_[gen_id] = true;
    gen_slack_w_[gen_id] = weight;
}
void GridModel::remove_gen_slackbus(int gen_id) {
    if (gen_id < 0 || gen_id >= (int)gen_bus_.size()) throw std::runtime_error("GridModel::remove_gen_slackbus: bad id");
    gen_slack_[gen_id] = false;
    gen_slack_w_[gen_id] = 0.0;
}
void GridModel::deactivate_powerline(int id) { line_status_.at(id) = false; }
void GridModel::reactivate_powerline(int id) { line_status_.at(id) = true; }

void GridModel::branch_admittance_dc(int k, cplx& yff, cplx& yft, cplx& ytf, cplx& ytt) const {
    // TwoSidesContainer_rxh_A::_update_model_coeffs_one_el, DC part: 1 / x, no shunt
    const double y = 1.0 / line_x_[k];
    yff = ytt = cplx(y, 0.0);
    yft = ytf = cplx(-y, 0.0);
}

void GridModel::branch_admittance(int k, cplx& yff, cplx& yft, cplx& ytf, cplx& ytt) const {
    // TwoSidesContainer_rxh_A::_update_model_coeffs_one_el with h split 0.5 / 0.5 (LineContainer.cpp:13-34);
    // ys = 1 / (r + j x) evaluated as (r - j x) / (r^2 + x^2), the operations (and their order) of the
    // device Ybus assembly (pfgpu_ybus_values_device)
    const double r = line_r_[k], x = line_x_[k];
    const double d = r * r + x * x;
    const cplx ys(r / d, -x / d);
    const cplx hh(0.5 * line_h_[k].real(), 0.5 * line_h_[k].imag());
    yff = cplx(ys.real() + hh.real(), ys.imag() + hh.imag());
    ytt = yff;
    yft = -ys;
    ytf = -ys;
}

// GridModel::pre_process_solver (GridModel.cpp:450-571) for this subset.
SolverInput GridModel::pre_process_solver(const CplxVect& Vinit, bool ac) const {
    const int nb = total_bus();
    if ((int)Vinit.size() != nb) {
        std::ostringstream os;
        os << "GridModel::ac_pf: Size of the Vinit should be the same as the total number of buses. Currently:  "
           << "Vinit: " << Vinit.size() << " and there are " << nb << " buses.";
        throw std::runtime_error(os.str());
    }
    SolverInput in;
    // active buses: touched by a connected element
    std::vector<char> active(nb, 0);
    for (size_t k = 0; k < line_r_.size(); ++k)
        if (line_status_[k]) active[line_from_[k]] = active[line_to_[k]] = 1;
    for (int b : gen_bus_) active[b] = 1;
    for (int b : load_bus_) active[b] = 1;
    for (int b : sgen_bus_) active[b] = 1;
    in.id_me_to_solver.assign(nb, -1);
    for (int b = 0; b < nb; ++b)
        if (active[b]) { in.id_me_to_solver[b] = (int)in.id_solver_to_me.size(); in.id_solver_to_me.push_back(b); }
    const int ns = (int)in.id_solver_to_me.size();
    // Ybus
    std::vector<std::vector<std::pair<int, cplx>>> cols(ns);
    auto add = [&](int r, int c, cplx v) { cols[c].push_back({r, v}); };
    for (size_t k = 0; k < line_r_.size(); ++k) {
        if (!line_status_[k]) continue;
        cplx yff, yft, ytf, ytt;
        if (ac) branch_admittance((int)k, yff, yft, ytf, ytt);
        else branch_admittance_dc((int)k, yff, yft, ytf, ytt);
        const int f = in.id_me_to_solver[line_from_[k]], t = in.id_me_to_solver[line_to_[k]];
        add(f, f, yff); add(t, t, ytt); add(f, t, yft); add(t, f, ytf);
    }
    in.Ybus.nrow = in.Ybus.ncol = ns;
    in.Ybus.p.assign(ns + 1, 0);
    for (int c = 0; c < ns; ++c) {
        auto& v = cols[c];
        bool diag = false;
        for (auto& e : v) diag |= (e.first == c);
        if (!diag) v.push_back({c, cplx(0, 0)});  // keep every diagonal entry structurally present
        // stable: duplicates (parallel branches, the diagonal) sum in branch order
        std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t q = 0; q < v.size(); ++q) {
            if (!in.Ybus.i.empty() && (int)in.Ybus.i.size() > in.Ybus.p[c] && in.Ybus.i.back() == v[q].first) {
                in.Ybus.x.back() += v[q].second;  // parallel elements sum up (setFromTriplets)
            } else {
                in.Ybus.i.push_back(v[q].first);
                in.Ybus.x.push_back(v[q].second);
            }
        }
        in.Ybus.p[c + 1] = (int)in.Ybus.i.size();
    }
    // Sbus (MW -> pu)
    in.Sbus.assign(ns, cplx(0, 0));
    for (size_t g = 0; g < gen_p_.size(); ++g) in.Sbus[in.id_me_to_solver[gen_bus_[g]]] += gen_p_[g];
    for (size_t g = 0; g < sgen_p_.size(); ++g) in.Sbus[in.id_me_to_solver[sgen_bus_[g]]] += cplx(sgen_p_[g], sgen_q_[g]);
    for (size_t l = 0; l < load_p_.size(); ++l) in.Sbus[in.id_me_to_solver[load_bus_[l]]] -= cplx(load_p_[l], load_q_[l]);
    if (std::fabs(sn_mva_ - 1.0) > kTolEqualFloat)
        for (auto& s : in.Sbus) s /= sn_mva_;
    // slack buses (slack generators' buses, generator order), weights normalised per solver bus
    std::vector<char> is_slack(ns, 0);
    in.slack_weights.assign(ns, 0.0);
    for (size_t g = 0; g < gen_bus_.size(); ++g) {
        if (!gen_slack_[g] || std::fabs(gen_slack_w_[g]) < kTolEqualFloat) continue;
        const int b = in.id_me_to_solver[gen_bus_[g]];
        if (!is_slack[b]) { is_slack[b] = 1; in.slack_ids.push_back(b); }
        in.slack_weights[b] += gen_slack_w_[g];
    }
    if (in.slack_ids.empty()) throw std::runtime_error("GridModel: no slack generator (use add_gen_slackbus)");
    const double wsum = std::accumulate(in.slack_weights.begin(), in.slack_weights.end(), 0.0);
    for (auto& w : in.slack_weights) w /= wsum;
    // PV buses: voltage-regulating generators not on a slack bus (GeneratorContainer::fillpv)
    std::vector<char> added(ns, 0);
    for (size_t g = 0; g < gen_bus_.size(); ++g) {
        const int b = in.id_me_to_solver[gen_bus_[g]];
        if (is_slack[b] || added[b]) continue;
        in.pv.push_back(b);
        added[b] = 1;
    }
    for (int b = 0; b < ns; ++b)
        if (!is_slack[b] && !added[b]) in.pq.push_back(b);
    // initial V and generator magnitude setpoints (GeneratorContainer::set_vm)
    in.V.assign(ns, cplx(init_vm_pu_, 0.0));
    for (int b = 0; b < ns; ++b) in.V[b] = Vinit[in.id_solver_to_me[b]];
    for (size_t g = 0; g < gen_bus_.size(); ++g) {
        const int b = in.id_me_to_solver[gen_bus_[g]];
        double m = std::abs(in.V[b]);
        if (std::fabs(m) < kTolEqualFloat) { in.V[b] = 1.0; m = 1.0; }
        in.V[b] *= gen_v_[g] / m;
    }
    return in;
}

CplxVect GridModel::dc_pf(const CplxVect& Vinit, int max_iter, double tol) {
    last_ = pre_process_solver(Vinit, false);
    CplxVect V = last_.V;
    const bool conv = dc_solver_->compute_pf(last_.Ybus, V, last_.Sbus, last_.slack_ids, last_.slack_weights, last_.pv,
                                             last_.pq, max_iter, tol / sn_mva_);
    CplxVect res;
    if (!conv) return res;
    V_solver_ = dc_solver_->get_V();
    res.assign(total_bus(), cplx(0, 0));
    for (size_t b = 0; b < last_.id_solver_to_me.size(); ++b) res[last_.id_solver_to_me[b]] = V_solver_[b];
    return res;
}

// GridModel::fillBp_Bpp for this subset (GridModel.cpp:1037-1060): per connected line,
// TwoSidesContainer_rxh_A::get_fdpf_coeffs (TwoSidesContainer_rxh_A.hpp:818-843) with ys_bp = 1/(jx)
// (XB) or 1/(r + jx) (BX) and ys_bpp the other one, the charging split half and half between the
// ends (LineContainer.cpp:13-34), placed as fillBp_Bpp does (:503-598): B' gets -Im(ys_bp) on both
// diagonals and +Im(ys_bp) off them, B'' -(Im(ys_bpp) + Im(h)/2) on the diagonals and +Im(ys_bpp)
// off them. Duplicates are summed and rows sorted (setFromTriplets).
void GridModel::fill_Bp_Bpp(const SolverInput& in, FDPFMethod method, CscR& Bp, CscR& Bpp) const {
    const int n = (int)in.id_solver_to_me.size();
    std::vector<std::vector<std::pair<int, double>>> cp(n), cpp(n);
    for (size_t k = 0; k < line_r_.size(); ++k) {
        if (!line_status_[k]) continue;
        const int f = in.id_me_to_solver[line_from_[k]], t = in.id_me_to_solver[line_to_[k]];
        if (f < 0 || t < 0) continue;
        const double r = line_r_[k], x = line_x_[k];
        const double y_jx = std::imag(1.0 / cplx(0.0, x)), y_rx = std::imag(1.0 / cplx(r, x));
        const double ybp = method == FDPFMethod::XB ? y_jx : y_rx;
        const double ybpp = method == FDPFMethod::XB ? y_rx : y_jx;
        const double half = 0.5 * line_h_[k].imag();
        cp[f].push_back({f, -ybp});
        cp[t].push_back({t, -ybp});
        cp[t].push_back({f, ybp});
        cp[f].push_back({t, ybp});
        cpp[f].push_back({f, -(ybpp + half)});
        cpp[t].push_back({t, -(ybpp + half)});
        cpp[t].push_back({f, ybpp});
        cpp[f].push_back({t, ybpp});
    }
    auto build = [n](std::vector<std::vector<std::pair<int, double>>>& cols, CscR& M) {
        M.nrow = M.ncol = n;
        M.p.assign(n + 1, 0);
        M.i.clear();
        M.x.clear();
        for (int c = 0; c < n; ++c) {
            auto& v = cols[c];
            std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            for (size_t q = 0; q < v.size(); ++q) {
                if ((int)M.i.size() > M.p[c] && M.i.back() == v[q].first) M.x.back() += v[q].second;
                else { M.i.push_back(v[q].first); M.x.push_back(v[q].second); }
            }
            M.p[c + 1] = (int)M.i.size();
        }
    };
    build(cp, Bp);
    build(cpp, Bpp);
}

CplxVect GridModel::ac_pf(const CplxVect& Vinit, int max_iter, double tol) {
    last_ = pre_process_solver(Vinit);
    if (const int kind = fdpf_kind(solver_->get_type())) {
        CscR Bp, Bpp;
        fill_Bp_Bpp(last_, (FDPFMethod)(kind - 1), Bp, Bpp);
        solver_->current().set_fdpf_matrices(std::move(Bp), std::move(Bpp));
    }
    CplxVect V = last_.V;
    const bool conv = solver_->compute_pf(last_.Ybus, V, last_.Sbus, last_.slack_ids, last_.slack_weights, last_.pv,
                                          last_.pq, max_iter, tol / sn_mva_);
    CplxVect res;
    if (!conv) return res;
    V_solver_ = solver_->current().get_V();
    res.assign(total_bus(), cplx(0, 0));
    for (size_t b = 0; b < last_.id_solver_to_me.size(); ++b) res[last_.id_solver_to_me[b]] = V_solver_[b];
    return res;
}

// ---------------------------------------------------------------------------
// Batch drivers
// ---------------------------------------------------------------------------
void BaseBatchSolverSynch::keep_device_rows(int ns, const IntVect& row_src) {
    ns_ = ns;
    d_row_src_.upload(row_src);
}

// Amps / active power at the origin side of every line for every result row
// (BaseBatchSolverSynch.hpp:86-192), on the device from the voltages the last compute left there:
// I = yff Ef + yft Et, S = Ef conj(I); A = |S| sn_mva / (sqrt(3) |Ef| vn_kv(from)), P = Re(S) sn_mva;
// after a DC compute P = (Re(yff) arg Ef + Re(yft) arg Et) sn_mva with the DC coefficients and
// A = |P| / (sqrt(3) |Ef| vn_kv(from)). Rows that did not converge have zero voltages (NaN amps);
// lines out of service give 0.
void BaseBatchSolverSynch::compute_flows_from_Vs(bool amps) {
    if (voltages_.empty() || ns_ == 0)
        throw std::runtime_error("BaseMultiplePowerflow::compute_flows_from_Vs: cannot compute the flows as the "
                                 "voltages are not set. Have you called compute(...) ? ");
    auto t0 = clk::now();
    const int nl = grid_model_.nb_powerline();
    IntVect f(nl, -1), t(nl, -1);
    CplxVect yff(nl), yft(nl);
    RealVect vn(nl, 1.0);
    for (int k = 0; k < nl; ++k) {
        if (!grid_model_.line_status(k)) continue;
        cplx ytf, ytt;
        if (dc_rows_) grid_model_.branch_admittance_dc(k, yff[k], yft[k], ytf, ytt);  // ydc_11, ydc_12
        else grid_model_.branch_admittance(k, yff[k], yft[k], ytf, ytt);
        f[k] = id_me_to_solver_[grid_model_.line_from(k)];
        t[k] = id_me_to_solver_[grid_model_.line_to(k)];
        vn[k] = grid_model_.bus_vn_kv(grid_model_.line_from(k));
    }
    IntVect cptr, cbr;
    clean_lists(cptr, cbr);
    const bool clean = !cbr.empty();
    int* df = d_tmp_[0].upload(f);
    int* dt = d_tmp_[1].upload(t);
    double* dyff = reinterpret_cast<double*>(d_tmp_[2].upload(yff));
    double* dyft = reinterpret_cast<double*>(d_tmp_[3].upload(yft));
    double* dvn = d_tmp_[4].upload(vn);
    int* dcp = clean ? d_tmp_[5].upload(cptr) : nullptr;
    int* dcb = clean ? d_tmp_[6].upload(cbr) : nullptr;
    RealVect& out = amps ? amps_ : pflows_;
    out.assign((size_t)nrows_ * nl, 0.0);
    double* dout = static_cast<double*>(d_tmp_[7].get(std::max<size_t>(out.size(), 1) * 8));
    auto flows = dc_rows_ ? pfgpu_batch_flows_dc_device : pfgpu_batch_flows_device;
    check_rc(flows(engine_ctx(), nrows_, ns_, d_V_.as<double>(), d_row_src_.as<int>(), nl, df, dt, dyff, dyft, dvn,
                   grid_model_.get_sn_mva(), dcp, dcb, amps ? dout : nullptr, amps ? nullptr : dout),
             dc_rows_ ? "pfgpu_batch_flows_dc_device" : "pfgpu_batch_flows_device");
    if (!out.empty()) check_rc(pfgpu_memcpy_d2h(engine_ctx(), out.data(), dout, out.size() * 8), "d2h");
    if (amps) timer_amps_ = since(t0);
}

void ContingencyAnalysis::add_all_n1() {
    for (int k = 0; k < grid_model_.nb_powerline(); ++k) defaults_.insert({k});
}
void ContingencyAnalysis::add_n1(int id) {
    if (id < 0 || id >= grid_model_.nb_powerline()) {
        std::ostringstream os;
        os << "ContingencyAnalysis::add_n1: impossible to add the line " << id << " there are only "
           << grid_model_.nb_powerline() << " branches on the grid.";
        throw std::runtime_error(os.str());
    }
    defaults_.insert({id});
}
void ContingencyAnalysis::add_nk(const IntVect& ids) {
    std::set<int> s;
    for (int id : ids) {
        if (id < 0 || id >= grid_model_.nb_powerline()) throw std::runtime_error("ContingencyAnalysis::add_nk: bad id");
        s.insert(id);
    }
    defaults_.insert(s);
}
void ContingencyAnalysis::add_multiple_n1(const IntVect& ids) { for (int id : ids) add_n1(id); }
void ContingencyAnalysis::remove_n1(int id) { defaults_.erase({id}); }
void ContingencyAnalysis::remove_nk(const IntVect& ids) { defaults_.erase(std::set<int>(ids.begin(), ids.end())); }
void ContingencyAnalysis::remove_multiple_n1(const IntVect& ids) { for (int id : ids) remove_n1(id); }
void ContingencyAnalysis::clear() {
    defaults_.clear();
    clear_results_only();
}
void ContingencyAnalysis::clear_results_only() {
    voltages_.clear(); amps_.clear(); pflows_.clear();
    nrows_ = 0;
    ns_ = 0;
    nb_solved_ = 0;
    timer_total_ = timer_solver_ = timer_pre_proc_ = timer_amps_ = timer_modif_ybus_ = 0;
}
std::vector<IntVect> ContingencyAnalysis::my_defaults() const {
    std::vector<IntVect> r;
    for (const auto& s : defaults_) r.emplace_back(s.begin(), s.end());
    return r;
}

namespace {
int csc_pos(const CscC& Y, int r, int c) {
    auto b = Y.i.begin() + Y.p[c], e = Y.i.begin() + Y.p[c + 1];
    auto it = std::lower_bound(b, e, r);
    if (it == e || *it != r) throw std::runtime_error("ContingencyAnalysis: coefficient not in the Ybus pattern");
    return (int)(it - Y.i.begin());
}
}  // namespace

// ContingencyAnalysis::init_li_coeffs (ContingencyAnalysis.cpp:47-110): the Ybus positions and values
// of each line's four coefficients, -1 for a line out of service
void ContingencyAnalysis::branch_tables(const SolverInput& in, bool dc, IntVect& pos, CplxVect& y) const {
    const int nl = grid_model_.nb_powerline();
    pos.assign(4 * (size_t)nl, -1);
    y.assign(4 * (size_t)nl, cplx(0, 0));
    for (int k = 0; k < nl; ++k) {
        if (!grid_model_.line_status(k)) continue;
        cplx yff, yft, ytf, ytt;
        if (dc) grid_model_.branch_admittance_dc(k, yff, yft, ytf, ytt);  // ydc_11 ... ydc_22
        else grid_model_.branch_admittance(k, yff, yft, ytf, ytt);
        const int f = in.id_me_to_solver[grid_model_.line_from(k)], t = in.id_me_to_solver[grid_model_.line_to(k)];
        pos[4 * k] = csc_pos(in.Ybus, f, f);
        pos[4 * k + 1] = csc_pos(in.Ybus, t, t);
        pos[4 * k + 2] = csc_pos(in.Ybus, f, t);
        pos[4 * k + 3] = csc_pos(in.Ybus, t, f);
        y[4 * k] = yff;
        y[4 * k + 1] = ytt;
        y[4 * k + 2] = yft;
        y[4 * k + 3] = ytf;
    }
}

void ContingencyAnalysis::clean_lists(IntVect& ptr, IntVect& br) const {
    ptr.assign(1, 0);
    br.clear();
    for (const auto& cont : defaults_) {
        for (int k : cont) br.push_back(k);
        ptr.push_back((int)br.size());
    }
}

namespace {
// contingency CSR (branch lists) and edit offsets (4 per in-service branch) of a list of contingencies
struct ContLists {
    IntVect ptr{0}, br, dptr{0};
};
void add_cont(ContLists& L, const std::set<int>& cont, const IntVect& pos) {
    int ne = 0;
    for (int k : cont) {
        L.br.push_back(k);
        if (pos[4 * k] >= 0) ne += 4;
    }
    L.ptr.push_back((int)L.br.size());
    L.dptr.push_back(L.dptr.back() + ne);
}
}  // namespace

IntVect ContingencyAnalysis::is_grid_connected_after_contingency() {
    // on the Ybus of the solver in use (ContingencyAnalysis.cpp:133-148): the DC one for a DC solver
    const bool dc = solver_.current().is_dc();
    SolverInput in = grid_model_.pre_process_solver(CplxVect(grid_model_.total_bus(), cplx(1, 0)), !dc);
    IntVect pos;
    CplxVect y;
    branch_tables(in, dc, pos, y);
    ContLists L;
    for (const auto& cont : defaults_) add_cont(L, cont, pos);
    const int ncont = (int)defaults_.size();
    IntVect res(ncont, 0);
    if (ncont == 0) return res;
    DevMem m[9];
    const int n = in.Ybus.ncol;
    int* dcp = m[0].upload(L.ptr);
    int* dcb = m[1].upload(L.br);
    int* dpos_br = m[2].upload(pos);
    double* dy = reinterpret_cast<double*>(m[3].upload(y));
    int* ddp = m[4].upload(L.dptr);
    int* ddq = static_cast<int*>(m[5].get(std::max(L.dptr.back(), 1) * sizeof(int)));
    double* ddv = static_cast<double*>(m[6].get(std::max(L.dptr.back(), 1) * 16));
    check_rc(pfgpu_contingency_edits_device(engine_ctx(), ncont, dcp, dcb, dpos_br, dy, ddp, ddq, ddv),
             "pfgpu_contingency_edits_device");
    int* dYp = m[7].upload(in.Ybus.p);
    int* dYi = m[8].upload(in.Ybus.i);
    DevMem mx, mc;
    double* dYx = reinterpret_cast<double*>(mx.upload(in.Ybus.x));
    int* dconn = static_cast<int*>(mc.get(ncont * sizeof(int)));
    check_rc(pfgpu_connectivity_device(engine_ctx(), ncont, n, dYp, dYi, dYx, ddp, ddq, ddv, dconn),
             "pfgpu_connectivity_device");
    check_rc(pfgpu_memcpy_d2h(engine_ctx(), res.data(), dconn, ncont * sizeof(int)), "d2h");
    return res;
}

// ContingencyAnalysis::compute (ContingencyAnalysis.cpp:166-233), on the device: the contingency Ybus
// edits and the connectivity check of every contingency, then one batched solve of the connected
// ones; the voltages stay on the device for the flows. With the DC solver (ac_solver_used() false)
// the DC Ybus and coefficients are used, and every contingency is solved: remove_from_Ybus reports
// the DC Ybus as always invertible (:113-131) and the solver's own failure decides.
void ContingencyAnalysis::compute(const CplxVect& Vinit, int max_iter, double tol) {
    auto t_all = clk::now();
    auto t_pre = clk::now();
    nb_solved_ = 0;
    timer_solver_ = timer_modif_ybus_ = 0;
    const double sn = grid_model_.get_sn_mva();
    const bool dc = solver_.current().is_dc();
    SolverInput in = grid_model_.pre_process_solver(Vinit, !dc);
    const int nb = grid_model_.total_bus(), ns = (int)in.id_solver_to_me.size();
    const int ncont = (int)defaults_.size();
    nrows_ = ncont;
    ns_ = 0;
    id_me_to_solver_ = in.id_me_to_solver;
    voltages_.assign((size_t)ncont * nb, cplx(0, 0));
    amps_.clear();
    pflows_.clear();
    GpuNRSolver& s = solver_.current();
    dc_rows_ = dc;
    s.reset();
    // base case ("powerflow in n", BaseBatchSolverSynch.hpp:296-313): tol as given
    CplxVect Vb = in.V;
    const bool base_conv = s.compute_pf(in.Ybus, Vb, in.Sbus, in.slack_ids, in.slack_weights, in.pv, in.pq, max_iter, tol);
    timer_solver_ += s.get_computation_time();
    const CplxVect Vstart = init_from_n_ ? s.get_V() : in.V;
    timer_pre_proc_ = since(t_pre);
    if (!base_conv) { timer_total_ = since(t_all); return; }
    // Ybus edits and connectivity of every contingency (device)
    auto t_y = clk::now();
    IntVect pos;
    CplxVect y;
    branch_tables(in, dc, pos, y);
    ContLists all;
    for (const auto& cont : defaults_) add_cont(all, cont, pos);
    const int n = in.Ybus.ncol;
    pfgpu_ctx* ctx = engine_ctx();
    IntVect conn(ncont, dc ? 1 : 0);
    DevMem m[9];
    double* dYx = reinterpret_cast<double*>(d_Y_.upload(in.Ybus.x));
    int* dpos_br = m[2].upload(pos);
    double* dy = reinterpret_cast<double*>(m[3].upload(y));
    if (ncont > 0 && !dc) {
        int* dcp = m[0].upload(all.ptr);
        int* dcb = m[1].upload(all.br);
        int* ddp = m[4].upload(all.dptr);
        int* ddq = static_cast<int*>(m[5].get(std::max(all.dptr.back(), 1) * sizeof(int)));
        double* ddv = static_cast<double*>(m[6].get(std::max(all.dptr.back(), 1) * 16));
        check_rc(pfgpu_contingency_edits_device(ctx, ncont, dcp, dcb, dpos_br, dy, ddp, ddq, ddv),
                 "pfgpu_contingency_edits_device");
        int* dYp = m[7].upload(in.Ybus.p);
        int* dYi = m[8].upload(in.Ybus.i);
        DevMem mc;
        int* dconn = static_cast<int*>(mc.get(ncont * sizeof(int)));
        check_rc(pfgpu_connectivity_device(ctx, ncont, n, dYp, dYi, dYx, ddp, ddq, ddv, dconn),
                 "pfgpu_connectivity_device");
        check_rc(pfgpu_memcpy_d2h(ctx, conn.data(), dconn, ncont * sizeof(int)), "d2h");
    }
    // the connected contingencies, in order, and their edits (device)
    IntVect which;
    ContLists sel;
    {
        int c = 0;
        for (const auto& cont : defaults_) {
            if (conn[c]) {
                which.push_back(c);
                add_cont(sel, cont, pos);
            }
            ++c;
        }
    }
    const int m_solve = (int)which.size();
    int* sdp = d_tmp_[0].upload(sel.dptr);
    int* sdq = static_cast<int*>(d_tmp_[1].get(std::max(sel.dptr.back(), 1) * sizeof(int)));
    double* sdv = static_cast<double*>(d_tmp_[2].get(std::max(sel.dptr.back(), 1) * 16));
    if (m_solve > 0) {
        int* scp = m[0].upload(sel.ptr);
        int* scb = m[1].upload(sel.br);
        check_rc(pfgpu_contingency_edits_device(ctx, m_solve, scp, scb, dpos_br, dy, sdp, sdq, sdv),
                 "pfgpu_contingency_edits_device");
    }
    timer_modif_ybus_ = since(t_y);
    // batched solve, tol / sn_mva (ContingencyAnalysis.cpp:223)
    auto t_s = clk::now();
    double* dS = reinterpret_cast<double*>(d_S_.upload(in.Sbus));
    double* dV0 = reinterpret_cast<double*>(d_V0_.upload(Vstart));
    double* dV = static_cast<double*>(d_V_.get((size_t)std::max(m_solve, 1) * ns * 16));
    int* dIt = static_cast<int*>(d_iters_.get((size_t)std::max(m_solve, 1) * sizeof(int)));
    int* dSt = static_cast<int*>(d_status_.get((size_t)std::max(m_solve, 1) * sizeof(int)));
    pfgpu_stats st{};
    s.solve_batch_device(in.Ybus, in.slack_ids, in.slack_weights, in.pv, in.pq, m_solve, dYx, sdp, sdq, sdv, dS, false,
                         dV0, false, dV, max_iter, tol / sn, dIt, dSt, &st);
    CplxVect Vout((size_t)m_solve * ns);
    IntVect status(m_solve);
    if (m_solve > 0) {
        check_rc(pfgpu_memcpy_d2h(ctx, Vout.data(), dV, Vout.size() * 16), "d2h");
        check_rc(pfgpu_memcpy_d2h(ctx, status.data(), dSt, status.size() * sizeof(int)), "d2h");
    }
    timer_solver_ += since(t_s);
    nb_solved_ = m_solve;
    IntVect row_src(ncont, -1);
    for (int w = 0; w < m_solve; ++w) {
        if (status[w] != 0) continue;
        const int row = which[w];
        row_src[row] = w;
        for (int b = 0; b < ns; ++b) voltages_[(size_t)row * nb + in.id_solver_to_me[b]] = Vout[(size_t)w * ns + b];
    }
    keep_device_rows(ns, row_src);
    timer_total_ = since(t_all);
}

void TimeSeries::clear() {
    voltages_.clear(); amps_.clear(); pflows_.clear(); sbuses_.clear();
    nrows_ = 0;
    ns_ = 0;
    nb_solved_ = 0;
    status_ = 1;
    timer_total_ = timer_solver_ = timer_pre_proc_ = timer_amps_ = 0;
}

// TimeSeries::compute_Vs (TimeSeries.cpp:14-109) on the device: the Sbus rows are built there from
// the injections, then the warm-started step loop runs inside one launch (pfgpu_solve_chains):
// by default as the reference's single chain; with set_chain_length(L), chains of L steps side by
// side, each chain head starting from the initial V. With the DC solver every step is solved from
// the initial V side by side: a DC solve sets the angles from B theta = P alone and keeps the
// magnitudes of its initial V, so a step's result does not depend on the step before it.
int TimeSeries::compute_Vs(int T, const RealVect& gen_p, const RealVect& sgen_p, const RealVect& load_p,
                           const RealVect& load_q, const CplxVect& Vinit, int max_iter, double tol) {
    auto t_all = clk::now();
    auto t_pre = clk::now();
    status_ = 0;
    nb_solved_ = 0;
    timer_solver_ = 0;
    const int ng = grid_model_.nb_gen(), nsg = grid_model_.nb_sgen(), nl = grid_model_.nb_load();
    if ((int)gen_p.size() != T * ng || (int)sgen_p.size() != T * nsg || (int)load_p.size() != T * nl ||
        (int)load_q.size() != T * nl)
        throw std::runtime_error("TimeSeries::compute_Vs: injection matrices must be [T x nb_element]");
    const double sn = grid_model_.get_sn_mva();
    const bool dc = solver_.current().is_dc();
    SolverInput in = grid_model_.pre_process_solver(Vinit, !dc);
    const int nb = grid_model_.total_bus(), ns = (int)in.id_solver_to_me.size();
    nbus_solver_ = ns;
    id_me_to_solver_ = in.id_me_to_solver;
    ns_ = 0;
    // element lists per solver bus (element order), for the device Sbus rows
    auto lists = [&](int count, auto bus_of, IntVect& ptr, IntVect& el) {
        std::vector<IntVect> per(ns);
        for (int e = 0; e < count; ++e) per[in.id_me_to_solver[bus_of(e)]].push_back(e);
        ptr.assign(1, 0);
        el.clear();
        for (int b = 0; b < ns; ++b) {
            el.insert(el.end(), per[b].begin(), per[b].end());
            ptr.push_back((int)el.size());
        }
    };
    IntVect gp, ge, sp, se, lp, le;
    lists(ng, [&](int e) { return grid_model_.gen_bus(e); }, gp, ge);
    lists(nsg, [&](int e) { return grid_model_.sgen_bus(e); }, sp, se);
    lists(nl, [&](int e) { return grid_model_.load_bus(e); }, lp, le);
    pfgpu_ctx* ctx = engine_ctx();
    DevMem m[10];
    double* dS = static_cast<double*>(d_S_.get((size_t)std::max(T, 1) * ns * 16));
    check_rc(pfgpu_sbus_rows_device(ctx, T, ns, m[0].upload(gp), m[1].upload(ge), ng, m[2].upload(gen_p),
                                    m[3].upload(sp), m[4].upload(se), nsg, m[5].upload(sgen_p), m[6].upload(lp),
                                    m[7].upload(le), nl, m[8].upload(load_p), m[9].upload(load_q), sn, dS),
             "pfgpu_sbus_rows_device");
    sbuses_.assign((size_t)T * ns, cplx(0, 0));
    if (T > 0) check_rc(pfgpu_memcpy_d2h(ctx, sbuses_.data(), dS, sbuses_.size() * 16), "d2h");
    nrows_ = T;
    voltages_.assign((size_t)T * nb, cplx(0, 0));
    amps_.clear();
    pflows_.clear();
    GpuNRSolver& s = solver_.current();
    dc_rows_ = dc;
    s.reset();
    CplxVect Vb = in.V;
    const bool base_conv = s.compute_pf(in.Ybus, Vb, in.Sbus, in.slack_ids, in.slack_weights, in.pv, in.pq, max_iter, tol);
    timer_solver_ += s.get_computation_time();
    timer_pre_proc_ = since(t_pre);
    if (!base_conv) { status_ = 0; timer_total_ = since(t_all); return -1; }
    const CplxVect Vstart = init_from_n_ ? s.get_V() : in.V;
    const int L = dc ? 1 : std::max(1, std::min(chain_len_ > 0 ? chain_len_ : T, std::max(T, 1)));
    auto t_s = clk::now();
    double* dYx = reinterpret_cast<double*>(d_Y_.upload(in.Ybus.x));
    double* dV0 = reinterpret_cast<double*>(d_V0_.upload(Vstart));
    double* dV = static_cast<double*>(d_V_.get((size_t)std::max(T, 1) * ns * 16));
    int* dIt = static_cast<int*>(d_iters_.get((size_t)std::max(T, 1) * sizeof(int)));
    int* dSt = static_cast<int*>(d_status_.get((size_t)std::max(T, 1) * sizeof(int)));
    pfgpu_stats st{};
    s.solve_chains_device(in.Ybus, in.slack_ids, in.slack_weights, in.pv, in.pq, T, L, dYx, dS, dV0, dV, max_iter,
                          tol / sn, dIt, dSt, &st);
    CplxVect Vout((size_t)T * ns);
    IntVect stt(T);
    if (T > 0) {
        check_rc(pfgpu_memcpy_d2h(ctx, Vout.data(), dV, Vout.size() * 16), "d2h");
        check_rc(pfgpu_memcpy_d2h(ctx, stt.data(), dSt, stt.size() * sizeof(int)), "d2h");
    }
    timer_solver_ += since(t_s);
    // the reference stops at the first divergence (TimeSeries.cpp:94-99)
    int first_fail = T;
    for (int t = 0; t < T; ++t)
        if (stt[t] != 0) { first_fail = t; break; }
    IntVect row_src(T, -1);
    for (int t = 0; t < first_fail; ++t) {
        row_src[t] = t;
        for (int b = 0; b < ns; ++b) voltages_[(size_t)t * nb + in.id_solver_to_me[b]] = Vout[(size_t)t * ns + b];
    }
    keep_device_rows(ns, row_src);
    nb_solved_ = std::min(T, first_fail + 1);
    status_ = first_fail == T ? 1 : 0;
    timer_total_ = since(t_all);
    return status_;
}

}  // namespace ls2g
