// Host-side mirror of lightsim2grid's solver surface for the GPU engine.
//
// Everything here is plain C++ over the C ABI (include/pfgpu.h); the HIP
// kernels run behind it. Names, argument meaning and error behaviour follow
// the reference:
//   SolverType / ChooseSolver       src/ChooseSolver.hpp:17-25, :37-557
//   ErrorType                       src/Utils.hpp:47-56
//   solver classes (compute_pf ...) src/powerflow_algorithm/BaseAlgo.hpp:43-266, src/main.cpp:228-240
//   GridModel (subset)              src/GridModel.hpp:207-354, GridModel.cpp:301-349, :450-759
//   ContingencyAnalysis             src/batch_algorithm/ContingencyAnalysis.{hpp,cpp}
//   TimeSeries                      src/batch_algorithm/TimeSeries.{hpp,cpp}
//   BaseBatchSolverSynch            src/batch_algorithm/BaseBatchSolverSynch.{hpp,cpp}
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "pfgpu.h"

namespace ls2g {

using cplx = std::complex<double>;
using RealVect = std::vector<double>;
using CplxVect = std::vector<cplx>;
using IntVect = std::vector<int>;

// The reference's values keep their order; the GPU engine's values follow.
enum class SolverType : int {
    SparseLU, KLU, GaussSeidel, DC, GaussSeidelSynch, NICSLU,
    SparseLUSingleSlack, KLUSingleSlack, NICSLUSingleSlack,
    KLUDC, NICSLUDC,
    CKTSO, CKTSOSingleSlack, CKTSODC,
    FDPF_XB_SparseLU, FDPF_BX_SparseLU, FDPF_XB_KLU, FDPF_BX_KLU,
    FDPF_XB_NICSLU, FDPF_BX_NICSLU, FDPF_XB_CKTSO, FDPF_BX_CKTSO,
    // MI355X batched Newton-Raphson (distributed / single slack; FP64 or FP32 factors), and the DC
    // and fast-decoupled (XB, BX) power flows on the same factor / solve kernels
    AMDGPU, AMDGPUSingleSlack, AMDGPUMixed, AMDGPUMixedSingleSlack, AMDGPUDC, AMDGPUFDPF_XB, AMDGPUFDPF_BX
};

// FDPFMethod (the reference's enum: Utils.hpp)
enum class FDPFMethod : int { XB, BX };

enum class ErrorType : int {
    NoError, SingularMatrix, TooManyIterations, InifiniteValue, SolverAnalyze, SolverFactor,
    SolverReFactor, SolverSolve, NotInitError, LicenseError
};

std::string solver_type_name(SolverType t);
bool is_gpu_type(SolverType t);
bool is_dc_type(SolverType t);
// 0: not a fast-decoupled type; 1 + FDPFMethod otherwise
int fdpf_kind(SolverType t);

// CSC matrix as the solvers take it (column pointers, sorted row indices, values)
struct CscC {
    int nrow = 0, ncol = 0;
    IntVect p, i;
    CplxVect x;
};
struct CscR {
    int nrow = 0, ncol = 0;
    IntVect p, i;
    RealVect x;
};

// The process-wide engine context (device 0), created on first use.
pfgpu_ctx* engine_ctx();
void check_rc(int rc, const char* what);

// Grow-only device buffer on the engine context.
class DevMem {
public:
    DevMem() = default;
    ~DevMem() { release(); }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    void* get(size_t bytes);
    template <class T>
    T* as() const { return static_cast<T*>(p_); }
    // copies v into the buffer (grown as needed) and returns its device address
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* d = static_cast<T*>(get(std::max<size_t>(v.size(), 1) * sizeof(T)));
        if (!v.empty()) check_rc(pfgpu_memcpy_h2d(engine_ctx(), d, v.data(), v.size() * sizeof(T)), "h2d");
        return d;
    }
    void release();

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// One GPU NR solver (the reference's BaseNRAlgo / BaseNRSingleSlackAlgo
// counterpart). It caches the loaded grid (J pattern, ordering, schedule)
// while the Ybus pattern and the slack/pv/pq sets stay the same, like the
// reference keeps J_, value_map_ and the KLU symbolic analysis.
// ---------------------------------------------------------------------------
class GpuNRSolver {
public:
    // dc: the DC power flow (BaseDCAlgo) on the same kernels; max_iter and tol are ignored.
    // fdpf: 1 + FDPFMethod for the fast-decoupled power flow (BaseFDPFAlgo); its B' and B'' come
    // from set_fdpf_matrices (GridModel::ac_pf builds them, as BaseFDPFAlgo asks its GridModel)
    GpuNRSolver(bool single_slack, int precision, bool dc = false, int fdpf = 0)
        : single_(single_slack), precision_(precision), dc_(dc), fdpf_(fdpf) {}
    ~GpuNRSolver();
    GpuNRSolver(const GpuNRSolver&) = delete;
    GpuNRSolver& operator=(const GpuNRSolver&) = delete;

    bool compute_pf(const CscC& Ybus, CplxVect& V, const CplxVect& Sbus, const IntVect& slack_ids,
                    const RealVect& slack_weights, const IntVect& pv, const IntVect& pq, int max_iter, double tol);
    // Batched solve: per-instance Ybus deltas (CSC value positions) or per-instance Sbus rows.
    // V0 is one row (shared) or ninst rows; outputs V (ninst x n), iters, status.
    void solve_batch(const CscC& Ybus, const IntVect& slack_ids, const RealVect& slack_weights, const IntVect& pv,
                     const IntVect& pq, int ninst, const IntVect* dptr, const IntVect* dpos, const CplxVect* dval,
                     const CplxVect& Sbus, bool sbus_per_instance, const CplxVect& V0, bool v0_per_instance,
                     int max_iter, double tol, CplxVect& Vout, IntVect& iters, IntVect& status, pfgpu_stats* st);
    // The same on device arrays of the engine context (pfgpu_solve_batch_device); the grid is
    // (re)loaded from the pattern and the slack / pv / pq sets as for compute_pf.
    void solve_batch_device(const CscC& Ybus, const IntVect& slack_ids, const RealVect& slack_weights,
                            const IntVect& pv, const IntVect& pq, int ninst, const double* d_Yx, const int* d_dptr,
                            const int* d_dpos, const double* d_dval, const double* d_S, bool sbus_per_instance,
                            const double* d_V0, bool v0_per_instance, double* d_V, int max_iter, double tol,
                            int* d_iters, int* d_status, pfgpu_stats* st);
    // Time-series chains on device arrays (pfgpu_solve_chains_device).
    void solve_chains_device(const CscC& Ybus, const IntVect& slack_ids, const RealVect& slack_weights,
                             const IntVect& pv, const IntVect& pq, int nrows, int chain_len, const double* d_Yx,
                             const double* d_S, const double* d_V0, double* d_V, int max_iter, double tol,
                             int* d_iters, int* d_status, pfgpu_stats* st);

    void reset();
    const CplxVect& get_V() const { return V_; }
    RealVect get_Va() const;
    RealVect get_Vm() const;
    int get_nb_iter() const { return nr_iter_; }
    ErrorType get_error() const { return err_; }
    bool converged() const { return err_ == ErrorType::NoError; }
    double get_computation_time() const { return timer_total_; }
    std::tuple<double, double, double, double> get_timers() const {
        return {timer_Fx_, timer_solve_, timer_check_, timer_total_};
    }
    // (Fx, solve, refactor, initialize, check, dSbus, fillJ, VaVm, pre_proc, total), BaseNRAlgo.hpp:47-62,
    // from the kernel's phase clocks: Fx = mismatch phase (the convergence test is part of it, check = 0),
    // solve = backward solve (the forward solve is fused into the factor phase), refactor = factor +
    // forward solve + dense tail, fillJ = the J fill (it computes dS/dV in the same pass: dSbus = 0),
    // VaVm = voltage update; initialize = grid load (analysis), pre_proc = host-to-device copies.
    std::tuple<double, double, double, double, double, double, double, double, double, double> get_timers_jacobian() const {
        return {timer_Fx_, timer_solve_, timer_refactor_, timer_initialize_, timer_check_, 0.0, timer_fillJ_,
                timer_VaVm_, timer_pre_proc_, timer_total_};
    }
    CscR get_J() const;
    bool single_slack() const { return single_; }
    bool is_dc() const { return dc_; }
    int fdpf() const { return fdpf_; }
    // B' and B'' (n x n, solver bus order) of the next fast-decoupled solves
    void set_fdpf_matrices(CscR Bp, CscR Bpp) {
        Bp_ = std::move(Bp);
        Bpp_ = std::move(Bpp);
    }
    int precision() const { return precision_; }
    const pfgpu_grid_info* grid_info() const { return grid_ ? &info_ : nullptr; }

private:
    void ensure_grid(const CscC& Y, const IntVect& slack_ids, const RealVect& w, const IntVect& pv, const IntVect& pq);

    bool single_;
    int precision_;
    bool dc_ = false;
    int fdpf_ = 0;
    CscR Bp_, Bpp_, key_Bp_, key_Bpp_;
    pfgpu_grid* grid_ = nullptr;
    pfgpu_grid_info info_{};
    IntVect key_p_, key_i_, key_slack_, key_pv_, key_pq_;
    RealVect key_w_;
    CplxVect V_;
    int nr_iter_ = 0;
    ErrorType err_ = ErrorType::NotInitError;
    RealVect Jx_;
    double timer_Fx_ = 0, timer_solve_ = 0, timer_check_ = 0, timer_total_ = 0, timer_initialize_ = 0, timer_pre_proc_ = 0;
    double timer_refactor_ = 0, timer_fillJ_ = 0, timer_VaVm_ = 0;
};

// ChooseSolver: holds one solver per available type and forwards to the current one.
class ChooseSolver {
public:
    ChooseSolver();
    std::vector<SolverType> available_solvers() const;
    void change_solver(SolverType t);
    SolverType get_type() const { return type_; }
    bool ac_solver_used() const { return !is_dc_type(type_); }
    GpuNRSolver& current();
    const GpuNRSolver& current() const;
    bool compute_pf(const CscC& Ybus, CplxVect& V, const CplxVect& Sbus, const IntVect& slack_ids,
                    const RealVect& slack_weights, const IntVect& pv, const IntVect& pq, int max_iter, double tol) {
        return current().compute_pf(Ybus, V, Sbus, slack_ids, slack_weights, pv, pq, max_iter, tol);
    }
    void reset() { current().reset(); }

private:
    SolverType type_;
    std::unique_ptr<GpuNRSolver> s_[7];
};

// ---------------------------------------------------------------------------
// GridModel subset: buses, powerlines, generators (PV), static generators,
// loads, slack generators, sn_mva / init_vm_pu, AC power flow.
// ---------------------------------------------------------------------------
struct SolverInput {
    CscC Ybus;
    CplxVect Sbus;
    IntVect pv, pq, slack_ids;
    RealVect slack_weights;
    CplxVect V;                 // initial V in solver bus order
    IntVect id_me_to_solver;    // -1 for inactive buses
    IntVect id_solver_to_me;
};

class GridModel {
public:
    void init_bus(int n_sub, int n_busbar_per_sub, const RealVect& bus_vn_kv, int nb_line, int nb_trafo);
    void set_init_vm_pu(double v) { init_vm_pu_ = v; }
    double get_init_vm_pu() const { return init_vm_pu_; }
    void set_sn_mva(double v) { sn_mva_ = v; }
    double get_sn_mva() const { return sn_mva_; }
    void init_powerlines(const RealVect& r, const RealVect& x, const CplxVect& h, const IntVect& from, const IntVect& to);
    void init_generators(const RealVect& p, const RealVect& v, const RealVect& min_q, const RealVect& max_q,
                         const IntVect& bus);
    void init_loads(const RealVect& p, const RealVect& q, const IntVect& bus);
    void init_sgens(const RealVect& p, const RealVect& q, const RealVect& pmin, const RealVect& pmax,
                    const RealVect& qmin, const RealVect& qmax, const IntVect& bus);
    void add_gen_slackbus(int gen_id, double weight);
    void remove_gen_slackbus(int gen_id);
    void deactivate_powerline(int id);
    void reactivate_powerline(int id);
    void change_solver(SolverType t) { solver_->change_solver(t); }
    std::vector<SolverType> available_solvers() const { return solver_->available_solvers(); }
    SolverType get_solver_type() const { return solver_->get_type(); }
    ChooseSolver& get_solver() { return *solver_; }
    CplxVect ac_pf(const CplxVect& Vinit, int max_iter, double tol);
    // DC power flow (GridModel::dc_pf, GridModel.cpp:676-759): DC Ybus from the lines' 1 / x
    // stamps, solved by the GPU DC solver; empty when it fails
    CplxVect dc_pf(const CplxVect& Vinit, int max_iter, double tol);
    // B' and B'' of the fast-decoupled power flow (GridModel::fillBp_Bpp, GridModel.cpp:1037-1060),
    // n x n in the solver bus order of `in`
    void fill_Bp_Bpp(const SolverInput& in, FDPFMethod method, CscR& Bp, CscR& Bpp) const;
    GpuNRSolver& get_dc_solver() { return *dc_solver_; }

    int total_bus() const { return (int)bus_vn_kv_.size(); }
    int nb_powerline() const { return (int)line_r_.size(); }
    // ac false: the DC Ybus (line stamps 1 / x, no shunts; TwoSidesContainer_rxh_A.hpp:934-940)
    SolverInput pre_process_solver(const CplxVect& Vinit, bool ac = true) const;
    const SolverInput& last_solver_input() const { return last_; }
    const CplxVect& get_V_solver() const { return V_solver_; }

    // branch data used by the batch classes (lines only; this subset has no trafos)
    void branch_admittance(int k, cplx& yff, cplx& yft, cplx& ytf, cplx& ytt) const;
    void branch_admittance_dc(int k, cplx& yff, cplx& yft, cplx& ytf, cplx& ytt) const;
    int line_from(int k) const { return line_from_[k]; }
    int line_to(int k) const { return line_to_[k]; }
    bool line_status(int k) const { return line_status_[k]; }
    double bus_vn_kv(int b) const { return bus_vn_kv_[b]; }
    int gen_bus(int g) const { return gen_bus_[g]; }
    int sgen_bus(int g) const { return sgen_bus_[g]; }
    int load_bus(int l) const { return load_bus_[l]; }
    int nb_gen() const { return (int)gen_bus_.size(); }
    int nb_sgen() const { return (int)sgen_bus_.size(); }
    int nb_load() const { return (int)load_bus_.size(); }

private:
    RealVect bus_vn_kv_;
    double init_vm_pu_ = 1.04, sn_mva_ = 1.0;
    RealVect line_r_, line_x_;
    CplxVect line_h_;
    IntVect line_from_, line_to_;
    std::vector<bool> line_status_;
    RealVect gen_p_, gen_v_;
    IntVect gen_bus_;
    std::vector<bool> gen_slack_;
    RealVect gen_slack_w_;
    RealVect sgen_p_, sgen_q_;
    IntVect sgen_bus_;
    RealVect load_p_, load_q_;
    IntVect load_bus_;
    std::shared_ptr<ChooseSolver> solver_ = std::make_shared<ChooseSolver>();
    std::shared_ptr<GpuNRSolver> dc_solver_ = std::make_shared<GpuNRSolver>(true, PFGPU_FP64, true);
    SolverInput last_;
    CplxVect V_solver_;
};

// ---------------------------------------------------------------------------
// Batch drivers (BaseBatchSolverSynch + ContingencyAnalysis / TimeSeries)
// ---------------------------------------------------------------------------
class BaseBatchSolverSynch {
public:
    explicit BaseBatchSolverSynch(const GridModel& gm) : grid_model_(gm) {
        solver_.change_solver(SolverType::AMDGPU);
    }
    std::vector<SolverType> available_solvers() const { return solver_.available_solvers(); }
    void change_solver(SolverType t) { solver_.change_solver(t); }
    SolverType get_solver_type() const { return solver_.get_type(); }
    bool get_init_from_n_powerflow() const { return init_from_n_; }
    void set_init_from_n_powerflow(bool v) { init_from_n_ = v; }
    double total_time() const { return timer_total_; }
    double solver_time() const { return timer_solver_; }
    double preprocessing_time() const { return timer_pre_proc_; }
    double amps_computation_time() const { return timer_amps_; }
    int nb_solved() const { return nb_solved_; }
    // results, row-major [nsteps x n]
    const CplxVect& get_voltages() const { return voltages_; }
    const RealVect& get_flows() const { return amps_; }
    const RealVect& get_power_flows() const { return pflows_; }
    int nb_rows() const { return nrows_; }
    int nb_bus_total() const { return grid_model_.total_bus(); }
    int nb_branch() const { return grid_model_.nb_powerline(); }
    // Flows of every result row on the device, from the solved voltages kept there
    // (pfgpu_batch_flows_device, or pfgpu_batch_flows_dc_device after a DC compute); clean lists
    // the branches zeroed per row (contingencies).
    void compute_flows_from_Vs(bool amps);

protected:
    // device results of the last compute: solved voltages (solver bus order), and per result
    // row the solved row it comes from (-1: zero voltages)
    void keep_device_rows(int ns, const IntVect& row_src);
    virtual void clean_lists(IntVect& ptr, IntVect& br) const { ptr.assign(nrows_ + 1, 0); br.clear(); }
    GridModel grid_model_;
    ChooseSolver solver_;
    bool init_from_n_ = false;
    double timer_total_ = 0, timer_solver_ = 0, timer_pre_proc_ = 0, timer_amps_ = 0;
    int nb_solved_ = 0;
    int nrows_ = 0;
    CplxVect voltages_;
    RealVect amps_, pflows_;
    IntVect id_me_to_solver_;  // of the last compute
    int ns_ = 0;               // solver buses of the last compute
    bool dc_rows_ = false;     // the last compute used the DC solver (ac_solver_used() false)
    DevMem d_V_, d_row_src_, d_iters_, d_status_, d_S_, d_Y_, d_V0_, d_tmp_[8];
};

class ContingencyAnalysis : public BaseBatchSolverSynch {
public:
    using BaseBatchSolverSynch::BaseBatchSolverSynch;
    void add_all_n1();
    void add_n1(int line_id);
    void add_nk(const IntVect& line_ids);
    void add_multiple_n1(const IntVect& line_ids);
    void remove_n1(int line_id);
    void remove_nk(const IntVect& line_ids);
    void remove_multiple_n1(const IntVect& line_ids);
    void clear();
    void clear_results_only();
    std::vector<IntVect> my_defaults() const;
    IntVect is_grid_connected_after_contingency();
    void compute(const CplxVect& Vinit, int max_iter, double tol);
    void compute_flows() { compute_flows_from_Vs(true); }
    void compute_power_flows() { compute_flows_from_Vs(false); }
    double modif_Ybus_time() const { return timer_modif_ybus_; }

private:
    // ContingencyAnalysis::clean_flows: each row zeroes the finite flows of its own branches
    void clean_lists(IntVect& ptr, IntVect& br) const override;
    // per-branch Ybus positions (solver CSC: ff, tt, ft, tf; -1 out of service) and stamps
    // (the AC coefficients, or the DC ones for the DC solver)
    void branch_tables(const SolverInput& in, bool dc, IntVect& pos, CplxVect& y) const;
    std::set<std::set<int>> defaults_;
    double timer_modif_ybus_ = 0;
};

class TimeSeries : public BaseBatchSolverSynch {
public:
    using BaseBatchSolverSynch::BaseBatchSolverSynch;
    // gen_p [T x nb_gen], sgen_p [T x nb_sgen], load_p/load_q [T x nb_load], row-major
    int compute_Vs(int T, const RealVect& gen_p, const RealVect& sgen_p, const RealVect& load_p, const RealVect& load_q,
                   const CplxVect& Vinit, int max_iter, double tol);
    void compute_flows() { compute_flows_from_Vs(true); }
    void compute_power_flows() { compute_flows_from_Vs(false); }
    int get_status() const { return status_; }
    const CplxVect& get_sbuses() const { return sbuses_; }
    int nb_bus_solver() const { return nbus_solver_; }
    void clear();
    // Steps run as warm-started chains of `chain_length` steps side by side; 0 (the default) or
    // chain_length >= T is the reference's single sequential chain.
    void set_chain_length(int L) { chain_len_ = L; }
    int get_chain_length() const { return chain_len_; }

private:
    int status_ = 1;
    CplxVect sbuses_;
    int nbus_solver_ = 0;
    int chain_len_ = 0;
};

}  // namespace ls2g
