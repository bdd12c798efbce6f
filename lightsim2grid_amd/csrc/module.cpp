// pybind11 module `lightsim2grid_amd_cpp`: the drop-in surface of lightsim2grid_cpp
// (src/main.cpp:94-1252) for the GPU engine: SolverType / ErrorType, the GPU
// solver classes with the KLUSolver method set (main.cpp:228-240), AnySolver,
// a GridModel subset, TimeSeriesCPP and ContingencyAnalysisCPP.
// The GIL is released around every solve, as in the reference.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "host.hpp"
#include "serve.hpp"

namespace py = pybind11;
using namespace ls2g;

namespace {

using carr = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>;
using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using iarr = py::array_t<int, py::array::c_style | py::array::forcecast>;

CplxVect cvec(const carr& a) { return CplxVect(a.data(), a.data() + a.size()); }
RealVect rvec(const darr& a) { return RealVect(a.data(), a.data() + a.size()); }
IntVect ivec(const iarr& a) { return IntVect(a.data(), a.data() + a.size()); }

py::array_t<std::complex<double>> to_np(const CplxVect& v) {
    py::array_t<std::complex<double>> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}
py::array_t<double> to_np(const RealVect& v) {
    py::array_t<double> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}
py::array_t<int> to_np(const IntVect& v) {
    py::array_t<int> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}
template <class T>
py::array_t<T> to_np2(const std::vector<T>& v, int rows, int cols) {
    py::array_t<T> a({rows, cols});
    std::copy(v.begin(), v.begin() + (size_t)rows * cols, a.mutable_data());
    return a;
}

// scipy.sparse matrix -> CSC (sorted indices, duplicates summed)
CscC csc_from_py(py::object Y) {
    py::module_ sps = py::module_::import("scipy.sparse");
    py::object M = sps.attr("csc_matrix")(Y, py::arg("dtype") = py::module_::import("numpy").attr("complex128"));
    M.attr("sum_duplicates")();
    M.attr("sort_indices")();
    CscC c;
    py::tuple shp = M.attr("shape");
    c.nrow = shp[0].cast<int>();
    c.ncol = shp[1].cast<int>();
    c.p = ivec(M.attr("indptr").cast<iarr>());
    c.i = ivec(M.attr("indices").cast<iarr>());
    c.x = cvec(M.attr("data").cast<carr>());
    return c;
}
py::object csc_to_py(const CscR& J) {
    py::module_ sps = py::module_::import("scipy.sparse");
    return sps.attr("csc_matrix")(py::make_tuple(to_np(J.x), to_np(J.i), to_np(J.p)),
                                  py::arg("shape") = py::make_tuple(J.nrow, J.ncol));
}
py::object cscc_to_py(const CscC& Y) {
    py::module_ sps = py::module_::import("scipy.sparse");
    return sps.attr("csc_matrix")(py::make_tuple(to_np(Y.x), to_np(Y.i), to_np(Y.p)),
                                  py::arg("shape") = py::make_tuple(Y.nrow, Y.ncol));
}

// the solver bus order of the last power flow, or of a flat start when none ran yet
SolverInput bus_order_of(const GridModel& g) {
    if (!g.last_solver_input().id_solver_to_me.empty()) return g.last_solver_input();
    return g.pre_process_solver(CplxVect(g.total_bus(), cplx(1.0, 0.0)));
}

template <bool SINGLE, int PREC, bool DC = false>
struct SolverT : GpuNRSolver {
    SolverT() : GpuNRSolver(SINGLE, PREC, DC) {}
};

template <class S>
void bind_solver(py::module_& m, const char* name, const char* doc) {
    auto compute = [](S& s, py::object Ybus, const carr& V, const carr& Sbus, const iarr& slack_ids,
                      const darr& slack_weights, const iarr& pv, const iarr& pq, int max_iter, double tol) {
        CscC Y = csc_from_py(Ybus);
        CplxVect Vv = cvec(V), Sv = cvec(Sbus);
        IntVect sl = ivec(slack_ids), pvv = ivec(pv), pqv = ivec(pq);
        RealVect w = rvec(slack_weights);
        py::gil_scoped_release nogil;
        return s.compute_pf(Y, Vv, Sv, sl, w, pvv, pqv, max_iter, tol);
    };
    py::class_<S>(m, name, doc)
        .def(py::init<>())
        .def("get_J", [](const S& s) { return csc_to_py(s.get_J()); }, "Jacobian of the last iteration (CSC).")
        .def("get_Va", [](const S& s) { return to_np(s.get_Va()); })
        .def("get_Vm", [](const S& s) { return to_np(s.get_Vm()); })
        .def("get_V", [](const S& s) { return to_np(s.get_V()); })
        .def("get_error", &S::get_error)
        .def("get_nb_iter", &S::get_nb_iter)
        .def("reset", &S::reset)
        .def("converged", &S::converged)
        .def("compute_pf", compute, "Newton-Raphson power flow on the GPU (same arguments as KLUSolver.compute_pf).")
        .def("get_timers", &S::get_timers)
        .def("get_timers_jacobian", &S::get_timers_jacobian)
        .def("get_computation_time", &S::get_computation_time)
        .def("solve", compute, "Alias of compute_pf (pandapower-compatible name).");
}

}  // namespace

PYBIND11_MODULE(lightsim2grid_amd_cpp, m) {
    m.doc() = "MI355X batched Newton-Raphson power flow behind lightsim2grid's solver surface";
    m.attr("klu_solver_available") = py::bool_(false);
    m.attr("nicslu_solver_available") = py::bool_(false);
    m.attr("cktso_solver_available") = py::bool_(false);
    m.attr("amdgpu_solver_available") = py::bool_(true);
    m.attr("compiled_march_native") = py::bool_(false);
    m.attr("compiled_o3_optim") = py::bool_(true);
    m.attr("version") = py::str("0.1.0");

    py::enum_<SolverType> st(m, "SolverType", "This enum controls the solver you want to use.");
    for (int k = 0; k <= (int)SolverType::AMDGPUFDPF_BX; ++k)
        st.value(solver_type_name((SolverType)k).c_str(), (SolverType)k);
    st.export_values();

    py::enum_<FDPFMethod>(m, "FDPFMethod", "Fast-decoupled variant (XB or BX).")
        .value("XB", FDPFMethod::XB)
        .value("BX", FDPFMethod::BX)
        .export_values();

    py::enum_<ErrorType>(m, "ErrorType", "This enum controls the error encountered in the solver")
        .value("NoError", ErrorType::NoError)
        .value("SingularMatrix", ErrorType::SingularMatrix)
        .value("TooManyIterations", ErrorType::TooManyIterations)
        .value("InifiniteValue", ErrorType::InifiniteValue)
        .value("SolverAnalyze", ErrorType::SolverAnalyze)
        .value("SolverFactor", ErrorType::SolverFactor)
        .value("SolverReFactor", ErrorType::SolverReFactor)
        .value("SolverSolve", ErrorType::SolverSolve)
        .value("NotInitError", ErrorType::NotInitError)
        .value("LicenseError", ErrorType::LicenseError)
        .export_values();

    bind_solver<SolverT<false, PFGPU_FP64>>(m, "AMDGPUSolver",
                                            "GPU Newton-Raphson, distributed slack, FP64 factors (drop-in for KLUSolver).");
    bind_solver<SolverT<true, PFGPU_FP64>>(m, "AMDGPUSolverSingleSlack",
                                           "GPU Newton-Raphson, single slack, FP64 factors (drop-in for KLUSolverSingleSlack).");
    bind_solver<SolverT<false, PFGPU_FP32>>(m, "AMDGPUMixedSolver",
                                            "GPU Newton-Raphson, distributed slack, FP32 factors with FP64 mismatch/update.");
    bind_solver<SolverT<true, PFGPU_FP32>>(m, "AMDGPUMixedSolverSingleSlack",
                                           "GPU Newton-Raphson, single slack, FP32 factors with FP64 mismatch/update.");
    bind_solver<SolverT<true, PFGPU_FP64, true>>(m, "AMDGPUDCSolver",
                                                 "GPU DC power flow on the same factor / solve kernels (drop-in for KLUDCSolver).");

    py::class_<ChooseSolver>(m, "AnySolver", "Dispatch over the available solver types (ChooseSolver).")
        .def(py::init<>())
        .def("get_type", &ChooseSolver::get_type)
        .def("available_solvers", &ChooseSolver::available_solvers)
        .def("change_solver", &ChooseSolver::change_solver)
        .def("get_Va", [](const ChooseSolver& s) { return to_np(s.current().get_Va()); })
        .def("get_Vm", [](const ChooseSolver& s) { return to_np(s.current().get_Vm()); })
        .def("get_V", [](const ChooseSolver& s) { return to_np(s.current().get_V()); })
        .def("get_J", [](const ChooseSolver& s) { return csc_to_py(s.current().get_J()); })
        .def("get_error", [](const ChooseSolver& s) { return s.current().get_error(); })
        .def("get_nb_iter", [](const ChooseSolver& s) { return s.current().get_nb_iter(); })
        .def("converged", [](const ChooseSolver& s) { return s.current().converged(); })
        .def("get_computation_time", [](const ChooseSolver& s) { return s.current().get_computation_time(); })
        .def("get_timers", [](const ChooseSolver& s) { return s.current().get_timers(); })
        .def("get_timers_jacobian", [](const ChooseSolver& s) { return s.current().get_timers_jacobian(); });

    py::class_<GridModel>(m, "GridModel", "Grid data model subset feeding the GPU solvers and the batch classes.")
        .def(py::init<>())
        .def("init_bus", [](GridModel& g, int n_sub, int n_bb, const darr& vn, int nl, int nt) {
            g.init_bus(n_sub, n_bb, rvec(vn), nl, nt);
        })
        .def("set_init_vm_pu", &GridModel::set_init_vm_pu)
        .def("get_init_vm_pu", &GridModel::get_init_vm_pu)
        .def("set_sn_mva", &GridModel::set_sn_mva)
        .def("get_sn_mva", &GridModel::get_sn_mva)
        .def("init_powerlines", [](GridModel& g, const darr& r, const darr& x, const carr& h, const iarr& f, const iarr& t) {
            g.init_powerlines(rvec(r), rvec(x), cvec(h), ivec(f), ivec(t));
        })
        .def("init_generators", [](GridModel& g, const darr& p, const darr& v, const darr& qmin, const darr& qmax,
                                   const iarr& bus) { g.init_generators(rvec(p), rvec(v), rvec(qmin), rvec(qmax), ivec(bus)); })
        .def("init_loads", [](GridModel& g, const darr& p, const darr& q, const iarr& bus) {
            g.init_loads(rvec(p), rvec(q), ivec(bus));
        })
        .def("init_sgens", [](GridModel& g, const darr& p, const darr& q, const darr& pmin, const darr& pmax,
                              const darr& qmin, const darr& qmax, const iarr& bus) {
            g.init_sgens(rvec(p), rvec(q), rvec(pmin), rvec(pmax), rvec(qmin), rvec(qmax), ivec(bus));
        })
        .def("add_gen_slackbus", &GridModel::add_gen_slackbus)
        .def("remove_gen_slackbus", &GridModel::remove_gen_slackbus)
        .def("deactivate_powerline", &GridModel::deactivate_powerline)
        .def("reactivate_powerline", &GridModel::reactivate_powerline)
        .def("change_solver", &GridModel::change_solver)
        .def("available_solvers", &GridModel::available_solvers)
        .def("get_solver_type", &GridModel::get_solver_type)
        .def("get_solver", &GridModel::get_solver, py::return_value_policy::reference_internal)
        .def("total_bus", &GridModel::total_bus)
        .def("ac_pf", [](GridModel& g, const carr& Vinit, int max_iter, double tol) {
            CplxVect v = cvec(Vinit);
            CplxVect r;
            {
                py::gil_scoped_release nogil;
                r = g.ac_pf(v, max_iter, tol);
            }
            return to_np(r);
        })
        .def("dc_pf", [](GridModel& g, const carr& Vinit, int max_iter, double tol) {
            CplxVect v = cvec(Vinit);
            CplxVect r;
            {
                py::gil_scoped_release nogil;
                r = g.dc_pf(v, max_iter, tol);
            }
            return to_np(r);
        })
        .def("debug_get_Bp_python", [](const GridModel& g, FDPFMethod m) {
            CscR Bp, Bpp;
            g.fill_Bp_Bpp(bus_order_of(g), m, Bp, Bpp);
            return csc_to_py(Bp);
        }, "B' of the fast-decoupled power flow in the solver bus order (fillBp_Bpp).")
        .def("debug_get_Bpp_python", [](const GridModel& g, FDPFMethod m) {
            CscR Bp, Bpp;
            g.fill_Bp_Bpp(bus_order_of(g), m, Bp, Bpp);
            return csc_to_py(Bpp);
        }, "B'' of the fast-decoupled power flow in the solver bus order (fillBp_Bpp).")
        .def("get_Ybus_solver", [](const GridModel& g) { return cscc_to_py(g.last_solver_input().Ybus); })
        .def("get_Sbus_solver", [](const GridModel& g) { return to_np(g.last_solver_input().Sbus); })
        .def("get_pv_solver", [](const GridModel& g) { return to_np(g.last_solver_input().pv); })
        .def("get_pq_solver", [](const GridModel& g) { return to_np(g.last_solver_input().pq); })
        .def("get_slack_ids_solver", [](const GridModel& g) { return to_np(g.last_solver_input().slack_ids); })
        .def("get_slack_weights_solver", [](const GridModel& g) { return to_np(g.last_solver_input().slack_weights); })
        .def("get_V_solver", [](const GridModel& g) { return to_np(g.get_V_solver()); });

    py::class_<TimeSeries>(m, "TimeSeriesCPP", "Batched time series of AC power flows (TimeSeriesCPP).")
        .def(py::init<const GridModel&>())
        .def_property("init_from_n_powerflow", &TimeSeries::get_init_from_n_powerflow,
                      &TimeSeries::set_init_from_n_powerflow)
        .def("change_solver", &TimeSeries::change_solver)
        .def("available_solvers", &TimeSeries::available_solvers)
        .def("get_solver_type", &TimeSeries::get_solver_type)
        .def("total_time", &TimeSeries::total_time)
        .def("solver_time", &TimeSeries::solver_time)
        .def("preprocessing_time", &TimeSeries::preprocessing_time)
        .def("amps_computation_time", &TimeSeries::amps_computation_time)
        .def("nb_solved", &TimeSeries::nb_solved)
        .def("get_status", &TimeSeries::get_status)
        .def("clear", &TimeSeries::clear)
        .def("close", &TimeSeries::clear)
        .def("set_chain_length", &TimeSeries::set_chain_length)
        .def("get_chain_length", &TimeSeries::get_chain_length)
        .def("compute_Vs", [](TimeSeries& ts, const darr& gen_p, const darr& sgen_p, const darr& load_p, const darr& load_q,
                              const carr& Vinit, int max_iter, double tol) {
            const int T = gen_p.ndim() == 2 ? (int)gen_p.shape(0) : (load_p.ndim() == 2 ? (int)load_p.shape(0) : 0);
            RealVect gp = rvec(gen_p), sg = rvec(sgen_p), lp = rvec(load_p), lq = rvec(load_q);
            CplxVect v = cvec(Vinit);
            py::gil_scoped_release nogil;
            return ts.compute_Vs(T, gp, sg, lp, lq, v, max_iter, tol);
        })
        .def("compute_flows", &TimeSeries::compute_flows)
        .def("compute_power_flows", &TimeSeries::compute_power_flows)
        .def("get_flows", [](const TimeSeries& t) { return to_np2(t.get_flows(), t.get_flows().empty() ? 0 : t.nb_rows(), t.nb_branch()); })
        .def("get_power_flows", [](const TimeSeries& t) {
            return to_np2(t.get_power_flows(), t.get_power_flows().empty() ? 0 : t.nb_rows(), t.nb_branch());
        })
        .def("get_voltages", [](const TimeSeries& t) { return to_np2(t.get_voltages(), t.nb_rows(), t.nb_bus_total()); })
        .def("get_sbuses", [](const TimeSeries& t) { return to_np2(t.get_sbuses(), t.nb_rows(), t.nb_bus_solver()); });

    py::class_<ServeLoop>(m, "ServeLoop",
                          "Continuous-batching request loop over grids resident on the GPU (one worker and stream per grid).")
        .def(py::init<size_t, int, int>(), py::arg("hbm_budget") = 0, py::arg("max_instances_per_launch") = 4096,
             py::arg("precision") = 1)
        .def("add_grid", [](ServeLoop& sl, py::object Ybus, const carr& Sbus, const carr& V0, const iarr& slack_ids,
                            const darr& w, const iarr& pv, const iarr& pq, int max_iter, double tol) {
            CscC Y = csc_from_py(Ybus);
            CplxVect S = cvec(Sbus), V = cvec(V0);
            IntVect sl_ = ivec(slack_ids), pv_ = ivec(pv), pq_ = ivec(pq);
            RealVect w_ = rvec(w);
            py::gil_scoped_release nogil;
            return sl.add_grid(Y, S, V, sl_, w_, pv_, pq_, max_iter, tol);
        }, py::arg("Ybus"), py::arg("Sbus"), py::arg("V0"), py::arg("slack_ids"), py::arg("slack_weights"),
           py::arg("pv"), py::arg("pq"), py::arg("max_iter") = 10, py::arg("tol") = 1e-8)
        .def("submit_pf", [](ServeLoop& sl, int grid, py::object Sbus) {
            CplxVect S;
            if (!Sbus.is_none()) S = cvec(Sbus.cast<carr>());
            return sl.submit(grid, 0, 1, S, {}, {}, {});
        }, py::arg("grid"), py::arg("Sbus") = py::none())
        .def("submit_contingencies", [](ServeLoop& sl, int grid, const iarr& dptr, const iarr& dpos, const carr& dval) {
            IntVect p = ivec(dptr);
            return sl.submit(grid, 1, (int)p.size() - 1, {}, p, ivec(dpos), cvec(dval));
        })
        .def("submit_time_series", [](ServeLoop& sl, int grid, const carr& Sbus_rows) {
            const int T = Sbus_rows.ndim() == 2 ? (int)Sbus_rows.shape(0) : 1;
            return sl.submit(grid, 2, T, cvec(Sbus_rows), {}, {}, {});
        })
        .def("ready", &ServeLoop::ready)
        .def("result", [](ServeLoop& sl, uint64_t id, double timeout) -> py::object {
            ServeResult r;
            bool ok;
            {
                py::gil_scoped_release nogil;
                ok = sl.wait(id, timeout, r);
            }
            if (!ok) return py::none();
            if (!r.ok) throw std::runtime_error("ServeLoop request failed: " + r.error);
            py::dict d;
            const int n = r.ninst ? (int)(r.V.size() / r.ninst) : 0;
            d["V"] = to_np2(r.V, r.ninst, n);
            d["iters"] = to_np(r.iters);
            d["status"] = to_np(r.status);
            d["latency_s"] = r.latency_s;
            d["queue_s"] = r.queue_s;
            d["launch_instances"] = r.launch_instances;
            return d;
        }, py::arg("id"), py::arg("timeout") = -1.0)
        .def("launches", &ServeLoop::launches)
        .def("budget", &ServeLoop::budget)
        .def("in_use", &ServeLoop::in_use)
        .def("close", [](ServeLoop& sl) {
            py::gil_scoped_release nogil;
            sl.close();
        });

    py::class_<ContingencyAnalysis>(m, "ContingencyAnalysisCPP", "Batched N-1 / N-k security analysis (ContingencyAnalysisCPP).")
        .def(py::init<const GridModel&>())
        .def_property("init_from_n_powerflow", &ContingencyAnalysis::get_init_from_n_powerflow,
                      &ContingencyAnalysis::set_init_from_n_powerflow)
        .def("change_solver", &ContingencyAnalysis::change_solver)
        .def("available_solvers", &ContingencyAnalysis::available_solvers)
        .def("get_solver_type", &ContingencyAnalysis::get_solver_type)
        .def("add_all_n1", &ContingencyAnalysis::add_all_n1)
        .def("add_n1", &ContingencyAnalysis::add_n1)
        .def("add_nk", [](ContingencyAnalysis& c, const iarr& ids) { c.add_nk(ivec(ids)); })
        .def("add_multiple_n1", [](ContingencyAnalysis& c, const iarr& ids) { c.add_multiple_n1(ivec(ids)); })
        .def("reset", &ContingencyAnalysis::clear)
        .def("clear", &ContingencyAnalysis::clear)
        .def("clear_results_only", &ContingencyAnalysis::clear_results_only)
        .def("close", &ContingencyAnalysis::clear)
        .def("remove_n1", &ContingencyAnalysis::remove_n1)
        .def("remove_nk", [](ContingencyAnalysis& c, const iarr& ids) { c.remove_nk(ivec(ids)); })
        .def("remove_multiple_n1", [](ContingencyAnalysis& c, const iarr& ids) { c.remove_multiple_n1(ivec(ids)); })
        .def("my_defaults", &ContingencyAnalysis::my_defaults)
        .def("is_grid_connected_after_contingency",
             [](ContingencyAnalysis& c) { return to_np(c.is_grid_connected_after_contingency()); })
        .def("compute", [](ContingencyAnalysis& c, const carr& Vinit, int max_iter, double tol) {
            CplxVect v = cvec(Vinit);
            py::gil_scoped_release nogil;
            c.compute(v, max_iter, tol);
        })
        .def("compute_flows", &ContingencyAnalysis::compute_flows)
        .def("compute_power_flows", &ContingencyAnalysis::compute_power_flows)
        .def("get_flows", [](const ContingencyAnalysis& c) {
            return to_np2(c.get_flows(), c.get_flows().empty() ? 0 : c.nb_rows(), c.nb_branch());
        })
        .def("get_voltages", [](const ContingencyAnalysis& c) { return to_np2(c.get_voltages(), c.nb_rows(), c.nb_bus_total()); })
        .def("get_power_flows", [](const ContingencyAnalysis& c) {
            return to_np2(c.get_power_flows(), c.get_power_flows().empty() ? 0 : c.nb_rows(), c.nb_branch());
        })
        .def("total_time", &ContingencyAnalysis::total_time)
        .def("solver_time", &ContingencyAnalysis::solver_time)
        .def("preprocessing_time", &ContingencyAnalysis::preprocessing_time)
        .def("amps_computation_time", &ContingencyAnalysis::amps_computation_time)
        .def("modif_Ybus_time", &ContingencyAnalysis::modif_Ybus_time)
        .def("nb_solved", &ContingencyAnalysis::nb_solved);
}
