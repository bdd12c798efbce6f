"""The pybind11 host mirror (lightsim2grid_amd_cpp): names and behaviour of the
reference's surface. CPU tests check the surface; gpu tests run the solvers.

Reference tests mirrored: test_KLUSolver.py:76-127 (fixtures, J blocks, Va/Vm at
1e-4 L1, iteration count), test_SecurityAnlysis_cpp.py (contingency semantics,
non-connected grids give zero voltages and NaN flows), test_Computers.py (time
series voltages)."""
import sys

import numpy as np
import pytest

from conftest import FIXTURES, load_fixture

REF_SOLVER_TYPES = ["SparseLU", "KLU", "GaussSeidel", "DC", "GaussSeidelSynch", "NICSLU", "SparseLUSingleSlack",
                    "KLUSingleSlack", "NICSLUSingleSlack", "KLUDC", "NICSLUDC", "CKTSO", "CKTSOSingleSlack", "CKTSODC",
                    "FDPF_XB_SparseLU", "FDPF_BX_SparseLU", "FDPF_XB_KLU", "FDPF_BX_KLU", "FDPF_XB_NICSLU",
                    "FDPF_BX_NICSLU", "FDPF_XB_CKTSO", "FDPF_BX_CKTSO"]


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def test_solver_type_keeps_reference_values():
    from lightsim2grid_amd import SolverType
    for k, nm in enumerate(REF_SOLVER_TYPES):
        assert int(getattr(SolverType, nm)) == k
    for nm in ("AMDGPU", "AMDGPUSingleSlack", "AMDGPUMixed", "AMDGPUMixedSingleSlack"):
        assert hasattr(SolverType, nm)


def test_solver_classes_have_reference_methods():
    from lightsim2grid_amd import solver
    for cls in (solver.AMDGPUSolver, solver.AMDGPUSolverSingleSlack, solver.AMDGPUMixedSolver,
                solver.AMDGPUMixedSolverSingleSlack):
        for m in ("get_J", "get_Va", "get_Vm", "get_V", "get_error", "get_nb_iter", "reset", "converged",
                  "compute_pf", "get_timers", "solve"):
            assert hasattr(cls, m), (cls, m)


def test_unavailable_solver_raises_like_the_reference():
    import lightsim2grid_amd as L
    a = L.solver.AnySolver()
    assert L.SolverType.AMDGPU in a.available_solvers()
    assert L.SolverType.KLU not in a.available_solvers()
    with pytest.raises(RuntimeError):
        a.change_solver(L.SolverType.KLU)
    ca_methods = ["add_all_n1", "add_n1", "add_nk", "add_multiple_n1", "remove_n1", "remove_nk", "remove_multiple_n1",
                  "clear", "clear_results_only", "reset", "close", "my_defaults", "is_grid_connected_after_contingency",
                  "compute", "compute_flows", "compute_power_flows", "get_flows", "get_voltages", "get_power_flows",
                  "total_time", "solver_time", "preprocessing_time", "amps_computation_time", "modif_Ybus_time",
                  "nb_solved", "init_from_n_powerflow", "change_solver", "available_solvers", "get_solver_type"]
    for m in ca_methods:
        assert hasattr(L.ContingencyAnalysisCPP, m), m
    ts_methods = ["compute_Vs", "compute_flows", "compute_power_flows", "get_flows", "get_power_flows", "get_voltages",
                  "get_sbuses", "change_solver", "available_solvers", "get_solver_type", "total_time", "solver_time",
                  "preprocessing_time", "amps_computation_time", "nb_solved", "get_status", "clear", "close",
                  "init_from_n_powerflow"]
    for m in ts_methods:
        assert hasattr(L.TimeSeriesCPP, m), m


def test_gridmodel_input_validation():
    import lightsim2grid_amd as L
    gm = L.GridModel()
    gm.init_bus(3, 1, np.ones(3), 2, 0)
    with pytest.raises(RuntimeError):
        gm.init_powerlines(np.ones(2), np.ones(3), np.ones(2, complex), np.array([0, 1]), np.array([1, 2]))
    with pytest.raises(RuntimeError):
        gm.init_powerlines(np.ones(2), np.ones(2), np.ones(2, complex), np.array([0, 1]), np.array([1, 7]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(FIXTURES))
@pytest.mark.parametrize("cls_name,single", [("AMDGPUSolver", False), ("AMDGPUSolverSingleSlack", True),
                                             ("AMDGPUMixedSolver", False)])
def test_solver_class_on_reference_fixtures(name, cls_name, single):
    """test_KLUSolver.py:94-114 with the GPU solver classes."""
    from lightsim2grid_amd import solver
    d = load_fixture(name)
    _say("fixture", name, cls_name)
    s = getattr(solver, cls_name)()
    s.reset()
    assert s.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    if "Mixed" not in cls_name:
        assert s.get_nb_iter() == d["iters"]
        J = s.get_J().toarray()
        Jc = J if single else J[1:, 1:]
        npvpq = len(d["pv"]) + len(d["pq"])
        dJ = np.abs(Jc - d["J"])
        for blk in (dJ[:npvpq, :npvpq], dJ[npvpq:, :npvpq], dJ[:npvpq, npvpq:], dJ[npvpq:, npvpq:]):
            assert blk.sum() <= 1e-4
    assert np.abs(s.get_Va() - np.angle(d["V"])).sum() <= 1e-4
    assert np.abs(s.get_Vm() - np.abs(d["V"])).sum() <= 1e-4
    assert int(s.get_error()) == 0 and s.converged()


@pytest.mark.gpu
def test_gridmodel_ac_pf_matches_oracle(oracle_mod):
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    _say("gridmodel ac_pf")
    gm = gridgen.to_gridmodel(g)
    V = gm.ac_pf(g.v0(), 10, 1e-8)
    d = g.solver_inputs()
    o = oracle_mod.OracleNR(False)
    assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    assert np.max(np.abs(V - o.get_V())) <= 1e-8
    assert gm.get_solver().get_nb_iter() == o.get_nb_iter()


@pytest.mark.gpu
@pytest.mark.parametrize("solver_type", ["AMDGPU", "AMDGPUMixed"])
def test_contingency_analysis_cpp(oracle_mod, solver_type):
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    _say("ContingencyAnalysisCPP", solver_type)
    ca = L.ContingencyAnalysisCPP(gridgen.to_gridmodel(g))
    ca.change_solver(getattr(L.SolverType, solver_type))
    ca.init_from_n_powerflow = True
    ca.add_all_n1()
    conts = ca.my_defaults()
    assert len(conts) == g.nbr
    ca.compute(g.v0(), 10, 1e-8)
    Vg = ca.get_voltages()
    assert Vg.shape == (g.nbr, g.n)
    bridges = gridgen.bridges(g)
    assert np.array_equal(np.asarray(ca.is_grid_connected_after_contingency()) == 1, ~bridges)
    # oracle: same sweep, bridges give non-connected grids (zero rows)
    d = g.solver_inputs()
    o = oracle_mod.OracleNR(False)
    o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    deltas = g.ybus_deltas([c for c in conts])
    Vr, itr, _, convr, conn, _ = oracle_mod.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                        d["pv"], d["pq"], o.get_V(), deltas, 10, 1e-8, False, False, 4)
    assert np.array_equal(conn, ~bridges)
    okg = np.any(Vg != 0, axis=1)
    assert np.array_equal(okg, convr)
    vtol = 1e-8 if solver_type == "AMDGPU" else 5e-8  # FP32-factor form: stated bound
    assert np.max(np.abs(Vg[convr] - Vr[convr])) <= vtol
    assert ca.nb_solved() == int(conn.sum())
    ca.compute_flows()
    A = ca.get_flows()
    assert A.shape == (g.nbr, g.nbr)
    assert np.all(np.isnan(A[~convr]))  # zero voltages -> NaN flows (test_SecurityAnlysis_cpp.py:237-259)
    for c, cont in enumerate(conts):
        if convr[c]:
            assert A[c, cont[0]] == 0.0  # clean_flows
    # amps against a direct evaluation of the branch equations on the oracle voltages
    yff, ytt, yft, ytf = g.branch_admittances()
    Ef, Et = Vr[:, g.f], Vr[:, g.t]
    Aref = np.abs(Ef * np.conj(yff * Ef + yft * Et)) / (np.sqrt(3) * np.abs(Ef))
    m = convr[:, None] & np.ones((1, g.nbr), bool)
    for c, cont in enumerate(conts):
        m[c, cont[0]] = False
    assert np.max(np.abs(A[m] - Aref[m])) <= 1e-7 if solver_type == "AMDGPU" else 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("solver_type", ["AMDGPU", "AMDGPUMixed"])
@pytest.mark.parametrize("chain", [0, 8])
def test_time_series_cpp(oracle_mod, chain, solver_type):
    """chain 0 (the default): the reference's single warm-started chain; 8: chains of 8 steps."""
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    T = 48
    _say("TimeSeriesCPP chain", chain, solver_type)
    Sb = gridgen.time_series_sbus(g, T, seed=3)
    gen_bus = np.r_[g.slack, g.pv]
    load_bus = g.pq
    gen_p = Sb[:, gen_bus].real * (gen_bus != g.slack)  # slack generator at 0 MW, as in the grid model
    load_p = -Sb[:, load_bus].real
    load_q = -Sb[:, load_bus].imag
    ts = L.TimeSeriesCPP(gridgen.to_gridmodel(g))
    ts.change_solver(getattr(L.SolverType, solver_type))
    assert ts.get_chain_length() == 0
    if chain:
        ts.set_chain_length(chain)
    st = ts.compute_Vs(gen_p, np.zeros((T, 0)), load_p, load_q, g.v0(), 10, 1e-8)
    assert st == 1 and ts.get_status() == 1
    assert ts.nb_solved() == T
    Vg = ts.get_voltages()
    d = g.solver_inputs()
    Vr, itr, _, convr, status, _ = oracle_mod.time_series(d["Ybus"], Sb, d["slack_ids"], d["slack_weights"], d["pv"],
                                                          d["pq"], g.v0(), 10, 1e-8, False, T // chain if chain else 1)
    assert status == 1 and convr.all()
    assert np.max(np.abs(Vg - Vr)) <= (1e-8 if solver_type == "AMDGPU" else 5e-8)
    assert np.allclose(ts.get_sbuses(), Sb, rtol=1e-14, atol=1e-14)
