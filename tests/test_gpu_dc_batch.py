"""ContingencyAnalysisCPP and TimeSeriesCPP with the DC solver (SolverType.AMDGPUDC). With a DC
solver the reference's batch classes use the DC Ybus and the branches' DC coefficients
(ContingencyAnalysis.cpp:47-110, :133-148), try every contingency (the DC Ybus counts as
invertible, :113-131, and the solver's failure decides) and compute DC flows
(BaseBatchSolverSynch.hpp:96-135, :151-186). Checked against the oracle's DC solve
(BaseDCAlgo.tpp:15-179), itself pinned on scipy's sparse solve, on the same inputs."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _dc_ybus(g):
    y = 1.0 / g.x
    Y = sp.csc_matrix((np.r_[y, y, -y, -y].astype(np.complex128),
                       (np.r_[g.f, g.t, g.f, g.t], np.r_[g.f, g.t, g.t, g.f])), shape=(g.n, g.n))
    Y.sum_duplicates()
    Y.sort_indices()
    return Y


def _without(Y, g, k):
    """Y with branch k's four DC coefficients subtracted (its entries stay, at zero)."""
    Yc = Y.copy()
    y = 1.0 / g.x[k]
    f, t = int(g.f[k]), int(g.t[k])
    for r, c, v in ((f, f, y), (t, t, y), (f, t, -y), (t, f, -y)):
        a, b = Yc.indptr[c], Yc.indptr[c + 1]
        Yc.data[a + int(np.searchsorted(Yc.indices[a:b], r))] -= v
    return Yc


def _dc_flows(g, V):
    """P = (theta_f - theta_t) / x and A = |P| / (sqrt(3) |V_f|) (vn_kv = 1, sn_mva = 1)."""
    th = np.angle(V)
    P = (th[:, g.f] - th[:, g.t]) / g.x
    return P, np.abs(P) / (np.sqrt(3) * np.abs(V[:, g.f]))


@pytest.mark.parametrize("name", ["case14_synth", "case118_synth"])
def test_contingency_analysis_dc(oracle_mod, name):
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make(name)
    d = g.solver_inputs()
    _say("ContingencyAnalysisCPP DC", name)
    ca = L.ContingencyAnalysisCPP(gridgen.to_gridmodel(g))
    ca.change_solver(L.SolverType.AMDGPUDC)
    ca.add_all_n1()
    conts = ca.my_defaults()
    assert len(conts) == g.nbr
    ca.compute(g.v0(), 10, 1e-8)
    Vg = np.asarray(ca.get_voltages())
    assert ca.nb_solved() == g.nbr  # every contingency is tried
    bridges = gridgen.bridges(g)
    assert np.array_equal(np.asarray(ca.is_grid_connected_after_contingency()) == 1, ~bridges)
    Y = _dc_ybus(g)
    Vr = np.zeros((g.nbr, g.n), np.complex128)
    convr = np.zeros(g.nbr, bool)
    for c, cont in enumerate(conts):
        V, err = oracle_mod.dc_pf(_without(Y, g, cont[0]), g.v0(), d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
        convr[c] = err == 0
        if convr[c]:
            Vr[c] = V
    assert np.array_equal(convr, ~bridges)  # a bridge leaves B singular: the DC solve fails
    okg = np.any(Vg != 0, axis=1)
    assert np.array_equal(okg, convr)
    assert np.max(np.abs(Vg[convr] - Vr[convr])) <= 1e-10
    # DC flows, and clean_flows on each row's own branch
    ca.compute_power_flows()
    P = np.asarray(ca.get_power_flows())
    ca.compute_flows()
    A = np.asarray(ca.get_flows())
    assert P.shape == A.shape == (g.nbr, g.nbr)
    assert np.all(np.isnan(A[~convr])) and np.all(P[~convr] == 0.0)
    Pref, Aref = _dc_flows(g, Vr)
    m = convr[:, None] & np.ones((1, g.nbr), bool)
    for c, cont in enumerate(conts):
        if convr[c]:
            assert P[c, cont[0]] == 0.0 and A[c, cont[0]] == 0.0
        m[c, cont[0]] = False
    assert np.max(np.abs(P[m] - Pref[m])) <= 1e-8
    assert np.max(np.abs(A[m] - Aref[m])) <= 1e-8


def test_time_series_dc(oracle_mod):
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    T = 48
    _say("TimeSeriesCPP DC")
    Sb = gridgen.time_series_sbus(g, T, seed=5)
    gen_bus = np.r_[g.slack, g.pv]
    gen_p = Sb[:, gen_bus].real * (gen_bus != g.slack)
    load_p, load_q = -Sb[:, g.pq].real, -Sb[:, g.pq].imag
    ts = L.TimeSeriesCPP(gridgen.to_gridmodel(g))
    ts.change_solver(L.SolverType.AMDGPUDC)
    st = ts.compute_Vs(gen_p, np.zeros((T, 0)), load_p, load_q, g.v0(), 10, 1e-8)
    assert st == 1 and ts.get_status() == 1 and ts.nb_solved() == T
    Vg = np.asarray(ts.get_voltages())
    # the reference's chain: every step from the previous step's voltages
    Y = _dc_ybus(g)
    d = g.solver_inputs()
    V = np.asarray(g.v0(), np.complex128)
    Vr = np.zeros((T, g.n), np.complex128)
    for t in range(T):
        V, err = oracle_mod.dc_pf(Y, V, Sb[t], d["slack_ids"], d["pv"], d["pq"])
        assert err == 0
        Vr[t] = V
    assert np.max(np.abs(Vg - Vr)) <= 1e-10
    ts.compute_power_flows()
    P = np.asarray(ts.get_power_flows())
    Pref, _ = _dc_flows(g, Vr)
    assert np.max(np.abs(P - Pref)) <= 1e-8
