"""DC power flow on the GPU factor / solve kernels (pfgpu_grid_load_dc, AMDGPUDCSolver,
GridModel.dc_pf) against the oracle's restatement of BaseDCAlgo<KLU>::compute_pf
(src/powerflow_algorithm/BaseDCAlgo.tpp:15-179), which is itself pinned on scipy's sparse solve
in test_oracle_fixtures.py."""
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


@pytest.mark.parametrize("name", ["case14_synth", "case118_synth", "case2000_synth", "case10000_synth"])
def test_dc_single_matches_oracle(gpu_ctx, oracle_mod, name):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make(name)
    d = g.solver_inputs()
    Y = _dc_ybus(g)
    V0 = d["V0"] * np.exp(1j * 0.05)
    _say("dc", name)
    Vr, err = oracle_mod.dc_pf(Y, V0, d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
    assert err == 0
    gr = abi.Grid(gpu_ctx, Y, d["slack_ids"], None, d["pv"], d["pq"], dc=True)
    assert gr.info["dc"] == 1
    r = gr.solve(gr.ybus_values(Y), d["Sbus"], V0, 10, 1e-8, precision=abi.FP64)
    assert r["status"][0] == 0 and r["iters"][0] == 1
    scale = max(1.0, np.max(np.abs(np.angle(Vr))))
    assert np.max(np.abs(r["V"][0] - Vr)) <= 1e-10 * scale


def test_dc_contingencies_match_oracle(gpu_ctx, oracle_mod):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    Y = _dc_ybus(g)
    _say("dc N-1")
    nb = np.nonzero(~gridgen.bridges(g))[0]
    Yc = Y.copy()
    y = 1.0 / g.x

    def pos(r, c):
        s, e = Y.indptr[c], Y.indptr[c + 1]
        return s + int(np.searchsorted(Y.indices[s:e], r))

    dptr, dpos, dval = [0], [], []
    for k in nb:
        f, t = int(g.f[k]), int(g.t[k])
        for (rr, cc, v) in ((f, f, y[k]), (t, t, y[k]), (f, t, -y[k]), (t, f, -y[k])):
            dpos.append(pos(rr, cc))
            dval.append(v)
        dptr.append(len(dpos))
    deltas = (np.array(dptr, np.int32), np.array(dpos, np.int32), np.array(dval, np.complex128))
    gr = abi.Grid(gpu_ctx, Y, d["slack_ids"], None, d["pv"], d["pq"], dc=True)
    r = gr.solve(gr.ybus_values(Y), d["Sbus"], d["V0"], 10, 1e-8, deltas=deltas)
    assert np.all(r["status"] == 0)
    for i in range(len(nb)):
        Yc = Y.copy()
        Yc.data[deltas[1][dptr[i]:dptr[i + 1]]] -= deltas[2][dptr[i]:dptr[i + 1]]
        Vr, err = oracle_mod.dc_pf(Yc, d["V0"], d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
        assert err == 0
        assert np.max(np.abs(r["V"][i] - Vr)) <= 1e-10 * max(1.0, np.max(np.abs(np.angle(Vr))))


def test_dc_errors_like_the_reference(gpu_ctx, oracle_mod):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case14_synth")
    d = g.solver_inputs()
    _say("dc errors")
    Y = _dc_ybus(g).tolil()
    b = int(d["pq"][2])
    Y[b, :] = 0.0
    Y[:, b] = 0.0
    Y[b, b] = 0.0
    Ys = sp.csc_matrix(Y)
    Ys.sort_indices()
    # keep the structural diagonal of the zeroed bus (explicit zero)
    Yfull = _dc_ybus(g)
    Yz = Yfull.copy()
    for c in range(g.n):
        for q in range(Yz.indptr[c], Yz.indptr[c + 1]):
            if c == b or Yz.indices[q] == b:
                Yz.data[q] = 0.0
    _, err = oracle_mod.dc_pf(Yz, d["V0"], d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
    assert err == 5
    gr = abi.Grid(gpu_ctx, Yz, d["slack_ids"], None, d["pv"], d["pq"], dc=True)
    assert gr.solve(gr.ybus_values(Yz), d["Sbus"], d["V0"])["status"][0] == 5
    Yh = Yfull * 1e-12
    _, err = oracle_mod.dc_pf(Yh, d["V0"], d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
    assert err == 7
    gr2 = abi.Grid(gpu_ctx, Yh, d["slack_ids"], None, d["pv"], d["pq"], dc=True)
    assert gr2.solve(gr2.ybus_values(Yh), d["Sbus"], d["V0"])["status"][0] == 7
    with pytest.raises(abi.PfgpuError):
        gr2.solve(gr2.ybus_values(Yh), d["Sbus"], d["V0"], precision=abi.FP32)


def test_dc_solver_class_and_gridmodel(oracle_mod):
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen, solver
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    Y = _dc_ybus(g)
    _say("dc solver class")
    s = solver.AMDGPUDCSolver()
    assert s.compute_pf(Y, d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    Vr, err = oracle_mod.dc_pf(Y, d["V0"], d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
    assert np.max(np.abs(s.get_V() - Vr)) <= 1e-10 * max(1.0, np.max(np.abs(np.angle(Vr))))
    assert s.get_nb_iter() == 1
    a = L.solver.AnySolver()
    assert L.SolverType.AMDGPUDC in a.available_solvers()
    gm = gridgen.to_gridmodel(g)
    V = gm.dc_pf(g.v0(), 10, 1e-8)
    Yg = gm.get_Ybus_solver()
    assert abs(Yg - Y).max() <= 1e-12 * abs(Y).max()
    Vr2, err = oracle_mod.dc_pf(Yg, g.v0(), gm.get_Sbus_solver(), gm.get_slack_ids_solver(), gm.get_pv_solver(),
                                gm.get_pq_solver())
    assert err == 0
    assert np.max(np.abs(V - Vr2)) <= 1e-10 * max(1.0, np.max(np.abs(np.angle(Vr2))))

