"""Fast-decoupled power flow on the GPU factor / solve kernels (pfgpu_grid_load_fdpf,
SolverType.AMDGPUFDPF_XB / _BX through GridModel.ac_pf) against the oracle's restatement of
BaseFDPFAlgo<KLU>::compute_pf (pinned in test_fdpf_oracle.py): same status, same iteration count,
|V - V_oracle| <= 1e-8."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["case14_synth", "case118_synth", "case2000_synth"])
@pytest.mark.parametrize("method", ["XB", "BX"])
def test_fdpf_matches_oracle(gpu_ctx, oracle_mod, name, method):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make(name)
    d = g.solver_inputs()
    _say("fdpf", name, method)
    Bp, Bpp = oracle_mod.fdpf_matrices(g.n, g.f, g.t, g.r, g.x, 1j * g.b, method)
    Vr, itr, err = oracle_mod.fdpf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"],
                                   Bp, Bpp, 30, 1e-8)
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], fdpf=(Bp, Bpp))
    assert gr.info["fdpf"] == 1
    r = gr.solve(gr.ybus_values(d["Ybus"]), d["Sbus"], d["V0"], 30, 1e-8, precision=abi.FP64)
    st = int(r["status"][0])
    assert (st == 0) == (err == 0)
    if err == 0:
        assert int(r["iters"][0]) == itr
        assert np.max(np.abs(r["V"][0] - Vr)) <= 1e-8
    else:  # the distributed-slack 2000-bus grid: the method does not converge there (nor the oracle)
        assert st in (2, 3)
    with pytest.raises(abi.PfgpuError):
        gr.solve(gr.ybus_values(d["Ybus"]), d["Sbus"], d["V0"], 30, 1e-8, precision=abi.FP32)


@pytest.mark.parametrize("kind", ["AMDGPUFDPF_XB", "AMDGPUFDPF_BX"])
def test_fdpf_through_gridmodel(oracle_mod, kind):
    """GridModel.change_solver(AMDGPUFDPF_*) then ac_pf: B' / B'' built by the host mirror, the
    solve on the GPU, equal to the oracle on the same inputs; a second call reuses the grid."""
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    _say("fdpf gridmodel", kind)
    gm = gridgen.to_gridmodel(g)
    gm.change_solver(getattr(L.SolverType, kind))
    method = "XB" if kind.endswith("XB") else "BX"
    Bp, Bpp = oracle_mod.fdpf_matrices(g.n, g.f, g.t, g.r, g.x, 1j * g.b, method)
    for _ in range(2):
        V = gm.ac_pf(g.v0(), 30, 1e-8)
        assert len(V) == g.n
        Vr, itr, err = oracle_mod.fdpf(gm.get_Ybus_solver(), g.v0(), gm.get_Sbus_solver(), gm.get_slack_ids_solver(),
                                       gm.get_slack_weights_solver(), gm.get_pv_solver(), gm.get_pq_solver(), Bp, Bpp,
                                       30, 1e-8)
        assert err == 0
        assert gm.get_solver().get_nb_iter() == itr
        assert np.max(np.abs(V - Vr)) <= 1e-8


def test_fdpf_single_chain_matches_stepwise_oracle(gpu_ctx, oracle_mod):
    """A 24-step time series on a fast-decoupled 118-bus grid, one warm-started chain (run in two
    batched passes): every step equals the oracle's FDPF solve started from the previous step's
    voltages (status, iterations, |dV| <= 1e-8)."""
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    T = 24
    Sb = gridgen.time_series_sbus(g, T, seed=2)
    _say("fdpf chain")
    Bp, Bpp = oracle_mod.fdpf_matrices(g.n, g.f, g.t, g.r, g.x, 1j * g.b, "XB")
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], fdpf=(Bp, Bpp))
    r = gr.solve_chains(gr.ybus_values(d["Ybus"]), Sb, g.v0(), chain_len=None, max_iter=30, tol=1e-8,
                        precision=abi.FP64)
    V = np.asarray(g.v0(), np.complex128)
    for t in range(T):
        Vr, itr, err = oracle_mod.fdpf(d["Ybus"], V, Sb[t], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"],
                                       Bp, Bpp, 30, 1e-8)
        assert err == 0 and r["status"][t] == 0
        assert int(r["iters"][t]) == itr
        assert np.max(np.abs(r["V"][t] - Vr)) <= 1e-8
        V = Vr
