"""Host re-pivot (the partial pivoting of klu_factor, KLUSolver.cpp:31-45): a flat start on a grid
whose Jacobian has an exactly zero diagonal block. The static diagonal pivots of the device plan
meet a zero pivot on the first factorisation; the library chooses pivot rows on that J, re-solves
on a plan with them and keeps the plan, as klu_refactor keeps klu_factor's pivots. The oracle's
first factorisation pivots off the diagonal where it is zero and converges; the GPU result must
equal it (same converged flag and iterations, |dV| <= 1e-8)."""
import sys

import numpy as np
import pytest

from conftest import zero_diagonal_grid

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("single", [False, True])
def test_zero_pivot_repivots_like_klu_factor(gpu_ctx, oracle_mod, prec, single):
    from lightsim2grid_amd import abi
    d = zero_diagonal_grid()
    _say("repivot", prec, single)
    a = (d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    o = oracle_mod.OracleNR(single)
    assert o.compute_pf(*a)
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], single_slack=single)
    Yx = gr.ybus_values(d["Ybus"])
    # the static plan alone (strict FP32 does no passes beyond the first) meets the zero pivot
    gr.reset()
    r = gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8, precision=abi.FP32_STRICT)
    assert int(r["status"][0]) == 5 and r["stats"]["repivot_rows"] == 0
    # a fresh solve: re-pivoted, converged like the oracle, three instances at once
    gr.reset()
    V0s = np.repeat(d["V0"][None, :], 3, axis=0)
    r = gr.solve(Yx, d["Sbus"], V0s, 10, 1e-8, precision=prec)
    assert r["stats"]["repivot_rows"] == 3
    assert np.all(r["status"] == 0)
    assert np.all(r["iters"] == o.get_nb_iter())
    assert np.max(np.abs(r["V"] - o.get_V()[None, :])) <= 1e-8
    # the plan is kept: a later solve refactors on the same pivots (no new re-pivot)
    r2 = gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8, precision=prec)
    assert r2["stats"]["repivot_rows"] == 0 and int(r2["status"][0]) == 0
    assert int(r2["iters"][0]) == o.get_nb_iter()
    assert np.max(np.abs(r2["V"][0] - o.get_V())) <= 1e-8
    # the host mirror: a fresh solver object converges, and so does its next call
    from lightsim2grid_amd import solver
    ctor = {(0, False): solver.AMDGPUSolver, (0, True): solver.AMDGPUSolverSingleSlack,
            (1, False): solver.AMDGPUMixedSolver, (1, True): solver.AMDGPUMixedSolverSingleSlack}[(prec, single)]
    sv = ctor()
    for _ in range(2):
        assert sv.compute_pf(*a)
        assert sv.get_nb_iter() == o.get_nb_iter()
        assert np.max(np.abs(sv.get_V() - o.get_V())) <= 1e-8


def test_zero_pivot_repivot_dc(gpu_ctx, oracle_mod):
    """The DC form (one factor + solve of B): a bus joined by lines of reactance x and -x has
    B_bb = 0; the oracle's restatement of BaseDCAlgo<KLU> pivots off the diagonal."""
    from lightsim2grid_amd import abi
    d = zero_diagonal_grid(dc=True)
    _say("repivot dc")
    V0 = d["V0"] * np.exp(1j * 0.05)
    Vr, err = oracle_mod.dc_pf(d["Ybus"], V0, d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
    assert err == 0
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], None, d["pv"], d["pq"], dc=True)
    r = gr.solve(gr.ybus_values(d["Ybus"]), d["Sbus"], V0, 10, 1e-8, precision=abi.FP64)
    assert r["stats"]["repivot_rows"] == 1
    assert r["status"][0] == 0 and r["iters"][0] == 1
    scale = max(1.0, np.max(np.abs(np.angle(Vr))))
    assert np.max(np.abs(r["V"][0] - Vr)) <= 1e-10 * scale
