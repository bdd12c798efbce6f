"""Work-queue order of large batches: the instances are taken longest first by the iteration
counts of the previous batch of the same size. The order must not change any result: the second
(reordered) sweep equals the first (index order) bit for bit, and both equal the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [0, 1])
def test_longest_first_order_keeps_results(gpu_ctx, oracle_mod, prec):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case2000_synth")
    d = g.solver_inputs()
    _say("queue order", prec)
    o = oracle_mod.OracleNR(False)
    assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    Vb = o.get_V()
    nb = np.nonzero(~gridgen.bridges(g))[0]
    deltas = g.ybus_deltas([[int(k)] for k in nb])
    ninst = len(nb)
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    dY = gpu_ctx.to_device(gr.ybus_values(d["Ybus"]))
    dS = gpu_ctx.to_device(np.ascontiguousarray(d["Sbus"], np.complex128))
    dV0 = gpu_ctx.to_device(np.ascontiguousarray(Vb, np.complex128))
    dd = tuple(gpu_ctx.to_device(x) for x in deltas)
    out = []
    for k in range(3):
        if k == 2:
            gr.reset()  # back to index order
        dV = gpu_ctx.empty((ninst, g.n), np.complex128)
        dIt, dSt = gpu_ctx.empty(ninst, np.int32), gpu_ctx.empty(ninst, np.int32)
        gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd,
                        instances_per_wg=1)
        out.append((dV.to_host(), dIt.to_host(), dSt.to_host()))
    for V, it, stt in out[1:]:
        assert np.array_equal(V, out[0][0]) and np.array_equal(it, out[0][1]) and np.array_equal(stt, out[0][2])
    # against the oracle on a sample of the sweep
    k = 256
    sub = (deltas[0][:k + 1], deltas[1][:deltas[0][k]], deltas[2][:deltas[0][k]])
    Vr, itr, _, convr, _, _ = oracle_mod.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"],
                                                      d["pq"], Vb, sub, 10, 1e-8, False, False, 4)
    V, it, stt = out[1]
    assert np.array_equal(stt[:k] == 0, convr)
    assert np.max(np.abs(V[:k][convr] - Vr[convr])) <= (1e-8 if prec == 0 else 5e-8)
