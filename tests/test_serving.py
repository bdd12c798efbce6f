"""Serving loop (continuous batching over resident grids): a request's result is
the same alone and in a mix of concurrent requests (same converged flag and
iterations, |dV| <= 1e-10 pu in FP64), it matches the oracle, and concurrent
requests share launches."""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _grid(name):
    from lightsim2grid_amd import gridgen
    g = gridgen.make(name)
    return g, g.solver_inputs()


@pytest.mark.parametrize("prec", [0, 1])
def test_serving_alone_vs_mix(oracle_mod, prec):
    """Alone and in a mix of concurrent requests on the 14-, 118- and 2000-bus grids, a request's
    result is bit for bit the same (the loop runs one instance per workgroup in every launch), in
    both precision forms; the FP64 results equal the oracle's."""
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    _say("serving alone vs mix", prec)
    sl = L.lightsim2grid_amd_cpp.ServeLoop(0, 4096, prec)
    grids = {}
    for name in ("case14_synth", "case118_synth", "case2000_synth"):
        g, d = _grid(name)
        o = oracle_mod.OracleNR(False)
        assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
        gid = sl.add_grid(d["Ybus"], d["Sbus"], o.get_V(), d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
        grids[name] = (gid, g, d, o.get_V())
    gid, g, d, Vb = grids["case118_synth"]
    nb = np.nonzero(~gridgen.bridges(g))[0][:40]
    dptr, dpos, dval = g.ybus_deltas([[int(k)] for k in nb])
    Sts = gridgen.time_series_sbus(g, 12, seed=5)
    gid2, g2, d2, Vb2 = grids["case2000_synth"]
    nb2 = np.nonzero(~gridgen.bridges(g2))[0][:24]
    deltas2 = g2.ybus_deltas([[int(k)] for k in nb2])
    # alone
    r_ca = sl.result(sl.submit_contingencies(gid, dptr, dpos, dval), 60.0)
    r_ts = sl.result(sl.submit_time_series(gid, Sts), 60.0)
    r_ca2 = sl.result(sl.submit_contingencies(gid2, *deltas2), 60.0)
    launches_alone = sl.launches()
    # in a mix: many concurrent requests on both grids from several threads
    ids = {}

    def client(k):
        out = []
        for j in range(6):
            gname = ("case14_synth", "case118_synth", "case2000_synth")[(k + j) % 3]
            gi, gg, dd, _ = grids[gname]
            out.append(sl.submit_pf(gi, None))
            out.append(sl.submit_time_series(gi, gridgen.time_series_sbus(gg, 3, seed=k * 10 + j)))
        ids[k] = out

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    id_ca = sl.submit_contingencies(gid, dptr, dpos, dval)
    id_ts = sl.submit_time_series(gid, Sts)
    id_ca2 = sl.submit_contingencies(gid2, *deltas2)
    for t in th:
        t.join()
    m_ca = sl.result(id_ca, 60.0)
    m_ts = sl.result(id_ts, 60.0)
    m_ca2 = sl.result(id_ca2, 60.0)
    others = [sl.result(i, 60.0) for k in ids for i in ids[k]]
    n_req = 3 + sum(len(v) for v in ids.values())
    assert sl.launches() - launches_alone < n_req  # requests were co-batched
    assert max(r["launch_instances"] for r in others + [m_ca, m_ts]) > 40
    for a, b in ((r_ca, m_ca), (r_ts, m_ts), (r_ca2, m_ca2)):
        assert np.array_equal(a["status"], b["status"])
        assert np.array_equal(a["iters"], b["iters"])
        assert np.array_equal(a["V"], b["V"])
    assert all(np.all(r["status"] == 0) for r in others)
    # against the oracle (FP64 form: equal iterations, 1e-8; FP32 factors: stated bound 5e-8, +-1)
    Vr, itr, _, convr, _, _ = oracle_mod.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"],
                                                      d["pq"], Vb, (dptr, dpos, dval), 10, 1e-8, False, False, 2)
    assert np.array_equal(m_ca["status"] == 0, convr)
    assert np.max(np.abs(m_ca["V"][convr] - Vr[convr])) <= (1e-8 if prec == 0 else 5e-8)
    assert np.max(np.abs(m_ca["iters"][convr] - itr[convr])) <= (0 if prec == 0 else 1)
    sl.close()


def test_serving_rejects_oversized_request():
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    _say("serving budget")
    g, d = _grid("case118_synth")
    sl = L.lightsim2grid_amd_cpp.ServeLoop(4 << 20, 4096, 1)  # 4 MiB budget
    gid = sl.add_grid(d["Ybus"], d["Sbus"], d["V0"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    with pytest.raises(RuntimeError):
        sl.submit_time_series(gid, gridgen.time_series_sbus(g, 2000, seed=1))
    r = sl.result(sl.submit_pf(gid, None), 60.0)
    assert r["status"][0] == 0
    sl.close()


def test_serving_c_abi_submit_poll(oracle_mod):
    """pfgpu_server_* / pfgpu_submit / pfgpu_poll through ctypes: results equal the oracle's,
    a poll without waiting returns not-ready or the result, and an unknown id is an error."""
    from lightsim2grid_amd import abi, gridgen
    _say("serving C ABI")
    g, d = _grid("case118_synth")
    o = oracle_mod.OracleNR(False)
    assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    Vb = o.get_V()
    srv = abi.Server(0, 4096, abi.FP64)
    gid = srv.add_grid(d["Ybus"], d["Sbus"], Vb, d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    nb = np.nonzero(~gridgen.bridges(g))[0][:30]
    deltas = g.ybus_deltas([[int(k)] for k in nb])
    ids = [srv.submit_contingencies(gid, *deltas), srv.submit_pf(gid), srv.submit_time_series(gid, gridgen.time_series_sbus(g, 5, seed=2))]
    first = srv.poll(ids[0], 0.0)
    res = [first if first is not None else srv.poll(ids[0], 60.0)] + [srv.poll(i, 60.0) for i in ids[1:]]
    Vr, itr, _, convr, _, _ = oracle_mod.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"],
                                                      d["pq"], Vb, deltas, 10, 1e-8, False, False, 2)
    assert np.array_equal(res[0]["status"] == 0, convr)
    assert np.max(np.abs(res[0]["V"][convr] - Vr[convr])) <= 1e-8
    assert res[1]["status"][0] == 0 and np.max(np.abs(res[1]["V"][0] - Vb)) <= 1e-8
    assert np.all(res[2]["status"] == 0)
    with pytest.raises(abi.PfgpuError):
        srv._rows[999999] = (1, g.n)
        srv.poll(999999, 0.0)
    srv.close()


def test_serving_close_fails_pending_requests():
    """close() with requests still queued: their results are errors, no caller waits forever, and a
    retired or unknown id is an error."""
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    _say("serving close with pending")
    g, d = _grid("case2000_synth")
    sl = L.lightsim2grid_amd_cpp.ServeLoop(0, 64, 0)
    gid = sl.add_grid(d["Ybus"], d["Sbus"], d["V0"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    ids = [sl.submit_time_series(gid, gridgen.time_series_sbus(g, 64, seed=k)) for k in range(6)]
    sl.close()
    outcomes = []
    for i in ids:
        try:
            outcomes.append(sl.result(i, 30.0) is not None)
        except RuntimeError as e:
            assert "closed" in str(e)
            outcomes.append("closed")
    assert len(outcomes) == len(ids)
    with pytest.raises(RuntimeError):
        sl.result(ids[0], 1.0)  # already retired
